// fgpe_mlp.hip -- the initial-guess surrogate: a fully connected network
// trained and evaluated with the project's own gfx950 kernels.
//
// Layers are y = act(x W + b) with W stored [K][N] (row major), FP32 master
// weights and BF16 copies feeding the matrix cores.  Every product goes
// through k_gemm: a 128x128 tile per 256-thread block, four waves each owning
// a 64x64 quarter computed with gfx950's v_mfma_f32_32x32x16_bf16 (BF16
// operands, FP32 accumulation), K staged through double-buffered LDS in chunks
// of 32 with 16 B operand loads, optional transposed operands and split-K with
// a deterministic second-stage sum.  The rest
// (bias + ReLU with the BF16 copy, the MSE loss gradient, the ReLU mask, bias
// gradients, Adam, minibatch gathering) are small elementwise kernels.
// No reference counterpart: the reference starts every run from its Gaussian
// (or Thomas-Fermi) guess (SRC/cgpe2d.f90:697-771).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../include/fgpe.h"

namespace fgpe_impl {
int fail(int code, const char* fmt, ...);
}
using fgpe_impl::fail;

#define MLPCHK(expr)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(FGPE_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

namespace fgpe_mlp_dev {

typedef uint16_t bf16;
typedef float floatx16 __attribute__((ext_vector_type(16)));

// round to nearest even
__host__ __device__ __forceinline__ bf16 to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(bf16 h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// Block tile BM x BN, K staged through LDS in chunks of BK (double-buffered).
// LDS holds both operands k-contiguous per row (A by m, B by n), rows padded
// to LDK elements (80 B) so a lane's 8 consecutive k are one 16 B read and
// the 32 lanes of an MFMA operand hit different banks.
constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;
constexpr int PPR = BK / 8;                  // 8-element pieces per K-contiguous row
constexpr int NPT = BM * BK / 8 / 256;       // pieces per thread per operand chunk
static_assert(BM == 128 && BN == 128 && BK % 16 == 0 && NPT >= 1, "tile shape");
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One 8-element piece of an operand tile: where it lies in global memory and
// in LDS.  `row` indexes the LDS rows (m for A, n for B), `k` the K offset in
// the chunk; `kmajor` pieces run along K in global memory (one 16 B LDS
// store), the others along m/n (eight 2 B LDS stores at consecutive rows).
struct Piece {
  int row, k;
};
// Operand X with LDS row index r (m or n) and K index k: element address
// X[k * ld + r] if the operand is stored K-major along r (transposed A,
// plain B), else X[r * ld + k].
template <bool KROWS>
__device__ __forceinline__ void load_piece(const bf16* __restrict__ X, int ld, int r0, int rmax, int k0, int kend,
                                           int v, bool vec, u32x4& out) {
  if constexpr (!KROWS) {
    // PPR pieces per row of BK: row r, k .. k+7 contiguous in memory
    const int r = v / PPR, kq = (v % PPR) * 8;
    const int gr = r0 + r, gk = k0 + kq;
    if (vec && gr < rmax && gk + 8 <= kend) {
      out = *reinterpret_cast<const u32x4*>(X + (long)gr * ld + gk);
    } else {
      unsigned short e[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) e[t] = (gr < rmax && gk + t < kend) ? X[(long)gr * ld + gk + t] : (unsigned short)0;
#pragma unroll
      for (int t = 0; t < 4; ++t) out[t] = (unsigned int)e[2 * t] | ((unsigned int)e[2 * t + 1] << 16);
    }
  } else {
    // 16 pieces per K row of 128: k fixed, rows r .. r+7 contiguous in memory
    const int k = v >> 4, rq = (v & 15) * 8;
    const int gk = k0 + k, gr = r0 + rq;
    if (vec && gk < kend && gr + 8 <= rmax) {
      out = *reinterpret_cast<const u32x4*>(X + (long)gk * ld + gr);
    } else {
      unsigned short e[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) e[t] = (gk < kend && gr + t < rmax) ? X[(long)gk * ld + gr + t] : (unsigned short)0;
#pragma unroll
      for (int t = 0; t < 4; ++t) out[t] = (unsigned int)e[2 * t] | ((unsigned int)e[2 * t + 1] << 16);
    }
  }
}
template <bool KROWS> __device__ __forceinline__ void store_piece(bf16* __restrict__ S, int v, const u32x4& in) {
  if constexpr (!KROWS) {
    const int r = v / PPR, kq = (v % PPR) * 8;
    *reinterpret_cast<u32x4*>(S + r * LDK + kq) = in;
  } else {
    const int k = v >> 4, rq = (v & 15) * 8;
#pragma unroll
    for (int t = 0; t < 8; ++t) S[(rq + t) * LDK + k] = (bf16)((in[t >> 1] >> (16 * (t & 1))) & 0xffffu);
  }
}

// C[M][N] (+)= op(A)[M][K] op(B)[K][N] over the K range of blockIdx.z.
// TA: A stored [K][lda] (element (m,k) at k*lda+m), else [M][lda].
// TB: B stored [N][ldb] (element (k,n) at n*ldb+k), else [K][ldb].
// vec: both operands 16 B aligned with leading dimensions a multiple of 8
// (then full 8-element pieces are single loads).  Split-K: block z writes
// its partial product to C + z*M*N.  256 threads: four waves, each a 64 x 64
// quarter of the tile as 2 x 2 products of v_mfma_f32_32x32x16_bf16 (BF16
// operands, FP32 accumulation).
template <int TA, int TB>
__global__ __launch_bounds__(256) void k_gemm(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb,
                                              float* __restrict__ C, int M, int N, int K, int kper, int vec) {
  __shared__ __align__(16) bf16 As[2][BM * LDK];
  __shared__ __align__(16) bf16 Bs[2][BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int k_begin = blockIdx.z * kper, k_end = min(K, k_begin + kper);
  float* Cz = C + (long)blockIdx.z * M * N;
  const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
  // A is K-rows when transposed, B when not
  constexpr bool AK = TA != 0, BKR = TB == 0;
  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  // BM * BK / 8 pieces of 8 per operand chunk: NPT per thread
  u32x4 ra[NPT], rb[NPT];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
      load_piece<AK>(A, lda, m0, M, k0, k_end, tid + 256 * p, vec, ra[p]);
      load_piece<BKR>(B, ldb, n0, N, k0, k_end, tid + 256 * p, vec, rb[p]);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
      store_piece<AK>(As[buf], tid + 256 * p, ra[p]);
      store_piece<BKR>(Bs[buf], tid + 256 * p, rb[p]);
    }
  };
  fetch(k_begin);
  stash(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = k_begin; k0 < k_end; k0 += BK) {
    const bool more = k0 + BK < k_end;
    if (more) fetch(k0 + BK);  // in flight during this chunk's products
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      // lane l: row (m or n) l%32 of its 32-row block, k = ks + 8(l/32) .. +7
      const int kk = ks + 8 * (lane >> 5);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const bf16x8*>(&As[buf][(wm + 32 * i + (lane & 31)) * LDK + kk]);
        b[i] = *reinterpret_cast<const bf16x8*>(&Bs[buf][(wn + 32 * i + (lane & 31)) * LDK + kk]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) stash(buf ^ 1);  // the other buffer's readers passed the last barrier
    __syncthreads();
    buf ^= 1;
  }
  // D layout: register r holds row 8(r/4) + 4(l/32) + r%4, column l%32
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = n0 + wn + 32 * j + (lane & 31);
        if (m < M && n < N) Cz[(long)m * N + n] = acc[i][j][r];
      }
}

// out[i] = sum_z part[z*n + i] (fixed order)
__global__ void k_splitk_sum(const float* __restrict__ part, int splits, long n, float* __restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(long)z * n + i];
    out[i] = s;
  }
}

// H = act(Z + b) in FP32 and BF16 (act 1: ReLU)
__global__ void k_bias_act(const float* __restrict__ Z, const float* __restrict__ bias, int M, int N, int relu,
                           float* __restrict__ H, bf16* __restrict__ Hb) {
  const long n = (long)M * N;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float v = Z[i] + bias[i % N];
    if (relu) v = fmaxf(v, 0.0f);
    H[i] = v;
    if (Hb) Hb[i] = to_bf16(v);
  }
}

// dY = 2 (Y - T) / (M N) as FP32 and BF16; per-block partial sums of (Y-T)^2
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, long n, float* __restrict__ dY,
                      bf16* __restrict__ dYb, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  const float scale = 2.0f / (float)n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    s += (double)d * d;
    const float g = scale * d;
    dY[i] = g;
    dYb[i] = to_bf16(g);
  }
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// dZ = dH * [H > 0] in FP32 and BF16
__global__ void k_relu_back(const float* __restrict__ dH, const float* __restrict__ H, long n, float* __restrict__ dZ,
                            bf16* __restrict__ dZb) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float g = H[i] > 0.0f ? dH[i] : 0.0f;
    dZ[i] = g;
    dZb[i] = to_bf16(g);
  }
}

// db[j] = sum_i dZ[i][j]
__global__ void k_colsum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int i = 0; i < M; ++i) s += dZ[(long)i * N + j];
    db[j] = s;
  }
}

// Adam on FP32 master weights, refreshing the BF16 copy
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2, bf16* __restrict__ wb) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float wi = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = to_bf16(wi);
  }
}

__global__ void k_to_bf16(const float* __restrict__ x, long n, bf16* __restrict__ y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = to_bf16(x[i]);
}

// rows idx[0..B) of a [rows][width] table into out[B][width]
__global__ void k_gather(const float* __restrict__ src, int width, const int* __restrict__ idx, int B,
                         float* __restrict__ out) {
  const long n = (long)B * width;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int r = (int)(i / width), c = (int)(i % width);
    out[i] = src[(long)idx[r] * width + c];
  }
}

// x[i] = a BF16 value in [-1, 1) from a hash of i
__global__ void k_fill_bf16(bf16* __restrict__ x, long n, unsigned int seed) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    unsigned int h = (unsigned int)i * 2654435761u ^ seed;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    x[i] = to_bf16((float)(h >> 8) * (1.0f / 8388608.0f) - 1.0f);
  }
}

}  // namespace fgpe_mlp_dev

using namespace fgpe_mlp_dev;

struct fgpe_mlp {
  int nl = 0;                 // layers
  std::vector<int> dims;      // nl + 1 widths
  int max_batch = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // per layer: master W [K][N], b [N], BF16 W, Adam moments, gradients
  std::vector<float*> W, b, gW, gb, mW, vW, mb, vb;
  std::vector<bf16*> Wb;
  // activations per layer boundary: H[l] [B][dims[l]] FP32 and BF16 (H[0] = input)
  std::vector<float*> H;
  std::vector<bf16*> Hb;
  float *Z = nullptr, *dA = nullptr, *dZ = nullptr, *part = nullptr;  // [B][max width], split-K partials
  bf16* dZb = nullptr;
  double* lpart = nullptr;
  float *T = nullptr;         // targets [B][out]
  float *dsX = nullptr, *dsY = nullptr;  // device dataset
  int ds_rows = 0;
  int* idx = nullptr;
  long step = 0;
  size_t maxw = 0;
  size_t part_elems = 0;
};

namespace {

int malloc_dev(fgpe_mlp* m, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) return fail(FGPE_EDEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  m->allocs.push_back(*p);
  return 0;
}

int grid1(long n) { return (int)std::min<long>((n + 255) / 256, 4096); }

// C[M][N] = op(A) op(B), split over K when the output has few tiles
int gemm(fgpe_mlp* m, int ta, int tb, const bf16* A, int lda, const bf16* B, int ldb, float* C, int M, int N, int K) {
  const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  int splits = 1;
  while (tiles * splits < 256 && K / (splits * 2) >= 4 * BK) splits *= 2;
  while (splits > 1 && (size_t)splits * M * N > m->part_elems) splits /= 2;  // partials must fit
  int kper = (K + splits - 1) / splits;
  kper = (kper + BK - 1) / BK * BK;
  splits = (K + kper - 1) / kper;
  float* out = splits > 1 ? m->part : C;
  const int vec = ((uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0 && lda % 8 == 0 && ldb % 8 == 0) ? 1 : 0;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, splits);
#define GEMM(a, b) \
  hipLaunchKernelGGL((k_gemm<a, b>), grid, dim3(256), 0, m->stream, A, lda, B, ldb, out, M, N, K, kper, vec)
  if (!ta && !tb)
    GEMM(0, 0);
  else if (ta && !tb)
    GEMM(1, 0);
  else if (!ta && tb)
    GEMM(0, 1);
  else
    GEMM(1, 1);
#undef GEMM
  MLPCHK(hipGetLastError());
  if (splits > 1) {
    hipLaunchKernelGGL(k_splitk_sum, dim3(grid1((long)M * N)), dim3(256), 0, m->stream, m->part, splits,
                       (long)M * N, C);
    MLPCHK(hipGetLastError());
  }
  return 0;
}

// forward over a batch whose BF16 input is in Hb[0]; output in H[nl]
int forward(fgpe_mlp* m, int B) {
  for (int l = 0; l < m->nl; ++l) {
    const int K = m->dims[l], N = m->dims[l + 1];
    int rc = gemm(m, 0, 0, m->Hb[l], K, m->Wb[l], N, m->Z, B, N, K);
    if (rc) return rc;
    const bool last = l == m->nl - 1;
    hipLaunchKernelGGL(k_bias_act, dim3(grid1((long)B * N)), dim3(256), 0, m->stream, m->Z, m->b[l], B, N,
                       last ? 0 : 1, m->H[l + 1], last ? (bf16*)nullptr : m->Hb[l + 1]);
    MLPCHK(hipGetLastError());
  }
  return 0;
}

// backward of the MSE loss against T; gradients into gW, gb; loss returned
int backward(fgpe_mlp* m, int B, double* loss) {
  const int nout = m->dims[m->nl];
  const long n = (long)B * nout;
  const int lb = std::min(grid1(n), 1024);
  hipLaunchKernelGGL(k_mse, dim3(lb), dim3(256), 0, m->stream, m->H[m->nl], m->T, n, m->dZ, m->dZb, m->lpart);
  MLPCHK(hipGetLastError());
  for (int l = m->nl - 1; l >= 0; --l) {
    const int K = m->dims[l], N = m->dims[l + 1];
    // dW = H[l]^T dZ, db = colsum(dZ)
    int rc = gemm(m, 1, 0, m->Hb[l], K, m->dZb, N, m->gW[l], K, N, B);
    if (rc) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(grid1(N)), dim3(256), 0, m->stream, m->dZ, B, N, m->gb[l]);
    MLPCHK(hipGetLastError());
    if (l == 0) break;
    // dH[l] = dZ W^T, then the ReLU mask of layer l-1's output
    if ((rc = gemm(m, 0, 1, m->dZb, N, m->Wb[l], N, m->dA, B, K, N))) return rc;
    hipLaunchKernelGGL(k_relu_back, dim3(grid1((long)B * K)), dim3(256), 0, m->stream, m->dA, m->H[l], (long)B * K,
                       m->dZ, m->dZb);
    MLPCHK(hipGetLastError());
  }
  std::vector<double> lp(lb);
  MLPCHK(hipMemcpyAsync(lp.data(), m->lpart, lb * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  double s = 0;
  for (double v : lp) s += v;
  if (loss) *loss = s / (double)n;
  return 0;
}

size_t param_count(const fgpe_mlp* m) {
  size_t n = 0;
  for (int l = 0; l < m->nl; ++l) n += (size_t)m->dims[l] * m->dims[l + 1] + m->dims[l + 1];
  return n;
}

int load_input(fgpe_mlp* m, const float* x_host, int B) {
  const int nin = m->dims[0];
  MLPCHK(hipMemcpyAsync(m->H[0], x_host, (size_t)B * nin * sizeof(float), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1((long)B * nin)), dim3(256), 0, m->stream, m->H[0], (long)B * nin, m->Hb[0]);
  MLPCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int fgpe_mlp_create(const int32_t* dims, int32_t nlayers, int32_t max_batch, uint64_t seed, fgpe_mlp** out) {
  if (!dims || !out || nlayers < 1 || nlayers > 16 || max_batch < 1) return fail(FGPE_EINVAL, "bad MLP shape");
  for (int l = 0; l <= nlayers; ++l)
    if (dims[l] < 1 || dims[l] > (1 << 20)) return fail(FGPE_EINVAL, "bad MLP width %d", dims[l]);
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FGPE_EDEVICE, "no HIP device");
  fgpe_mlp* m = new fgpe_mlp();
  auto bail = [&](int rc) {
    fgpe_mlp_destroy(m);
    return rc;
  };
  m->nl = nlayers;
  m->dims.assign(dims, dims + nlayers + 1);
  m->max_batch = max_batch;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipStreamCreate failed"));
  for (int d : m->dims) m->maxw = std::max(m->maxw, (size_t)d);
  int rc;
  const int L = nlayers;
  m->W.resize(L);
  m->b.resize(L);
  m->gW.resize(L);
  m->gb.resize(L);
  m->mW.resize(L);
  m->vW.resize(L);
  m->mb.resize(L);
  m->vb.resize(L);
  m->Wb.resize(L);
  m->H.resize(L + 1);
  m->Hb.resize(L + 1);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < L; ++l) {
    const size_t K = dims[l], N = dims[l + 1];
    float** f[] = {&m->W[l], &m->gW[l], &m->mW[l], &m->vW[l]};
    for (auto p : f)
      if ((rc = malloc_dev(m, (void**)p, K * N * sizeof(float)))) return bail(rc);
    float** g[] = {&m->b[l], &m->gb[l], &m->mb[l], &m->vb[l]};
    for (auto p : g)
      if ((rc = malloc_dev(m, (void**)p, N * sizeof(float)))) return bail(rc);
    if ((rc = malloc_dev(m, (void**)&m->Wb[l], K * N * sizeof(bf16)))) return bail(rc);
    // He-uniform init, zero bias and moments
    std::vector<float> w(K * N);
    const double lim = std::sqrt(6.0 / (double)K);
    std::uniform_real_distribution<double> U(-lim, lim);
    for (auto& v : w) v = (float)U(rng);
    std::vector<bf16> wb(K * N);
    for (size_t i = 0; i < w.size(); ++i) wb[i] = to_bf16(w[i]);
    MLPCHK(hipMemcpy(m->W[l], w.data(), K * N * sizeof(float), hipMemcpyHostToDevice));
    MLPCHK(hipMemcpy(m->Wb[l], wb.data(), K * N * sizeof(bf16), hipMemcpyHostToDevice));
    MLPCHK(hipMemset(m->b[l], 0, N * sizeof(float)));
    MLPCHK(hipMemset(m->mW[l], 0, K * N * sizeof(float)));
    MLPCHK(hipMemset(m->vW[l], 0, K * N * sizeof(float)));
    MLPCHK(hipMemset(m->mb[l], 0, N * sizeof(float)));
    MLPCHK(hipMemset(m->vb[l], 0, N * sizeof(float)));
  }
  for (int l = 0; l <= L; ++l) {
    if ((rc = malloc_dev(m, (void**)&m->H[l], (size_t)max_batch * dims[l] * sizeof(float)))) return bail(rc);
    if ((rc = malloc_dev(m, (void**)&m->Hb[l], (size_t)max_batch * dims[l] * sizeof(bf16)))) return bail(rc);
  }
  const size_t bw = (size_t)max_batch * m->maxw;
  m->part_elems = std::max(bw, (size_t)m->maxw * 512) * 8;
  if ((rc = malloc_dev(m, (void**)&m->Z, bw * sizeof(float))) || (rc = malloc_dev(m, (void**)&m->dA, bw * sizeof(float))) ||
      (rc = malloc_dev(m, (void**)&m->dZ, bw * sizeof(float))) || (rc = malloc_dev(m, (void**)&m->dZb, bw * sizeof(bf16))) ||
      (rc = malloc_dev(m, (void**)&m->T, (size_t)max_batch * dims[L] * sizeof(float))) ||
      (rc = malloc_dev(m, (void**)&m->part, m->part_elems * sizeof(float))) ||
      (rc = malloc_dev(m, (void**)&m->lpart, 1024 * sizeof(double))) ||
      (rc = malloc_dev(m, (void**)&m->idx, max_batch * sizeof(int))))
    return bail(rc);
  *out = m;
  return 0;
}

void fgpe_mlp_destroy(fgpe_mlp* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : m->allocs) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int fgpe_mlp_param_count(fgpe_mlp* m, int64_t* n) {
  if (!m || !n) return fail(FGPE_EINVAL, "null argument");
  *n = (int64_t)param_count(m);
  return 0;
}

int fgpe_mlp_get_params(fgpe_mlp* m, float* w) {
  if (!m || !w) return fail(FGPE_EINVAL, "null argument");
  MLPCHK(hipStreamSynchronize(m->stream));
  size_t o = 0;
  for (int l = 0; l < m->nl; ++l) {
    const size_t K = m->dims[l], N = m->dims[l + 1];
    MLPCHK(hipMemcpy(w + o, m->W[l], K * N * sizeof(float), hipMemcpyDeviceToHost));
    o += K * N;
    MLPCHK(hipMemcpy(w + o, m->b[l], N * sizeof(float), hipMemcpyDeviceToHost));
    o += N;
  }
  return 0;
}

int fgpe_mlp_set_params(fgpe_mlp* m, const float* w) {
  if (!m || !w) return fail(FGPE_EINVAL, "null argument");
  MLPCHK(hipStreamSynchronize(m->stream));
  size_t o = 0;
  for (int l = 0; l < m->nl; ++l) {
    const size_t K = m->dims[l], N = m->dims[l + 1];
    MLPCHK(hipMemcpy(m->W[l], w + o, K * N * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1((long)(K * N))), dim3(256), 0, m->stream, m->W[l], (long)(K * N), m->Wb[l]);
    MLPCHK(hipGetLastError());
    o += K * N;
    MLPCHK(hipMemcpy(m->b[l], w + o, N * sizeof(float), hipMemcpyHostToDevice));
    o += N;
  }
  MLPCHK(hipStreamSynchronize(m->stream));
  m->step = 0;
  return 0;
}

int fgpe_mlp_forward(fgpe_mlp* m, const float* x, int32_t batch, float* y) {
  if (!m || !x || batch < 1 || batch > m->max_batch) return fail(FGPE_EINVAL, "bad MLP forward argument");
  int rc;
  if ((rc = load_input(m, x, batch)) || (rc = forward(m, batch))) return rc;
  if (y)
    MLPCHK(hipMemcpyAsync(y, m->H[m->nl], (size_t)batch * m->dims[m->nl] * sizeof(float), hipMemcpyDeviceToHost,
                          m->stream));
  MLPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int fgpe_mlp_gradients(fgpe_mlp* m, const float* x, const float* t, int32_t batch, float* grad, double* loss) {
  if (!m || !x || !t || batch < 1 || batch > m->max_batch) return fail(FGPE_EINVAL, "bad MLP gradient argument");
  int rc;
  if ((rc = load_input(m, x, batch)) || (rc = forward(m, batch))) return rc;
  MLPCHK(hipMemcpyAsync(m->T, t, (size_t)batch * m->dims[m->nl] * sizeof(float), hipMemcpyHostToDevice, m->stream));
  if ((rc = backward(m, batch, loss))) return rc;
  if (grad) {
    size_t o = 0;
    for (int l = 0; l < m->nl; ++l) {
      const size_t K = m->dims[l], N = m->dims[l + 1];
      MLPCHK(hipMemcpy(grad + o, m->gW[l], K * N * sizeof(float), hipMemcpyDeviceToHost));
      o += K * N;
      MLPCHK(hipMemcpy(grad + o, m->gb[l], N * sizeof(float), hipMemcpyDeviceToHost));
      o += N;
    }
  }
  return 0;
}

int fgpe_mlp_set_dataset(fgpe_mlp* m, const float* x, const float* y, int32_t rows) {
  if (!m || !x || !y || rows < 1) return fail(FGPE_EINVAL, "bad MLP dataset");
  const size_t nin = m->dims[0], nout = m->dims[m->nl];
  if (m->dsX) {
    (void)hipFree(m->dsX);
    (void)hipFree(m->dsY);
    m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->dsX), m->allocs.end());
    m->allocs.erase(std::remove(m->allocs.begin(), m->allocs.end(), (void*)m->dsY), m->allocs.end());
    m->dsX = m->dsY = nullptr;
  }
  int rc;
  if ((rc = malloc_dev(m, (void**)&m->dsX, rows * nin * sizeof(float))) ||
      (rc = malloc_dev(m, (void**)&m->dsY, rows * nout * sizeof(float))))
    return rc;
  MLPCHK(hipMemcpy(m->dsX, x, rows * nin * sizeof(float), hipMemcpyHostToDevice));
  MLPCHK(hipMemcpy(m->dsY, y, rows * nout * sizeof(float), hipMemcpyHostToDevice));
  m->ds_rows = rows;
  return 0;
}

// One Adam step on the rows idx[0..batch) of the device dataset.
int fgpe_mlp_train_step(fgpe_mlp* m, const int32_t* idx, int32_t batch, float lr, double* loss) {
  if (!m || !idx || batch < 1 || batch > m->max_batch || !m->dsX) return fail(FGPE_EINVAL, "bad MLP train argument");
  for (int i = 0; i < batch; ++i)
    if (idx[i] < 0 || idx[i] >= m->ds_rows) return fail(FGPE_EINVAL, "row index %d out of range", idx[i]);
  const int nin = m->dims[0], nout = m->dims[m->nl];
  MLPCHK(hipMemcpyAsync(m->idx, idx, batch * sizeof(int), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_gather, dim3(grid1((long)batch * nin)), dim3(256), 0, m->stream, m->dsX, nin, m->idx, batch,
                     m->H[0]);
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1((long)batch * nin)), dim3(256), 0, m->stream, m->H[0], (long)batch * nin,
                     m->Hb[0]);
  hipLaunchKernelGGL(k_gather, dim3(grid1((long)batch * nout)), dim3(256), 0, m->stream, m->dsY, nout, m->idx, batch,
                     m->T);
  MLPCHK(hipGetLastError());
  int rc;
  if ((rc = forward(m, batch)) || (rc = backward(m, batch, loss))) return rc;
  ++m->step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  for (int l = 0; l < m->nl; ++l) {
    const long kn = (long)m->dims[l] * m->dims[l + 1];
    hipLaunchKernelGGL(k_adam, dim3(grid1(kn)), dim3(256), 0, m->stream, m->W[l], m->gW[l], m->mW[l], m->vW[l], kn, lr,
                       b1, b2, eps, c1, c2, m->Wb[l]);
    hipLaunchKernelGGL(k_adam, dim3(grid1(m->dims[l + 1])), dim3(256), 0, m->stream, m->b[l], m->gb[l], m->mb[l],
                       m->vb[l], (long)m->dims[l + 1], lr, b1, b2, eps, c1, c2, (bf16*)nullptr);
  }
  MLPCHK(hipGetLastError());
  MLPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

// C = op(A) op(B) for host BF16 operands (the GEMM kernel on its own, for tests)
int fgpe_mlp_gemm_test(int32_t ta, int32_t tb, const uint16_t* A, const uint16_t* B, float* C, int32_t M, int32_t N,
                       int32_t K) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1) return fail(FGPE_EINVAL, "bad GEMM argument");
  fgpe_mlp tmp;
  if (hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking) != hipSuccess)
    return fail(FGPE_EDEVICE, "hipStreamCreate failed");
  auto done = [&](int rc) {
    (void)hipStreamSynchronize(tmp.stream);
    for (void* p : tmp.allocs) (void)hipFree(p);
    (void)hipStreamDestroy(tmp.stream);
    return rc;
  };
  bf16 *dA, *dB;
  float* dC;
  // split-K partials for up to 16 splits, at most 2^26 floats (gemm() falls
  // back to fewer splits, or none, when they do not fit)
  tmp.part_elems = std::min<size_t>((size_t)M * N * 16, (size_t)1 << 26);
  int rc;
  if ((rc = malloc_dev(&tmp, (void**)&dA, (size_t)M * K * sizeof(bf16))) ||
      (rc = malloc_dev(&tmp, (void**)&dB, (size_t)K * N * sizeof(bf16))) ||
      (rc = malloc_dev(&tmp, (void**)&dC, (size_t)M * N * sizeof(float))) ||
      (rc = malloc_dev(&tmp, (void**)&tmp.part, tmp.part_elems * sizeof(float))))
    return done(rc);
  if (hipMemcpy(dA, A, (size_t)M * K * sizeof(bf16), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dB, B, (size_t)K * N * sizeof(bf16), hipMemcpyHostToDevice) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "copy failed"));
  const int lda = ta ? M : K, ldb = tb ? K : N;
  if ((rc = gemm(&tmp, ta, tb, dA, lda, dB, ldb, dC, M, N, K))) return done(rc);
  if (hipStreamSynchronize(tmp.stream) != hipSuccess ||
      hipMemcpy(C, dC, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "GEMM failed"));
  return done(0);
}

}  // extern "C"

int fgpe_mlp_gemm_time(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, int32_t reps, double* ms) {
  if (!ms || M < 1 || N < 1 || K < 1 || reps < 1) return fail(FGPE_EINVAL, "bad GEMM timing argument");
  fgpe_mlp tmp;
  if (hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking) != hipSuccess)
    return fail(FGPE_EDEVICE, "hipStreamCreate failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto done = [&](int rc) {
    (void)hipStreamSynchronize(tmp.stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void* p : tmp.allocs) (void)hipFree(p);
    (void)hipStreamDestroy(tmp.stream);
    return rc;
  };
  bf16 *dA, *dB;
  float* dC;
  // split-K partials for up to 16 splits, at most 2^26 floats (gemm() falls
  // back to fewer splits, or none, when they do not fit)
  tmp.part_elems = std::min<size_t>((size_t)M * N * 16, (size_t)1 << 26);
  int rc;
  if ((rc = malloc_dev(&tmp, (void**)&dA, (size_t)M * K * sizeof(bf16))) ||
      (rc = malloc_dev(&tmp, (void**)&dB, (size_t)K * N * sizeof(bf16))) ||
      (rc = malloc_dev(&tmp, (void**)&dC, (size_t)M * N * sizeof(float))) ||
      (rc = malloc_dev(&tmp, (void**)&tmp.part, tmp.part_elems * sizeof(float))))
    return done(rc);
  hipLaunchKernelGGL(k_fill_bf16, dim3(grid1((long)M * K)), dim3(256), 0, tmp.stream, dA, (long)M * K, 1u);
  hipLaunchKernelGGL(k_fill_bf16, dim3(grid1((long)K * N)), dim3(256), 0, tmp.stream, dB, (long)K * N, 2u);
  if (hipGetLastError() != hipSuccess) return done(fail(FGPE_EDEVICE, "fill failed"));
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "hipEventCreate failed"));
  const int lda = ta ? M : K, ldb = tb ? K : N;
  if ((rc = gemm(&tmp, ta, tb, dA, lda, dB, ldb, dC, M, N, K))) return done(rc);
  if (hipEventRecord(e0, tmp.stream) != hipSuccess) return done(fail(FGPE_EDEVICE, "hipEventRecord failed"));
  for (int r = 0; r < reps; ++r)
    if ((rc = gemm(&tmp, ta, tb, dA, lda, dB, ldb, dC, M, N, K))) return done(rc);
  float t = 0.0f;
  if (hipEventRecord(e1, tmp.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&t, e0, e1) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "GEMM timing failed"));
  *ms = (double)t / reps;
  return done(0);
}

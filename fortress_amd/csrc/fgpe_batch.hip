// fgpe_batch.hip -- many independent q2D instances in one persistent launch.
//
// Each block of k_batch takes an instance from a work queue (an atomic
// counter over a list sorted largest grid first) and runs every pass of that
// instance itself: the x-line pass (inverse x FFT, 1/N, SE, SP, NORMC,
// forward x FFT, kx factor) and the y-line pass (pending renormalisation, y
// FFT, ky factor, SOC, inverse y FFT) of the single-instance q2D step
// (run_steps in fgpe_ops.hpp, SRC/cgpe2d.f90:440-485), with the NORMC sum and
// the renormalisation scale formed inside the block, so no grid-wide
// synchronisation is ever needed.  At the end of a launch the block applies
// the last scale, measures max|PHI - PHI_OLD| against the state of the
// previous step and evaluates ENERGY/RAD (SRC/cgpe2d.f90:488-495, 1076-1281).
// The state of an instance lives in HBM/L2 between passes; a block touches
// only its own instance, and __syncthreads orders its own global writes and
// reads.
#include <numeric>

#include "fgpe_impl.hpp"

namespace fgpe {

enum : int { BM_NORMT = 1, BM_STEPS = 2, BM_OBS = 4, BM_CONV = 8 };

template <typename T> struct BInst {
  C_t<T>* phi;       // [ms][n][n]
  C_t<T>* old;       // state of the step before the last one (real space)
  int n, sched;
  Phys ph;
  FinArgs f;
  ObsArgs oa;
  const R_t<T>* tw;   // stage twiddles of (n, Q of sched)
  const R_t<T>* kex;  // exp(-i CDT kx^2) - 1        [n]
  const R_t<T>* kel;  // exp(-i CDT ky^2) - 1 per m  [ms][n]
  const T *kx, *ky, *vx, *vy;
  double* sig;        // NEWTON state [2]
};

constexpr int BT = 256;  // threads per block of the batch kernel

// block-wide sum of a[M] (result valid in thread 0); red: >= 4*M doubles
template <int M> __device__ __forceinline__ void bsum(double (&a)[M], double* red) {
  wave_sum<M>(a);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int m = 0; m < M; ++m) red[wid * M + m] = a[m];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int m = 0; m < M; ++m) a[m] = red[m] + red[M + m] + red[2 * M + m] + red[3 * M + m];
}

__device__ __forceinline__ double bmax(double v, double* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// One instance of line length N with Q points per thread.
template <typename T, int SPIN, int N, int Q>
__device__ __noinline__ void batch_item(const BInst<T>& I, fgpe_batch_result* R, int nsteps, int mode,
                                        unsigned char* smem) {
  using C = C_t<T>;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int NC = 3 * MS + 2;
  constexpr int P = N / Q;
  constexpr int LPI = BT / P;  // lines per iteration
  static_assert(N % LPI == 0 && BT % P == 0, "schedule does not tile the grid");
  constexpr int LS = line_stride(N);
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  constexpr long NP = (long)N * N;
  R_t<T>* lds_tw = reinterpret_cast<R_t<T>*>(smem);
  double* red = reinterpret_cast<double*>(lds_tw + TWS);  // 4*NC doubles
  double* bc = red + 4 * NC;                                // scales [MS], old scales [MS]
  C* lds = reinterpret_cast<C*>(bc + 16);
  const int tid = threadIdx.x;

  for (int i = tid; i < TWS; i += BT) lds_tw[i] = I.tw[i];
  if (tid < 16) bc[tid] = 1.0;
  __syncthreads();

  const Geom g{N, N, 1, NP};
  C* phi = I.phi;
  const T dt = (T)I.ph.dt;

  if (mode & BM_NORMT) {
    // NORMT (SRC/cgpe2d.f90:1025-1066): scale to total norm 1
    double acc[MS];
    SFor<0, MS>::run([&](auto mi) { acc[decltype(mi)::value] = 0.0; });
    for (long q = tid; q < NP; q += BT) {
      const double w = I.oa.wx[q % N] * I.oa.wy[q / N];
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        const C a = phi[m * NP + q];
        acc[m] += w * ((double)a.x * (double)a.x + (double)a.y * (double)a.y);
      });
    }
    bsum<MS>(acc, red);
    if (tid == 0) {
      double t = 0;
      for (int m = 0; m < MS; ++m) t += acc[m];
      bc[0] = 1.0 / sqrt(t);
    }
    __syncthreads();
    const T c = (T)bc[0];
    for (long q = tid; q < NP * MS; q += BT) phi[q] = cscale(phi[q], c);
    __syncthreads();
    if (tid < 16) bc[tid] = 1.0;
    __syncthreads();
  }

  double maxdiff = 0.0;
  if ((mode & BM_STEPS) && nsteps > 0) {
    if ((mode & BM_CONV) && nsteps == 1)  // PHI_OLD is the state the launch starts from
      for (long q = tid; q < NP * MS; q += BT) I.old[q] = phi[q];
    // x-line pass: flags as k_xpass; the forward FFT ends with the kx factor
    auto xphase = [&](int flags, bool keep_old, double (&acc)[MS]) __attribute__((always_inline)) {
      __syncthreads();  // the previous pass's global writes of this block are done
      const int tl = tid / P, j = tid % P;
      C* my = lds + tl * LS;
      for (int l0 = 0; l0 < N; l0 += LPI) {
        const int y = l0 + tl;
        C v[MS][Q];
        const unsigned int voff = (unsigned int)(((long)y * N + j) * (long)sizeof(C));
        SFor<0, MS>::run([&](auto mi) {
          constexpr int m = decltype(mi)::value;
          const CompBuf<C> buf(phi + m * NP, NP);
          SFor<0, Q>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            v[m][q] = buf.ld(voff, (unsigned int)(P * q * sizeof(C)));
          });
        });
        if (flags & XF_INV)
          SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, +1>(v[decltype(mi)::value], my, lds_tw, j); });
        if (flags & XF_PHYS) {
          const T vy = I.vy[y];
          const double wy = I.oa.wy[y];
          SFor<0, Q>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            const int x = j + P * q;
            C pt[MS];
            SFor<0, MS>::run([&](auto mi) {
              constexpr int m = decltype(mi)::value;
              pt[m] = scalec(v[m][q], I.ph.inv_npts);
            });
            if constexpr (SPIN == 1)
              se1_point<C, T>(pt, I.ph, dt, sizeof(T) == 4 ? (T)0.05 : (T)1e-3);
            else
              se2_point<C, T>(pt, I.ph, dt, (T)1, sizeof(T) == 4 ? (T)1e-8 : (T)1e-17);
            sp_point<SPIN, C, T>(pt, I.vx[x] + vy, I.ph, dt);
            const double w = I.oa.wx[x] * wy;
            SFor<0, MS>::run([&](auto mi) {
              constexpr int m = decltype(mi)::value;
              v[m][q] = pt[m];
              acc[m] += w * ((double)pt[m].x * (double)pt[m].x + (double)pt[m].y * (double)pt[m].y);
              if (keep_old) I.old[m * NP + (long)y * N + x] = pt[m];
            });
          });
        }
        if (flags & XF_FWD) {
          SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, -1>(v[decltype(mi)::value], my, lds_tw, j); });
          SFor<0, Q>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            const R_t<T> f = I.kex[j + P * q];
            SFor<0, MS>::run([&](auto mi) {
              constexpr int m = decltype(mi)::value;
              v[m][q] = rot1(v[m][q], f);
            });
          });
        }
        SFor<0, MS>::run([&](auto mi) {
          constexpr int m = decltype(mi)::value;
          const CompBuf<C> buf(phi + m * NP, NP);
          SFor<0, Q>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            buf.st(v[m][q], voff, (unsigned int)(P * q * sizeof(C)));
          });
        });
      }
    };
    // y-line pass: pending scale, forward y FFT, ky factor, SOC, inverse y FFT
    auto yphase = [&]() __attribute__((always_inline)) {
      __syncthreads();  // the previous pass's global writes of this block are done
      const int l = tid % LPI, j = tid / LPI;
      C* my = lds + l * LS;
      C sc[MS];
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        sc[m] = mk<C>((T)bc[m], (T)0);
      });
      for (int x0 = 0; x0 < N; x0 += LPI) {
        const int x = x0 + l;
        C v[MS][Q];
        const unsigned int voff = (unsigned int)(((long)j * N + x) * (long)sizeof(C));
        const unsigned int sstep = (unsigned int)((long)P * N * (long)sizeof(C));
        SFor<0, MS>::run([&](auto mi) {
          constexpr int m = decltype(mi)::value;
          const CompBuf<C> buf(phi + m * NP, NP);
          SFor<0, Q>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            v[m][q] = cmul(buf.ld(voff, q * sstep), sc[m]);
          });
        });
        SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, -1>(v[decltype(mi)::value], my, lds_tw, j); });
        SFor<0, Q>::run([&](auto qi) {
          constexpr int q = decltype(qi)::value;
          const int y = j + P * q;
          C e1, dm1[MS], pt[MS], cr[MS];
          soc_factors<SPIN, C, T>(I.kx[x], I.ky[y], I.ph, dt, e1, dm1);
          SFor<0, MS>::run([&](auto mi) {
            constexpr int m = decltype(mi)::value;
            v[m][q] = rot1(v[m][q], I.kel[m * N + y]);
            pt[m] = v[m][q];
          });
          soc_corr<SPIN, C, T>(pt, e1, dm1, cr);
          SFor<0, MS>::run([&](auto mi) {
            constexpr int m = decltype(mi)::value;
            v[m][q] = cadd(v[m][q], cr[m]);
          });
        });
        SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, +1>(v[decltype(mi)::value], my, lds_tw, j); });
        SFor<0, MS>::run([&](auto mi) {
          constexpr int m = decltype(mi)::value;
          const CompBuf<C> buf(phi + m * NP, NP);
          SFor<0, Q>::run([&](auto qi) {
            constexpr int q = decltype(qi)::value;
            buf.st(v[m][q], voff, q * sstep);
          });
        });
      }
    };

    double acc[MS];
    xphase(XF_FWD, false, acc);
    for (int s = 1; s <= nsteps; ++s) {
      yphase();
      SFor<0, MS>::run([&](auto mi) { acc[decltype(mi)::value] = 0.0; });
      const bool keep = (mode & BM_CONV) && nsteps >= 2 && s == nsteps - 1;
      xphase(XF_INV | XF_PHYS | (s < nsteps ? XF_FWD : 0), keep, acc);
      bsum<MS>(acc, red);
      if (tid == 0) {
        double c[5], nn[5] = {0, 0, 0, 0, 0};
        int err = 0;
        for (int m = 0; m < MS; ++m) nn[m] = acc[m];
        renorm_scale(I.f, nn, I.sig, &err, c);
        if (err) R->err = 1;
        for (int m = 0; m < MS; ++m) {
          if (keep) bc[8 + m] = c[m];  // scale that normalises PHI_OLD
          bc[m] = c[m];
          R->norm[m] = nn[m];
        }
      }
      __syncthreads();
    }
    // the last step ends in real space: apply its scale, compare with PHI_OLD
    C sc[MS], so[MS];
    SFor<0, MS>::run([&](auto mi) {
      constexpr int m = decltype(mi)::value;
      sc[m] = mk<C>((T)bc[m], (T)0);
      so[m] = mk<C>(nsteps >= 2 ? (T)bc[8 + m] : (T)1, (T)0);
    });
    for (long q = tid; q < NP; q += BT) {
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        const C a = cmul(phi[m * NP + q], sc[m]);
        phi[m * NP + q] = a;
        if (mode & BM_CONV) {
          const C b = cmul(I.old[m * NP + q], so[m]);
          const double dr = (double)a.x - (double)b.x, di = (double)a.y - (double)b.y;
          maxdiff = fmax(maxdiff, sqrt(dr * dr + di * di));
        }
      });
    }
    if (mode & BM_CONV) maxdiff = bmax(maxdiff, red);
    __syncthreads();
  }

  if (mode & BM_OBS) {
    double acc[NC];
    SFor<0, NC>::run([&](auto ci) { acc[decltype(ci)::value] = 0.0; });
    for (long q = tid; q < NP; q += BT) obs_point<T, SPIN, 2>(phi, g, I.oa, q, acc);
    bsum<NC>(acc, red);
    if (tid == 0) {
      double sumn = 0;
      for (int m = 0; m < MS; ++m) sumn += acc[2 * MS + 2 + m];
      R->en = acc[0] / sumn;
      R->mz = acc[1];
      for (int m = 0; m < MS; ++m) {
        R->mu[m] = acc[2 + m] / (acc[2 * MS + 2 + m] + 1.0e-40);
        R->rms[m] = sqrt(acc[MS + 2 + m]);
      }
    }
  }
  if (tid == 0) {
    if (mode & BM_STEPS) R->steps += nsteps;
    if (mode & BM_CONV) R->maxdiff = maxdiff;
  }
  __syncthreads();
}

template <typename T, int SPIN>
__global__ __launch_bounds__(BT) void k_batch(const BInst<T>* __restrict__ inst, const int* __restrict__ items,
                                              int nitems, unsigned int* counter, int nsteps, int mode,
                                              fgpe_batch_result* res) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = item;
    if (it >= nitems) return;
    const int id = items[it];
    const BInst<T>& I = inst[id];
    fgpe_batch_result* R = res + id;
    switch (I.sched) {
      case 0: batch_item<T, SPIN, 64, 8>(I, R, nsteps, mode, smem); break;
      case 1: batch_item<T, SPIN, 64, 4>(I, R, nsteps, mode, smem); break;
      case 2: batch_item<T, SPIN, 128, 8>(I, R, nsteps, mode, smem); break;
      case 3: batch_item<T, SPIN, 128, 16>(I, R, nsteps, mode, smem); break;
      case 4: batch_item<T, SPIN, 192, 12>(I, R, nsteps, mode, smem); break;
      case 5: batch_item<T, SPIN, 192, 6>(I, R, nsteps, mode, smem); break;
      case 6: batch_item<T, SPIN, 256, 8>(I, R, nsteps, mode, smem); break;
      default: batch_item<T, SPIN, 256, 16>(I, R, nsteps, mode, smem); break;
    }
  }
}

}  // namespace fgpe

using namespace fgpe_impl;

namespace {

constexpr int SCHED_N[FGPE_BATCH_NSCHED] = {64, 64, 128, 128, 192, 192, 256, 256};
constexpr int SCHED_Q[FGPE_BATCH_NSCHED] = {8, 4, 8, 16, 12, 6, 8, 16};
int size_class(int n) { return n == 64 ? 0 : n == 128 ? 1 : n == 192 ? 2 : n == 256 ? 3 : -1; }

template <typename T> size_t sched_lds(int s) {
  const int n = SCHED_N[s], q = SCHED_Q[s], p = n / q, lpi = BT / p;
  const int tws = std::max(twiddle_size(n, q), 1);
  return tws * sizeof(R_t<T>) + (4 * 17 + 16) * sizeof(double) + (size_t)lpi * line_stride(n) * sizeof(C<T>);
}

}  // namespace

struct fgpe_batch {
  int count = 0, spin = 1, ms = 3;
  bool fp32 = true;
  int dev = 0, ncu = 1;
  size_t lds_optin = 65536;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<fgpe_params> p;
  std::vector<int> n, sched;
  std::vector<long> off;
  void* phi = nullptr;
  void* old = nullptr;
  void* stage = nullptr;
  size_t stage_elems = 0;
  void* inst = nullptr;          // device BInst<T>[count]
  std::vector<unsigned char> inst_host;
  size_t inst_size = 0;
  fgpe_batch_result* res_dev = nullptr;
  std::vector<fgpe_batch_result> res;
  std::vector<double> en_old;
  unsigned int* counter = nullptr;
  int* items = nullptr;
  double last_ms = 0;
  const void* tw_sched[FGPE_BATCH_NSCHED] = {};
  std::vector<void*> allocs;
};

namespace {

int balloc(fgpe_batch* b, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(FGPE_EDEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  b->allocs.push_back(*p);
  return 0;
}
template <typename V> int bupload(fgpe_batch* b, const V** dst, const std::vector<V>& src) {
  void* d;
  int rc = balloc(b, &d, src.size() * sizeof(V));
  if (rc) return rc;
  HIPCHK(hipMemcpy(d, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice));
  *dst = (const V*)d;
  return 0;
}

template <typename T> BInst<T>* host_inst(fgpe_batch* b) { return reinterpret_cast<BInst<T>*>(b->inst_host.data()); }

template <typename T> int build_batch(fgpe_batch* b) {
  b->inst_size = sizeof(BInst<T>);
  b->inst_host.assign(b->inst_size * b->count, 0);
  BInst<T>* in = host_inst<T>(b);
  int rc;
  // twiddles per schedule (shared by all instances of that schedule)
  for (int s = 0; s < FGPE_BATCH_NSCHED; ++s) {
    const R_t<T>* t;
    if ((rc = bupload(b, &t, stage_twiddles<T>(SCHED_N[s], SCHED_Q[s])))) return rc;
    b->tw_sched[s] = (const void*)t;
  }
  double* sigd;
  if ((rc = balloc(b, (void**)&sigd, 2 * sizeof(double) * b->count))) return rc;
  HIPCHK(hipMemset(sigd, 0, 2 * sizeof(double) * b->count));
  for (int i = 0; i < b->count; ++i) {
    const fgpe_params& p = b->p[i];
    const int n = b->n[i];
    BInst<T>& I = in[i];
    I.phi = (C<T>*)b->phi + b->off[i];
    I.old = (C<T>*)b->old + b->off[i];
    I.n = n;
    I.sched = b->sched[i];
    I.ph.tau0 = p.tau[0];
    I.ph.tau1 = p.tau[1];
    I.ph.tau2 = p.tau[2];
    I.ph.dt = p.dt;
    I.ph.gx = p.gamma[0];
    I.ph.gy = p.gamma[1];
    I.ph.gz = p.gamma[2];
    I.ph.inv_npts = 1.0 / ((double)n * n);
    I.ph.imag = p.imag;
    I.ph.dim = 2;
    I.f.imag = p.imag;
    I.f.soc = p.soc;
    I.f.dim = 2;
    I.f.ms = b->ms;
    I.f.mag = p.mag;
    I.f.dt = p.dt;
    I.sig = sigd + 2 * i;
    auto kx = kgrid(n, p.d[0]), ky = kgrid(n, p.d[1]);
    auto xx = xgrid(n, p.d[0]), yy = xgrid(n, p.d[1]);
    std::vector<R_t<T>> kex(n), kel((size_t)b->ms * n);
    for (int k = 0; k < n; ++k) kex[k] = ke_factor<T>(kx[k] * kx[k], p.dt, p.imag);
    for (int m = 0; m < b->ms; ++m)
      for (int k = 0; k < n; ++k) kel[(size_t)m * n + k] = ke_factor<T>(ky[k] * ky[k], p.dt, p.imag);
    std::vector<T> kxt(kx.begin(), kx.end()), kyt(ky.begin(), ky.end());
    std::vector<double> x2(n), y2(n), vxd(n), vyd(n), zero(1, 0.0);
    for (int k = 0; k < n; ++k) {
      x2[k] = xx[k] * xx[k];
      y2[k] = yy[k] * yy[k];
      vxd[k] = p.alpha[0] * p.alpha[0] * x2[k];
      vyd[k] = p.alpha[1] * p.alpha[1] * y2[k];
    }
    std::vector<T> vx(vxd.begin(), vxd.end()), vy(vyd.begin(), vyd.end());
    if ((rc = bupload(b, &I.kex, kex)) || (rc = bupload(b, &I.kel, kel)) || (rc = bupload(b, &I.kx, kxt)) ||
        (rc = bupload(b, &I.ky, kyt)) || (rc = bupload(b, &I.vx, vx)) || (rc = bupload(b, &I.vy, vy)))
      return rc;
    ObsArgs& a = I.oa;
    a.tau0 = p.tau[0];
    a.tau1 = p.tau[1];
    a.tau2 = p.tau[2];
    a.gx = p.gamma[0];
    a.gy = p.gamma[1];
    a.gz = 0.0;
    a.dx = p.d[0];
    a.dy = p.d[1];
    a.dz = 1.0;
    if ((rc = bupload(b, &a.wx, simpson_weights(n, p.d[0]))) || (rc = bupload(b, &a.wy, simpson_weights(n, p.d[1]))) ||
        (rc = bupload(b, &a.wz, std::vector<double>(1, 1.0))) || (rc = bupload(b, &a.x2, x2)) ||
        (rc = bupload(b, &a.y2, y2)) || (rc = bupload(b, &a.z2, zero)) || (rc = bupload(b, &a.vxd, vxd)) ||
        (rc = bupload(b, &a.vyd, vyd)) || (rc = bupload(b, &a.vzd, zero)))
      return rc;
    I.tw = (const R_t<T>*)b->tw_sched[I.sched];
  }
  if ((rc = balloc(b, &b->inst, b->inst_size * b->count))) return rc;
  return 0;
}

// Launch k_batch over the listed instances.
template <typename T> int launch(fgpe_batch* b, const std::vector<int>& list, int nsteps, int mode, bool timed) {
  if (list.empty()) return 0;
  size_t lds = 0;
  for (int id : list) lds = std::max(lds, sched_lds<T>(b->sched[id]));
  if (lds > b->lds_optin) return fail(FGPE_EINVAL, "batch schedule needs %zu B of LDS", lds);
  auto kern = k_batch<T, 1>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BT, lds) != hipSuccess || per_cu < 1) per_cu = 1;
  const int blocks = (int)std::min<long>((long)list.size(), (long)b->ncu * per_cu);
  HIPCHK(hipMemcpyAsync(b->items, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemsetAsync(b->counter, 0, sizeof(unsigned int), b->stream));
  if (timed) HIPCHK(hipEventRecord(b->ev0, b->stream));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(BT), lds, b->stream, (const BInst<T>*)b->inst, b->items,
                     (int)list.size(), b->counter, nsteps, mode, b->res_dev);
  HIPCHK(hipGetLastError());
  if (timed) HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (timed) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
  }
  return 0;
}

int launch_any(fgpe_batch* b, const std::vector<int>& list, int nsteps, int mode, bool timed) {
  return b->fp32 ? launch<float>(b, list, nsteps, mode, timed) : launch<double>(b, list, nsteps, mode, timed);
}

// largest grids first, so the tail of the queue holds the short work items
std::vector<int> sorted_list(const fgpe_batch* b, bool only_active) {
  std::vector<int> l;
  for (int i = 0; i < b->count; ++i)
    if (!only_active || !b->res[i].done) l.push_back(i);
  std::stable_sort(l.begin(), l.end(), [b](int a, int c) { return b->n[a] > b->n[c]; });
  return l;
}

int pull_results(fgpe_batch* b) {
  HIPCHK(hipMemcpy(b->res.data(), b->res_dev, b->count * sizeof(fgpe_batch_result), hipMemcpyDeviceToHost));
  return 0;
}
int push_results(fgpe_batch* b) {
  HIPCHK(hipMemcpy(b->res_dev, b->res.data(), b->count * sizeof(fgpe_batch_result), hipMemcpyHostToDevice));
  return 0;
}

// point every descriptor at its schedule's twiddles and upload the array
template <typename T> int build_batch_sched(fgpe_batch* b) {
  BInst<T>* in = host_inst<T>(b);
  for (int i = 0; i < b->count; ++i) {
    in[i].sched = b->sched[i];
    in[i].tw = (const R_t<T>*)b->tw_sched[b->sched[i]];
  }
  HIPCHK(hipMemcpy(b->inst, b->inst_host.data(), b->inst_size * b->count, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

// The batch kernel faulted on its first device run (an illegal memory access
// inside the first launch of fgpe_batch_initial_guess) and the cause has not
// been found: every entry point refuses to run until it is.
static const bool kBatchEnabled = false;

int fgpe_batch_create(const fgpe_params* p, int32_t count, fgpe_batch** out) {
  if (!p || !out || count < 1) return fail(FGPE_EINVAL, "null argument or empty batch");
  *out = nullptr;
  if (!kBatchEnabled)
    return fail(FGPE_EINVAL, "batched launch disabled in this build: its kernel faulted on the device and the cause "
                             "is not yet found");
  for (int i = 0; i < count; ++i) {
    if (p[i].dim != 2) return fail(FGPE_EINVAL, "batch instance %d: only q2D instances are batched", i);
    if (p[i].spin != 1) return fail(FGPE_EINVAL, "batch instance %d: only spin-1 instances are batched", i);
    if (p[i].n[0] != p[i].n[1] || size_class(p[i].n[0]) < 0)
      return fail(FGPE_EINVAL, "batch instance %d: grid %dx%d (square 64, 128, 192 or 256)", i, p[i].n[0], p[i].n[1]);
    if (p[i].precision != p[0].precision) return fail(FGPE_EINVAL, "batch instance %d: mixed precision", i);
    if (!(p[i].dt > 0) || !(p[i].d[0] > 0) || !(p[i].d[1] > 0)) return fail(FGPE_EINVAL, "batch instance %d: bad DT/DX", i);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FGPE_EDEVICE, "no HIP device");
  fgpe_batch* b = new fgpe_batch();
  auto bail = [&](int rc) {
    fgpe_batch_destroy(b);
    return rc;
  };
  b->count = count;
  b->spin = 1;
  b->ms = 3;
  b->fp32 = p[0].precision == FGPE_FP32;
  b->p.assign(p, p + count);
  if (hipGetDevice(&b->dev) != hipSuccess) return bail(fail(FGPE_EDEVICE, "hipGetDevice failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->dev) == hipSuccess) {
    b->ncu = prop.multiProcessorCount;
    b->lds_optin = std::max<size_t>(prop.sharedMemPerBlock, 65536);
  }
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, b->dev) == hipSuccess && optin > 0)
    b->lds_optin = (size_t)optin;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipStreamCreate failed"));
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipEventCreate failed"));
  const size_t es = b->fp32 ? sizeof(float2) : sizeof(double2);
  long total = 0;
  b->n.resize(count);
  b->sched.resize(count);
  b->off.resize(count);
  for (int i = 0; i < count; ++i) {
    const int n = p[i].n[0];
    b->n[i] = n;
    b->sched[i] = 2 * size_class(n);  // default: the first schedule of the length
    b->off[i] = total;
    total += (long)n * n * b->ms;
    b->stage_elems = std::max(b->stage_elems, (size_t)n * n * b->ms);
  }
  int rc;
  if ((rc = balloc(b, &b->phi, total * es)) || (rc = balloc(b, &b->old, total * es)) ||
      (rc = balloc(b, &b->stage, b->stage_elems * sizeof(double2))) ||
      (rc = balloc(b, (void**)&b->res_dev, count * sizeof(fgpe_batch_result))) ||
      (rc = balloc(b, (void**)&b->counter, sizeof(unsigned int))) ||
      (rc = balloc(b, (void**)&b->items, count * sizeof(int))))
    return bail(rc);
  HIPCHK(hipMemset(b->phi, 0, total * es));
  HIPCHK(hipMemset(b->old, 0, total * es));
  b->res.assign(count, fgpe_batch_result{});
  b->en_old.assign(count, 0.0);
  if ((rc = push_results(b))) return bail(rc);
  rc = b->fp32 ? build_batch<float>(b) : build_batch<double>(b);
  if (rc) return bail(rc);
  HIPCHK(hipMemcpy(b->inst, b->inst_host.data(), b->inst_size * count, hipMemcpyHostToDevice));
  *out = b;
  return 0;
}

void fgpe_batch_destroy(fgpe_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (void* a : b->allocs) (void)hipFree(a);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int fgpe_batch_initial_guess(fgpe_batch* b) {
  if (!b) return fail(FGPE_EINVAL, "null batch");
  for (int i = 0; i < b->count; ++i) {
    const fgpe_params& p = b->p[i];
    const int n = b->n[i];
    auto xx = xgrid(n, p.d[0]), yy = xgrid(n, p.d[1]);
    std::vector<double> x2(n), y2(n), z2(1, 0.0);
    for (int k = 0; k < n; ++k) {
      x2[k] = xx[k] * xx[k];
      y2[k] = yy[k] * yy[k];
    }
    // small per-instance tables in the staging buffer: x2 | y2 | z2
    double* st = (double*)b->stage;
    HIPCHK(hipMemcpy(st, x2.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(st + n, y2.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(st + 2 * n, z2.data(), sizeof(double), hipMemcpyHostToDevice));
    GuessArgs a;
    a.dim = 2;
    a.spin = 1;
    a.fpc = p.option_fpc;
    a.ax = p.alpha[0];
    a.ay = p.alpha[1];
    a.az = p.alpha[2];
    a.t0 = p.tau[0];
    a.t1 = p.tau[1];
    a.t2 = p.tau[2];
    a.mag = p.mag;
    a.x2 = st;
    a.y2 = st + n;
    a.z2 = st + 2 * n;
    const Geom g{n, n, 1, (long)n * n};
    const int blocks = std::min(grid_for(g.npts, 256), 1024);
    if (b->fp32)
      hipLaunchKernelGGL(k_guess<float>, dim3(blocks), dim3(256), 0, b->stream, (float2*)b->phi + b->off[i], g, a);
    else
      hipLaunchKernelGGL(k_guess<double>, dim3(blocks), dim3(256), 0, b->stream, (double2*)b->phi + b->off[i], g, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  for (auto& r : b->res) r = fgpe_batch_result{};
  int rc = push_results(b);
  if (rc) return rc;
  if ((rc = launch_any(b, sorted_list(b, false), 0, BM_NORMT | BM_OBS, false))) return rc;
  if ((rc = pull_results(b))) return rc;
  for (int i = 0; i < b->count; ++i) b->en_old[i] = b->res[i].en;
  return 0;
}

int fgpe_batch_upload(fgpe_batch* b, int32_t i, const double* phi) {
  if (!b || !phi || i < 0 || i >= b->count) return fail(FGPE_EINVAL, "bad batch upload argument");
  const long n = (long)b->n[i] * b->n[i] * b->ms;
  if (!b->fp32) {
    HIPCHK(hipMemcpy((double2*)b->phi + b->off[i], phi, n * sizeof(double2), hipMemcpyHostToDevice));
    return 0;
  }
  HIPCHK(hipMemcpy(b->stage, phi, n * sizeof(double2), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_from64<float>, dim3(std::min(grid_for(n, 256), 4096)), dim3(256), 0, b->stream,
                     (const double2*)b->stage, (float2*)b->phi + b->off[i], n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fgpe_batch_download(fgpe_batch* b, int32_t i, double* phi) {
  if (!b || !phi || i < 0 || i >= b->count) return fail(FGPE_EINVAL, "bad batch download argument");
  const long n = (long)b->n[i] * b->n[i] * b->ms;
  if (!b->fp32) {
    HIPCHK(hipMemcpy(phi, (double2*)b->phi + b->off[i], n * sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
  }
  hipLaunchKernelGGL(k_to64<float>, dim3(std::min(grid_for(n, 256), 4096)), dim3(256), 0, b->stream,
                     (const float2*)b->phi + b->off[i], (double2*)b->stage, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(phi, b->stage, n * sizeof(double2), hipMemcpyDeviceToHost));
  return 0;
}

int fgpe_batch_run(fgpe_batch* b, int64_t nsteps) {
  if (!b || nsteps < 0 || nsteps > (1 << 30)) return fail(FGPE_EINVAL, "bad batch run argument");
  int rc = launch_any(b, sorted_list(b, false), (int)nsteps, BM_STEPS | BM_OBS | BM_CONV, true);
  if (rc) return rc;
  return pull_results(b);
}

int fgpe_batch_solve(fgpe_batch* b, int32_t stp, double tol, int64_t max_steps, int32_t* converged) {
  if (!b || stp < 1 || max_steps < 0) return fail(FGPE_EINVAL, "bad batch solve argument");
  double total_ms = 0;
  int rc;
  for (;;) {
    std::vector<int> list;
    for (int id : sorted_list(b, true))
      if (b->res[id].steps + stp <= max_steps) list.push_back(id);
    if (list.empty()) break;
    if ((rc = launch_any(b, list, stp, BM_STEPS | BM_OBS | BM_CONV, true))) return rc;
    total_ms += b->last_ms;
    if ((rc = pull_results(b))) return rc;
    bool changed = false;
    for (int id : list) {
      fgpe_batch_result& r = b->res[id];
      const double den = stp * 2.0 * b->p[id].dt;
      r.conv = r.maxdiff / den;
      r.conv_en = std::fabs(r.en - b->en_old[id]) / den;
      b->en_old[id] = r.en;
      if (r.conv <= tol || r.conv_en <= tol || r.err) {
        r.done = 1;
        changed = true;
      }
    }
    if (changed && (rc = push_results(b))) return rc;
  }
  b->last_ms = total_ms;
  int nc = 0;
  for (const auto& r : b->res) nc += r.done ? 1 : 0;
  if (converged) *converged = nc;
  return 0;
}

int fgpe_batch_result_get(fgpe_batch* b, int32_t i, fgpe_batch_result* r) {
  if (!b || !r || i < 0 || i >= b->count) return fail(FGPE_EINVAL, "bad batch result argument");
  *r = b->res[i];
  return 0;
}

int fgpe_batch_last_ms(fgpe_batch* b, double* ms) {
  if (!b || !ms) return fail(FGPE_EINVAL, "null argument");
  *ms = b->last_ms;
  return 0;
}

int fgpe_batch_tune(fgpe_batch* b, int32_t trial_steps, double* ms_out, int32_t* chosen) {
  if (!b || trial_steps < 1) return fail(FGPE_EINVAL, "bad batch tune argument");
  const size_t es = b->fp32 ? sizeof(float2) : sizeof(double2);
  long total = 0;
  for (int i = 0; i < b->count; ++i) total += (long)b->n[i] * b->n[i] * b->ms;
  // trials advance the state: keep a copy and restore it afterwards
  void* backup = nullptr;
  HIPCHK(hipMalloc(&backup, total * es));
  auto done = [&](int rc) {
    (void)hipFree(backup);
    return rc;
  };
  if (hipMemcpy(backup, b->phi, total * es, hipMemcpyDeviceToDevice) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "backup copy failed"));
  std::vector<fgpe_batch_result> saved = b->res;
  double ms[FGPE_BATCH_NSCHED] = {0};
  int best[4] = {-1, -1, -1, -1};
  for (int cls = 0; cls < 4; ++cls) {
    std::vector<int> list;
    for (int i = 0; i < b->count; ++i)
      if (size_class(b->n[i]) == cls) list.push_back(i);
    if (list.empty()) continue;
    double best_ms = 1e300;
    for (int s = 2 * cls; s < 2 * cls + 2; ++s) {
      for (int id : list) b->sched[id] = s;
      int rc = b->fp32 ? build_batch_sched<float>(b) : build_batch_sched<double>(b);
      if (rc) return done(rc);
      // one untimed launch to warm up, then the timed one
      if ((rc = launch_any(b, list, 1, BM_STEPS, false))) return done(rc);
      if ((rc = launch_any(b, list, trial_steps, BM_STEPS, true))) return done(rc);
      ms[s] = b->last_ms;
      if (b->last_ms < best_ms) {
        best_ms = b->last_ms;
        best[cls] = s;
      }
    }
    for (int id : list) b->sched[id] = best[cls];
  }
  int rc = b->fp32 ? build_batch_sched<float>(b) : build_batch_sched<double>(b);
  if (rc) return done(rc);
  if (hipMemcpy(b->phi, backup, total * es, hipMemcpyDeviceToDevice) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "restore copy failed"));
  b->res = saved;
  if ((rc = push_results(b))) return done(rc);
  if (ms_out)
    for (int s = 0; s < FGPE_BATCH_NSCHED; ++s) ms_out[s] = ms[s];
  if (chosen)
    for (int c = 0; c < 4; ++c) chosen[c] = best[c];
  return done(0);
}

}  // extern "C"

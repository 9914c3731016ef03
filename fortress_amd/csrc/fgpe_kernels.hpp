// fgpe_kernels.hpp -- the propagator passes of libfgpe.
//
// State layout in HBM: planar per component, phi[m][z][y][x] (x fastest),
// float2 (FP32 mode) or double2 (FP64 mode).  One time step of the reference
// loop (SRC/cgpe3d.f90:452-495) becomes line passes that keep whole FFT lines
// in registers/LDS:
//   3D: A  inverse-x + SE + SP + NORMC partials + forward-x, all comps
//       B  forward-y + the kx and ky factors of KE, per component
//       C  forward-z + the kz/GAMMAZ factor of KE + inverse-z, per component
//          (applies the pending renormalisation)
//       D  SOC in (kx, ky, z) space + inverse-y + 1/N, all comps
//   2D: K  forward-y + KE + SOC + inverse-y (all comps), then A
//   1D: one workgroup runs whole steps (k_line1d)
// followed by a one-block finalize that turns the NORMC partials into the
// norms and the per-component scale of the imaginary-time renormalisation.
#pragma once
#include "fgpe_device.hpp"

namespace fgpe {

enum : int { XF_INV = 1, XF_PHYS = 2, XF_FWD = 4, XF_SCALE = 8, XF_NORM = 16 };

// KE (SRC/cgpe3d.f90:875-899) is a product of one factor per axis, each
// diagonal in its wave number, so each factor is applied in a pass that has
// that axis in k space: 2D kx after the forward x FFT (pass A) and ky in the
// k pass; 3D kx and ky after the forward y FFT (pass B) and kz with the
// GAMMAZ term in pass C.  Each factor w is stored as w - 1 (rot_of format)
// and applied as v + v (w - 1).
template <typename T> struct Tables {
  const R_t<T>* twx;  // per-stage FFT twiddles, x axis
  const R_t<T>* twy;
  const R_t<T>* twz;
  const R_t<T>* twd;  // y axis for the 3D all-component pass D
  const R_t<T>* kex;  // exp(-i CDT kx^2) - 1                          [NX]
  const R_t<T>* key;  // exp(-i CDT ky^2) - 1 (3D)                     [NY or 1]
  const R_t<T>* kel;  // line axis: 3D exp(-i CDT (kz^2 - 2 L GAMMAZ kz)) - 1,
                      // 2D exp(-i CDT ky^2) - 1 (the same in every row;
                      // the 2D k pass reads row 0), per component  [MS][N_line]
  const T* kx;        // wave numbers (FFT order, positive Nyquist)
  const T* ky;
  const T* vx;  // ALPHAX^2 x^2
  const T* vy;
  const T* vz;
  const double* wx;  // Simpson weights (SRC/simpson.f90)
  const double* wy;
  const double* wz;
  const C_t<T>* cscale;  // pending per-component renormalisation [MS]
  const C_t<T>* soc2;    // 2D SOC factors of kx, ky >= 0: [NY/2+1][NX/2+1][1 + SPIN] (soc2_table)
};

struct Geom {
  int nx, ny, nz;
  long npts;
};

// ------------------------------------------------ renormalisation logic
struct FinArgs {
  int imag, soc, dim, ms;
  double mag, dt;
};

// NEWTON/USRFUN (SRC/cgpe3d.f90:1516-1588) on the device
__device__ __forceinline__ int newton_dev(double* x, const double* n, double mag, double dt) {
  double p0 = 0, p1 = 0;
  for (int k = 0; k < 1000; ++k) {
    double e1 = exp(2.0 * (x[0] + 2.0 * x[1]) * dt), e2 = exp(2.0 * (x[0] + x[1]) * dt), e3 = exp(2.0 * x[0] * dt),
           e4 = exp(2.0 * (x[0] - x[1]) * dt), e5 = exp(2.0 * (x[0] - 2.0 * x[1]) * dt);
    double f0 = e1 * n[0] + e2 * n[1] + e3 * n[2] + e4 * n[3] + e5 * n[4] - 1.0;
    double f1 = 2.0 * e1 * n[0] + e2 * n[1] - e4 * n[3] - 2.0 * e5 * n[4] - mag;
    double j11 = (e1 * n[0] + e2 * n[1] + e3 * n[2] + e4 * n[3] + e5 * n[4]) * 2.0 * dt;
    double j12 = e1 * n[0] * 4.0 * dt + e2 * n[1] * 2.0 * dt + e4 * n[3] * (-2.0 * dt) + e5 * n[4] * (-4.0 * dt);
    double j21 = (2.0 * e1 * n[0] + e2 * n[1] - e4 * n[3] - 2.0 * e5 * n[4]) * 2.0 * dt;
    double j22 = 2.0 * e1 * n[0] * 4.0 * dt + e2 * n[1] * 2.0 * dt - e4 * n[3] * (-2.0 * dt) -
                 2.0 * e5 * n[4] * (-4.0 * dt);
    p1 = (f1 * j11 - f0 * j21) / (j21 * j12 - j11 * j22);
    p0 = (-f0 - j12 * p1) / j11;
    x[0] += p0;
    x[1] += p1;
    if (fmax(fabs(p0), fabs(p1)) <= 1.0e-6) break;
  }
  // a step that is not <= TOL_NR fails, a NaN one included (the reference's
  // MAXVAL(ABS(P)).GT.TOL_NR lets NaN through and the state turns NaN)
  return !(fabs(p0) <= 1.0e-6 && fabs(p1) <= 1.0e-6);
}

// norms n[ms] -> per-component scale c[ms] applied after the step:
// real time 1; imaginary time sigma_m / TNORM with sigma from the SOC = 0
// projection (SRC/cgpe3d.f90:460-491) or 1 with SO coupling (493).
__device__ __forceinline__ void renorm_scale(const FinArgs& f, const double* n, double* sig, int* err, double* c) {
  const double EPS = 1.0e-40;
  if (!f.imag) {
    for (int m = 0; m < f.ms; ++m) c[m] = 1.0;
    return;
  }
  double sigma[5] = {1, 1, 1, 1, 1};
  if (!f.soc) {
    const double mag = f.mag;
    if (f.ms == 3) {
      double root = sqrt(4.0 * (1.0 - mag * mag) * n[0] * n[2] + mag * mag * n[1] * n[1]);
      sigma[1] = f.dim == 3 ? sqrt(1.0 - mag * mag) / sqrt(n[1] + root + EPS)
                            : sqrt(1.0 - mag * mag) / (sqrt(n[1] + root) + EPS);
      sigma[0] = sqrt(1.0 + mag - sigma[1] * sigma[1] * n[1]) / (sqrt(2.0 * n[0]) + EPS);
      sigma[2] = sqrt(1.0 - mag - sigma[1] * sigma[1] * n[1]) / (sqrt(2.0 * n[2]) + EPS);
    } else {
      if (newton_dev(sig, n, mag, f.dt)) *err = 1;
      const double dt = f.dt;
      sigma[0] = exp((sig[0] + 2.0 * sig[1]) * dt);
      sigma[1] = exp((sig[0] + sig[1]) * dt);
      sigma[2] = exp(sig[0] * dt);
      sigma[3] = exp((sig[0] - sig[1]) * dt);
      sigma[4] = exp((sig[0] - 2.0 * sig[1]) * dt);
    }
  }
  double t2 = 0;
  for (int m = 0; m < f.ms; ++m) t2 += sigma[m] * sigma[m] * n[m];
  const double tn = sqrt(t2);
  for (int m = 0; m < f.ms; ++m) c[m] = sigma[m] / tn;
}

// ----------------------------------------------------------- reductions
// butterfly sum over the active lanes of a wave; blocks narrower than a wave
// (small 1D lines) only exchange among their own lanes
template <int M> __device__ __forceinline__ void wave_sum(double (&a)[M]) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
    if (off < (int)blockDim.x)
#pragma unroll
      for (int m = 0; m < M; ++m) a[m] += __shfl_xor(a[m], off, 64);
}

// sum a[M] over the block and write out[blk*M + m] (blk: the block index, by
// default blockIdx.x); red: >= 16*M doubles
template <int M> __device__ __forceinline__ void block_sum_store(double (&a)[M], double* red, double* out, long blk) {
  wave_sum<M>(a);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int m = 0; m < M; ++m) red[wid * M + m] = a[m];
  __syncthreads();
  if (threadIdx.x < M) {
    double s = 0;
    for (int w = 0; w < nw; ++w) s += red[w * M + threadIdx.x];
    out[blk * M + threadIdx.x] = s;
  }
}
template <int M> __device__ __forceinline__ void block_sum_store(double (&a)[M], double* red, double* out) {
  block_sum_store<M>(a, red, out, (long)blockIdx.x);
}

// Sum the NORMC partials of a pass in a fixed order (deterministic), store
// the norms and the per-component renormalisation scale of the step.
template <typename T, int MS>
__device__ __forceinline__ void finalize_block(const double* __restrict__ partials, int nparts, const FinArgs& f,
                                               double* norm_out, C_t<T>* cscale, double* cscale64, double* sig,
                                               int* err) {
  __shared__ double red[16 * MS];
  double a[MS];
  SFor<0, MS>::run([&](auto mi) { a[decltype(mi)::value] = 0.0; });
#pragma unroll 8
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    SFor<0, MS>::run([&](auto mi) {
      constexpr int m = decltype(mi)::value;
      a[m] += partials[(long)i * MS + m];
    });
  wave_sum<MS>(a);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0)
    SFor<0, MS>::run([&](auto mi) { red[wid * MS + decltype(mi)::value] = a[decltype(mi)::value]; });
  __syncthreads();
  if (threadIdx.x == 0) {
    double n[5] = {0, 0, 0, 0, 0}, c[5];
    for (int w = 0; w < nw; ++w)
      for (int m = 0; m < MS; ++m) n[m] += red[w * MS + m];
    renorm_scale(f, n, sig, err, c);
    for (int m = 0; m < MS; ++m) {
      norm_out[m] = n[m];
      cscale64[m] = c[m];
      cscale[m] = mk<C_t<T>>((T)c[m], (T)0);
    }
  }
}
template <typename T, int MS>
__global__ __launch_bounds__(1024) void k_finalize(const double* __restrict__ partials, int nparts, FinArgs f,
                                                   double* norm_out, C_t<T>* cscale, double* cscale64, double* sig,
                                                   int* err) {
  finalize_block<T, MS>(partials, nparts, f, norm_out, cscale, cscale64, sig, err);
}

// ------------------------------------------------------------ pass A (x)
// Lines along x, all components.  flags: XF_INV inverse FFT first, XF_PHYS
// real-space physics (1/N, SE, SP, NORMC partials), XF_FWD forward FFT last,
// XF_SCALE multiply by the pending scale on load, XF_NORM (with XF_PHYS) write
// the block's NORMC partials.
// FULL: the pass also applies the 1/N of BFT and the kx factor of KE (2D);
// in 3D passes D and B apply them (they commute with the y transforms), so the
// VALU-bound pass A carries less.  MULTI: the MS components' lines are
// transformed together (line_fft_multi, MS LDS lines per x line), for blocks
// of one line whose time is the transforms' barriers (q2D).
template <typename T, int SPIN, int N, int Q, bool FULL = true, bool MULTI = false>
__device__ __forceinline__ void xpass_block(int blk, C_t<T>* __restrict__ phi, const Geom& g, const Tables<T>& tb,
                                            const Phys& ph, int flags, double* __restrict__ partials) {
  using C = C_t<T>;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int P = N / Q;
  constexpr int LS = line_stride(N);
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R_t<T>* lds_tw = reinterpret_cast<R_t<T>*>(smem_raw);
  R_t<T>* lds_kex = lds_tw + TWS;
  double* red = reinterpret_cast<double*>(lds_kex + N);
  C* lds = reinterpret_cast<C*>(red + 16 * MS);

  // a line's P threads are consecutive: inside one wavefront when P divides 64
  constexpr bool WAVE = 64 % P == 0;
  const int lpw = blockDim.x / P;
  const int tl = threadIdx.x / P, j = threadIdx.x % P;
  const long line = (long)blk * lpw + tl;
  const long nlines = (long)g.ny * g.nz;
  const bool valid = line < nlines;
  const int y = valid ? (int)(line % g.ny) : 0, z = valid ? (int)(line / g.ny) : 0;
  C* my = lds + tl * (MULTI ? MS : 1) * LS;

  for (int i = threadIdx.x; i < TWS; i += blockDim.x) lds_tw[i] = tb.twx[i];
  if (flags & XF_FWD)
    if (FULL)
      for (int i = threadIdx.x; i < N; i += blockDim.x) lds_kex[i] = tb.kex[i];

  C v[MS][Q];
  const unsigned int voff = (unsigned int)((line * N + j) * (long)sizeof(C));
  SFor<0, MS>::run([&](auto mi) {
    constexpr int m = decltype(mi)::value;
    const CompBuf<C> buf(phi + m * g.npts, g.npts);
    const C sc = (flags & XF_SCALE) ? tb.cscale[m] : mk<C>((T)1, (T)0);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      const C a = buf.ld(voff, (unsigned int)(P * q * sizeof(C)));
      v[m][q] = (flags & XF_SCALE) ? cmul(a, sc) : a;
    });
  });
  __syncthreads();
  if (flags & XF_INV) {
    if constexpr (MULTI)
      line_fft_multi<N, Q, +1, MS, WAVE>(v, my, LS, lds_tw, j);
    else
      SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, +1, WAVE>(v[decltype(mi)::value], my, lds_tw, j); });
  }
  if (flags & XF_PHYS) {
    const T dt = (T)ph.dt;
    const T vyz = tb.vy[y] + tb.vz[z];
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      const int x = j + P * q;
      C pt[MS];
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        pt[m] = FULL ? scalec(v[m][q], ph.inv_npts) : v[m][q];
      });
      if constexpr (SPIN == 1)
        se1_point<C, T>(pt, ph, dt, sizeof(T) == 4 ? (T)0.05 : (T)1e-3);
      else
        se2_point<C, T>(pt, ph, dt, ph.dim == 1 ? (T)0.5 : (T)1, sizeof(T) == 4 ? (T)1e-8 : (T)1e-17);
      sp_point<SPIN, C, T>(pt, tb.vx[x] + vyz, ph, dt);
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        v[m][q] = pt[m];
      });
      // one point at a time: keeps the SE temporaries of a single point live
      __builtin_amdgcn_sched_barrier(0);
    });
    if (flags & XF_NORM) {
      // NORMC partials after the point loop, so the FP64 sums are not live
      // across SE and SP
      const double wyz = tb.wy[y] * tb.wz[z];
      double acc[MS];
      SFor<0, MS>::run([&](auto mi) { acc[decltype(mi)::value] = 0.0; });
      SFor<0, Q>::run([&](auto qi) {
        constexpr int q = decltype(qi)::value;
        const double wx = tb.wx[j + P * q];
        SFor<0, MS>::run([&](auto mi) {
          constexpr int m = decltype(mi)::value;
          acc[m] += wx * ((double)v[m][q].x * (double)v[m][q].x + (double)v[m][q].y * (double)v[m][q].y);
        });
      });
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        acc[m] = valid ? acc[m] * wyz : 0.0;
      });
      block_sum_store<MS>(acc, red, partials, (long)blk);
    }
  }
  auto store = [&](auto mi) {
    constexpr int m = decltype(mi)::value;
    const CompBuf<C> buf(phi + m * g.npts, g.npts);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      buf.st(v[m][q], voff, (unsigned int)(P * q * sizeof(C)));
    });
  };
  if (flags & XF_FWD) {
    // forward FFT, the kx factor of the next step's KE, store: one component
    // at a time, so the stores overlap the remaining transforms (all at once
    // with MULTI)
    auto kx_store = [&](auto mi) {
      constexpr int m = decltype(mi)::value;
      if constexpr (FULL)
        SFor<0, Q>::run([&](auto qi) {
          constexpr int q = decltype(qi)::value;
          v[m][q] = rot1(v[m][q], lds_kex[j + P * q]);
        });
      store(mi);
    };
    if constexpr (MULTI) {
      line_fft_multi<N, Q, -1, MS, WAVE>(v, my, LS, lds_tw, j);
      SFor<0, MS>::run(kx_store);
    } else {
      SFor<0, MS>::run([&](auto mi) {
        line_fft<N, Q, -1, WAVE>(v[decltype(mi)::value], my, lds_tw, j);
        kx_store(mi);
      });
    }
  } else {
    SFor<0, MS>::run(store);
  }
}
// blk0: index of the launch's first block in the whole-grid numbering, so a
// launch over a z-chunk keeps each block's lines and NORMC partial slot
template <typename T, int SPIN, int N, int Q, bool FULL, bool MULTI>
__global__ __launch_bounds__(256) void k_xpass(C_t<T>* __restrict__ phi, Geom g, Tables<T> tb, Phys ph, int flags,
                                               double* __restrict__ partials, int blk0) {
  xpass_block<T, SPIN, N, Q, FULL, MULTI>(blk0 + (int)blockIdx.x, phi, g, tb, ph, flags, partials);
}

// ------------------------------------------------------- pass B/D (y)
// Lines along y for one component; a block owns TX consecutive x at one z
// (x = x0 + TX blockIdx.x, z = z0 + blockIdx.y, so a launch may cover an x- or
// z-chunk).  dir -1 forward (then the ky and kx factors of KE), +1 inverse
// (then 1/N); scale applies the pending renormalisation.  3D only.
template <typename T, int N, int Q, int TX>
__global__ __launch_bounds__(TX *(N / Q)) void k_ypass(C_t<T>* __restrict__ phi, Geom g, Tables<T> tb, int dir, int scale,
                                                       double inv_npts, int x0, int z0) {
  using C = C_t<T>;
  constexpr int P = N / Q;
  constexpr int LS = line_stride(N);
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R_t<T>* lds_tw = reinterpret_cast<R_t<T>*>(smem_raw);
  C* lds = reinterpret_cast<C*>(lds_tw + TWS);
  const int l = threadIdx.x % TX, j = threadIdx.x / TX;
  const int x = x0 + (int)blockIdx.x * TX + l;
  const int z = z0 + (int)blockIdx.y, m = blockIdx.z;
  for (int i = threadIdx.x; i < TWS; i += blockDim.x) lds_tw[i] = tb.twy[i];
  const CompBuf<C> buf(phi + m * g.npts, g.npts);
  const unsigned int voff = (unsigned int)(((long)z * g.nx * g.ny + x + (long)j * g.nx) * (long)sizeof(C));
  const unsigned int sstep = (unsigned int)(P * g.nx * sizeof(C));
  C v[Q];
  const C sc = scale ? tb.cscale[m] : mk<C>((T)1, (T)0);
  SFor<0, Q>::run([&](auto qi) {
    constexpr int q = decltype(qi)::value;
    const C a = buf.ld(voff, q * sstep);
    v[q] = scale ? cmul(a, sc) : a;
  });
  __syncthreads();
  if (dir < 0) {
    line_fft<N, Q, -1>(v, lds + l * LS, lds_tw, j);
    // the ky and kx factors of KE (kx taken over from pass A: it is constant
    // along the line and commutes with the y transform)
    const R_t<T> fx = tb.kex[x];
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      v[q] = rot1(rot1(v[q], tb.key[j + P * q]), fx);
    });
  } else {
    line_fft<N, Q, +1>(v, lds + l * LS, lds_tw, j);
    // the 1/N of BFT, taken over from pass A (it commutes with the x transform)
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      v[q] = scalec(v[q], inv_npts);
    });
  }
  SFor<0, Q>::run([&](auto qi) {
    constexpr int q = decltype(qi)::value;
    buf.st(v[q], voff, q * sstep);
  });
}

// ------------------------------------------------- 3D pass C (z lines)
// Lines along z for one component m = blockIdx.z; a block owns TX consecutive
// x (x = x0 + TX blockIdx.x) at one y.  Forward z FFT, the kz factor of KE
// with the GAMMAZ term exp(-i CDT (kz^2 - 2 L GAMMAZ kz)) (SRC/cgpe3d.f90:
// 875-899; stored minus one, applied as v + v (w - 1)), inverse z FFT, so the
// line returns to z space: the SO coupling of SRC/cgpe3d.f90:904-1004 depends
// on (kx, ky) only and mixes components, so it commutes with the z
// transforms and follows in pass D, after this KE factor as in the
// reference's order.  scale applies the pending renormalisation.
template <typename T, int N, int Q, int TX>
__global__ __launch_bounds__(TX *(N / Q)) void k_zpass(C_t<T>* __restrict__ phi, Geom g, Tables<T> tb, int scale, int x0) {
  using C = C_t<T>;
  constexpr int P = N / Q;
  constexpr int LS = line_stride(N);
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R_t<T>* lds_tw = reinterpret_cast<R_t<T>*>(smem_raw);
  C* lds = reinterpret_cast<C*>(lds_tw + TWS);
  const int l = threadIdx.x % TX, j = threadIdx.x / TX;
  const int x = x0 + (int)blockIdx.x * TX + l;
  const int y = blockIdx.y, m = blockIdx.z;
  for (int i = threadIdx.x; i < TWS; i += blockDim.x) lds_tw[i] = tb.twz[i];
  const long plane = (long)g.nx * g.ny;
  const CompBuf<C> buf(phi + m * g.npts, g.npts);
  const unsigned int voff = (unsigned int)(((long)y * g.nx + x + (long)j * plane) * (long)sizeof(C));
  const unsigned int sstep = (unsigned int)(P * plane * (long)sizeof(C));
  C v[Q];
  const C sc = scale ? tb.cscale[m] : mk<C>((T)1, (T)0);
  SFor<0, Q>::run([&](auto qi) {
    constexpr int q = decltype(qi)::value;
    const C a = buf.ld(voff, q * sstep);
    v[q] = scale ? cmul(a, sc) : a;
  });
  __syncthreads();
  C* my = lds + l * LS;
  line_fft<N, Q, -1>(v, my, lds_tw, j);
  const R_t<T>* ke = tb.kel + (long)m * N;
  SFor<0, Q>::run([&](auto qi) {
    constexpr int q = decltype(qi)::value;
    v[q] = rot1(v[q], ke[j + P * q]);
  });
  line_fft<N, Q, +1>(v, my, lds_tw, j);
  SFor<0, Q>::run([&](auto qi) {
    constexpr int q = decltype(qi)::value;
    buf.st(v[q], voff, q * sstep);
  });
}

// ------------------------------------------------ 3D pass D (y lines + SOC)
// Lines along y, all components, in (kx, ky, z) space; a block owns TX
// consecutive x at one z (x = x0 + TX bx, z = z0 + by).  The SOC factor of
// (kx, ky) (SRC/cgpe3d.f90:904-1004, from the FP64-built table soc2) is
// applied at every point as its change v + corr, then each component gets the
// inverse y FFT and the 1/N of BFT.  TX * sizeof(C) = 64 B rows use the
// paired block order of kpass_coords.
template <typename T, int SPIN, int N, int Q, int TX> constexpr size_t dpass_lds_bytes() {
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  return (size_t)TWS * sizeof(R_t<T>) + (size_t)TX * line_stride(N) * sizeof(C_t<T>);
}
template <int ROW_BYTES> __device__ __forceinline__ void kpass_coords(int& bx, int& by);
template <typename T, int SPIN, int N, int Q, int TX, bool SCALE>
__global__ __launch_bounds__(TX *(N / Q)) void k_dpass(C_t<T>* __restrict__ phi, Geom g, Tables<T> tb, Phys ph,
                                                       double inv_npts, int bx0, int z0) {
  using C = C_t<T>;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int P = N / Q;
  constexpr int LS = line_stride(N);
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R_t<T>* lds_tw = reinterpret_cast<R_t<T>*>(smem_raw);
  C* lds = reinterpret_cast<C*>(lds_tw + TWS);
  int bx, by;
  kpass_coords<TX * (int)sizeof(C)>(bx, by);
  const int l = threadIdx.x % TX, j = threadIdx.x / TX;
  const int x = (bx0 + bx) * TX + l;
  const int z = z0 + by;
  for (int i = threadIdx.x; i < TWS; i += blockDim.x) lds_tw[i] = tb.twd[i];
  const unsigned int voff = (unsigned int)(((long)z * g.nx * g.ny + x + (long)j * g.nx) * (long)sizeof(C));
  const unsigned int sstep = (unsigned int)((long)P * g.nx * (long)sizeof(C));
  C v[MS][Q];
  SFor<0, MS>::run([&](auto mi) {
    constexpr int m = decltype(mi)::value;
    const CompBuf<C> buf(phi + m * g.npts, g.npts);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      const C a = buf.ld(voff, q * sstep);
      if constexpr (SCALE)
        v[m][q] = cmul(a, tb.cscale[m]);
      else
        v[m][q] = a;
    });
  });
  SFor<0, Q>::run([&](auto qi) {
    constexpr int q = decltype(qi)::value;
    const int ky = j + P * q;
    C e1, dm1[MS], pt[MS], cr[MS];
    soc_from_quarter<SPIN, C>(tb.soc2, x, ky, g.nx, g.ny, ph.imag, e1, dm1);
    SFor<0, MS>::run([&](auto mi) { pt[decltype(mi)::value] = v[decltype(mi)::value][q]; });
    soc_corr<SPIN, C, T>(pt, e1, dm1, cr);
    SFor<0, MS>::run([&](auto mi) {
      constexpr int m = decltype(mi)::value;
      v[m][q] = cadd(v[m][q], cr[m]);
    });
    __builtin_amdgcn_sched_barrier(0);
  });
  __syncthreads();
  C* my = lds + l * LS;
  // each component is stored as soon as its inverse FFT is done
  SFor<0, MS>::run([&](auto mi) {
    constexpr int m = decltype(mi)::value;
    line_fft<N, Q, +1>(v[m], my, lds_tw, j);
    const CompBuf<C> buf(phi + m * g.npts, g.npts);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      buf.st(scalec(v[m][q], inv_npts), voff, q * sstep);
    });
  });
}

// ------------------------------------------------- 2D k pass (y lines)
// Lines along y, all components: forward FFT, the ky factor of KE
// (SRC/cgpe2d.f90:812-829), SOC (SRC/cgpe2d.f90:1554-1650, factors from the
// FP64-built table soc2), inverse FFT.  A block owns TX consecutive x.
// LDS: stage twiddles | KE factors [N] (in 2D the ky^2 factor is the same
// for every component: row 0 of the per-component table) | one padded line
// per column (MS per column with MULTI: the components' lines transformed
// together, as in the q2D x pass).
template <typename T, int SPIN, int N, int Q, int TX, bool MULTI = false> constexpr size_t kpass_lds_bytes() {
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  return (size_t)TWS * sizeof(R_t<T>) + (size_t)N * sizeof(R_t<T>) +
         (size_t)TX * (MULTI ? 2 * SPIN + 1 : 1) * line_stride(N) * sizeof(C_t<T>);
}
template <typename T, int SPIN, int N, int Q, int TX, bool MULTI = false>
__device__ __forceinline__ void kpass_block(int bx, C_t<T>* __restrict__ phi, const Geom& g, const Tables<T>& tb,
                                            const Phys& ph, int scale) {
  using C = C_t<T>;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int P = N / Q;
  constexpr int LS = line_stride(N);
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R_t<T>* lds_tw = reinterpret_cast<R_t<T>*>(smem_raw);
  R_t<T>* lds_ke = lds_tw + TWS;
  C* lds = reinterpret_cast<C*>(lds_ke + N);
  const int l = threadIdx.x % TX, j = threadIdx.x / TX;
  const int x = bx * TX + l;
  for (int i = threadIdx.x; i < TWS; i += blockDim.x) lds_tw[i] = tb.twy[i];
  for (int i = threadIdx.x; i < N; i += blockDim.x) lds_ke[i] = tb.kel[i];
  const unsigned int voff = (unsigned int)((x + (long)j * g.nx) * (long)sizeof(C));
  const unsigned int sstep = (unsigned int)((long)P * g.nx * (long)sizeof(C));
  C v[MS][Q];
  SFor<0, MS>::run([&](auto mi) {
    constexpr int m = decltype(mi)::value;
    const CompBuf<C> buf(phi + m * g.npts, g.npts);
    const C sc = scale ? tb.cscale[m] : mk<C>((T)1, (T)0);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      const C a = buf.ld(voff, q * sstep);
      v[m][q] = scale ? cmul(a, sc) : a;
    });
  });
  __syncthreads();
  C* my = lds + l * (MULTI ? MS : 1) * LS;
  if constexpr (MULTI)
    line_fft_multi<N, Q, -1, MS>(v, my, LS, lds_tw, j);
  else
    SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, -1>(v[decltype(mi)::value], my, lds_tw, j); });
  SFor<0, Q>::run([&](auto qi) {
    constexpr int q = decltype(qi)::value;
    const int pos = j + P * q;
    C e1, dm1[MS];
    soc_from_quarter<SPIN, C>(tb.soc2, x, pos, g.nx, g.ny, ph.imag, e1, dm1);
    // line-axis KE factor, then the SOC change of the result
    C pt[MS], cr[MS];
    const R_t<T> ke = lds_ke[pos];
    SFor<0, MS>::run([&](auto mi) {
      constexpr int m = decltype(mi)::value;
      v[m][q] = rot1(v[m][q], ke);
      pt[m] = v[m][q];
    });
    soc_corr<SPIN, C, T>(pt, e1, dm1, cr);
    SFor<0, MS>::run([&](auto mi) {
      constexpr int m = decltype(mi)::value;
      v[m][q] = cadd(v[m][q], cr[m]);
    });
    __builtin_amdgcn_sched_barrier(0);
  });
  // each component is stored as soon as its inverse FFT is done, so the
  // stores overlap the remaining transforms (all at once with MULTI)
  auto store = [&](auto mi) {
    constexpr int m = decltype(mi)::value;
    const CompBuf<C> buf(phi + m * g.npts, g.npts);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      buf.st(v[m][q], voff, q * sstep);
    });
  };
  if constexpr (MULTI) {
    line_fft_multi<N, Q, +1, MS>(v, my, LS, lds_tw, j);
    SFor<0, MS>::run(store);
  } else {
    SFor<0, MS>::run([&](auto mi) {
      line_fft<N, Q, +1>(v[decltype(mi)::value], my, lds_tw, j);
      store(mi);
    });
  }
}
// Block coordinates of the all-component passes.  When a block's rows are 64 B (half of a
// 128 B line), the two blocks sharing each line are given workgroup ids 8
// apart: workgroups go to the 8 XCDs round-robin (id mod 8), each XCD with its
// own L2, so the pair runs on one XCD at the same time and the line is fetched
// from memory once instead of once per half.
template <int ROW_BYTES> __device__ __forceinline__ void kpass_coords(int& bx, int& by) {
  const int nbx = gridDim.x, total = gridDim.x * gridDim.y;
  if (ROW_BYTES != 64 || (nbx & 1) || (total & 15)) {
    bx = blockIdx.x;
    by = blockIdx.y;
    return;
  }
  const int lin = blockIdx.x + blockIdx.y * nbx;
  const int l = (lin & ~15) + ((lin & 7) << 1) + ((lin >> 3) & 1);
  bx = l % nbx;
  by = l / nbx;
}

template <typename T, int SPIN, int N, int Q, int TX, bool MULTI>
__global__ __launch_bounds__(TX *(N / Q)) void k_kpass(C_t<T>* __restrict__ phi, Geom g, Tables<T> tb, Phys ph, int scale) {
  int bx, by;
  kpass_coords<TX * (int)sizeof(C_t<T>)>(bx, by);
  kpass_block<T, SPIN, N, Q, TX, MULTI>(bx, phi, g, tb, ph, scale);
}

// ------------------------------------------------------------------- 1D
// One workgroup advances the whole q1D problem nsteps times: forward FFT,
// KE, SOC with the constant basis, inverse FFT, SE, SP, NORMC, projection /
// NORMT (SRC/cgpe1d.f90:433-474).  N = NX points, P = N/Q threads.
template <typename T, int SPIN, int N, int Q>
__global__ __launch_bounds__(N / Q) void k_line1d(C_t<T>* __restrict__ phi, Tables<T> tb, Phys ph, FinArgs f,
                                                 long nsteps, double* norm_out, double* sig, int* err) {
  using C = C_t<T>;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int P = N / Q;
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R_t<T>* lds_tw = reinterpret_cast<R_t<T>*>(smem_raw);
  double* red = reinterpret_cast<double*>(lds_tw + TWS);
  double* bc = red + 16 * MS;  // broadcast slots
  C* lds = reinterpret_cast<C*>(bc + 8);
  const int j = threadIdx.x;
  for (int i = threadIdx.x; i < TWS; i += blockDim.x) lds_tw[i] = tb.twx[i];
  C v[MS][Q];
  const unsigned int voff = (unsigned int)(j * sizeof(C));
  SFor<0, MS>::run([&](auto mi) {
    constexpr int m = decltype(mi)::value;
    const CompBuf<C> buf(phi + m * N, N);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      v[m][q] = buf.ld(voff, (unsigned int)(P * q * sizeof(C)));
    });
  });
  __syncthreads();
  const T dt = (T)ph.dt;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  for (long it = 0; it < nsteps; ++it) {
    SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, -1>(v[decltype(mi)::value], lds, lds_tw, j); });
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      const int x = j + P * q;
      C e1, dm1[MS];
      soc_factors<SPIN, C, T>(tb.kx[x], (T)0, ph, dt, e1, dm1);
      C pt[MS], cr[MS];
      const R_t<T> fk = tb.kex[x];
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        v[m][q] = rot1(v[m][q], fk);
        pt[m] = v[m][q];
      });
      soc_corr<SPIN, C, T>(pt, e1, dm1, cr);
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        v[m][q] = cadd(v[m][q], cr[m]);
      });
      __builtin_amdgcn_sched_barrier(0);
    });
    SFor<0, MS>::run([&](auto mi) { line_fft<N, Q, +1>(v[decltype(mi)::value], lds, lds_tw, j); });
    double acc[MS];
    SFor<0, MS>::run([&](auto mi) { acc[decltype(mi)::value] = 0.0; });
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      const int x = j + P * q;
      C pt[MS];
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        pt[m] = scalec(v[m][q], ph.inv_npts);
      });
      if constexpr (SPIN == 1)
        se1_point<C, T>(pt, ph, dt, sizeof(T) == 4 ? (T)0.05 : (T)1e-3);
      else
        se2_point<C, T>(pt, ph, dt, (T)0.5, sizeof(T) == 4 ? (T)1e-8 : (T)1e-17);
      sp_point<SPIN, C, T>(pt, tb.vx[x], ph, dt);
      const double wx = tb.wx[x];
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        v[m][q] = pt[m];
        acc[m] += wx * ((double)pt[m].x * (double)pt[m].x + (double)pt[m].y * (double)pt[m].y);
      });
      // one point at a time: keeps the SE temporaries of a single point live
      __builtin_amdgcn_sched_barrier(0);
    });
    wave_sum<MS>(acc);
    __syncthreads();
    if (lane == 0)
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        red[wid * MS + m] = acc[m];
      });
    __syncthreads();
    if (threadIdx.x == 0) {
      double n[5] = {0, 0, 0, 0, 0}, c[5];
      for (int w = 0; w < nw; ++w)
        for (int m = 0; m < MS; ++m) n[m] += red[w * MS + m];
      renorm_scale(f, n, sig, err, c);
      for (int m = 0; m < MS; ++m) {
        bc[m] = c[m];
        norm_out[m] = n[m];
      }
    }
    __syncthreads();
    if (f.imag) {
      SFor<0, MS>::run([&](auto mi) {
        constexpr int m = decltype(mi)::value;
        const T c = (T)bc[m];
        SFor<0, Q>::run([&](auto qi) { v[m][decltype(qi)::value] = cscale(v[m][decltype(qi)::value], c); });
      });
    }
  }
  SFor<0, MS>::run([&](auto mi) {
    constexpr int m = decltype(mi)::value;
    const CompBuf<C> buf(phi + m * N, N);
    SFor<0, Q>::run([&](auto qi) {
      constexpr int q = decltype(qi)::value;
      buf.st(v[m][q], voff, (unsigned int)(P * q * sizeof(C)));
    });
  });
}

}  // namespace fgpe

// fgpe_impl.hpp -- shared host-side declarations of libfgpe: the handle,
// error reporting, per-precision launch configuration and the templated
// entry points implemented in the per-precision/per-spin translation units
// (fgpe_ops_*.hip) and used by the C ABI (fgpe.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fgpe.h"
#include "fgpe_point.hpp"

namespace fgpe_impl {
using namespace fgpe;

int fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return ::fgpe_impl::fail(FGPE_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double EPS = 1.0e-40;

// Points per thread (Q) of each line pass and the x-width (TX) of the
// strided-line tiles.  Q sets the register footprint: the all-component
// passes hold 2*SPIN+1 complex values per point in registers.
template <typename T> struct Cfg;
template <> struct Cfg<float> {
  // pass A (x lines, all components): 4 points per thread keep it at <= 128
  // VGPRs (4 waves per SIMD); its SE/SP work is latency-bound, not VALU-bound
  static constexpr int QX = 4;
  static constexpr int QY = 16;  // pass B (y lines, one component)
  static constexpr int QZ = 16;  // 3D pass C (z lines, one component)
  static constexpr int QK = 8;   // 2D k pass and 3D pass D (all components)
  static constexpr int Q1 = 16;  // q1D whole-step kernel
  static constexpr int TX = 16;  // x columns per strided-line block (128 B rows)
  // x columns per pass-D block (3D y lines, all components): 16 columns,
  // whole 128 B rows, 512 threads (349 -> 321 us at 256^3 spin-2 against 8
  // columns; 4 points per thread was slower)
  static constexpr int TXK = 16;
  // q2D: a grid has only NY x lines, so pass A runs one line per block to
  // spread over the CUs, at 4 points per thread (radix-4 stages: half the
  // roundings of radix 2); the k pass runs 8 points per thread in tiles of 8
  // columns.  4 points in tiles of 4 ran the 256^2 spin-2 step in 34.4 us
  // instead of 39.4, but its fourth radix stage's roundings took the FP32
  // density error of a q2D spin-2 imaginary-time case over its 1e-5 bar
  // (1.05e-5)
  static constexpr int QX2 = 4;
  static constexpr int QK2 = 8;
  static constexpr int TXK2 = 8;
};
template <> struct Cfg<double> {
  // pass A at 4 points per thread: at 8 the five components' lines held
  // 256 VGPRs + 161 AGPRs (one wave per SIMD); 256^3 spin-2 1745 -> 924 us
  static constexpr int QX = 4;
  static constexpr int QY = 8;
  static constexpr int QZ = 8;
  // pass D at 8 points per thread in blocks of at most 512 threads (256
  // VGPRs); 4 points per thread made 1024-thread blocks capped at 128 VGPRs
  // that spilled (256^3 spin-2: 743 -> 630 us)
  static constexpr int QK = 8;
  static constexpr int Q1 = 8;
  static constexpr int TX = 8;
  // pass D in 16-column blocks (two 128 B rows; 8 for lines of 384 and 512
  // points, where 16 columns would exceed 512 threads).  8-column blocks of
  // 96-point lines (128 threads, each line spread over both waves) gave
  // results that changed from run to run on grids of 32x96x192 and larger;
  // 16 columns are bit-reproducible on every grid tried
  static constexpr int TXK = 16;
  // q2D: pass A at 2 points per thread, one line per block; the k pass at 4
  // points per thread in tiles of 4 columns (256^2 spin-2: 13.0k -> 21.3k
  // steps/s with tiles of 2, then 38.7 -> 36.9 us per step with tiles of 4
  // once the components' lines were transformed together; 2 points per
  // thread in the k pass was slower)
  static constexpr int QX2 = 2;
  static constexpr int QK2 = 4;
  static constexpr int TXK2 = 4;
};

// Line lengths with compiled passes: 2D/3D axes, and the q1D line.  A
// length is 2^a or 3 * 2^a; its FFT is planned from radix 16/12/8/6/4/3/2
// stages (pick_radix).
#define FGPE_LINE_SIZES(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#define FGPE_1D_SIZES(X) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072) X(4096)

// Points per thread of a pass over lines of n points: the pass's default
// (Cfg) for powers of two; for n = 3 * 2^a a multiple of 6, so the radix-3
// factor has a stage (12 where the default is 16, else 6).
__host__ __device__ constexpr int q_for(int n, int qdef) { return (n & (n - 1)) == 0 ? qdef : (qdef >= 12 ? 12 : 6); }

// Points per thread of pass A: QX in 3D, QX2 in q2D.
template <typename T> constexpr int x_pass_q(int dim, int nx) {
  return q_for(nx, dim == 2 ? Cfg<T>::QX2 : Cfg<T>::QX);
}
// Points per thread of 3D passes B (y lines) and C (z lines): the default,
// but at most 8 for lines of 512 points and more, whose 16-column blocks need
// 70 KB of LDS (two blocks per CU): 8 points per thread gives each block twice
// the waves in the same LDS (512^3 FP32 pass C: 4.20 -> 2.89 ms).
template <typename T> constexpr int long_line_q(int n, int qdef) { return q_for(n, (n >= 512 && qdef > 8) ? 8 : qdef); }
template <typename T> constexpr int y_pass_q(int ny) { return long_line_q<T>(ny, Cfg<T>::QY); }
template <typename T> constexpr int z_pass_q(int nz) { return long_line_q<T>(nz, Cfg<T>::QZ); }
// Threads of the one-block NORMC finalisation over nparts partials: about 16
// per thread, in whole waves, at most 1024 (a q2D grid's few hundred partials
// then take one wave, which launches and retires sooner than sixteen).
inline int fin_threads(long nparts) {
  const long t = ((nparts + 16 * 64 - 1) / (16 * 64)) * 64;
  return (int)std::min<long>(1024, std::max<long>(64, t));
}
// Blocks of one pass-A launch: P = NX/Q threads per x line, lines packed into
// blocks of at least 256 threads (3D) or of one wave (q2D, where the grid has
// few lines).  Each block writes MS NORMC partials.
template <typename T> inline int x_pass_lpw(int dim, int nx) {
  const int p = nx / x_pass_q<T>(dim, nx);
  if (dim == 2) return p >= 64 ? 1 : 64 / p;
  return (p >= 256 ? p : 256) / p;
}
template <typename T> inline long x_pass_blocks(int dim, int nx, long nlines) {
  const int lpw = x_pass_lpw<T>(dim, nx);
  return (nlines + lpw - 1) / lpw;
}

// ---------------------------------------------------- host-side tables
// wave numbers in FFT order with positive Nyquist (SRC/cgpe3d.f90:702-726)
inline std::vector<double> kgrid(int n, double d) {
  std::vector<double> k(n, 0.0);
  if (n == 1) return k;
  double l = n * d;
  for (int i = 1; i <= 1 + n / 2; ++i) k[i - 1] = (i - 1.0) * 2.0 * PI / l;
  for (int i = 1; i <= n / 2 - 1; ++i) k[i + n / 2] = -k[n / 2 - i];
  return k;
}
// x_i = (-1 + 2(i-1)/N) L/2 (SRC/cgpe3d.f90:729-744)
inline std::vector<double> xgrid(int n, double d) {
  std::vector<double> x(n, 0.0);
  if (n == 1) return x;
  double l = n * d;
  for (int i = 1; i <= n; ++i) x[i - 1] = (-1.0 + 2.0 * (i - 1.0) / n) * l / 2.0;
  return x;
}
// Simpson weights of SRC/simpson.f90:4-22 as a vector: sum_i w_i f_i equals
// SIMPSON(f, d) for every f.
inline std::vector<double> simpson_weights(int n, double d) {
  std::vector<double> w(n, 0.0);
  if (n == 1) {
    w[0] = 1.0;
    return w;
  }
  // 1-based: w(1)=d/3; even 2..N-4: 4d/3; odd 3..N-5: 2d/3; w(N-3)=d/3+3d/8;
  // w(N-2)=w(N-1)=9d/8; w(N)=3d/8
  w[0] = d / 3.0;
  for (int i = 2; i <= n - 4; i += 2) w[i - 1] = 4.0 * d / 3.0;
  for (int i = 3; i <= n - 5; i += 2) w[i - 1] = 2.0 * d / 3.0;
  w[n - 4] = d / 3.0 + 3.0 * d / 8.0;
  w[n - 3] = 9.0 * d / 8.0;
  w[n - 2] = 9.0 * d / 8.0;
  w[n - 1] = 3.0 * d / 8.0;
  return w;
}

// A unit-modulus constant re + i im in the kernels' rotation format: FP32 as
// hi + lo parts (their sum is the FP64 value), FP64 as is.
template <typename T> R_t<T> rot_from(double re, double im);
template <> inline R_t<float> rot_from<float>(double re, double im) {
  R_t<float> w;
  w.x = (float)re;
  w.y = (float)im;
  w.z = (float)(re - (double)w.x);
  w.w = (float)(im - (double)w.y);
  return w;
}
template <> inline R_t<double> rot_from<double>(double re, double im) {
  R_t<double> w;
  w.x = re;
  w.y = im;
  return w;
}

template <typename T> std::vector<R_t<T>> stage_twiddles(int n, int q) {
  std::vector<R_t<T>> t;
  int ns = 1;
  while (ns < n) {
    int r = pick_radix(n / ns, q);
    if (ns > 1)
      for (int rr = 0; rr < r; ++rr)
        for (int k = 0; k < ns; ++k) {
          double a = -2.0 * PI * (double)(k * rr) / (double)(ns * r);
          t.push_back(rot_from<T>(std::cos(a), std::sin(a)));
        }
    ns *= r;
  }
  if (t.empty()) t.push_back(rot_from<T>(1.0, 0.0));
  return t;
}

// exp(-i CDT a) - 1 for real a, formed without cancellation: imaginary time
// expm1(-DT a), real time -2 sin^2(DT a/2) - i sin(DT a)
template <typename T> R_t<T> ke_factor(double a, double dt, int imag) {
  if (imag) return rot_from<T>(std::expm1(-dt * a), 0.0);
  const double sh = std::sin(0.5 * dt * a);
  return rot_from<T>(-2.0 * sh * sh, -std::sin(dt * a));
}

// 2D SOC factors for soc_from_quarter, in FP64 then rounded: ALPHA, BETA and
// the factors of SRC/cgpe2d.f90 SOC (the 2D form of SRC/cgpe3d.f90:904-1004):
// e1 = exp(-i BETA) = (GX kx - i GY ky)/|.| (1 when ALPHA <= 1e-30),
// th = DT L ALPHA with ALPHA = 2 |(GX kx, GY ky)|.  Only the quarter kx, ky >= 0
// is stored ([NY/2 + 1][NX/2 + 1][1 + SPIN]; indices 0..N/2 of the FFT-order
// wave numbers, N/2 the positive Nyquist): ALPHA depends on |kx|, |ky| alone,
// and negating kx or ky negates the real or imaginary part of e1, which the
// reader applies (exactly: rounding is symmetric in sign).
template <typename T>
std::vector<fgpe::C_t<T>> soc2_table(const std::vector<double>& kx, const std::vector<double>& ky, double gx, double gy,
                             double dt, int imag, int spin) {
  const size_t nx = kx.size() / 2 + 1, ny = ky.size() / 2 + 1;
  std::vector<fgpe::C_t<T>> t(nx * ny * (spin + 1));
  auto put = [&](size_t i, double a, double b) {
    t[i].x = (T)a;
    t[i].y = (T)b;
  };
  for (size_t y = 0; y < ny; ++y)
    for (size_t x = 0; x < nx; ++x) {
      const size_t o = (y * nx + x) * (spin + 1);
      const double ax = gx * kx[x], ay = gy * ky[y], r = std::sqrt(ax * ax + ay * ay), alpha = 2.0 * r;
      if (alpha <= 1e-30)
        put(o, 1.0, 0.0);
      else
        put(o, ax / r, -ay / r);
      for (int L = 1; L <= spin; ++L) {
        const double th = dt * L * alpha;
        if (imag) {
          put(o + L, std::expm1(-th), std::expm1(th));
        } else {
          const double sh = std::sin(0.5 * th), ch = std::cos(0.5 * th);
          put(o + L, -2.0 * sh * sh, 2.0 * sh * ch);
        }
      }
    }
  return t;
}

}  // namespace fgpe_impl

struct fgpe_handle {
  fgpe_params p;
  int dim, spin, ms, nx, ny, nz;
  long npts;
  bool fp32;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // 3D schedule (fgpe_set_schedule): z-planes per chunk of passes A and B,
  // x-columns per chunk of passes C and D (0: the whole grid in one launch per
  // pass) and streams (1-3); the extra streams and the fork/join events are
  // created with the handle
  int zchunk = 0, xchunk = 0, nstreams = 1;
  hipStream_t aux[2] = {nullptr, nullptr};
  hipEvent_t evs[4] = {nullptr, nullptr, nullptr, nullptr};  // fork/join of the 3D schedule; [3]: NORMC finalised
  int dev = 0;
  size_t lds_max = 65536;
  // state
  void* phi = nullptr;      // [ms][npts] complex T
  void* phi_old = nullptr;  // kept copy for CONV
  void* stage64 = nullptr;  // FP64 staging for upload/download in FP32 mode
  // tables (T precision unless noted; twiddles and KE factors rot_of<T>)
  void *twx = nullptr, *twy = nullptr, *twz = nullptr;
  void* twd = nullptr;  // 3D pass D: y axis at QK points per thread
  void *kex = nullptr, *key = nullptr, *kel = nullptr;  // KE factors (rot_of<T>)
  void* soc2 = nullptr;  // 2D SOC factor table (soc2_table)
  void *kx = nullptr, *ky = nullptr;
  void *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *wx = nullptr, *wy = nullptr, *wz = nullptr;
  double *x2 = nullptr, *y2 = nullptr, *z2 = nullptr;
  double *vxd = nullptr, *vyd = nullptr, *vzd = nullptr;
  void* cscale = nullptr;        // [ms] complex T
  double* cscale64 = nullptr;    // [ms]
  double* dnorm = nullptr;       // [8]: norms, TNORM
  double* obs_sums = nullptr;    // [32]: the observables' column integrals
  double* dsig = nullptr;        // Newton state
  int* derr = nullptr;
  double* partials = nullptr;    // reduction scratch
  size_t partials_n = 0;
  bool noalloc = false;          // partials is the no-allocation scratch: never grow it
  double last_ms = 0;
  bool timed = false;
  // per-kernel timing (fgpe_pass_times): events recorded around launches
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> ev_kind;
  std::vector<void*> allocs;
};


namespace fgpe_impl {
template <typename T> using C = fgpe::C_t<T>;

int dalloc(fgpe_handle* h, void** p, size_t bytes);
int ensure_partials(fgpe_handle* h, size_t n);
int grid_for(long n, int threads);
fgpe::Geom geom(fgpe_handle* h);
fgpe::Phys phys(fgpe_handle* h);
fgpe::FinArgs finargs(fgpe_handle* h);
template <typename T> inline fgpe::Tables<T> tables(fgpe_handle* h) {
  fgpe::Tables<T> t;
  t.twx = (const fgpe::R_t<T>*)h->twx;
  t.twy = (const fgpe::R_t<T>*)h->twy;
  t.twz = (const fgpe::R_t<T>*)h->twz;
  t.twd = (const fgpe::R_t<T>*)h->twd;
  t.kex = (const fgpe::R_t<T>*)h->kex;
  t.key = (const fgpe::R_t<T>*)h->key;
  t.kel = (const fgpe::R_t<T>*)h->kel;
  t.kx = (const T*)h->kx;
  t.ky = (const T*)h->ky;
  t.vx = (const T*)h->vx;
  t.vy = (const T*)h->vy;
  t.vz = (const T*)h->vz;
  t.wx = h->wx;
  t.wy = h->wy;
  t.wz = h->wz;
  t.cscale = (const fgpe::C_t<T>*)h->cscale;
  t.soc2 = (const fgpe::C_t<T>*)h->soc2;
  return t;
}



// z-planes per chunk of the 3D schedule's first phase: the handle's choice
// when it divides NZ and its lines fill whole pass-A blocks, else the whole grid
template <typename T> inline int chunk_planes(const fgpe_handle* h) {
  const int zc = h->zchunk;
  if (zc <= 0 || zc >= h->nz || h->nz % zc) return h->nz;
  if (((long)zc * h->ny) % x_pass_lpw<T>(h->dim, h->nx)) return h->nz;
  return zc;
}
// x-columns per chunk of the second phase: the handle's choice when it
// divides NX into whole y-pass blocks and pairs of k-pass blocks (the pairs
// that share a 128 B row), else the whole grid
template <typename T> inline int chunk_cols(const fgpe_handle* h) {
  const int xc = h->xchunk;
  const int q = std::max(Cfg<T>::TX, 2 * Cfg<T>::TXK);
  if (xc <= 0 || xc >= h->nx || h->nx % xc || xc % q) return h->nx;
  return xc;
}

// Tiles of the observables line pass along `axis` (lines per tile from the
// axis' line length), and the blocks that loop over them.
template <typename T, int SPIN> long obs_axis_tiles(const fgpe_handle* h, int axis) {
  const int n = axis == 0 ? h->nx : axis == 1 ? h->ny : h->nz;
  int L = 1;
  switch (n) {
#define FGPE_CASE(s) \
  case s: L = obs_lines<T, SPIN, s>(); break;
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  const long nbx = (h->nx + L - 1) / L;
  return axis == 0 ? ((long)h->ny * h->nz + L - 1) / L : axis == 1 ? nbx * h->nz : nbx * h->ny;
}
template <typename T, int SPIN> long obs_axis_blocks(const fgpe_handle* h, int axis) {
  return std::min(obs_axis_tiles<T, SPIN>(h, axis), OBS_MAX_BLOCKS);
}

// begin/end markers of one timed kernel launch (no-ops unless profiling)
int prof_begin(fgpe_handle* h);
void prof_end(fgpe_handle* h, int kind);

// implemented per (T, SPIN) in fgpe_ops_*.hip
template <typename T, int SPIN> int run_steps(fgpe_handle* h, long nsteps);
template <typename T, int SPIN> int observables_t(fgpe_handle* h, double* en, double* mu, double* mz, double* rms);
// implemented per T in fgpe_ops_*_s1.hip
template <typename T> int normc_t(fgpe_handle* h, double* norm);
template <typename T> int normalize_t(fgpe_handle* h, double* tnorm);
template <typename T> int conv_t(fgpe_handle* h, double* out);
template <typename T> int guess_t(fgpe_handle* h);
template <typename T> int from64_t(fgpe_handle* h, long n);
template <typename T> int to64_t(fgpe_handle* h, long n);
}  // namespace fgpe_impl

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
  static constexpr int QX = 8;  // pass A (x lines, all components)
  static constexpr int QY = 16;  // passes B/D (y lines, one component)
  static constexpr int QK = 8;   // pass C (k-space lines, all components)
  static constexpr int Q1 = 16;  // q1D whole-step kernel
  static constexpr int TX = 16;  // x columns per strided-line block (128 B rows)
};
template <> struct Cfg<double> {
  static constexpr int QX = 8;
  static constexpr int QY = 8;
  static constexpr int QK = 4;
  static constexpr int Q1 = 8;
  static constexpr int TX = 8;
};

}  // namespace fgpe_impl

struct fgpe_handle {
  fgpe_params p;
  int dim, spin, ms, nx, ny, nz;
  long npts;
  bool fp32;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int dev = 0;
  size_t lds_max = 65536;
  // state
  void* phi = nullptr;      // [ms][npts] complex T
  void* phi_old = nullptr;  // kept copy for CONV
  void* stage64 = nullptr;  // FP64 staging for upload/download in FP32 mode
  // tables (T precision unless noted; twiddles and KE factors rot_of<T>)
  void *twx = nullptr, *twy = nullptr, *twz = nullptr;
  void *kex = nullptr, *key = nullptr, *kel = nullptr;  // KE factors (rot_of<T>)
  void *kx = nullptr, *ky = nullptr;
  void *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *wx = nullptr, *wy = nullptr, *wz = nullptr;
  double *x2 = nullptr, *y2 = nullptr, *z2 = nullptr;
  double *vxd = nullptr, *vyd = nullptr, *vzd = nullptr;
  void* cscale = nullptr;        // [ms] complex T
  double* cscale64 = nullptr;    // [ms]
  double* dnorm = nullptr;       // [8]: norms, TNORM
  double* dsig = nullptr;        // Newton state
  int* derr = nullptr;
  double* partials = nullptr;    // reduction scratch
  size_t partials_n = 0;
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
template <typename T> using R_t = fgpe::R_t<T>;

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
  return t;
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

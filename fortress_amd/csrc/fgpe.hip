// fgpe.hip -- libfgpe: C ABI (include/fgpe.h) over the gfx950 propagator.
//
// Host side of the boundary: device allocation, the tables that the
// reference's INITIALIZE/KE/SOC/SP recompute every step (SRC/cgpe3d.f90:666-754,
// 875-1065) and the entry points.  The kernels and their launch code live in
// fgpe_ops_{f32,f64}_{s1,s2}.hip.  No CPU fallback: every compute entry point
// launches HIP kernels and fails with an error when no device is usable.
#include "fgpe_impl.hpp"

#include <mutex>

// Scratch of fgpe_run_noalloc: the NORMC block partials of pass A, held in a
// fixed device array instead of memory the library allocates.  The largest
// supported grid needs 655,360 doubles (512^3 spin-2 FP32: 131,072 x-pass
// blocks x 5 components), so 2^20 doubles (8 MiB per device) covers every
// grid fgpe_create accepts.  Each device holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (scratch_ptr below), never through the
// array's name.
static constexpr size_t NOALLOC_SCRATCH_DOUBLES = size_t(1) << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace fgpe_impl {

thread_local std::string g_err;

// per-device address of `scratch`, looked up once per device by fgpe_create
// with that device current
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[MAX_DEVICES];
static bool g_scratch_done[MAX_DEVICES];
static std::string g_scratch_err[MAX_DEVICES];

static void scratch_setup(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[dev]) return;
  g_scratch_done[dev] = true;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_err[dev] = hipGetErrorString(e);
    return;
  }
  g_scratch_ptr[dev] = (double*)p;
}

static double* scratch_ptr(int dev, std::string* why) {
  if (dev < 0 || dev >= MAX_DEVICES) {
    *why = "device index out of range";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) *why = g_scratch_done[dev] ? g_scratch_err[dev] : "not set up on this device";
  return g_scratch_ptr[dev];
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int dalloc(fgpe_handle* h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(FGPE_EDEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  h->allocs.push_back(*p);
  return 0;
}

template <typename V> int upload_vec(fgpe_handle* h, void** dst, const std::vector<V>& src) {
  int rc = dalloc(h, dst, src.size() * sizeof(V));
  if (rc) return rc;
  HIPCHK(hipMemcpy(*dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice));
  return 0;
}

template <typename T> int build_tables(fgpe_handle* h) {
  const fgpe_params& p = h->p;
  // each axis' twiddle tables follow the points-per-thread (radix plan) of
  // the pass that transforms it: x by pass A (or the q1D kernel), y by pass B
  // in 3D and by the k-space pass in 2D, y again by the all-component pass D
  // in 3D (its own table, QK points per thread), z by pass C
  const int qx = h->dim == 1 ? q_for(h->nx, Cfg<T>::Q1) : x_pass_q<T>(h->dim, h->nx);
  const int qy = h->dim == 3 ? y_pass_q<T>(h->ny) : q_for(h->ny, Cfg<T>::QK2);
  const int qz = z_pass_q<T>(h->nz);
  auto kx = kgrid(h->nx, p.d[0]), ky = kgrid(h->ny, p.d[1]), kz = kgrid(h->nz, p.d[2]);
  auto xx = xgrid(h->nx, p.d[0]), yy = xgrid(h->ny, p.d[1]), zz = xgrid(h->nz, p.d[2]);
  int rc;
  if ((rc = upload_vec(h, &h->twx, stage_twiddles<T>(h->nx, qx)))) return rc;
  if ((rc = upload_vec(h, &h->twy, stage_twiddles<T>(h->ny > 1 ? h->ny : qy, qy)))) return rc;
  if ((rc = upload_vec(h, &h->twz, stage_twiddles<T>(h->nz > 1 ? h->nz : qz, qz)))) return rc;
  if (h->dim == 3 && (rc = upload_vec(h, &h->twd, stage_twiddles<T>(h->ny, q_for(h->ny, Cfg<T>::QK))))) return rc;
  const double dt = p.dt;
  const int imag = p.imag;
  // KE factors minus one, per axis (SRC/cgpe3d.f90:875-899; 2D
  // SRC/cgpe2d.f90:812-829; 1D SRC/cgpe1d.f90:826-842): kx in pass A, ky in
  // pass B (3D), line axis per component in pass C (3D kz^2 - 2 L GAMMAZ kz,
  // 2D ky^2)
  std::vector<R_t<T>> kex(h->nx), key(h->dim == 3 ? h->ny : 1, rot_from<T>(0.0, 0.0));
  for (int i = 0; i < h->nx; ++i) kex[i] = ke_factor<T>(kx[i] * kx[i], dt, imag);
  if (h->dim == 3)
    for (int j = 0; j < h->ny; ++j) key[j] = ke_factor<T>(ky[j] * ky[j], dt, imag);
  int nl = h->dim == 3 ? h->nz : (h->dim == 2 ? h->ny : 1);
  std::vector<R_t<T>> kel((size_t)h->ms * nl);
  for (int m = 0; m < h->ms; ++m) {
    const int L = m - h->spin;
    for (int t = 0; t < nl; ++t) {
      double a;
      if (h->dim == 3)  // kz^2 - 2 L GAMMAZ kz (SRC/cgpe3d.f90:892-893)
        a = kz[t] * kz[t] - 2.0 * L * p.gamma[2] * kz[t];
      else if (h->dim == 2)
        a = ky[t] * ky[t];
      else
        a = 0.0;
      kel[(size_t)m * nl + t] = ke_factor<T>(a, dt, imag);
    }
  }
  if ((rc = upload_vec(h, &h->kex, kex))) return rc;
  if ((rc = upload_vec(h, &h->key, key))) return rc;
  if ((rc = upload_vec(h, &h->kel, kel))) return rc;
  // SOC factors per (ky, kx): the 2D k pass and the 3D pass D read them
  if (h->dim >= 2 &&
      (rc = upload_vec(h, &h->soc2, soc2_table<T>(kx, ky, p.gamma[0], p.gamma[1], dt, imag, h->spin))))
    return rc;
  std::vector<T> kxt(kx.begin(), kx.end()), kyt(ky.begin(), ky.end());
  if ((rc = upload_vec(h, &h->kx, kxt))) return rc;
  if ((rc = upload_vec(h, &h->ky, kyt))) return rc;
  // V = ALPHAX^2 x^2 + ALPHAY^2 y^2 + ALPHAZ^2 z^2 (SRC/cgpe3d.f90:752)
  std::vector<double> vxd(h->nx), vyd(h->ny, 0.0), vzd(h->nz, 0.0), x2(h->nx), y2(h->ny, 0.0), z2(h->nz, 0.0);
  for (int i = 0; i < h->nx; ++i) {
    x2[i] = xx[i] * xx[i];
    vxd[i] = p.alpha[0] * p.alpha[0] * x2[i];
  }
  if (h->dim >= 2)
    for (int j = 0; j < h->ny; ++j) {
      y2[j] = yy[j] * yy[j];
      vyd[j] = p.alpha[1] * p.alpha[1] * y2[j];
    }
  if (h->dim >= 3)
    for (int k = 0; k < h->nz; ++k) {
      z2[k] = zz[k] * zz[k];
      vzd[k] = p.alpha[2] * p.alpha[2] * z2[k];
    }
  std::vector<T> vx(vxd.begin(), vxd.end()), vy(vyd.begin(), vyd.end()), vz(vzd.begin(), vzd.end());
  if ((rc = upload_vec(h, &h->vx, vx))) return rc;
  if ((rc = upload_vec(h, &h->vy, vy))) return rc;
  if ((rc = upload_vec(h, &h->vz, vz))) return rc;
  if ((rc = upload_vec(h, (void**)&h->vxd, vxd))) return rc;
  if ((rc = upload_vec(h, (void**)&h->vyd, vyd))) return rc;
  if ((rc = upload_vec(h, (void**)&h->vzd, vzd))) return rc;
  if ((rc = upload_vec(h, (void**)&h->x2, x2))) return rc;
  if ((rc = upload_vec(h, (void**)&h->y2, y2))) return rc;
  if ((rc = upload_vec(h, (void**)&h->z2, z2))) return rc;
  if ((rc = upload_vec(h, (void**)&h->wx, simpson_weights(h->nx, p.d[0])))) return rc;
  if ((rc = upload_vec(h, (void**)&h->wy, simpson_weights(h->ny, p.d[1])))) return rc;
  if ((rc = upload_vec(h, (void**)&h->wz, simpson_weights(h->nz, p.d[2])))) return rc;
  return 0;
}

Phys phys(fgpe_handle* h) {
  Phys ph;
  ph.tau0 = h->p.tau[0];
  ph.tau1 = h->p.tau[1];
  ph.tau2 = h->p.tau[2];
  ph.dt = h->p.dt;
  ph.gx = h->p.gamma[0];
  ph.gy = h->p.gamma[1];
  ph.gz = h->p.gamma[2];
  ph.inv_npts = 1.0 / (double)h->npts;
  ph.imag = h->p.imag;
  ph.dim = h->dim;
  return ph;
}

FinArgs finargs(fgpe_handle* h) {
  FinArgs f;
  f.imag = h->p.imag;
  f.soc = h->p.soc;
  f.dim = h->dim;
  f.ms = h->ms;
  f.mag = h->p.mag;
  f.dt = h->p.dt;
  return f;
}

Geom geom(fgpe_handle* h) { return Geom{h->nx, h->ny, h->nz, h->npts}; }

int ensure_partials(fgpe_handle* h, size_t n) {
  if (h->partials_n >= n) return 0;
  if (h->noalloc)
    return fail(FGPE_EINVAL, "the no-allocation run needs %zu doubles of scratch; its limit is %zu", n,
                h->partials_n);
  void* p;
  int rc = dalloc(h, &p, n * sizeof(double));
  if (rc) return rc;
  h->partials = (double*)p;
  h->partials_n = n;
  return 0;
}


int prof_begin(fgpe_handle* h) {
  if (!h->profiling) return 0;
  if (h->ev_used + 2 > h->ev_pool.size()) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      h->profiling = false;
      return fail(FGPE_EDEVICE, "hipEventCreate failed while profiling");
    }
    h->ev_pool.push_back(e0);
    h->ev_pool.push_back(e1);
  }
  (void)hipEventRecord(h->ev_pool[h->ev_used], h->stream);
  return 0;
}

void prof_end(fgpe_handle* h, int kind) {
  if (!h->profiling) return;
  (void)hipEventRecord(h->ev_pool[h->ev_used + 1], h->stream);
  h->ev_used += 2;
  h->ev_kind.push_back(kind);
}

int grid_for(long n, int threads) {
  long b = (n + threads - 1) / threads;
  return (int)std::min<long>(std::max<long>(b, 1), 8192);
}

static bool supported_line(int n, int dim1) {
#define FGPE_CASE(s) || n == s
  if (dim1) return false FGPE_1D_SIZES(FGPE_CASE);
  return false FGPE_LINE_SIZES(FGPE_CASE);
#undef FGPE_CASE
}

static size_t elem_size(const fgpe_handle* h) { return h->fp32 ? sizeof(float2) : sizeof(double2); }

template <typename T> static int run_t(fgpe_handle* h, long nsteps) {
  return h->spin == 1 ? run_steps<T, 1>(h, nsteps) : run_steps<T, 2>(h, nsteps);
}
template <typename T> static int obs_t(fgpe_handle* h, double* en, double* mu, double* mz, double* rms) {
  return h->spin == 1 ? observables_t<T, 1>(h, en, mu, mz, rms) : observables_t<T, 2>(h, en, mu, mz, rms);
}

}  // namespace fgpe_impl

using namespace fgpe_impl;

// ============================================================== C ABI
extern "C" {

int fgpe_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return FGPE_EINVAL;
  snprintf(buf, len, "%s", g_err.c_str());
  return 0;
}

int fgpe_derive(const fgpe_input* in, int32_t precision, fgpe_params* o, double* aosc_out) {
  if (!in || !o) return fail(FGPE_EINVAL, "null argument");
  if (in->dim < 1 || in->dim > 3 || (in->spin != 1 && in->spin != 2))
    return fail(FGPE_EINVAL, "dim must be 1..3 and spin 1 or 2");
  if (in->nu[0] <= 0.0 || in->mass <= 0.0) return fail(FGPE_EINVAL, "NUX and mass must be positive");
  const double AMU = 1.66e-27, HBAR = 1.054560653e-34, AU = 0.529177208e-10;
  memset(o, 0, sizeof(*o));
  o->dim = in->dim;
  o->spin = in->spin;
  o->precision = precision;
  o->imag = in->switch_im == 1 ? 1 : 0;
  o->mag = in->mag;
  int soc = 0;
  for (int d = 0; d < 3; ++d) {
    o->n[d] = d < in->dim ? in->n[d] : 1;
    o->d[d] = d < in->dim ? in->d[d] : 1.0;
    o->gamma[d] = d < in->dim ? in->gamma[d] : 0.0;
    if (d < in->dim && std::fabs(in->gamma[d]) > EPS) soc = 1;  // SRC/cgpe3d.f90:256
  }
  o->dt = in->dt;
  o->soc = soc;
  o->option_fpc = soc ? 4 : in->option_fpc;  // read only without SOC (SRC/cgpe3d.f90:257-261)
  const double nux = in->nu[0];
  for (int d = 0; d < 3; ++d) o->alpha[d] = in->nu[d] / nux;  // SRC/cgpe3d.f90:277-279
  const double aosc = std::sqrt(HBAR / (in->mass * AMU * (2.0 * PI * nux)));  // 334-335
  if (aosc_out) *aosc_out = aosc;
  const double a0 = in->a[0], a2 = in->a[1], a4 = in->a[2];
  // 3D SRC/cgpe3d.f90:337-345; 2D x sqrt(ALPHAZ/(2pi)) SRC/cgpe2d.f90:328-336;
  // 1D x sqrt(ALPHAY*ALPHAZ)/(2pi) SRC/cgpe1d.f90:324-332
  const double f = in->dim == 3 ? 1.0
                   : in->dim == 2 ? std::sqrt(o->alpha[2] / (2.0 * PI))
                                  : std::sqrt(o->alpha[1] * o->alpha[2]) / (2.0 * PI);
  if (in->spin == 1) {
    const double na = (double)(float)in->natoms;  // REAL(NATOMS)
    o->tau[0] = 2.0 * f * (4.0 * PI * na * (a0 * AU + 2.0 * a2 * AU) / (3.0 * aosc));
    o->tau[1] = 2.0 * f * (4.0 * PI * na * (a2 * AU - a0 * AU) / (3.0 * aosc));
    o->tau[2] = 0.0;
  } else {
    const double na = (double)in->natoms;
    o->tau[0] = 2.0 * (f * 4.0 * PI * na * (4.0 * a2 * AU + 3.0 * a4 * AU) / (7.0 * aosc));
    o->tau[1] = 2.0 * (f * 4.0 * PI * na * (a4 * AU - a2 * AU) / (7.0 * aosc));
    o->tau[2] = 2.0 * (f * 4.0 * PI * na * (7.0 * a0 * AU - 10.0 * a2 * AU + 3.0 * a4 * AU) / (7.0 * aosc));
  }
  return 0;
}

int fgpe_create(const fgpe_params* p, fgpe_handle** out) {
  if (!p || !out) return fail(FGPE_EINVAL, "null argument");
  *out = nullptr;
  if (p->dim < 1 || p->dim > 3 || (p->spin != 1 && p->spin != 2))
    return fail(FGPE_EINVAL, "dim must be 1..3 and spin 1 or 2");
  if (p->precision != FGPE_FP32 && p->precision != FGPE_FP64) return fail(FGPE_EINVAL, "bad precision");
  const bool fp32 = p->precision == FGPE_FP32;
  const int TX = fp32 ? Cfg<float>::TX : Cfg<double>::TX;
  const int nx = p->n[0], ny = p->dim >= 2 ? p->n[1] : 1, nz = p->dim >= 3 ? p->n[2] : 1;
  if (p->dim == 1) {
    if (!supported_line(nx, 1))
      return fail(FGPE_EINVAL, "NX=%d not supported in 1D (2^a or 3*2^a, 64..4096)", nx);
  } else {
    if (!supported_line(nx, 0) || !supported_line(ny, 0) || (p->dim == 3 && !supported_line(nz, 0)))
      return fail(FGPE_EINVAL, "grid %dx%dx%d not supported (each axis 2^a or 3*2^a, 32..512)", nx, ny, nz);
    if (nx % TX) return fail(FGPE_EINVAL, "NX must be a multiple of %d", TX);
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) return fail(FGPE_EDEVICE, "no HIP device available (%s)", hipGetErrorString(e));
  fgpe_handle* h = new fgpe_handle();
  h->p = *p;
  if (p->dim < 3) {
    h->p.n[2] = 1;
    h->p.gamma[2] = 0.0;
  }
  if (p->dim < 2) {
    h->p.n[1] = 1;
    h->p.gamma[1] = 0.0;
  }
  h->dim = p->dim;
  h->spin = p->spin;
  h->ms = 2 * p->spin + 1;
  h->nx = nx;
  h->ny = ny;
  h->nz = nz;
  h->npts = (long)nx * ny * nz;
  h->fp32 = fp32;
  if (p->dim == 3) {  // default 3D schedule: eighths of each axis on two streams (fgpe_tune_schedule refines)
    h->zchunk = nz / 8;
    h->xchunk = nx / 8;
    h->nstreams = 2;
  }
  int rc = 0;
  auto bail = [&](int code) {
    fgpe_destroy(h);
    return code;
  };
  if (hipGetDevice(&h->dev) != hipSuccess) return bail(fail(FGPE_EDEVICE, "hipGetDevice failed"));
  scratch_setup(h->dev);  // once per device, this device current
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->dev) == hipSuccess) h->lds_max = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipStreamCreate failed"));
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipEventCreate failed"));
  if (p->dim == 3) {  // streams and fork/join events of the chunked 3D schedule, default priority
    for (auto& s : h->aux)
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(FGPE_EDEVICE, "hipStreamCreate failed"));
    for (auto& e : h->evs)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        return bail(fail(FGPE_EDEVICE, "hipEventCreate failed"));
  }
  const size_t es = elem_size(h);
  const size_t bytes = es * (size_t)h->ms * h->npts;
  if ((rc = dalloc(h, &h->phi, bytes))) return bail(rc);
  if ((rc = dalloc(h, &h->phi_old, bytes))) return bail(rc);
  if (hipMemset(h->phi_old, 0, bytes) != hipSuccess) return bail(fail(FGPE_EDEVICE, "hipMemset failed"));
  if (hipMemset(h->phi, 0, bytes) != hipSuccess) return bail(fail(FGPE_EDEVICE, "hipMemset failed"));
  if ((rc = dalloc(h, &h->cscale, 5 * sizeof(double2)))) return bail(rc);
  if ((rc = dalloc(h, (void**)&h->cscale64, 8 * sizeof(double)))) return bail(rc);
  if ((rc = dalloc(h, (void**)&h->dnorm, 8 * sizeof(double)))) return bail(rc);
  if ((rc = dalloc(h, (void**)&h->obs_sums, 32 * sizeof(double)))) return bail(rc);
  if ((rc = dalloc(h, (void**)&h->dsig, 2 * sizeof(double)))) return bail(rc);
  if ((rc = dalloc(h, (void**)&h->derr, sizeof(int)))) return bail(rc);
  const double sig0[2] = {20.0, 20.0};  // SRC/cgpe3d.f90:443-444
  if (hipMemcpy(h->dsig, sig0, sizeof sig0, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipMemcpy failed"));
  if (hipMemset(h->derr, 0, sizeof(int)) != hipSuccess || hipMemset(h->dnorm, 0, 8 * sizeof(double)) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipMemset failed"));
  rc = fp32 ? build_tables<float>(h) : build_tables<double>(h);
  if (rc) return bail(rc);
  *out = h;
  return 0;
}

void fgpe_destroy(fgpe_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void* p : h->allocs) (void)hipFree(p);
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  for (hipEvent_t e : h->evs)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t s : h->aux)
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int fgpe_initial_guess(fgpe_handle* h) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  return h->fp32 ? guess_t<float>(h) : guess_t<double>(h);
}

int fgpe_upload(fgpe_handle* h, const double* phi) {
  if (!h || !phi) return fail(FGPE_EINVAL, "null argument");
  const long n = h->npts * h->ms;
  if (!h->fp32) {
    HIPCHK(hipMemcpy(h->phi, phi, n * sizeof(double2), hipMemcpyHostToDevice));
    return 0;
  }
  if (!h->stage64) {
    int rc = dalloc(h, &h->stage64, n * sizeof(double2));
    if (rc) return rc;
  }
  HIPCHK(hipMemcpy(h->stage64, phi, n * sizeof(double2), hipMemcpyHostToDevice));
  int rc = from64_t<float>(h, n);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fgpe_download(fgpe_handle* h, double* phi) {
  if (!h || !phi) return fail(FGPE_EINVAL, "null argument");
  const long n = h->npts * h->ms;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (!h->fp32) {
    HIPCHK(hipMemcpy(phi, h->phi, n * sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
  }
  if (!h->stage64) {
    int rc = dalloc(h, &h->stage64, n * sizeof(double2));
    if (rc) return rc;
  }
  int rc = to64_t<float>(h, n);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(phi, h->stage64, n * sizeof(double2), hipMemcpyDeviceToHost));
  return 0;
}

int fgpe_normalize(fgpe_handle* h, double* tnorm) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  return h->fp32 ? normalize_t<float>(h, tnorm) : normalize_t<double>(h, tnorm);
}

int fgpe_run(fgpe_handle* h, int64_t nsteps) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  if (nsteps < 0) return fail(FGPE_EINVAL, "negative step count");
  if (nsteps == 0) return 0;
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  int rc = h->fp32 ? run_t<float>(h, nsteps) : run_t<double>(h, nsteps);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  h->timed = true;
  int err = 0;
  HIPCHK(hipMemcpyAsync(&err, h->derr, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (err)
    return fail(FGPE_ENEWTON, "Newton-Raphson could not calculate the roots within 1e-6 tolerance");
  return 0;
}

int fgpe_set_schedule(fgpe_handle* h, int32_t zchunk, int32_t xchunk, int32_t nstreams) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  if (zchunk < 0 || xchunk < 0 || nstreams < 1 || nstreams > 3)
    return fail(FGPE_EINVAL, "bad schedule (%d, %d, %d)", zchunk, xchunk, nstreams);
  HIPCHK(hipStreamSynchronize(h->stream));
  h->zchunk = h->dim == 3 ? zchunk : 0;
  h->xchunk = h->dim == 3 ? xchunk : 0;
  h->nstreams = h->dim == 3 ? nstreams : 1;
  return 0;
}

int fgpe_get_schedule(fgpe_handle* h, int32_t* zchunk, int32_t* xchunk, int32_t* nstreams) {
  if (!h || !zchunk || !xchunk || !nstreams) return fail(FGPE_EINVAL, "null argument");
  *zchunk = h->dim == 3 ? (h->fp32 ? chunk_planes<float>(h) : chunk_planes<double>(h)) : 0;
  *xchunk = h->dim == 3 ? (h->fp32 ? chunk_cols<float>(h) : chunk_cols<double>(h)) : 0;
  *nstreams = h->nstreams;
  return 0;
}

int fgpe_tune_schedule(fgpe_handle* h, int32_t trial_steps, double* ms_out, int32_t* chosen) {
  if (!h || trial_steps < 1 || !ms_out || !chosen) return fail(FGPE_EINVAL, "bad tuning argument");
  const int nz = h->nz, nx = h->nx;
  const int cand[FGPE_NSCHED_CAND][3] = {{0, 0, 1},           {nz / 8, nx / 8, 2},  {nz / 4, nx / 4, 2},
                                         {nz / 8, nx / 4, 2}, {nz / 4, nx / 8, 2},  {nz / 8, nx / 8, 3},
                                         {nz / 16, nx / 8, 3}, {nz / 8, 0, 2}};
  for (int i = 0; i < FGPE_NSCHED_CAND; ++i) ms_out[i] = 0.0;
  chosen[0] = h->zchunk;
  chosen[1] = h->xchunk;
  chosen[2] = h->nstreams;
  if (h->dim != 3) return 0;
  // the trial steps advance the state: keep it, and the scalars a run
  // updates, and put them back afterwards
  const size_t bytes = elem_size(h) * (size_t)h->ms * h->npts;
  void* keep = nullptr;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMalloc(&keep, bytes));
  double sc[8 + 2];
  auto restore = [&]() -> int {
    HIPCHK(hipMemcpyAsync(h->phi, keep, bytes, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->dnorm, sc, 8 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->dsig, sc + 8, 2 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
  };
  int rc = 0;
  if (hipMemcpyAsync(keep, h->phi, bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(sc, h->dnorm, 8 * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipMemcpyAsync(sc + 8, h->dsig, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    rc = fail(FGPE_EDEVICE, "saving the state for tuning failed");
  int best = -1;
  for (int i = 0; i < FGPE_NSCHED_CAND && !rc; ++i) {
    h->zchunk = cand[i][0];
    h->xchunk = cand[i][1];
    h->nstreams = cand[i][2];
    if ((rc = fgpe_run(h, 1)) || (rc = fgpe_run(h, trial_steps))) break;
    double ms = 0;
    if ((rc = fgpe_last_run_ms(h, &ms))) break;
    ms_out[i] = ms / trial_steps;
    if (best < 0 || ms_out[i] < ms_out[best]) best = i;
    if ((rc = restore())) break;
  }
  (void)hipFree(keep);
  if (rc) return rc;
  h->zchunk = cand[best][0];
  h->xchunk = cand[best][1];
  h->nstreams = cand[best][2];
  return fgpe_get_schedule(h, &chosen[0], &chosen[1], &chosen[2]);
}

int fgpe_run_noalloc(fgpe_handle* h, int64_t nsteps) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  if (nsteps < 0) return fail(FGPE_EINVAL, "negative step count");
  std::string why;
  double* ptr = scratch_ptr(h->dev, &why);
  if (!ptr) return fail(FGPE_EDEVICE, "no-allocation scratch unavailable on device %d: %s", h->dev, why.c_str());
  // pass A writes MS partials per block (q1D keeps its norms inside its one block)
  const size_t need = h->dim == 1 ? 0
                      : (size_t)h->ms * (size_t)(h->fp32 ? x_pass_blocks<float>(h->dim, h->nx, (long)h->ny * h->nz)
                                                         : x_pass_blocks<double>(h->dim, h->nx, (long)h->ny * h->nz));
  if (need > NOALLOC_SCRATCH_DOUBLES)
    return fail(FGPE_EINVAL, "grid %dx%dx%d needs %zu doubles of scratch; the no-allocation limit is %zu", h->nx,
                h->ny, h->nz, need, NOALLOC_SCRATCH_DOUBLES);
  double* own = h->partials;
  const size_t own_n = h->partials_n;
  h->partials = ptr;
  h->partials_n = NOALLOC_SCRATCH_DOUBLES;
  h->noalloc = true;
  const int rc = fgpe_run(h, nsteps);
  h->noalloc = false;
  h->partials = own;
  h->partials_n = own_n;
  return rc;
}

int fgpe_last_norms(fgpe_handle* h, double* norm) {
  if (!h || !norm) return fail(FGPE_EINVAL, "null argument");
  double n[8];
  HIPCHK(hipMemcpyAsync(n, h->dnorm, sizeof n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int m = 0; m < h->ms; ++m) norm[m] = n[m];
  return 0;
}

int fgpe_normc(fgpe_handle* h, double* norm) {
  if (!h || !norm) return fail(FGPE_EINVAL, "null argument");
  return h->fp32 ? normc_t<float>(h, norm) : normc_t<double>(h, norm);
}

int fgpe_observables(fgpe_handle* h, double* en, double* mu, double* mz, double* rms) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  return h->fp32 ? obs_t<float>(h, en, mu, mz, rms) : obs_t<double>(h, en, mu, mz, rms);
}

int fgpe_keep_old(fgpe_handle* h) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  HIPCHK(hipMemcpyAsync(h->phi_old, h->phi, elem_size(h) * h->ms * h->npts, hipMemcpyDeviceToDevice, h->stream));
  return 0;
}

int fgpe_conv_max(fgpe_handle* h, double* maxdiff) {
  if (!h || !maxdiff) return fail(FGPE_EINVAL, "null argument");
  return h->fp32 ? conv_t<float>(h, maxdiff) : conv_t<double>(h, maxdiff);
}

int fgpe_center_abs(fgpe_handle* h, double* absphi) {
  if (!h || !absphi) return fail(FGPE_EINVAL, "null argument");
  const long c = (long)(h->nx / 2) + (long)h->nx * ((h->dim >= 2 ? h->ny / 2 : 0) + (long)h->ny * (h->dim >= 3 ? h->nz / 2 : 0));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int m = 0; m < h->ms; ++m) {
    const long q = (long)m * h->npts + c;
    if (h->fp32) {
      float2 v;
      HIPCHK(hipMemcpy(&v, (float2*)h->phi + q, sizeof v, hipMemcpyDeviceToHost));
      absphi[m] = std::hypot((double)v.x, (double)v.y);
    } else {
      double2 v;
      HIPCHK(hipMemcpy(&v, (double2*)h->phi + q, sizeof v, hipMemcpyDeviceToHost));
      absphi[m] = std::hypot(v.x, v.y);
    }
  }
  return 0;
}

int fgpe_sync(fgpe_handle* h) {
  if (!h) return fail(FGPE_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int fgpe_pass_times(fgpe_handle* h, int64_t nsteps, double* ms_out, int32_t* count_out) {
  if (!h || !ms_out || !count_out) return fail(FGPE_EINVAL, "null argument");
  // whole-grid launches on one stream, so every line-pass launch moves the
  // whole spinor once each way
  const int zc = h->zchunk, xc = h->xchunk;
  h->zchunk = h->xchunk = 0;
  h->profiling = true;
  h->ev_used = 0;
  h->ev_kind.clear();
  int rc = fgpe_run(h, nsteps);
  h->profiling = false;
  h->zchunk = zc;
  h->xchunk = xc;
  if (rc) return rc;
  double sum[FGPE_NPASS] = {0};
  int cnt[FGPE_NPASS] = {0};
  for (size_t i = 0; i < h->ev_kind.size(); ++i) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, h->ev_pool[2 * i], h->ev_pool[2 * i + 1]));
    sum[h->ev_kind[i]] += t;
    cnt[h->ev_kind[i]] += 1;
  }
  for (int k = 0; k < FGPE_NPASS; ++k) {
    ms_out[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
    count_out[k] = cnt[k];
  }
  return 0;
}

int fgpe_last_run_ms(fgpe_handle* h, double* ms) {
  if (!h || !ms) return fail(FGPE_EINVAL, "null argument");
  if (!h->timed) return fail(FGPE_ESTATE, "no timed run yet");
  HIPCHK(hipEventSynchronize(h->ev1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
  *ms = t;
  return 0;
}

}  // extern "C"

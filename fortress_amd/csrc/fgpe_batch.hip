// fgpe_batch.hip -- many independent q2D instances advanced together (spin 1
// or spin 2, one spin per batch; grids of 64, 128, 192 or 256 points per axis,
// square or not).
//
// Instances are grouped by grid shape (x length, y length).  Every pass of the
// q2D step (run_steps in fgpe_ops.hpp, SRC/cgpe2d.f90:440-485) is ONE launch
// per group over all of the group's instances: the grid is instances x
// blocks-per-instance, a block reads its instance's descriptor (state, tables,
// parameters) and runs the same x-line or k-line block code as the
// single-instance kernels (xpass_block / kpass_block), so a batched instance
// computes what fgpe_run computes under the same schedule.  NORMC and the
// renormalisation scale come from one finalize launch per group (a block per
// instance); observables and the convergence maximum from one launch per
// group (several blocks per instance, partial sums reduced on the host in a
// fixed order).  A step of B instances in G shape groups costs 2G launches
// (3G in imaginary time), independent of B.
//
// Schedules: each line length has two (x-pass points per thread, k-pass
// points per thread) choices; the x pass of a group follows the choice of its
// x-line length, the k pass that of its y-line length.  fgpe_batch_tune times
// both choices of every length on the batch's own instances and keeps the
// faster.
#include <numeric>

#include "fgpe_impl.hpp"

namespace fgpe {

template <typename T> struct GInst {
  C_t<T>* phi;       // [ms][ny][nx]
  C_t<T>* old;       // PHI_OLD for the convergence test
  Tables<T> tb;      // twx / twy: the twiddles of the instance's schedule
  Phys ph;
  FinArgs f;
  ObsArgs oa;
  double* partials;  // NORMC partials of the x pass
  double* norm;      // NORMC of the last step [ms]
  double* cscale64;  // renormalisation scale [ms]
  C_t<T>* cscale;    // the same in the propagation precision (= tb.cscale)
  double* sig;       // NEWTON state [2] (spin 2 only; finalize takes it)
  int* err;
};

constexpr int GB_OBS = 16;  // blocks per instance of the norm / observables / max-diff launches

// x lines of NX points (NY lines per instance)
template <typename T, int SPIN, int NX, int Q>
__global__ __launch_bounds__(256) void kg_xpass(const GInst<T>* __restrict__ gi, const int* __restrict__ list, int nb,
                                                int ny, int flags) {
  const GInst<T>& I = gi[list[blockIdx.x / nb]];
  const Geom g{NX, ny, 1, (long)NX * ny};
  xpass_block<T, SPIN, NX, Q>((int)(blockIdx.x % nb), I.phi, g, I.tb, I.ph, flags, I.partials);
}

// y lines of NY points (NX columns per instance)
template <typename T, int SPIN, int NY, int Q, int TX>
__global__ __launch_bounds__(TX *(NY / Q)) void kg_kpass(const GInst<T>* __restrict__ gi,
                                                         const int* __restrict__ list, int nb, int nx, int scale) {
  // blocks (instance, x block) laid out as pass C's grid: the two 64 B halves
  // of a 128 B row go to workgroups 8 apart, on one XCD (kpass_coords)
  int bx, by;
  kpass_coords<TX * (int)sizeof(C_t<T>)>(bx, by);
  const GInst<T>& I = gi[list[bx / nb]];
  const Geom g{nx, NY, 1, (long)nx * NY};
  kpass_block<T, SPIN, NY, Q, TX>(bx % nb, I.phi, g, I.tb, I.ph, scale);
}

template <typename T, int MS>
__global__ __launch_bounds__(1024) void kg_finalize(const GInst<T>* __restrict__ gi, const int* __restrict__ list,
                                                    int nparts) {
  const GInst<T>& I = gi[list[blockIdx.x]];
  finalize_block<T, MS>(I.partials, nparts, I.f, I.norm, I.cscale, I.cscale64, I.sig, I.err);
}

template <typename T>
__global__ void kg_set_scale(const GInst<T>* __restrict__ gi, const int* __restrict__ list, int ms) {
  const GInst<T>& I = gi[list[blockIdx.x]];
  if ((int)threadIdx.x < ms) I.cscale[threadIdx.x] = mk<C_t<T>>((T)1, (T)0);
}

// phi *= cscale per component; blocks: instances x nb over np points each
template <typename T>
__global__ __launch_bounds__(256) void kg_apply_scale(const GInst<T>* __restrict__ gi, const int* __restrict__ list,
                                                      int nb, long np, int ms) {
  const GInst<T>& I = gi[list[blockIdx.x / nb]];
  const long b = blockIdx.x % nb;
  for (long q = b * 256 + threadIdx.x; q < np; q += (long)nb * 256)
    for (int m = 0; m < ms; ++m) I.phi[m * np + q] = cmul(I.phi[m * np + q], I.cscale[m]);
}

// old = phi over n3 values
template <typename T>
__global__ __launch_bounds__(256) void kg_keep_old(const GInst<T>* __restrict__ gi, const int* __restrict__ list,
                                                   int nb, long n3) {
  const GInst<T>& I = gi[list[blockIdx.x / nb]];
  const long b = blockIdx.x % nb;
  for (long q = b * 256 + threadIdx.x; q < n3; q += (long)nb * 256) I.old[q] = I.phi[q];
}

// NORMC partials (Simpson-weighted |phi_m|^2): out[(li*GB_OBS + b)*MS + m]
template <typename T, int MS>
__global__ __launch_bounds__(256) void kg_norm(const GInst<T>* __restrict__ gi, const int* __restrict__ list, int nx,
                                               int ny, double* __restrict__ out) {
  __shared__ double red[16 * MS];
  const int li = blockIdx.x / GB_OBS, b = blockIdx.x % GB_OBS;
  const GInst<T>& I = gi[list[li]];
  const long np = (long)nx * ny;
  double acc[MS];
#pragma unroll
  for (int m = 0; m < MS; ++m) acc[m] = 0.0;
  for (long q = (long)b * 256 + threadIdx.x; q < np; q += (long)GB_OBS * 256) {
    const double w = I.oa.wx[q % nx] * I.oa.wy[q / nx];
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      const C_t<T> a = I.phi[m * np + q];
      acc[m] += w * ((double)a.x * (double)a.x + (double)a.y * (double)a.y);
    }
  }
  block_sum_store<MS>(acc, red, out, (long)blockIdx.x);
}

// ENERGY/RAD/NORMC columns (3 MS + 2): out[(li*GB_OBS + b)*NC + c]
template <typename T, int SPIN>
__global__ __launch_bounds__(256) void kg_obs(const GInst<T>* __restrict__ gi, const int* __restrict__ list, int nx,
                                              int ny, double* __restrict__ out) {
  constexpr int NC = 3 * (2 * SPIN + 1) + 2;
  __shared__ double red[16 * NC];
  const int li = blockIdx.x / GB_OBS, b = blockIdx.x % GB_OBS;
  const GInst<T>& I = gi[list[li]];
  const Geom g{nx, ny, 1, (long)nx * ny};
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (long q = (long)b * 256 + threadIdx.x; q < g.npts; q += (long)GB_OBS * 256)
    obs_point<T, SPIN, 2>(I.phi, g, I.oa, q, acc);
  block_sum_store<NC>(acc, red, out, (long)blockIdx.x);
}

// max |phi - old| over n3 values: out[li*GB_OBS + b]
template <typename T>
__global__ __launch_bounds__(256) void kg_maxdiff(const GInst<T>* __restrict__ gi, const int* __restrict__ list,
                                                  long n3, double* __restrict__ out) {
  __shared__ double red[4];
  const int li = blockIdx.x / GB_OBS, b = blockIdx.x % GB_OBS;
  const GInst<T>& I = gi[list[li]];
  double mx = 0.0;
  for (long q = (long)b * 256 + threadIdx.x; q < n3; q += (long)GB_OBS * 256) {
    const C_t<T> a = I.phi[q], c = I.old[q];
    const double dr = (double)a.x - (double)c.x, di = (double)a.y - (double)c.y;
    mx = fmax(mx, sqrt(dr * dr + di * di));
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = mx;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

}  // namespace fgpe

using namespace fgpe_impl;

namespace {

constexpr int NCLS = 4;
constexpr int NGRP = NCLS * NCLS;  // shape groups: (x-length class, y-length class)
constexpr int CLS_N[NCLS] = {64, 128, 192, 256};
// schedule s = 2*length class + variant: points per thread of the x pass and
// of the k pass (FP64 holds twice the registers per point: smaller Q)
constexpr int SQX32[FGPE_BATCH_NSCHED] = {4, 8, 4, 4, 12, 6, 4, 4};
constexpr int SQK32[FGPE_BATCH_NSCHED] = {8, 4, 8, 4, 12, 6, 8, 4};
constexpr int SQX64[FGPE_BATCH_NSCHED] = {8, 4, 8, 4, 6, 12, 8, 4};
constexpr int SQK64[FGPE_BATCH_NSCHED] = {4, 8, 4, 8, 6, 6, 4, 8};
// x columns per k-pass block: whole 128 B rows in FP64 (8); 16 in FP32 (the
// square spin-1 sweep 981k -> 1,026k instance-steps/s against 8, the others
// within 2 %; 16 in FP64 made the FP64 solve of the sweep 2.5 % slower)
template <typename T> constexpr int btx() { return sizeof(T) == 4 ? 16 : 8; }
int size_class(int n) { return n == 64 ? 0 : n == 128 ? 1 : n == 192 ? 2 : n == 256 ? 3 : -1; }
template <typename T> int sqx(int s) { return sizeof(T) == 4 ? SQX32[s] : SQX64[s]; }
template <typename T> int sqk(int s) { return sizeof(T) == 4 ? SQK32[s] : SQK64[s]; }
// x-pass lines per block (as launch_x) and blocks per instance of NY lines
inline int x_lpw(int nx, int q) { return std::max(256 / (nx / q), 1); }
inline int x_blocks(int nx, int ny, int q) { return (ny + x_lpw(nx, q) - 1) / x_lpw(nx, q); }

}  // namespace

struct fgpe_batch {
  int count = 0;
  int spin = 1, ms = 3;
  bool fp32 = true;
  int dev = 0;
  size_t lds_max = 65536;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // the shape groups' independent step chains run on gstream[group % 4]
  // (gstream[0] is the batch's stream), forked from and joined back to it by gev
  hipStream_t gstream[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t gev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<fgpe_params> p;
  std::vector<int> nx, ny, cx, cy;  // grid and length classes per instance
  int csched[NCLS] = {0, 2, 4, 6};  // schedule per line length (the first variant until tuned)
  std::vector<long> off, poff;
  void* phi = nullptr;
  void* old = nullptr;
  void* stage = nullptr;
  size_t stage_elems = 0;
  void* ginst = nullptr;  // device GInst<T>[count]
  std::vector<unsigned char> ginst_host;
  size_t ginst_size = 0;
  double* norms = nullptr;    // [count][ms]
  double* scale64 = nullptr;  // [count][ms]
  void* cscale = nullptr;     // [count][ms] complex
  int* errs = nullptr;        // [count]
  double* obs = nullptr;      // [count][GB_OBS][3 ms + 3]: observables columns, then max-diff
  int* dlist = nullptr;       // [count]: the listed instances, grouped by shape
  std::vector<fgpe_batch_result> res;
  std::vector<double> en_old;
  double last_ms = 0;
  const void* twx[FGPE_BATCH_NSCHED] = {};
  const void* twk[FGPE_BATCH_NSCHED] = {};
  std::vector<void*> allocs;
  long npts(int i) const { return (long)nx[i] * ny[i]; }
  int ncol() const { return 3 * ms + 2; }
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

template <typename T> GInst<T>* host_inst(fgpe_batch* b) { return reinterpret_cast<GInst<T>*>(b->ginst_host.data()); }

// point every descriptor at the twiddles of its lengths' schedules and upload
// the array
template <typename T> int upload_inst(fgpe_batch* b) {
  GInst<T>* in = host_inst<T>(b);
  for (int i = 0; i < b->count; ++i) {
    in[i].tb.twx = (const R_t<T>*)b->twx[b->csched[b->cx[i]]];
    in[i].tb.twy = (const R_t<T>*)b->twk[b->csched[b->cy[i]]];
  }
  HIPCHK(hipMemcpy(b->ginst, b->ginst_host.data(), b->ginst_size * b->count, hipMemcpyHostToDevice));
  return 0;
}

template <typename T> int build_batch(fgpe_batch* b) {
  b->ginst_size = sizeof(GInst<T>);
  b->ginst_host.assign(b->ginst_size * b->count, 0);
  GInst<T>* in = host_inst<T>(b);
  const int ms = b->ms;
  int rc;
  for (int s = 0; s < FGPE_BATCH_NSCHED; ++s) {
    const R_t<T>* t;
    if ((rc = bupload(b, &t, stage_twiddles<T>(CLS_N[s / 2], sqx<T>(s))))) return rc;
    b->twx[s] = (const void*)t;
    if ((rc = bupload(b, &t, stage_twiddles<T>(CLS_N[s / 2], sqk<T>(s))))) return rc;
    b->twk[s] = (const void*)t;
  }
  long ptotal = 0;
  b->poff.resize(b->count);
  for (int i = 0; i < b->count; ++i) {
    const int c = b->cx[i];
    const int nb = std::max(x_blocks(b->nx[i], b->ny[i], sqx<T>(2 * c)), x_blocks(b->nx[i], b->ny[i], sqx<T>(2 * c + 1)));
    b->poff[i] = ptotal;
    ptotal += (long)nb * ms;
  }
  double *partials, *sig;
  if ((rc = balloc(b, (void**)&partials, ptotal * sizeof(double))) ||
      (rc = balloc(b, (void**)&sig, 2 * sizeof(double) * b->count)))
    return rc;
  {
    // the spin-2 Newton start (20, 20) of every instance (SRC/cgpe2d.f90:432-433)
    std::vector<double> sig0((size_t)2 * b->count, 20.0);
    HIPCHK(hipMemcpy(sig, sig0.data(), sig0.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  for (int i = 0; i < b->count; ++i) {
    const fgpe_params& p = b->p[i];
    const int nx = b->nx[i], ny = b->ny[i];
    GInst<T>& I = in[i];
    I.phi = (C<T>*)b->phi + b->off[i];
    I.old = (C<T>*)b->old + b->off[i];
    // Phys and FinArgs as phys()/finargs() of a q2D handle
    I.ph.tau0 = p.tau[0];
    I.ph.tau1 = p.tau[1];
    I.ph.tau2 = p.tau[2];
    I.ph.dt = p.dt;
    I.ph.gx = p.gamma[0];
    I.ph.gy = p.gamma[1];
    I.ph.gz = 0.0;
    I.ph.inv_npts = 1.0 / ((double)nx * ny);
    I.ph.imag = p.imag;
    I.ph.dim = 2;
    I.f.imag = p.imag;
    I.f.soc = p.soc;
    I.f.dim = 2;
    I.f.ms = ms;
    I.f.mag = p.mag;
    I.f.dt = p.dt;
    I.partials = partials + b->poff[i];
    I.norm = b->norms + (size_t)ms * i;
    I.cscale64 = b->scale64 + (size_t)ms * i;
    I.cscale = (C<T>*)b->cscale + (size_t)ms * i;
    I.sig = sig + 2 * i;
    I.err = b->errs + i;
    // tables as build_tables (fgpe.hip) builds them for a q2D handle
    auto kx = kgrid(nx, p.d[0]), ky = kgrid(ny, p.d[1]);
    auto xx = xgrid(nx, p.d[0]), yy = xgrid(ny, p.d[1]);
    std::vector<R_t<T>> kex(nx), kel((size_t)ms * ny);
    for (int k = 0; k < nx; ++k) kex[k] = ke_factor<T>(kx[k] * kx[k], p.dt, p.imag);
    for (int m = 0; m < ms; ++m)
      for (int k = 0; k < ny; ++k) kel[(size_t)m * ny + k] = ke_factor<T>(ky[k] * ky[k], p.dt, p.imag);
    std::vector<T> kxt(kx.begin(), kx.end()), kyt(ky.begin(), ky.end()), vz(1, (T)0);
    std::vector<double> x2(nx), y2(ny), vxd(nx), vyd(ny), zero(1, 0.0), one(1, 1.0);
    for (int k = 0; k < nx; ++k) {
      x2[k] = xx[k] * xx[k];
      vxd[k] = p.alpha[0] * p.alpha[0] * x2[k];
    }
    for (int k = 0; k < ny; ++k) {
      y2[k] = yy[k] * yy[k];
      vyd[k] = p.alpha[1] * p.alpha[1] * y2[k];
    }
    std::vector<T> vx(vxd.begin(), vxd.end()), vy(vyd.begin(), vyd.end());
    auto wx = simpson_weights(nx, p.d[0]), wy = simpson_weights(ny, p.d[1]);
    Tables<T>& t = I.tb;
    if ((rc = bupload(b, &t.kex, kex)) || (rc = bupload(b, &t.kel, kel)) || (rc = bupload(b, &t.kx, kxt)) ||
        (rc = bupload(b, &t.ky, kyt)) || (rc = bupload(b, &t.vx, vx)) || (rc = bupload(b, &t.vy, vy)) ||
        (rc = bupload(b, &t.vz, vz)) || (rc = bupload(b, &t.wx, wx)) || (rc = bupload(b, &t.wy, wy)) ||
        (rc = bupload(b, &t.wz, one)))
      return rc;
    {
      const fgpe::C_t<T>* st;
      if ((rc = bupload(b, &st, soc2_table<T>(kx, ky, p.gamma[0], p.gamma[1], p.dt, p.imag, b->spin)))) return rc;
      t.soc2 = st;
    }
    t.key = t.kex;  // not read by the 2D passes
    t.twz = nullptr;
    t.twd = nullptr;
    t.cscale = I.cscale;
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
    a.wx = t.wx;
    a.wy = t.wy;
    a.wz = t.wz;
    if ((rc = bupload(b, &a.x2, x2)) || (rc = bupload(b, &a.y2, y2)) || (rc = bupload(b, &a.z2, zero)) ||
        (rc = bupload(b, &a.vxd, vxd)) || (rc = bupload(b, &a.vyd, vyd)) || (rc = bupload(b, &a.vzd, zero)))
      return rc;
  }
  if ((rc = balloc(b, &b->ginst, b->ginst_size * b->count))) return rc;
  return upload_inst<T>(b);
}

// the listed instances grouped by shape (group 4 cy + cx), uploaded to dlist
struct Groups {
  int start[NGRP] = {}, cnt[NGRP] = {};
  int total = 0;
  std::vector<int> ids;  // dlist contents
};
inline int gx_of(int gr) { return gr % NCLS; }
inline int gy_of(int gr) { return gr / NCLS; }

int upload_list(fgpe_batch* b, const std::vector<int>& list, Groups& g) {
  g = Groups();
  for (int gr = 0; gr < NGRP; ++gr) {
    g.start[gr] = (int)g.ids.size();
    for (int id : list)
      if (b->cx[id] == gx_of(gr) && b->cy[id] == gy_of(gr)) g.ids.push_back(id);
    g.cnt[gr] = (int)g.ids.size() - g.start[gr];
  }
  g.total = (int)g.ids.size();
  HIPCHK(hipStreamSynchronize(b->stream));  // the previous list is no longer read
  if (g.total) HIPCHK(hipMemcpy(b->dlist, g.ids.data(), g.ids.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

template <typename K> void allow(K kernel, size_t bytes) {
  if (bytes > 65536) (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// one x-pass launch over the cnt instances of x length NX (ny lines each) at
// dlist + start
template <typename T, int SPIN, int NX, int Q>
int xlaunch(fgpe_batch* b, hipStream_t st, int start, int cnt, int ny, int flags) {
  constexpr int P = NX / Q, MS = 2 * SPIN + 1;
  constexpr int TWS = twiddle_size(NX, Q) > 0 ? twiddle_size(NX, Q) : 1;
  const int threads = P >= 256 ? P : 256, lpw = threads / P, nb = (ny + lpw - 1) / lpw;
  const size_t lds = (TWS + NX) * sizeof(R_t<T>) + 16 * MS * sizeof(double) + (size_t)lpw * line_stride(NX) * sizeof(C<T>);
  if (lds > b->lds_max) return fail(FGPE_EINVAL, "batch x pass needs %zu B of LDS", lds);
  auto kern = kg_xpass<T, SPIN, NX, Q>;
  allow(kern, lds);
  hipLaunchKernelGGL(kern, dim3(cnt * nb), dim3(threads), lds, st, (const GInst<T>*)b->ginst, b->dlist + start,
                     nb, ny, flags);
  HIPCHK(hipGetLastError());
  if ((flags & XF_PHYS) && (flags & XF_NORM)) {
    hipLaunchKernelGGL((kg_finalize<T, MS>), dim3(cnt), dim3(fin_threads(nb)), 0, st, (const GInst<T>*)b->ginst,
                       b->dlist + start, nb);
    HIPCHK(hipGetLastError());
  }
  return 0;
}
// one k-pass launch (2D: y lines of NY points, KE, SOC) over the cnt instances
// (nx columns each)
template <typename T, int SPIN, int NY, int Q> int klaunch(fgpe_batch* b, hipStream_t st, int start, int cnt, int nx) {
  constexpr int P = NY / Q;
  constexpr int BTX = btx<T>();
  static_assert(BTX * P <= 1024, "batch k pass block shape");
  const int nb = nx / BTX;
  const size_t lds = kpass_lds_bytes<T, SPIN, NY, Q, BTX>();
  if (lds > b->lds_max) return fail(FGPE_EINVAL, "batch k pass needs %zu B of LDS", lds);
  auto kern = kg_kpass<T, SPIN, NY, Q, BTX>;
  allow(kern, lds);
  hipLaunchKernelGGL(kern, dim3(cnt * nb), dim3(BTX * P), lds, st, (const GInst<T>*)b->ginst, b->dlist + start,
                     nb, nx, 1);
  HIPCHK(hipGetLastError());
  return 0;
}

// schedule dispatch (compile-time Q per schedule)
template <typename T, int SPIN> int xpass_s(fgpe_batch* b, hipStream_t st, int s, int start, int cnt, int ny, int flags) {
  constexpr bool F = sizeof(T) == 4;
  switch (s) {
    case 0: return xlaunch<T, SPIN, 64, F ? SQX32[0] : SQX64[0]>(b, st, start, cnt, ny, flags);
    case 1: return xlaunch<T, SPIN, 64, F ? SQX32[1] : SQX64[1]>(b, st, start, cnt, ny, flags);
    case 2: return xlaunch<T, SPIN, 128, F ? SQX32[2] : SQX64[2]>(b, st, start, cnt, ny, flags);
    case 3: return xlaunch<T, SPIN, 128, F ? SQX32[3] : SQX64[3]>(b, st, start, cnt, ny, flags);
    case 4: return xlaunch<T, SPIN, 192, F ? SQX32[4] : SQX64[4]>(b, st, start, cnt, ny, flags);
    case 5: return xlaunch<T, SPIN, 192, F ? SQX32[5] : SQX64[5]>(b, st, start, cnt, ny, flags);
    case 6: return xlaunch<T, SPIN, 256, F ? SQX32[6] : SQX64[6]>(b, st, start, cnt, ny, flags);
    case 7: return xlaunch<T, SPIN, 256, F ? SQX32[7] : SQX64[7]>(b, st, start, cnt, ny, flags);
  }
  return fail(FGPE_EINVAL, "bad schedule %d", s);
}
template <typename T, int SPIN> int kpass_s(fgpe_batch* b, hipStream_t st, int s, int start, int cnt, int nx) {
  constexpr bool F = sizeof(T) == 4;
  switch (s) {
    case 0: return klaunch<T, SPIN, 64, F ? SQK32[0] : SQK64[0]>(b, st, start, cnt, nx);
    case 1: return klaunch<T, SPIN, 64, F ? SQK32[1] : SQK64[1]>(b, st, start, cnt, nx);
    case 2: return klaunch<T, SPIN, 128, F ? SQK32[2] : SQK64[2]>(b, st, start, cnt, nx);
    case 3: return klaunch<T, SPIN, 128, F ? SQK32[3] : SQK64[3]>(b, st, start, cnt, nx);
    case 4: return klaunch<T, SPIN, 192, F ? SQK32[4] : SQK64[4]>(b, st, start, cnt, nx);
    case 5: return klaunch<T, SPIN, 192, F ? SQK32[5] : SQK64[5]>(b, st, start, cnt, nx);
    case 6: return klaunch<T, SPIN, 256, F ? SQK32[6] : SQK64[6]>(b, st, start, cnt, nx);
    case 7: return klaunch<T, SPIN, 256, F ? SQK32[7] : SQK64[7]>(b, st, start, cnt, nx);
  }
  return fail(FGPE_EINVAL, "bad schedule %d", s);
}

// nsteps steps of every listed instance, from real space back to real space:
// the pass schedule of run_steps (fgpe_ops.hpp) for dim 2.  Shape groups
// share nothing, so each group's chain of launches runs on one of four
// streams: the small groups' launches fill the GPU beside the large ones'.
template <typename T, int SPIN> int steps_t(fgpe_batch* b, const Groups& g, long nsteps) {
  if (nsteps <= 0 || g.total == 0) return 0;
  int rc;
  hipLaunchKernelGGL(kg_set_scale<T>, dim3(g.total), dim3(64), 0, b->stream, (const GInst<T>*)b->ginst, b->dlist,
                     b->ms);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->gev[0], b->stream));
  for (int c = 1; c < 4; ++c) HIPCHK(hipStreamWaitEvent(b->gstream[c], b->gev[0], 0));
  const int imag = b->p[0].imag;
  for (int gr = 0, k = 0; gr < NGRP; ++gr) {
    if (!g.cnt[gr]) continue;
    hipStream_t st = b->gstream[k++ % 4];
    const int nx = CLS_N[gx_of(gr)], ny = CLS_N[gy_of(gr)];
    const int sx = b->csched[gx_of(gr)], sk = b->csched[gy_of(gr)];
    if ((rc = xpass_s<T, SPIN>(b, st, sx, g.start[gr], g.cnt[gr], ny, XF_FWD))) return rc;
    for (long s = 1; s <= nsteps; ++s) {
      if ((rc = kpass_s<T, SPIN>(b, st, sk, g.start[gr], g.cnt[gr], nx))) return rc;
      const int flags = XF_INV | XF_PHYS | (s < nsteps ? XF_FWD : 0) | ((imag || s == nsteps) ? XF_NORM : 0);
      if ((rc = xpass_s<T, SPIN>(b, st, sx, g.start[gr], g.cnt[gr], ny, flags))) return rc;
    }
    if (imag) {
      const long np = (long)nx * ny;
      const int nb = (int)std::min<long>((np + 255) / 256, 64);
      hipLaunchKernelGGL(kg_apply_scale<T>, dim3(g.cnt[gr] * nb), dim3(256), 0, st, (const GInst<T>*)b->ginst,
                         b->dlist + g.start[gr], nb, np, b->ms);
      HIPCHK(hipGetLastError());
    }
  }
  for (int c = 1; c < 4; ++c) {
    HIPCHK(hipEventRecord(b->gev[c], b->gstream[c]));
    HIPCHK(hipStreamWaitEvent(b->stream, b->gev[c], 0));
  }
  return 0;
}

template <typename T> int keep_old_t(fgpe_batch* b, const Groups& g) {
  for (int gr = 0; gr < NGRP; ++gr)
    if (g.cnt[gr]) {
      const long n3 = (long)b->ms * CLS_N[gx_of(gr)] * CLS_N[gy_of(gr)];
      const int nb = (int)std::min<long>((n3 + 255) / 256, 64);
      hipLaunchKernelGGL(kg_keep_old<T>, dim3(g.cnt[gr] * nb), dim3(256), 0, b->stream, (const GInst<T>*)b->ginst,
                         b->dlist + g.start[gr], nb, n3);
      HIPCHK(hipGetLastError());
    }
  return 0;
}

// observables (and the convergence maximum) of the listed instances into res
template <typename T, int SPIN> int events_t(fgpe_batch* b, const Groups& g, bool conv) {
  constexpr int MS = 2 * SPIN + 1, NC = 3 * MS + 2;
  const size_t ncol = (size_t)g.total * GB_OBS * NC;
  std::vector<double> part(ncol), md((size_t)g.total * GB_OBS);
  for (int gr = 0; gr < NGRP; ++gr)
    if (g.cnt[gr]) {
      const int nx = CLS_N[gx_of(gr)], ny = CLS_N[gy_of(gr)];
      hipLaunchKernelGGL((kg_obs<T, SPIN>), dim3(g.cnt[gr] * GB_OBS), dim3(256), 0, b->stream,
                         (const GInst<T>*)b->ginst, b->dlist + g.start[gr], nx, ny,
                         b->obs + (size_t)g.start[gr] * GB_OBS * NC);
      HIPCHK(hipGetLastError());
      if (conv) {
        hipLaunchKernelGGL(kg_maxdiff<T>, dim3(g.cnt[gr] * GB_OBS), dim3(256), 0, b->stream, (const GInst<T>*)b->ginst,
                           b->dlist + g.start[gr], (long)MS * nx * ny, b->obs + ncol + (size_t)g.start[gr] * GB_OBS);
        HIPCHK(hipGetLastError());
      }
    }
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(part.data(), b->obs, ncol * sizeof(double), hipMemcpyDeviceToHost));
  if (conv) HIPCHK(hipMemcpy(md.data(), b->obs + ncol, md.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int li = 0; li < g.total; ++li) {
    fgpe_batch_result& r = b->res[g.ids[li]];
    double s[NC] = {0};
    for (int k = 0; k < GB_OBS; ++k)
      for (int c = 0; c < NC; ++c) s[c] += part[((size_t)li * GB_OBS + k) * NC + c];
    double sumn = 0;
    for (int m = 0; m < MS; ++m) sumn += s[2 * MS + 2 + m];
    // EN = int col1 / sum(NORM), MU = int col / (NORM + EPS), RMS = sqrt(int
    // r^2 |phi|^2) (SRC/cgpe2d.f90:1276-1279, 1069), MZ = int sum m |phi_m|^2
    r.en = s[0] / sumn;
    r.mz = s[1];
    for (int m = 0; m < MS; ++m) {
      r.mu[m] = s[2 + m] / (s[2 * MS + 2 + m] + EPS);
      r.rms[m] = std::sqrt(s[MS + 2 + m]);
    }
    if (conv) {
      double mx = 0;
      for (int k = 0; k < GB_OBS; ++k) mx = std::max(mx, md[(size_t)li * GB_OBS + k]);
      r.maxdiff = mx;
    }
  }
  return 0;
}

int pull_norms(fgpe_batch* b, const Groups& g) {
  const int ms = b->ms;
  std::vector<double> nn((size_t)b->count * ms);
  std::vector<int> er(b->count);
  HIPCHK(hipMemcpy(nn.data(), b->norms, nn.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(er.data(), b->errs, er.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (int id : g.ids) {
    for (int m = 0; m < ms; ++m) b->res[id].norm[m] = nn[(size_t)id * ms + m];
    if (er[id]) b->res[id].err = 1;
  }
  return 0;
}

// advance the listed instances by nsteps (PHI_OLD kept one step before the
// end when conv), then their observables; device time of the steps into
// last_ms when timed
template <typename T, int SPIN>
int advance_t(fgpe_batch* b, const std::vector<int>& list, long nsteps, bool conv, bool timed) {
  Groups g;
  int rc = upload_list(b, list, g);
  if (rc || g.total == 0) return rc;
  if (timed) HIPCHK(hipEventRecord(b->ev0, b->stream));
  if (conv && nsteps >= 1) {
    if ((rc = steps_t<T, SPIN>(b, g, nsteps - 1)) || (rc = keep_old_t<T>(b, g)) || (rc = steps_t<T, SPIN>(b, g, 1)))
      return rc;
  } else if ((rc = steps_t<T, SPIN>(b, g, nsteps))) {
    return rc;
  }
  if (timed) HIPCHK(hipEventRecord(b->ev1, b->stream));
  if ((rc = events_t<T, SPIN>(b, g, conv))) return rc;
  if (timed) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
  }
  if (nsteps > 0 && (rc = pull_norms(b, g))) return rc;
  for (int id : g.ids) b->res[id].steps += nsteps;
  return 0;
}

int advance(fgpe_batch* b, const std::vector<int>& list, long nsteps, bool conv, bool timed) {
  if (b->spin == 1)
    return b->fp32 ? advance_t<float, 1>(b, list, nsteps, conv, timed) : advance_t<double, 1>(b, list, nsteps, conv, timed);
  return b->fp32 ? advance_t<float, 2>(b, list, nsteps, conv, timed) : advance_t<double, 2>(b, list, nsteps, conv, timed);
}

std::vector<int> all_ids(const fgpe_batch* b) {
  std::vector<int> all(b->count);
  std::iota(all.begin(), all.end(), 0);
  return all;
}

// NORMT of every instance (as fgpe_normalize): total norm 1
template <typename T, int MS> int normt_t(fgpe_batch* b) {
  Groups g;
  int rc = upload_list(b, all_ids(b), g);
  if (rc) return rc;
  std::vector<double> part((size_t)g.total * GB_OBS * MS);
  for (int gr = 0; gr < NGRP; ++gr)
    if (g.cnt[gr]) {
      hipLaunchKernelGGL((kg_norm<T, MS>), dim3(g.cnt[gr] * GB_OBS), dim3(256), 0, b->stream,
                         (const GInst<T>*)b->ginst, b->dlist + g.start[gr], CLS_N[gx_of(gr)], CLS_N[gy_of(gr)],
                         b->obs + (size_t)g.start[gr] * GB_OBS * MS);
      HIPCHK(hipGetLastError());
    }
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(part.data(), b->obs, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<C<T>> sc((size_t)b->count * MS);
  for (int li = 0; li < g.total; ++li) {
    double t = 0;
    for (int m = 0; m < MS; ++m) {
      double s = 0;
      for (int k = 0; k < GB_OBS; ++k) s += part[((size_t)li * GB_OBS + k) * MS + m];
      t += s;
    }
    const T c = (T)(1.0 / std::sqrt(t));
    for (int m = 0; m < MS; ++m) {
      sc[(size_t)g.ids[li] * MS + m].x = c;
      sc[(size_t)g.ids[li] * MS + m].y = (T)0;
    }
  }
  HIPCHK(hipMemcpy(b->cscale, sc.data(), sc.size() * sizeof(C<T>), hipMemcpyHostToDevice));
  for (int gr = 0; gr < NGRP; ++gr)
    if (g.cnt[gr]) {
      const long np = (long)CLS_N[gx_of(gr)] * CLS_N[gy_of(gr)];
      const int nb = (int)std::min<long>((np + 255) / 256, 64);
      hipLaunchKernelGGL(kg_apply_scale<T>, dim3(g.cnt[gr] * nb), dim3(256), 0, b->stream, (const GInst<T>*)b->ginst,
                         b->dlist + g.start[gr], nb, np, MS);
      HIPCHK(hipGetLastError());
    }
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int normt(fgpe_batch* b) {
  if (b->spin == 1) return b->fp32 ? normt_t<float, 3>(b) : normt_t<double, 3>(b);
  return b->fp32 ? normt_t<float, 5>(b) : normt_t<double, 5>(b);
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int fgpe_batch_create(const fgpe_params* p, int32_t count, fgpe_batch** out) {
  if (!p || !out || count < 1) return fail(FGPE_EINVAL, "null argument or empty batch");
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    if (p[i].dim != 2) return fail(FGPE_EINVAL, "batch instance %d: only q2D instances are batched", i);
    if (p[i].spin != 1 && p[i].spin != 2) return fail(FGPE_EINVAL, "batch instance %d: spin %d", i, p[i].spin);
    if (p[i].spin != p[0].spin) return fail(FGPE_EINVAL, "batch instance %d: mixed spin", i);
    if (size_class(p[i].n[0]) < 0 || size_class(p[i].n[1]) < 0)
      return fail(FGPE_EINVAL, "batch instance %d: grid %dx%d (64, 128, 192 or 256 points per axis)", i, p[i].n[0],
                  p[i].n[1]);
    if (p[i].precision != p[0].precision) return fail(FGPE_EINVAL, "batch instance %d: mixed precision", i);
    if (p[i].imag != p[0].imag) return fail(FGPE_EINVAL, "batch instance %d: mixed real and imaginary time", i);
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
  b->spin = p[0].spin;
  b->ms = 2 * b->spin + 1;
  b->fp32 = p[0].precision == FGPE_FP32;
  b->p.assign(p, p + count);
  if (hipGetDevice(&b->dev) != hipSuccess) return bail(fail(FGPE_EDEVICE, "hipGetDevice failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->dev) == hipSuccess) b->lds_max = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipStreamCreate failed"));
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(FGPE_EDEVICE, "hipEventCreate failed"));
  b->gstream[0] = b->stream;
  for (int c = 0; c < 4; ++c) {
    // default priority, like the batch's own stream
    if ((c > 0 && hipStreamCreateWithFlags(&b->gstream[c], hipStreamNonBlocking) != hipSuccess) ||
        hipEventCreateWithFlags(&b->gev[c], hipEventDisableTiming) != hipSuccess)
      return bail(fail(FGPE_EDEVICE, "hipStreamCreate/hipEventCreate failed"));
  }
  const size_t es = b->fp32 ? sizeof(float2) : sizeof(double2);
  const int ms = b->ms;
  long total = 0;
  b->nx.resize(count);
  b->ny.resize(count);
  b->cx.resize(count);
  b->cy.resize(count);
  b->off.resize(count);
  for (int i = 0; i < count; ++i) {
    b->nx[i] = p[i].n[0];
    b->ny[i] = p[i].n[1];
    b->cx[i] = size_class(b->nx[i]);
    b->cy[i] = size_class(b->ny[i]);
    b->off[i] = total;
    total += b->npts(i) * ms;
    b->stage_elems = std::max(b->stage_elems, (size_t)b->npts(i) * ms);
  }
  int rc;
  if ((rc = balloc(b, &b->phi, total * es)) || (rc = balloc(b, &b->old, total * es)) ||
      (rc = balloc(b, &b->stage, b->stage_elems * sizeof(double2))) ||
      (rc = balloc(b, (void**)&b->norms, ms * sizeof(double) * count)) ||
      (rc = balloc(b, (void**)&b->scale64, ms * sizeof(double) * count)) ||
      (rc = balloc(b, &b->cscale, ms * es * count)) || (rc = balloc(b, (void**)&b->errs, sizeof(int) * count)) ||
      (rc = balloc(b, (void**)&b->obs, sizeof(double) * (size_t)count * GB_OBS * (b->ncol() + 1))) ||
      (rc = balloc(b, (void**)&b->dlist, sizeof(int) * count)))
    return bail(rc);
  HIPCHK(hipMemset(b->phi, 0, total * es));
  HIPCHK(hipMemset(b->old, 0, total * es));
  HIPCHK(hipMemset(b->norms, 0, ms * sizeof(double) * count));
  HIPCHK(hipMemset(b->errs, 0, sizeof(int) * count));
  b->res.assign(count, fgpe_batch_result{});
  b->en_old.assign(count, 0.0);
  rc = b->fp32 ? build_batch<float>(b) : build_batch<double>(b);
  if (rc) return bail(rc);
  *out = b;
  return 0;
}

void fgpe_batch_destroy(fgpe_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (void* a : b->allocs) (void)hipFree(a);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (int c = 0; c < 4; ++c) {
    if (c > 0 && b->gstream[c]) {
      (void)hipStreamSynchronize(b->gstream[c]);
      (void)hipStreamDestroy(b->gstream[c]);
    }
    if (b->gev[c]) (void)hipEventDestroy(b->gev[c]);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int fgpe_batch_initial_guess(fgpe_batch* b) {
  if (!b) return fail(FGPE_EINVAL, "null batch");
  for (int i = 0; i < b->count; ++i) {
    const fgpe_params& p = b->p[i];
    GuessArgs a;
    a.dim = 2;
    a.spin = b->spin;
    a.fpc = p.option_fpc;
    a.ax = p.alpha[0];
    a.ay = p.alpha[1];
    a.az = p.alpha[2];
    a.t0 = p.tau[0];
    a.t1 = p.tau[1];
    a.t2 = p.tau[2];
    a.mag = p.mag;
    const Geom g{b->nx[i], b->ny[i], 1, b->npts(i)};
    const int blocks = std::min(grid_for(g.npts, 256), 1024);
    if (b->fp32) {
      const GInst<float>& I = host_inst<float>(b)[i];
      a.x2 = I.oa.x2;
      a.y2 = I.oa.y2;
      a.z2 = I.oa.z2;
      hipLaunchKernelGGL(k_guess<float>, dim3(blocks), dim3(256), 0, b->stream, I.phi, g, a);
    } else {
      const GInst<double>& I = host_inst<double>(b)[i];
      a.x2 = I.oa.x2;
      a.y2 = I.oa.y2;
      a.z2 = I.oa.z2;
      hipLaunchKernelGGL(k_guess<double>, dim3(blocks), dim3(256), 0, b->stream, I.phi, g, a);
    }
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  int rc = normt(b);
  if (rc) return rc;
  for (auto& r : b->res) r = fgpe_batch_result{};
  if ((rc = advance(b, all_ids(b), 0, false, false))) return rc;
  for (int i = 0; i < b->count; ++i) b->en_old[i] = b->res[i].en;
  return 0;
}

int fgpe_batch_normalize(fgpe_batch* b) {
  if (!b) return fail(FGPE_EINVAL, "null batch");
  int rc = normt(b);
  if (rc) return rc;
  for (auto& r : b->res) r = fgpe_batch_result{};
  if ((rc = advance(b, all_ids(b), 0, false, false))) return rc;
  for (int i = 0; i < b->count; ++i) b->en_old[i] = b->res[i].en;
  return 0;
}

int fgpe_batch_upload(fgpe_batch* b, int32_t i, const double* phi) {
  if (!b || !phi || i < 0 || i >= b->count) return fail(FGPE_EINVAL, "bad batch upload argument");
  const long n = b->npts(i) * b->ms;
  HIPCHK(hipStreamSynchronize(b->stream));
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
  const long n = b->npts(i) * b->ms;
  HIPCHK(hipStreamSynchronize(b->stream));
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
  return advance(b, all_ids(b), (long)nsteps, nsteps >= 1, true);
}

int fgpe_batch_solve(fgpe_batch* b, int32_t stp, double tol, int64_t max_steps, int32_t* converged) {
  if (!b || stp < 1 || max_steps < 0) return fail(FGPE_EINVAL, "bad batch solve argument");
  double total_ms = 0;
  int rc;
  for (;;) {
    std::vector<int> list;
    for (int id = 0; id < b->count; ++id)
      if (!b->res[id].done && b->res[id].steps + stp <= max_steps) list.push_back(id);
    if (list.empty()) break;
    if ((rc = advance(b, list, stp, true, true))) return rc;
    total_ms += b->last_ms;
    for (int id : list) {
      fgpe_batch_result& r = b->res[id];
      const double den = stp * 2.0 * b->p[id].dt;
      r.conv = r.maxdiff / den;
      r.conv_en = std::fabs(r.en - b->en_old[id]) / den;
      b->en_old[id] = r.en;
      if (r.conv <= tol || r.conv_en <= tol || r.err) r.done = 1;
    }
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
  for (int i = 0; i < b->count; ++i) total += b->npts(i) * b->ms;
  // trials advance the state: keep a copy and restore it afterwards
  void* backup = nullptr;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMalloc(&backup, total * es));
  auto done = [&](int rc) {
    (void)hipFree(backup);
    return rc;
  };
  if (hipMemcpy(backup, b->phi, total * es, hipMemcpyDeviceToDevice) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "backup copy failed"));
  std::vector<fgpe_batch_result> saved = b->res;
  double ms[FGPE_BATCH_NSCHED] = {0};
  int best[NCLS] = {-1, -1, -1, -1};
  int rc;
  // one line length at a time: its two schedules timed on the instances of
  // its square group, or, without one, on every instance with lines of that
  // length (the other length keeping its current schedule)
  for (int c = 0; c < NCLS; ++c) {
    std::vector<int> sq, any;
    for (int i = 0; i < b->count; ++i) {
      if (b->cx[i] == c && b->cy[i] == c) sq.push_back(i);
      if (b->cx[i] == c || b->cy[i] == c) any.push_back(i);
    }
    const std::vector<int>& list = sq.empty() ? any : sq;
    if (list.empty()) continue;
    double best_ms = 1e300;
    for (int s = 2 * c; s < 2 * c + 2; ++s) {
      b->csched[c] = s;
      if ((rc = b->fp32 ? upload_inst<float>(b) : upload_inst<double>(b))) return done(rc);
      // one untimed step to warm up, then the timed trial
      if ((rc = advance(b, list, 1, false, false)) || (rc = advance(b, list, trial_steps, false, true))) return done(rc);
      ms[s] = b->last_ms;
      if (b->last_ms < best_ms) {
        best_ms = b->last_ms;
        best[c] = s;
      }
    }
    b->csched[c] = best[c];
  }
  if ((rc = b->fp32 ? upload_inst<float>(b) : upload_inst<double>(b))) return done(rc);
  HIPCHK(hipStreamSynchronize(b->stream));
  if (hipMemcpy(b->phi, backup, total * es, hipMemcpyDeviceToDevice) != hipSuccess)
    return done(fail(FGPE_EDEVICE, "restore copy failed"));
  b->res = saved;
  if (ms_out)
    for (int s = 0; s < FGPE_BATCH_NSCHED; ++s) ms_out[s] = ms[s];
  if (chosen)
    for (int c = 0; c < NCLS; ++c) chosen[c] = best[c];
  return done(0);
}

}  // extern "C"

// fgpe_launch.hpp -- the pass launches of libfgpe for one line length N:
// each checks its LDS budget and launches its kernel.  fgpe_lines.hip
// instantiates them once per (precision, spin, N), so every line length is
// its own translation unit and they compile in parallel; fgpe_ops.hpp only
// declares them and dispatches on the runtime grid size.
#pragma once
#include "fgpe_impl.hpp"

namespace fgpe_impl {

template <typename K> void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// ------------------------------------------------------------- launches
// Pass A over the z-planes [z0, z0 + zc) on stream st, Q points per thread.
// The chunk's lines fill whole blocks (chunk_planes), so each block keeps its
// whole-grid index and NORMC partial slot.  FULL: q2D (the pass also applies
// 1/N and the kx factor; in 3D passes B and D do).
template <typename T, int SPIN, int N, int Q, bool FULL>
int launch_x_q(fgpe_handle* h, int flags, hipStream_t st, int z0, int zc) {
  constexpr int P = N / Q;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  const int lpw = x_pass_lpw<T>(h->dim, N);
  const int threads = lpw * P;
  const long lines0 = (long)z0 * h->ny, lines = (long)zc * h->ny;
  const int blocks = (int)x_pass_blocks<T>(h->dim, N, lines);
  const int blk0 = (int)(lines0 / lpw);
  // q2D (FULL) transforms the MS components' lines together: MS LDS lines per x line
  constexpr bool MULTI = FULL;
  const size_t lds = (TWS + N) * sizeof(R_t<T>) + 16 * MS * sizeof(double) +
                     (size_t)lpw * (MULTI ? MS : 1) * line_stride(N) * sizeof(C<T>);
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "x pass needs %zu B of LDS", lds);
  if (flags & XF_PHYS) {
    int rc = ensure_partials(h, (size_t)x_pass_blocks<T>(h->dim, N, (long)h->ny * h->nz) * MS);
    if (rc) return rc;
  }
  auto kern = k_xpass<T, SPIN, N, Q, FULL, MULTI>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, st, (C<T>*)h->phi, geom(h), tables<T>(h), phys(h),
                     flags, h->partials, blk0);
  HIPCHK(hipGetLastError());
  prof_end(h, (flags & XF_PHYS) ? FGPE_PASS_X : FGPE_PASS_X0);
  return 0;
}
template <typename T, int SPIN, int N> int launch_x(fgpe_handle* h, int flags, hipStream_t st, int z0, int zc) {
  if (h->dim == 3) return launch_x_q<T, SPIN, N, q_for(N, Cfg<T>::QX), false>(h, flags, st, z0, zc);
  return launch_x_q<T, SPIN, N, q_for(N, Cfg<T>::QX2), true>(h, flags, st, z0, zc);
}

// NORMC partials of a whole pass A -> norms and the pending renormalisation
template <typename T, int SPIN, int N> int launch_fin(fgpe_handle* h) {
  constexpr int MS = 2 * SPIN + 1;
  const int blocks = (int)x_pass_blocks<T>(h->dim, N, (long)h->ny * h->nz);
  prof_begin(h);
  hipLaunchKernelGGL((k_finalize<T, MS>), dim3(1), dim3(fin_threads(blocks)), 0, h->stream, (const double*)h->partials, blocks,
                     finargs(h), h->dnorm, (C<T>*)h->cscale, h->cscale64, h->dsig, h->derr);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_FIN);
  return 0;
}

// y lines of one component (3D pass B) over x in [x0, x0 + xw) and z in
// [z0, z0 + zc) on stream st.
template <typename T, int N>
int launch_y(fgpe_handle* h, int dir, int scale, hipStream_t st, int x0, int xw, int z0, int zc) {
  constexpr int Q = y_pass_q<T>(N), P = N / Q, TX = Cfg<T>::TX;
  static_assert(TX * P <= 1024, "pass B block shape");
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  const size_t lds = TWS * sizeof(R_t<T>) + (size_t)TX * line_stride(N) * sizeof(C<T>);
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "y pass needs %zu B of LDS", lds);
  auto kern = k_ypass<T, N, Q, TX>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(xw / TX, zc, h->ms), dim3(TX * P), lds, st, (C<T>*)h->phi, geom(h), tables<T>(h),
                     dir, scale, 1.0 / (double)h->npts, x0, z0);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_Y);
  return 0;
}

// 2D k pass (y lines, KE, SOC), all of the grid.
template <typename T, int SPIN, int N> int launch_k(fgpe_handle* h, int scale) {
  constexpr int Q = q_for(N, Cfg<T>::QK2), P = N / Q, TX = Cfg<T>::TXK2;
  static_assert(TX * P <= 1024, "2D k pass block shape");
  // the components' lines transformed together where their LDS lines fit in
  // 128 KB
  constexpr bool MULTI = kpass_lds_bytes<T, SPIN, N, Q, TX, true>() <= 131072;
  const size_t lds = kpass_lds_bytes<T, SPIN, N, Q, TX, MULTI>();
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "k pass needs %zu B of LDS", lds);
  auto kern = k_kpass<T, SPIN, N, Q, TX, MULTI>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(h->nx / TX), dim3(TX * P), lds, h->stream, (C<T>*)h->phi, geom(h), tables<T>(h),
                     phys(h), scale);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_K);
  return 0;
}

// 3D pass C (z lines, KE) over x in [x0, x0 + xw) on stream st.
template <typename T, int N> int launch_z(fgpe_handle* h, int scale, hipStream_t st, int x0, int xw) {
  constexpr int Q = z_pass_q<T>(N), P = N / Q, TX = Cfg<T>::TX;
  static_assert(TX * P <= 1024, "pass C block shape");
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  const size_t lds = TWS * sizeof(R_t<T>) + (size_t)TX * line_stride(N) * sizeof(C<T>);
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "z pass needs %zu B of LDS", lds);
  auto kern = k_zpass<T, N, Q, TX>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(xw / TX, h->ny, h->ms), dim3(TX * P), lds, st, (C<T>*)h->phi, geom(h), tables<T>(h),
                     scale, x0);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_Z);
  return 0;
}

// 3D pass D (SOC, inverse y, 1/N) over x in [x0, x0 + xw) on stream st.
template <typename T, int SPIN, int N> int launch_d(fgpe_handle* h, hipStream_t st, int x0, int xw, int scale) {
  // columns per block: TXK, fewer where P = N/Q threads per line would make
  // the block exceed its thread limit (FP64: 512, so the five components'
  // lines keep 256 VGPRs; FP32: 1024)
  constexpr int Q = q_for(N, Cfg<T>::QK), P = N / Q;
  constexpr int MAXT = sizeof(T) == 8 ? 512 : 1024;
  constexpr int TX = Cfg<T>::TXK * P <= MAXT ? Cfg<T>::TXK : MAXT / P;
  static_assert(TX * P <= MAXT && TX >= 8, "pass D block shape");
  const size_t lds = dpass_lds_bytes<T, SPIN, N, Q, TX>();
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "d pass needs %zu B of LDS", lds);
  // scale: multiply by the pending renormalisation on load (a separate
  // instantiation, so the kernel without it keeps its registers)
  auto kern = scale ? k_dpass<T, SPIN, N, Q, TX, true> : k_dpass<T, SPIN, N, Q, TX, false>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(xw / TX, h->nz), dim3(TX * P), lds, st, (C<T>*)h->phi, geom(h), tables<T>(h), phys(h),
                     1.0 / (double)h->npts, x0 / TX, 0);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_K);
  return 0;
}

template <typename T, int SPIN, int N> int launch_1d(fgpe_handle* h, long nsteps) {
  constexpr int Q = q_for(N, Cfg<T>::Q1), P = N / Q;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  const size_t lds = TWS * sizeof(R_t<T>) + 16 * MS * sizeof(double) + 8 * sizeof(double) + line_stride(N) * sizeof(C<T>);
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "1D step needs %zu B of LDS", lds);
  auto kern = k_line1d<T, SPIN, N, Q>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(1), dim3(P), lds, h->stream, (C<T>*)h->phi, tables<T>(h), phys(h), finargs(h),
                     nsteps, h->dnorm, h->dsig, h->derr);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_1D);
  return 0;
}


// One axis of the 2D/3D observables (k_obs_line) for lines of N points;
// returns the number of blocks launched, whose sums start at partials + off
// (the caller has sized the partials for every axis).
template <typename T, int SPIN, int DIM, int AXIS, int N>
int launch_obs_axis(fgpe_handle* h, const ObsArgs& a, size_t off, int* nblk) {
  const size_t lds = obs_line_lds<T, SPIN, N>();
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "observables need %zu B of LDS", lds);
  const long blocks = obs_axis_blocks<T, SPIN>(h, AXIS);
  auto kern = k_obs_line<T, SPIN, DIM, AXIS, N>;
  allow_lds(kern, lds);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(OBS_TPB), lds, h->stream, (const C<T>*)h->phi, geom(h), a,
                     obs_axis_tiles<T, SPIN>(h, AXIS), h->partials + off);
  HIPCHK(hipGetLastError());
  *nblk = (int)blocks;
  return 0;
}
}  // namespace fgpe_impl

// The 2D/3D passes of one (T, SPIN, N); the y and z passes do not depend on
// the spin and are instantiated with spin 1.
#define FGPE_INSTANTIATE_LINES(T, SPIN, N)                                                             \
  template int fgpe_impl::launch_x<T, SPIN, N>(fgpe_handle*, int, hipStream_t, int, int);              \
  template int fgpe_impl::launch_fin<T, SPIN, N>(fgpe_handle*);                                        \
  template int fgpe_impl::launch_k<T, SPIN, N>(fgpe_handle*, int);                                     \
  template int fgpe_impl::launch_d<T, SPIN, N>(fgpe_handle*, hipStream_t, int, int, int);              \
  template int fgpe_impl::launch_obs_axis<T, SPIN, 2, 0, N>(fgpe_handle*, const ObsArgs&, size_t, int*); \
  template int fgpe_impl::launch_obs_axis<T, SPIN, 2, 1, N>(fgpe_handle*, const ObsArgs&, size_t, int*); \
  template int fgpe_impl::launch_obs_axis<T, SPIN, 3, 0, N>(fgpe_handle*, const ObsArgs&, size_t, int*); \
  template int fgpe_impl::launch_obs_axis<T, SPIN, 3, 1, N>(fgpe_handle*, const ObsArgs&, size_t, int*); \
  template int fgpe_impl::launch_obs_axis<T, SPIN, 3, 2, N>(fgpe_handle*, const ObsArgs&, size_t, int*);
#define FGPE_INSTANTIATE_YZ(T, N)                                                                      \
  template int fgpe_impl::launch_y<T, N>(fgpe_handle*, int, int, hipStream_t, int, int, int, int);      \
  template int fgpe_impl::launch_z<T, N>(fgpe_handle*, int, hipStream_t, int, int);
#define FGPE_INSTANTIATE_1D(T, SPIN, N) template int fgpe_impl::launch_1d<T, SPIN, N>(fgpe_handle*, long);

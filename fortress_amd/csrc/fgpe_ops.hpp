// fgpe_ops.hpp -- templated launch code of libfgpe: pass launches with their
// LDS budgets, runtime size dispatch, the per-run pass schedule and the
// event-time reductions.  Instantiated per precision and spin in
// fgpe_ops_*.hip so the kernels compile in parallel.
#pragma once
#include "fgpe_impl.hpp"

namespace fgpe_impl {

template <typename K> void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// ------------------------------------------------------------- launches
// Pass A over the z-planes [z0, z0 + zc) on stream st.  The chunk's lines
// fill whole blocks (chunk_planes), so each block keeps its whole-grid index
// and NORMC partial slot.
template <typename T, int SPIN, int N> int launch_x(fgpe_handle* h, int flags, hipStream_t st, int z0, int zc) {
  constexpr int Q = Cfg<T>::QX, P = N / Q;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int TWS = twiddle_size(N, Q) > 0 ? twiddle_size(N, Q) : 1;
  const int threads = P >= 256 ? P : 256;
  const int lpw = threads / P;
  const long lines0 = (long)z0 * h->ny, lines = (long)zc * h->ny;
  const int blocks = (int)x_pass_blocks<T>(N, lines);
  const int blk0 = (int)(lines0 / lpw);
  const size_t lds = (TWS + N) * sizeof(R_t<T>) + 16 * MS * sizeof(double) + (size_t)lpw * line_stride(N) * sizeof(C<T>);
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "x pass needs %zu B of LDS", lds);
  if (flags & XF_PHYS) {
    int rc = ensure_partials(h, (size_t)x_pass_blocks<T>(N, (long)h->ny * h->nz) * MS);
    if (rc) return rc;
  }
  // 3D: passes B and D apply the kx factor and 1/N
  auto kern = h->dim == 3 ? k_xpass<T, SPIN, N, Q, false> : k_xpass<T, SPIN, N, Q, true>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, st, (C<T>*)h->phi, geom(h), tables<T>(h), phys(h),
                     flags, h->partials, blk0);
  HIPCHK(hipGetLastError());
  prof_end(h, (flags & XF_PHYS) ? FGPE_PASS_X : FGPE_PASS_X0);
  return 0;
}

// NORMC partials of a whole pass A -> norms and the pending renormalisation
template <typename T, int SPIN, int N> int launch_fin(fgpe_handle* h) {
  constexpr int MS = 2 * SPIN + 1;
  const int blocks = (int)x_pass_blocks<T>(N, (long)h->ny * h->nz);
  prof_begin(h);
  hipLaunchKernelGGL((k_finalize<T, MS>), dim3(1), dim3(1024), 0, h->stream, (const double*)h->partials, blocks,
                     finargs(h), h->dnorm, (C<T>*)h->cscale, h->cscale64, h->dsig, h->derr);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_FIN);
  return 0;
}

// y lines of one component (3D pass B) over x in [x0, x0 + xw) and z in
// [z0, z0 + zc) on stream st.
template <typename T, int N>
int launch_y(fgpe_handle* h, int dir, int scale, hipStream_t st, int x0, int xw, int z0, int zc) {
  constexpr int Q = Cfg<T>::QY, P = N / Q, TX = Cfg<T>::TX;
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
  constexpr int Q = Cfg<T>::QK, P = N / Q, TX = Cfg<T>::TXK;
  const size_t lds = kpass_lds_bytes<T, SPIN, N, Q, TX>();
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "k pass needs %zu B of LDS", lds);
  auto kern = k_kpass<T, SPIN, N, Q, TX>;
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
  constexpr int Q = Cfg<T>::QZ, P = N / Q, TX = Cfg<T>::TX;
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
template <typename T, int SPIN, int N> int launch_d(fgpe_handle* h, hipStream_t st, int x0, int xw) {
  constexpr int Q = Cfg<T>::QK, P = N / Q, TX = Cfg<T>::TXK;
  const size_t lds = dpass_lds_bytes<T, SPIN, N, Q, TX>();
  if (lds > h->lds_max) return fail(FGPE_EINVAL, "d pass needs %zu B of LDS", lds);
  auto kern = k_dpass<T, SPIN, N, Q, TX>;
  allow_lds(kern, lds);
  prof_begin(h);
  hipLaunchKernelGGL(kern, dim3(xw / TX, h->nz), dim3(TX * P), lds, st, (C<T>*)h->phi, geom(h), tables<T>(h), phys(h),
                     1.0 / (double)h->npts, x0 / TX, 0);
  HIPCHK(hipGetLastError());
  prof_end(h, FGPE_PASS_K);
  return 0;
}

template <typename T, int SPIN, int N> int launch_1d(fgpe_handle* h, long nsteps) {
  constexpr int Q = Cfg<T>::Q1, P = N / Q;
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

// runtime size dispatch
template <typename T, int SPIN> int dispatch_x(fgpe_handle* h, int flags, hipStream_t st, int z0, int zc) {
  switch (h->nx) {
    case 32: return launch_x<T, SPIN, 32>(h, flags, st, z0, zc);
    case 64: return launch_x<T, SPIN, 64>(h, flags, st, z0, zc);
    case 128: return launch_x<T, SPIN, 128>(h, flags, st, z0, zc);
    case 256: return launch_x<T, SPIN, 256>(h, flags, st, z0, zc);
    case 512: return launch_x<T, SPIN, 512>(h, flags, st, z0, zc);
  }
  return fail(FGPE_EINVAL, "unsupported NX=%d", h->nx);
}
template <typename T, int SPIN> int dispatch_x(fgpe_handle* h, int flags) {
  return dispatch_x<T, SPIN>(h, flags, h->stream, 0, h->nz);
}
template <typename T, int SPIN> int dispatch_fin(fgpe_handle* h) {
  switch (h->nx) {
    case 32: return launch_fin<T, SPIN, 32>(h);
    case 64: return launch_fin<T, SPIN, 64>(h);
    case 128: return launch_fin<T, SPIN, 128>(h);
    case 256: return launch_fin<T, SPIN, 256>(h);
    case 512: return launch_fin<T, SPIN, 512>(h);
  }
  return fail(FGPE_EINVAL, "unsupported NX=%d", h->nx);
}
template <typename T>
int dispatch_y(fgpe_handle* h, int dir, int scale, hipStream_t st, int x0, int xw, int z0, int zc) {
  switch (h->ny) {
    case 32: return launch_y<T, 32>(h, dir, scale, st, x0, xw, z0, zc);
    case 64: return launch_y<T, 64>(h, dir, scale, st, x0, xw, z0, zc);
    case 128: return launch_y<T, 128>(h, dir, scale, st, x0, xw, z0, zc);
    case 256: return launch_y<T, 256>(h, dir, scale, st, x0, xw, z0, zc);
    case 512: return launch_y<T, 512>(h, dir, scale, st, x0, xw, z0, zc);
  }
  return fail(FGPE_EINVAL, "unsupported NY=%d", h->ny);
}
template <typename T, int SPIN> int dispatch_k(fgpe_handle* h, int scale) {
  switch (h->ny) {
    case 32: return launch_k<T, SPIN, 32>(h, scale);
    case 64: return launch_k<T, SPIN, 64>(h, scale);
    case 128: return launch_k<T, SPIN, 128>(h, scale);
    case 256: return launch_k<T, SPIN, 256>(h, scale);
    case 512: return launch_k<T, SPIN, 512>(h, scale);
  }
  return fail(FGPE_EINVAL, "unsupported NY=%d", h->ny);
}
template <typename T> int dispatch_z(fgpe_handle* h, int scale, hipStream_t st, int x0, int xw) {
  switch (h->nz) {
    case 32: return launch_z<T, 32>(h, scale, st, x0, xw);
    case 64: return launch_z<T, 64>(h, scale, st, x0, xw);
    case 128: return launch_z<T, 128>(h, scale, st, x0, xw);
    case 256: return launch_z<T, 256>(h, scale, st, x0, xw);
    case 512: return launch_z<T, 512>(h, scale, st, x0, xw);
  }
  return fail(FGPE_EINVAL, "unsupported NZ=%d", h->nz);
}
template <typename T, int SPIN> int dispatch_d(fgpe_handle* h, hipStream_t st, int x0, int xw) {
  switch (h->ny) {
    case 32: return launch_d<T, SPIN, 32>(h, st, x0, xw);
    case 64: return launch_d<T, SPIN, 64>(h, st, x0, xw);
    case 128: return launch_d<T, SPIN, 128>(h, st, x0, xw);
    case 256: return launch_d<T, SPIN, 256>(h, st, x0, xw);
    case 512: return launch_d<T, SPIN, 512>(h, st, x0, xw);
  }
  return fail(FGPE_EINVAL, "unsupported NY=%d", h->ny);
}
template <typename T, int SPIN> int dispatch_1d(fgpe_handle* h, long nsteps) {
  switch (h->nx) {
    case 64: return launch_1d<T, SPIN, 64>(h, nsteps);
    case 128: return launch_1d<T, SPIN, 128>(h, nsteps);
    case 256: return launch_1d<T, SPIN, 256>(h, nsteps);
    case 512: return launch_1d<T, SPIN, 512>(h, nsteps);
    case 1024: return launch_1d<T, SPIN, 1024>(h, nsteps);
    case 2048: return launch_1d<T, SPIN, 2048>(h, nsteps);
    case 4096: return launch_1d<T, SPIN, 4096>(h, nsteps);
  }
  return fail(FGPE_EINVAL, "unsupported 1D NX=%d", h->nx);
}

// 3D step schedule, two phases per step, each a pipeline over chunks:
//   phase 1, by z-chunks: A (inverse-x, SE, SP, NORMC, forward-x) then the
//     next step's B (forward-y, kx and ky factors of KE) -- both act within
//     z-planes;
//   phase 2, by x-chunks: C (forward-z, kz/GAMMAZ factor of KE, inverse-z,
//     one component per block), applying the pending renormalisation, then D
//     (SOC in (kx, ky, z) space, inverse-y, 1/N, all components) -- both act
//     within x-slabs.
// The SO coupling depends on (kx, ky) only (SRC/cgpe3d.f90:904-1004), so it
// commutes with the z transforms and is applied after the inverse z FFT,
// still after the whole KE factor as in the reference's order: pass C then
// moves one component at a time and D transforms once per component.
// Chunks go round-robin to up to three streams, so the VALU-bound A overlaps
// the memory-bound B of neighbouring chunks, C the heavier D; the phases are
// joined on the handle's stream (C needs every z of B, A every x of D).  The
// renormalisation (uniform with SO coupling, per component without it, where
// no pass mixes components) commutes with the FFTs, KE and SOC, so C applies
// it once NORMC of every chunk of A is in.  Every chunking gives
// bit-identical results: each line is transformed by the same code and each
// pass-A block keeps its NORMC partial slot.
template <typename T, int SPIN> int run_steps_3d(fgpe_handle* h, long nsteps) {
  const int zc = chunk_planes<T>(h), xc = chunk_cols<T>(h);
  const int nst = h->profiling ? 1 : std::max(1, std::min(h->nstreams, 3));
  hipStream_t st[3] = {h->stream, h->aux[0], h->aux[1]};
  auto fork = [&]() -> int {
    if (nst == 1) return 0;
    HIPCHK(hipEventRecord(h->evs[0], h->stream));
    for (int i = 1; i < nst; ++i) HIPCHK(hipStreamWaitEvent(st[i], h->evs[0], 0));
    return 0;
  };
  auto join = [&]() -> int {
    for (int i = 1; i < nst; ++i) {
      HIPCHK(hipEventRecord(h->evs[i], st[i]));
      HIPCHK(hipStreamWaitEvent(h->stream, h->evs[i], 0));
    }
    return 0;
  };
  auto phase1 = [&](int flags, bool fwd_y) -> int {
    int rc;
    if ((rc = fork())) return rc;
    for (int z0 = 0, c = 0; z0 < h->nz; z0 += zc, ++c) {
      if ((rc = dispatch_x<T, SPIN>(h, flags, st[c % nst], z0, zc))) return rc;
      if (fwd_y && (rc = dispatch_y<T>(h, -1, 0, st[c % nst], 0, h->nx, z0, zc))) return rc;
    }
    return join();
  };
  auto phase2 = [&]() -> int {
    int rc;
    if ((rc = fork())) return rc;
    for (int x0 = 0, c = 0; x0 < h->nx; x0 += xc, ++c) {
      if ((rc = dispatch_z<T>(h, 1, st[c % nst], x0, xc))) return rc;
      if ((rc = dispatch_d<T, SPIN>(h, st[c % nst], x0, xc))) return rc;
    }
    return join();
  };
  int rc;
  if ((rc = phase1(XF_FWD, true))) return rc;  // the run's first forward-x and forward-y
  for (long s = 1; s <= nsteps; ++s) {
    if ((rc = phase2())) return rc;
    // NORMC every step in imaginary time (it sets the renormalisation); in
    // real time the scale is 1 and only the last step's norms are reported
    const int flags = XF_INV | XF_PHYS | (s < nsteps ? XF_FWD : 0) | ((h->p.imag || s == nsteps) ? XF_NORM : 0);
    if ((rc = phase1(flags, s < nsteps))) return rc;
    if ((flags & XF_NORM) && (rc = dispatch_fin<T, SPIN>(h))) return rc;
  }
  return 0;
}

template <typename T, int SPIN> int run_steps(fgpe_handle* h, long nsteps) {
  if (h->dim == 1) {
    long done = 0;
    while (done < nsteps) {
      long n = std::min<long>(nsteps - done, 4096);
      int rc = dispatch_1d<T, SPIN>(h, n);
      if (rc) return rc;
      done += n;
    }
    return 0;
  }
  int rc;
  hipLaunchKernelGGL(k_set_scale_one<T>, dim3(1), dim3(64), 0, h->stream, (C<T>*)h->cscale, h->ms);
  if (h->dim == 3) {
    if ((rc = run_steps_3d<T, SPIN>(h, nsteps))) return rc;
  } else {
    if ((rc = dispatch_x<T, SPIN>(h, XF_FWD))) return rc;
    for (long s = 1; s <= nsteps; ++s) {
      if ((rc = dispatch_k<T, SPIN>(h, 1))) return rc;
      // NORMC every step in imaginary time (it sets the renormalisation); in
      // real time the scale is 1 and only the last step's norms are reported
      const int flags = XF_INV | XF_PHYS | (s < nsteps ? XF_FWD : 0) | ((h->p.imag || s == nsteps) ? XF_NORM : 0);
      if ((rc = dispatch_x<T, SPIN>(h, flags))) return rc;
      if ((flags & XF_NORM) && (rc = dispatch_fin<T, SPIN>(h))) return rc;
    }
  }
  if (h->p.imag) {
    hipLaunchKernelGGL(k_apply_scale<T>, dim3(grid_for(h->npts * h->ms, 256)), dim3(256), 0, h->stream,
                       (C<T>*)h->phi, h->npts, h->ms, (const C<T>*)h->cscale);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

template <typename T> int normc_t(fgpe_handle* h, double* norm) {
  const int blocks = grid_for(h->npts, 256) > 2048 ? 2048 : grid_for(h->npts, 256);
  int rc = ensure_partials(h, (size_t)blocks * h->ms);
  if (rc) return rc;
  if (h->ms == 3)
    hipLaunchKernelGGL((k_norm_reduce<T, 3>), dim3(blocks), dim3(256), 0, h->stream, (const C<T>*)h->phi, geom(h),
                       h->wx, h->wy, h->wz, h->partials);
  else
    hipLaunchKernelGGL((k_norm_reduce<T, 5>), dim3(blocks), dim3(256), 0, h->stream, (const C<T>*)h->phi, geom(h),
                       h->wx, h->wy, h->wz, h->partials);
  HIPCHK(hipGetLastError());
  std::vector<double> part((size_t)blocks * h->ms);
  HIPCHK(hipMemcpyAsync(part.data(), h->partials, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int m = 0; m < h->ms; ++m) {
    double s = 0;
    for (int b = 0; b < blocks; ++b) s += part[(size_t)b * h->ms + m];
    norm[m] = s;
  }
  return 0;
}

template <typename T> int normalize_t(fgpe_handle* h, double* tnorm) {
  double n[5];
  int rc = normc_t<T>(h, n);
  if (rc) return rc;
  double t = 0;
  for (int m = 0; m < h->ms; ++m) t += n[m];
  t = std::sqrt(t);
  std::vector<C<T>> c(h->ms);
  for (int m = 0; m < h->ms; ++m) {
    c[m].x = (T)(1.0 / t);
    c[m].y = (T)0;
  }
  HIPCHK(hipMemcpyAsync(h->cscale, c.data(), c.size() * sizeof(C<T>), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_apply_scale<T>, dim3(grid_for(h->npts * h->ms, 256)), dim3(256), 0, h->stream, (C<T>*)h->phi,
                     h->npts, h->ms, (const C<T>*)h->cscale);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  if (tnorm) *tnorm = t;
  return 0;
}

// Tiles of the observables line pass along `axis` (lines per tile from the
// axis' line length), and the blocks that loop over them.
template <typename T, int SPIN> long obs_axis_tiles(const fgpe_handle* h, int axis) {
  const int n = axis == 0 ? h->nx : axis == 1 ? h->ny : h->nz;
  const int L = n == 32    ? obs_lines<T, SPIN, 32>()
                : n == 64  ? obs_lines<T, SPIN, 64>()
                : n == 128 ? obs_lines<T, SPIN, 128>()
                : n == 256 ? obs_lines<T, SPIN, 256>()
                           : obs_lines<T, SPIN, 512>();
  const long nbx = (h->nx + L - 1) / L;
  return axis == 0 ? ((long)h->ny * h->nz + L - 1) / L : axis == 1 ? nbx * h->nz : nbx * h->ny;
}
template <typename T, int SPIN> long obs_axis_blocks(const fgpe_handle* h, int axis) {
  return std::min(obs_axis_tiles<T, SPIN>(h, axis), OBS_MAX_BLOCKS);
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
template <typename T, int SPIN, int DIM, int AXIS>
int dispatch_obs_axis(fgpe_handle* h, const ObsArgs& a, size_t off, int* nblk) {
  const int n = AXIS == 0 ? h->nx : AXIS == 1 ? h->ny : h->nz;
  switch (n) {
    case 32: return launch_obs_axis<T, SPIN, DIM, AXIS, 32>(h, a, off, nblk);
    case 64: return launch_obs_axis<T, SPIN, DIM, AXIS, 64>(h, a, off, nblk);
    case 128: return launch_obs_axis<T, SPIN, DIM, AXIS, 128>(h, a, off, nblk);
    case 256: return launch_obs_axis<T, SPIN, DIM, AXIS, 256>(h, a, off, nblk);
    case 512: return launch_obs_axis<T, SPIN, DIM, AXIS, 512>(h, a, off, nblk);
  }
  return fail(FGPE_EINVAL, "unsupported line length %d", n);
}

template <typename T, int SPIN> int observables_t(fgpe_handle* h, double* en, double* mu, double* mz, double* rms) {
  int blocks = std::min(grid_for(h->npts, 256), 4096);
  const int nc = 3 * h->ms + 2;
  long need = blocks;
  if (h->dim >= 2) {
    need = 0;
    for (int ax = 0; ax < h->dim; ++ax) need += obs_axis_blocks<T, SPIN>(h, ax);
  }
  int rc = ensure_partials(h, (size_t)need * nc);
  if (rc) return rc;
  ObsArgs a;
  a.tau0 = h->p.tau[0];
  a.tau1 = h->p.tau[1];
  a.tau2 = h->p.tau[2];
  a.gx = h->p.gamma[0];
  a.gy = h->p.gamma[1];
  a.gz = h->p.gamma[2];
  a.dx = h->p.d[0];
  a.dy = h->p.d[1];
  a.dz = h->p.d[2];
  a.wx = h->wx;
  a.wy = h->wy;
  a.wz = h->wz;
  a.x2 = h->x2;
  a.y2 = h->y2;
  a.z2 = h->z2;
  a.vxd = h->vxd;
  a.vyd = h->vyd;
  a.vzd = h->vzd;
  const C<T>* ph = (const C<T>*)h->phi;
  Geom g = geom(h);
  if (h->dim == 1) {
    hipLaunchKernelGGL((k_observables<T, SPIN, 1>), dim3(blocks), dim3(256), 0, h->stream, ph, g, a, h->partials);
    HIPCHK(hipGetLastError());
  } else {
    // one read-only line pass per axis; the x pass adds the pointwise terms
    int b0 = 0, b1 = 0, b2 = 0;
    if (h->dim == 2) {
      if ((rc = dispatch_obs_axis<T, SPIN, 2, 0>(h, a, 0, &b0))) return rc;
      if ((rc = dispatch_obs_axis<T, SPIN, 2, 1>(h, a, (size_t)b0 * nc, &b1))) return rc;
    } else {
      if ((rc = dispatch_obs_axis<T, SPIN, 3, 0>(h, a, 0, &b0))) return rc;
      if ((rc = dispatch_obs_axis<T, SPIN, 3, 1>(h, a, (size_t)b0 * nc, &b1))) return rc;
      if ((rc = dispatch_obs_axis<T, SPIN, 3, 2>(h, a, (size_t)(b0 + b1) * nc, &b2))) return rc;
    }
    blocks = b0 + b1 + b2;
  }
  // the block sums reduced on the device in a fixed order; NC values come back
  constexpr int NC = 3 * (2 * SPIN + 1) + 2;
  hipLaunchKernelGGL((k_sum_rows<NC>), dim3(1), dim3(1024), 0, h->stream, (const double*)h->partials, (long)blocks,
                     h->obs_sums);
  HIPCHK(hipGetLastError());
  double s[NC];
  HIPCHK(hipMemcpyAsync(s, h->obs_sums, sizeof s, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const int ms = h->ms;
  double sumn = 0;
  for (int m = 0; m < ms; ++m) sumn += s[2 * ms + 2 + m];
  // EN = int col1 / sum(NORM), MZ = int col2, MU = int col/(NORM+EPS),
  // RMS = sqrt(int r^2 |phi|^2) (SRC/cgpe3d.f90:1506-1510,1265-1267)
  if (en) *en = s[0] / sumn;
  if (mz) *mz = s[1];
  for (int m = 0; m < ms; ++m) {
    if (mu) mu[m] = s[2 + m] / (s[2 * ms + 2 + m] + EPS);
    if (rms) rms[m] = std::sqrt(s[ms + 2 + m]);
  }
  return 0;
}

template <typename T> int conv_t(fgpe_handle* h, double* out) {
  const long n = h->npts * h->ms;
  const int blocks = std::min(grid_for(n, 256), 4096);
  int rc = ensure_partials(h, blocks);
  if (rc) return rc;
  hipLaunchKernelGGL(k_conv_max<T>, dim3(blocks), dim3(256), 0, h->stream, (const C<T>*)h->phi,
                     (const C<T>*)h->phi_old, n, h->partials);
  HIPCHK(hipGetLastError());
  std::vector<double> part(blocks);
  HIPCHK(hipMemcpyAsync(part.data(), h->partials, blocks * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double mx = 0;
  for (double v : part) mx = std::max(mx, v);
  *out = mx;
  return 0;
}

template <typename T> int guess_t(fgpe_handle* h) {
  GuessArgs a;
  a.dim = h->dim;
  a.spin = h->spin;
  a.fpc = h->p.option_fpc;
  a.ax = h->p.alpha[0];
  a.ay = h->p.alpha[1];
  a.az = h->p.alpha[2];
  a.t0 = h->p.tau[0];
  a.t1 = h->p.tau[1];
  a.t2 = h->p.tau[2];
  a.mag = h->p.mag;
  a.x2 = h->x2;
  a.y2 = h->y2;
  a.z2 = h->z2;
  hipLaunchKernelGGL(k_guess<T>, dim3(grid_for(h->npts, 256)), dim3(256), 0, h->stream, (C<T>*)h->phi, geom(h), a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}



template <typename T> int from64_t(fgpe_handle* h, long n) {
  hipLaunchKernelGGL(k_from64<T>, dim3(grid_for(n, 256)), dim3(256), 0, h->stream, (const double2*)h->stage64,
                     (C<T>*)h->phi, n);
  HIPCHK(hipGetLastError());
  return 0;
}
template <typename T> int to64_t(fgpe_handle* h, long n) {
  hipLaunchKernelGGL(k_to64<T>, dim3(grid_for(n, 256)), dim3(256), 0, h->stream, (const C<T>*)h->phi,
                     (double2*)h->stage64, n);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace fgpe_impl

#define FGPE_INSTANTIATE_SPIN(T, S)                                                                    \
  template int fgpe_impl::run_steps<T, S>(fgpe_handle*, long);                                         \
  template int fgpe_impl::observables_t<T, S>(fgpe_handle*, double*, double*, double*, double*);
#define FGPE_INSTANTIATE_PREC(T)                                         \
  template int fgpe_impl::normc_t<T>(fgpe_handle*, double*);             \
  template int fgpe_impl::normalize_t<T>(fgpe_handle*, double*);         \
  template int fgpe_impl::conv_t<T>(fgpe_handle*, double*);              \
  template int fgpe_impl::guess_t<T>(fgpe_handle*);                      \
  template int fgpe_impl::from64_t<T>(fgpe_handle*, long);               \
  template int fgpe_impl::to64_t<T>(fgpe_handle*, long);

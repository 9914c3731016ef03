// fgpe_ops.hpp -- templated host code of libfgpe: runtime size dispatch to
// the pass launches (fgpe_launch.hpp, one translation unit per line length),
// the per-run pass schedule and the event-time reductions.  Instantiated per
// precision and spin in fgpe_ops_*.hip.
#pragma once
#include "fgpe_impl.hpp"

namespace fgpe_impl {

// pass launches, defined in fgpe_launch.hpp and instantiated per line length
// N in fgpe_lines.hip
template <typename T, int SPIN, int N> int launch_x(fgpe_handle* h, int flags, hipStream_t st, int z0, int zc);
template <typename T, int SPIN, int N> int launch_fin(fgpe_handle* h);
template <typename T, int N>
int launch_y(fgpe_handle* h, int dir, int scale, hipStream_t st, int x0, int xw, int z0, int zc);
template <typename T, int SPIN, int N> int launch_k(fgpe_handle* h, int scale);
template <typename T, int N> int launch_z(fgpe_handle* h, int scale, hipStream_t st, int x0, int xw);
template <typename T, int SPIN, int N> int launch_d(fgpe_handle* h, hipStream_t st, int x0, int xw, int scale);
template <typename T, int SPIN, int N> int launch_1d(fgpe_handle* h, long nsteps);
template <typename T, int SPIN, int DIM, int AXIS, int N>
int launch_obs_axis(fgpe_handle* h, const ObsArgs& a, size_t off, int* nblk);

// runtime size dispatch over the line lengths of fgpe_impl.hpp
template <typename T, int SPIN> int dispatch_x(fgpe_handle* h, int flags, hipStream_t st, int z0, int zc) {
  switch (h->nx) {
#define FGPE_CASE(n) \
  case n: return launch_x<T, SPIN, n>(h, flags, st, z0, zc);
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  return fail(FGPE_EINVAL, "unsupported NX=%d", h->nx);
}
template <typename T, int SPIN> int dispatch_x(fgpe_handle* h, int flags) {
  return dispatch_x<T, SPIN>(h, flags, h->stream, 0, h->nz);
}
template <typename T, int SPIN> int dispatch_fin(fgpe_handle* h) {
  switch (h->nx) {
#define FGPE_CASE(n) \
  case n: return launch_fin<T, SPIN, n>(h);
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  return fail(FGPE_EINVAL, "unsupported NX=%d", h->nx);
}
template <typename T>
int dispatch_y(fgpe_handle* h, int dir, int scale, hipStream_t st, int x0, int xw, int z0, int zc) {
  switch (h->ny) {
#define FGPE_CASE(n) \
  case n: return launch_y<T, n>(h, dir, scale, st, x0, xw, z0, zc);
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  return fail(FGPE_EINVAL, "unsupported NY=%d", h->ny);
}
template <typename T, int SPIN> int dispatch_k(fgpe_handle* h, int scale) {
  switch (h->ny) {
#define FGPE_CASE(n) \
  case n: return launch_k<T, SPIN, n>(h, scale);
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  return fail(FGPE_EINVAL, "unsupported NY=%d", h->ny);
}
template <typename T> int dispatch_z(fgpe_handle* h, int scale, hipStream_t st, int x0, int xw) {
  switch (h->nz) {
#define FGPE_CASE(n) \
  case n: return launch_z<T, n>(h, scale, st, x0, xw);
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  return fail(FGPE_EINVAL, "unsupported NZ=%d", h->nz);
}
template <typename T, int SPIN> int dispatch_d(fgpe_handle* h, hipStream_t st, int x0, int xw, int scale) {
  switch (h->ny) {
#define FGPE_CASE(n) \
  case n: return launch_d<T, SPIN, n>(h, st, x0, xw, scale);
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  return fail(FGPE_EINVAL, "unsupported NY=%d", h->ny);
}
template <typename T, int SPIN> int dispatch_1d(fgpe_handle* h, long nsteps) {
  switch (h->nx) {
#define FGPE_CASE(n) \
  case n: return launch_1d<T, SPIN, n>(h, nsteps);
    FGPE_1D_SIZES(FGPE_CASE)
#undef FGPE_CASE
  }
  return fail(FGPE_EINVAL, "unsupported 1D NX=%d", h->nx);
}

// 3D step schedule, two phases per step, each a pipeline over chunks:
//   phase 1, by z-chunks: A (inverse-x, SE, SP, NORMC, forward-x) then the
//     next step's B (forward-y, kx and ky factors of KE) -- both act within
//     z-planes;
//   phase 2, by x-chunks: C (forward-z, kz/GAMMAZ factor of KE, inverse-z,
//     one component per block), then D (SOC in (kx, ky, z) space,
//     inverse-y, 1/N, all components), one of them applying the pending
//     renormalisation -- both act within x-slabs.
// The SO coupling depends on (kx, ky) only (SRC/cgpe3d.f90:904-1004), so it
// commutes with the z transforms and is applied after the inverse z FFT,
// still after the whole KE factor as in the reference's order: pass C then
// moves one component at a time and D transforms once per component.
// Chunks go round-robin to up to three streams, so the VALU-bound A overlaps
// the memory-bound B of neighbouring chunks, C the heavier D; the phases are
// joined on the handle's stream (C needs every z of B, A every x of D).  The
// renormalisation (uniform with SO coupling, per component without it, where
// no pass mixes components) commutes with the FFTs, KE and SOC, so in FP32 D
// applies it and the NORMC finalisation that sets it runs beside the first C
// launches instead of between the phases (FP64: C applies it, after the
// finalisation).  Every chunking gives
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
  // FP32: D applies the renormalisation, and fin (the NORMC finalisation of
  // the last pass A) runs first on the handle's stream, beside the other
  // streams' first C launches (C does not need the scale); those streams wait
  // for it before their first D.  FP64: C applies it after the finalisation
  // (D's extra FP64 products cost more than the overlap saves: 256^3 spin-2
  // imaginary time 2490 -> 2526 us per step with D applying it; FP32
  // 1221 -> 1194).
  constexpr bool LATE = sizeof(T) == 4;
  auto phase2 = [&](bool fin) -> int {
    int rc;
    if ((rc = fork())) return rc;
    if (fin) {
      if ((rc = dispatch_fin<T, SPIN>(h))) return rc;
      if (nst > 1) HIPCHK(hipEventRecord(h->evs[3], h->stream));
    }
    bool waits[3] = {false, fin && nst > 1, fin && nst > 2};
    for (int x0 = 0, c = 0; x0 < h->nx; x0 += xc, ++c) {
      const int i = c % nst;
      if ((rc = dispatch_z<T>(h, LATE ? 0 : 1, st[i], x0, xc))) return rc;
      if (waits[i]) {
        HIPCHK(hipStreamWaitEvent(st[i], h->evs[3], 0));
        waits[i] = false;
      }
      if ((rc = dispatch_d<T, SPIN>(h, st[i], x0, xc, LATE ? 1 : 0))) return rc;
    }
    for (int i = 1; i < nst; ++i)  // a stream with no chunk still orders after the finalisation
      if (waits[i]) HIPCHK(hipStreamWaitEvent(st[i], h->evs[3], 0));
    return join();
  };
  int rc;
  if ((rc = phase1(XF_FWD, true))) return rc;  // the run's first forward-x and forward-y
  bool fin = false;
  for (long s = 1; s <= nsteps; ++s) {
    if ((rc = phase2(fin))) return rc;
    // NORMC every step in imaginary time (it sets the renormalisation); in
    // real time the scale is 1 and only the last step's norms are reported
    const int flags = XF_INV | XF_PHYS | (s < nsteps ? XF_FWD : 0) | ((h->p.imag || s == nsteps) ? XF_NORM : 0);
    if ((rc = phase1(flags, s < nsteps))) return rc;
    // finalised at the start of the next phase 2 (FP32), or here
    fin = LATE && (flags & XF_NORM) && s < nsteps;
    if ((flags & XF_NORM) && !fin && (rc = dispatch_fin<T, SPIN>(h))) return rc;
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

template <typename T, int SPIN, int DIM, int AXIS>
int dispatch_obs_axis(fgpe_handle* h, const ObsArgs& a, size_t off, int* nblk) {
  const int n = AXIS == 0 ? h->nx : AXIS == 1 ? h->ny : h->nz;
  switch (n) {
#define FGPE_CASE(n) \
  case n: return launch_obs_axis<T, SPIN, DIM, AXIS, n>(h, a, off, nblk);
    FGPE_LINE_SIZES(FGPE_CASE)
#undef FGPE_CASE
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

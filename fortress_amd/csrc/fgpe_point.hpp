// fgpe_point.hpp -- point-parallel kernels of libfgpe: observables (ENERGY,
// RAD, NORMC) with the nine-point stencil and Simpson weights, the
// convergence maximum, the initial guess, scaling and precision conversion.
// These run at STP/NSTP events only, in FP64 arithmetic.
#pragma once
#include "fgpe_kernels.hpp"

namespace fgpe {

// small FP64 complex with operators, for the long observable formulas
struct Z {
  double r, i;
  __host__ __device__ Z() : r(0), i(0) {}
  __host__ __device__ constexpr Z(double a, double b = 0) : r(a), i(b) {}
};
__host__ __device__ inline Z operator+(Z a, Z b) { return Z(a.r + b.r, a.i + b.i); }
__host__ __device__ inline Z operator-(Z a, Z b) { return Z(a.r - b.r, a.i - b.i); }
__host__ __device__ inline Z operator-(Z a) { return Z(-a.r, -a.i); }
__host__ __device__ inline Z operator*(Z a, Z b) { return Z(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r); }
__host__ __device__ inline Z operator*(double s, Z a) { return Z(s * a.r, s * a.i); }
__host__ __device__ inline Z operator*(Z a, double s) { return Z(s * a.r, s * a.i); }
__host__ __device__ inline Z operator/(Z a, double s) { return Z(a.r / s, a.i / s); }
__host__ __device__ inline Z cj(Z a) { return Z(a.r, -a.i); }
__host__ __device__ inline double ab2(Z a) { return a.r * a.r + a.i * a.i; }

template <typename T> __device__ __forceinline__ Z ld(const C_t<T>* p, long q) {
  C_t<T> a = p[q];
  return Z((double)a.x, (double)a.y);
}

// DIFF (SRC/diff.f90:4-33) at 0-based index i of a line of n points:
// periodically wrapped five-point at the four points of each end, nine-point
// inside.  f(k) returns the value at 0-based k.
template <typename F> __device__ __forceinline__ Z diff_at(F f, int i, int n, double d) {
  if (i >= 4 && i <= n - 5) {
    return (3.0 * f(i - 4) - 32.0 * f(i - 3) + 168.0 * f(i - 2) - 672.0 * f(i - 1) + 672.0 * f(i + 1) -
            168.0 * f(i + 2) + 32.0 * f(i + 3) - 3.0 * f(i + 4)) /
           (840.0 * d);
  }
  auto w = [n](int k) { return k < 0 ? k + n : (k >= n ? k - n : k); };
  return (f(w(i - 2)) - 8.0 * f(w(i - 1)) + 8.0 * f(w(i + 1)) - f(w(i + 2))) / (12.0 * d);
}

struct ObsArgs {
  double tau0, tau1, tau2, gx, gy, gz, dx, dy, dz;
  const double *wx, *wy, *wz;  // Simpson weights
  const double *x2, *y2, *z2;
  const double *vxd, *vyd, *vzd;  // potential parts in FP64
};

// Columns of ENERGY (3D SRC/cgpe3d.f90:1381-1503, 2D SRC/cgpe2d.f90:1145-1267,
// 1D SRC/cgpe1d.f90:1150-1233) at one point, then RAD's r^2|phi|^2 and NORMC's
// |phi|^2, all weighted by the separable Simpson weight and summed.
// One point q of the grid g: adds its Simpson-weighted contributions to
// acc[3 MS + 2] (ENERGY's EN and MZ integrands, the MU numerators, RAD's
// r^2 |phi_m|^2 and NORMC's |phi_m|^2).
template <typename T, int SPIN, int DIM>
__device__ __forceinline__ void obs_point(const C_t<T>* __restrict__ phi, const Geom& g, const ObsArgs& a, long q,
                                          double* acc) {
  constexpr int MS = 2 * SPIN + 1;
  const double s2 = 1.41421356237309504880, s6 = 2.44948974278317809820, s15 = 1.22474487139158904910;
  const double t0 = a.tau0, t1 = a.tau1, t2 = a.tau2, gx = a.gx, gy = a.gy, gz = a.gz;
  const Z IM(0.0, 1.0);
  (void)s6;
  (void)gz;
  {
    const int i = (int)(q % g.nx);
    const int jy = (int)((q / g.nx) % g.ny);
    const int k = (int)(q / ((long)g.nx * g.ny));
    Z ph[MS], px[MS], py[MS], pz[MS], cph[MS];
    double dp2[MS], am[MS];
    double rho = 0;
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      const C_t<T>* pm = phi + m * g.npts;
      ph[m] = ld<T>(pm, q);
      cph[m] = cj(ph[m]);
      am[m] = ab2(ph[m]);
      rho += am[m];
      const long lx = q - i;
      px[m] = diff_at([&](int t) { return ld<T>(pm, lx + t); }, i, g.nx, a.dx);
      dp2[m] = ab2(px[m]);
      if (DIM >= 2) {
        const long ly = q - (long)jy * g.nx;
        py[m] = diff_at([&](int t) { return ld<T>(pm, ly + (long)t * g.nx); }, jy, g.ny, a.dy);
        dp2[m] += ab2(py[m]);
      }
      if (DIM >= 3) {
        const long lz = q - (long)k * g.nx * g.ny;
        pz[m] = diff_at([&](int t) { return ld<T>(pm, lz + (long)t * g.nx * g.ny); }, k, g.nz, a.dz);
        dp2[m] += ab2(pz[m]);
      }
    }
    const double v = a.vxd[i] + (DIM >= 2 ? a.vyd[jy] : 0.0) + (DIM >= 3 ? a.vzd[k] : 0.0);
    double c[MS + 2];
    if constexpr (SPIN == 1) {
      const double s5 = 0.70710678118654752440;
      Z sxp[3] = {s5 * ph[1], s5 * (ph[0] + ph[2]), s5 * ph[1]};
      Z syp[3] = {(IM / s2) * ph[1], (IM / s2) * (-ph[0] + ph[2]), (IM / s2) * (-ph[1])};
      Z ax, ay;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        ax = ax + cph[m] * sxp[m];
        ay = ay + cph[m] * syp[m];
      }
      const double fx = ax.r, fy = ay.r, fz = am[0] - am[2];
      const double base1 =
          v * rho + t0 * rho * rho / 2.0 + t1 * (fx * fx + fy * fy + fz * fz) / 2.0 + dp2[0] + dp2[1] + dp2[2];
      Z so1, so3, so4, so5;
      if constexpr (DIM == 1) {
        so1 = s2 * gx * (-IM * cph[0] * px[1] - IM * cph[2] * px[1] - IM * cph[1] * (px[0] + px[2]));
        so3 = s2 * gx * (-IM * cph[0] * px[1]);
        so4 = s2 * gx * (-IM * cph[1] * (px[0] + px[2]));
        so5 = s2 * gx * (-IM * cph[2] * px[1]);
      } else {
        Z xy = cph[0] * (-IM * gx * px[1] - gy * py[1]) + cph[1] * (-IM * gx * px[2] - gy * py[2]) +
               cph[1] * (-IM * gx * px[0] + gy * py[0]) + cph[2] * (-IM * gx * px[1] + gy * py[1]);
        if constexpr (DIM == 3) {
          so1 = 2.0 * s5 * xy + 2.0 * (cph[0] * (-IM * gz * pz[0]) + cph[2] * (IM * gz * pz[2]));
          so3 = s2 * (cph[0] * (-IM * gx * px[1] - gy * py[1] - IM * s2 * gz * pz[0]));
          so5 = s2 * (cph[2] * (-IM * gx * px[1] + gy * py[1] + IM * gz * s2 * pz[2]));
        } else {
          so1 = s2 * xy;
          so3 = s2 * (cph[0] * (-IM * gx * px[1] - gy * py[1]));
          so5 = s2 * (cph[2] * (-IM * gx * px[1] + gy * py[1]));
        }
        so4 = s2 * (cph[1] * (-IM * gx * px[2] - gy * py[2]) + cph[1] * (-IM * gx * px[0] + gy * py[0]));
      }
      c[0] = (Z(base1) + so1).r;
      c[1] = am[0] - am[2];
      c[2] = (Z(v * am[0] + dp2[0] + t0 * (am[0] + am[1] + am[2]) * am[0] + t1 * (am[0] + am[1] - am[2]) * am[0]) +
              t1 * cph[2] * ph[1] * ph[1] * cph[0] + so3)
                 .r;
      c[3] = (Z(v * am[1] + dp2[1] + t0 * (am[0] + am[1] + am[2]) * am[1] + t1 * (am[0] + am[2]) * am[1]) +
              2.0 * t1 * ph[0] * ph[2] * cph[1] * cph[1] + so4)
                 .r;
      c[4] = (Z(v * am[2] + dp2[2] + t0 * (am[0] + am[1] + am[2]) * am[2] + t1 * (am[1] + am[2] - am[0]) * am[2]) +
              t1 * cph[0] * ph[1] * ph[1] * cph[2] + so5)
                 .r;
    } else {
      const double r = s15;
      Z sxp[5] = {ph[1], ph[0] + r * ph[2], r * (ph[1] + ph[3]), r * ph[2] + ph[4], ph[3]};
      Z syp[5] = {IM * (-ph[1]), IM * (ph[0] - r * ph[2]), IM * (r * ph[1] - r * ph[3]), IM * (r * ph[2] - ph[4]),
                  IM * ph[3]};
      Z ax, ay;
      double fz = 0;
      const double sz[5] = {2, 1, 0, -1, -2};
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        ax = ax + cph[m] * sxp[m];
        ay = ay + cph[m] * syp[m];
        fz += sz[m] * am[m];
      }
      const double fx = ax.r, fy = ay.r;
      Z fm = Z(fx, -fy);
      Z cc[6];
      cc[0] = t1 * fm - 0.4 * t2 * ph[3] * cj(ph[4]);
      cc[1] = 0.2 * t2 * ph[2] * cj(ph[4]);
      cc[2] = s15 * t1 * fm - 0.2 * t2 * ph[2] * cj(ph[3]);
      cc[3] = s15 * t1 * fm - 0.2 * t2 * ph[1] * cj(ph[2]);
      cc[4] = 0.2 * t2 * ph[0] * cj(ph[2]);
      cc[5] = t1 * fm - 0.4 * t2 * ph[0] * cj(ph[1]);
      Z th = (2.0 * ph[0] * ph[4] - 2.0 * ph[1] * ph[3] + ph[2] * ph[2]) / 2.23606797749978969641;
      const double base1 = v * rho + dp2[0] + dp2[1] + dp2[2] + dp2[3] + dp2[4] + t0 * rho * rho / 2.0 +
                           t1 * (fx * fx + fy * fy + fz * fz) / 2.0 + t2 * ab2(th) / 2.0;
      Z so1, so3, so4, so5, so6, so7;
      if constexpr (DIM == 1) {
        so1 = gx * (-IM * 2.0 * (cph[0] * px[1]) - IM * 2.0 * (cph[1] * px[0]) -
                    IM * s6 * (cph[1] * px[2] + cph[2] * px[1] + cph[2] * px[3] + cph[3] * px[2]) -
                    IM * 2.0 * (cph[3] * px[4] + cph[4] * px[3]));
        so3 = -IM * 2.0 * gx * (cph[0] * px[1]);
        so4 = gx * (-IM * 2.0 * (cph[1] * px[0]) - IM * s6 * cph[1] * px[2]);
        so5 = -IM * gx * s6 * (cph[2] * px[1] + cph[2] * px[3]);
        so6 = gx * (-IM * s6 * cph[3] * px[2] - IM * 2.0 * cph[3] * px[4]);
        so7 = -IM * 2.0 * gx * cph[4] * px[3];
      } else if constexpr (DIM == 2) {
        so1 = -IM * 2.0 *
              (cph[0] * gx * px[1] - IM * cph[0] * gy * py[1] + cph[1] * gx * px[0] + IM * cph[1] * gy * py[0] +
               s15 * (cph[1] * gx * px[2] - IM * cph[1] * gy * py[2] + cph[2] * gx * px[1] + IM * cph[2] * gy * py[1] +
                      cph[2] * gx * px[3] - IM * cph[2] * gy * py[3] + cph[3] * gx * px[2] + IM * cph[3] * gy * py[2]) +
               cph[3] * gx * px[4] - IM * cph[3] * gy * py[4] + cph[4] * gx * px[3] + IM * cph[4] * gy * py[3]);
        so3 = -IM * 2.0 * (cph[0] * gx * px[1] - IM * cph[0] * gy * py[1]);
        so4 = -IM * 2.0 *
              (cph[1] * gx * px[0] + IM * cph[1] * gy * py[0] + s15 * (cph[1] * gx * px[2] - IM * cph[1] * gy * py[2]));
        so5 = -IM * 2.0 *
              (s15 * (cph[2] * gx * px[1] + IM * cph[2] * gy * py[1] + cph[2] * gx * px[3] - IM * cph[2] * gy * py[3]));
        so6 = -IM * 2.0 *
              (cph[3] * gx * px[4] - IM * cph[3] * gy * py[4] + s15 * (cph[3] * gx * px[2] + IM * cph[3] * gy * py[2]));
        so7 = -IM * 2.0 * (cph[4] * gx * px[3] + IM * cph[4] * gy * py[3]);
      } else {
        so1 = -IM * cph[0] * (2.0 * gx * px[1] - IM * 2.0 * gy * py[1] + 4.0 * gz * pz[0]) -
              IM * cph[1] * (gx * (2.0 * px[0] + s6 * px[2]) + IM * gy * (2.0 * py[0] - s6 * py[2]) + 2.0 * gz * pz[1]) -
              IM * cph[2] * (gx * s6 * (px[1] + px[3]) + IM * gy * s6 * (py[1] - py[3])) -
              IM * cph[3] * (gx * (s6 * px[2] + 2.0 * px[4]) + IM * gy * (s6 * py[2] - 2.0 * py[4]) - 2.0 * gz * pz[3]) -
              IM * cph[4] * (2.0 * gx * px[3] + IM * 2.0 * gy * py[3] - 4.0 * gz * pz[4]);
        so3 = -IM * 2.0 * (cph[0] * gx * px[1] - IM * cph[0] * gy * py[1] + 2.0 * cph[0] * gz * pz[0]);
        so4 = -IM * 2.0 *
              (cph[1] * gx * px[0] + IM * cph[1] * gy * py[0] + s15 * (cph[1] * gx * px[2] - IM * cph[1] * gy * py[2]) +
               cph[1] * gz * pz[1]);
        so5 = -IM * 2.0 *
              (s15 * (cph[2] * gx * px[1] + IM * cph[2] * gy * py[1] + cph[2] * gx * px[3] - IM * cph[2] * gy * py[3]));
        so6 = -IM * 2.0 *
              (cph[3] * gx * px[4] - IM * cph[3] * gy * py[4] + s15 * (cph[3] * gx * px[2] + IM * cph[3] * gy * py[2]) -
               cph[3] * gz * pz[3]);
        so7 = -IM * 2.0 * (cph[4] * gx * px[3] + IM * cph[4] * gy * py[3] - 2.0 * cph[4] * gz * pz[4]);
      }
      c[0] = (Z(base1) + so1).r;
      c[1] = 2.0 * am[0] + am[1] - am[3] - 2.0 * am[4];
      c[2] = (Z((v + t0 * rho + 2.0 * t1 * fz + 0.4 * t2 * am[4]) * am[0]) + (cc[0] * ph[1] + cc[1] * ph[2]) * cph[0] +
              so3 + Z(dp2[0]))
                 .r;
      c[3] = (Z((v + t0 * rho + t1 * fz + 0.4 * t2 * am[3]) * am[1]) + (cj(cc[0]) * ph[0] + cc[2] * ph[2]) * cph[1] +
              so4 + Z(dp2[1]))
                 .r;
      c[4] = (Z((v + t0 * rho + 0.2 * t2 * am[2]) * am[2]) +
              (cj(cc[1]) * ph[0] + cj(cc[2]) * ph[1] + cc[3] * ph[3] + cc[4] * ph[4]) * cph[2] + so5 + Z(dp2[2]))
                 .r;
      c[5] = (Z((v + t0 * rho - t1 * fz + 0.4 * t2 * am[1]) * am[3]) + (cj(cc[3]) * ph[2] + cc[5] * ph[4]) * cph[3] +
              so6 + Z(dp2[3]))
                 .r;
      c[6] = (Z((v + t0 * rho - 2.0 * t1 * fz + 0.4 * t2 * am[0]) * am[4]) +
              (cj(cc[4]) * ph[2] + cj(cc[5]) * ph[3]) * cph[4] + so7 + Z(dp2[4]))
                 .r;
    }
    const double w = a.wx[i] * (DIM >= 2 ? a.wy[jy] : 1.0) * (DIM >= 3 ? a.wz[k] : 1.0);
    const double r2 = a.x2[i] + (DIM >= 2 ? a.y2[jy] : 0.0) + (DIM >= 3 ? a.z2[k] : 0.0);
#pragma unroll
    for (int cidx = 0; cidx < MS + 2; ++cidx) acc[cidx] += w * c[cidx];
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      acc[MS + 2 + m] += w * r2 * am[m];
      acc[2 * MS + 2 + m] += w * am[m];
    }
  }
}

template <typename T, int SPIN, int DIM>
__global__ __launch_bounds__(256) void k_observables(const C_t<T>* __restrict__ phi, Geom g, ObsArgs a,
                                                     double* __restrict__ partials) {
  constexpr int MS = 2 * SPIN + 1;
  constexpr int NC = 3 * MS + 2;
  __shared__ double red[16 * NC];
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < g.npts; q += (long)gridDim.x * blockDim.x)
    obs_point<T, SPIN, DIM>(phi, g, a, q, acc);
  block_sum_store<NC>(acc, red, partials);
}

// ------------------------------------------- observables as line passes
// ENERGY's integrand splits into a pointwise part and one part per axis d
// that involves only the derivatives along d (SRC/cgpe3d.f90:1381-1503,
// SRC/cgpe2d.f90:1145-1267, SRC/cgpe1d.f90:1150-1233):
//   kinetic  sum_m |d_d phi_m|^2           (EN and MU_m columns)
//   SO       Re(-i gamma_d phi^H (2 S_d) d_d phi) = gamma_d Im(phi^H u)
//            with u = 2 S_d d_d phi, and per component gamma_d Im(phi_m^* u_m)
//            for MU_m (standard S_x, S_y; S_z diagonal, 3D only).
// So each axis is one read-only pass over lines along it: the lines of all
// components sit in LDS, DIFF (SRC/diff.f90:4-33) reads its stencil from
// there, and the x pass also adds the pointwise terms, RAD and NORMC.

// u = 2 S_d v for axis d (0 x, 1 y, 2 z)
template <int SPIN, int AXIS> __device__ __forceinline__ void two_s(const Z* v, Z* u) {
  const Z I(0.0, 1.0);
  if constexpr (SPIN == 1) {
    const double r2 = 1.41421356237309504880;
    if constexpr (AXIS == 0) {
      u[0] = r2 * v[1];
      u[1] = r2 * (v[0] + v[2]);
      u[2] = r2 * v[1];
    } else if constexpr (AXIS == 1) {
      u[0] = (-r2) * (I * v[1]);
      u[1] = r2 * (I * (v[0] - v[2]));
      u[2] = r2 * (I * v[1]);
    } else {
      u[0] = 2.0 * v[0];
      u[1] = Z();
      u[2] = (-2.0) * v[2];
    }
  } else {
    const double r6 = 2.44948974278317809820;
    if constexpr (AXIS == 0) {
      u[0] = 2.0 * v[1];
      u[1] = 2.0 * v[0] + r6 * v[2];
      u[2] = r6 * (v[1] + v[3]);
      u[3] = r6 * v[2] + 2.0 * v[4];
      u[4] = 2.0 * v[3];
    } else if constexpr (AXIS == 1) {
      u[0] = (-2.0) * (I * v[1]);
      u[1] = I * (2.0 * v[0] - r6 * v[2]);
      u[2] = r6 * (I * (v[1] - v[3]));
      u[3] = I * (r6 * v[2] - 2.0 * v[4]);
      u[4] = 2.0 * (I * v[3]);
    } else {
      u[0] = 4.0 * v[0];
      u[1] = 2.0 * v[1];
      u[2] = Z();
      u[3] = (-2.0) * v[3];
      u[4] = (-4.0) * v[4];
    }
  }
}

// Pointwise columns at one point (no derivatives): EN's V rho + interaction
// energy, MZ, the MU_m numerators' potential/interaction parts, RAD's
// r^2 |phi_m|^2 and NORMC's |phi_m|^2, as in obs_point.
template <int SPIN> __device__ __forceinline__ void obs_pointwise(const Z* ph, double v, double r2, const ObsArgs& a,
                                                                  double* c) {
  constexpr int MS = 2 * SPIN + 1;
  const double t0 = a.tau0, t1 = a.tau1, t2 = a.tau2;
  const double s2 = 1.41421356237309504880, s15 = 1.22474487139158904910;
  const Z IM(0.0, 1.0);
  Z cph[MS];
  double am[MS], rho = 0;
#pragma unroll
  for (int m = 0; m < MS; ++m) {
    cph[m] = cj(ph[m]);
    am[m] = ab2(ph[m]);
    rho += am[m];
  }
  if constexpr (SPIN == 1) {
    const double s5 = 0.70710678118654752440;
    Z sxp[3] = {s5 * ph[1], s5 * (ph[0] + ph[2]), s5 * ph[1]};
    Z syp[3] = {(IM / s2) * ph[1], (IM / s2) * (-ph[0] + ph[2]), (IM / s2) * (-ph[1])};
    Z ax, ay;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      ax = ax + cph[m] * sxp[m];
      ay = ay + cph[m] * syp[m];
    }
    const double fx = ax.r, fy = ay.r, fz = am[0] - am[2];
    c[0] += v * rho + t0 * rho * rho / 2.0 + t1 * (fx * fx + fy * fy + fz * fz) / 2.0;
    c[1] += am[0] - am[2];
    c[2] += (Z(v * am[0] + t0 * (am[0] + am[1] + am[2]) * am[0] + t1 * (am[0] + am[1] - am[2]) * am[0]) +
             t1 * cph[2] * ph[1] * ph[1] * cph[0]).r;
    c[3] += (Z(v * am[1] + t0 * (am[0] + am[1] + am[2]) * am[1] + t1 * (am[0] + am[2]) * am[1]) +
             2.0 * t1 * ph[0] * ph[2] * cph[1] * cph[1]).r;
    c[4] += (Z(v * am[2] + t0 * (am[0] + am[1] + am[2]) * am[2] + t1 * (am[1] + am[2] - am[0]) * am[2]) +
             t1 * cph[0] * ph[1] * ph[1] * cph[2]).r;
  } else {
    const double r = s15;
    Z sxp[5] = {ph[1], ph[0] + r * ph[2], r * (ph[1] + ph[3]), r * ph[2] + ph[4], ph[3]};
    Z syp[5] = {IM * (-ph[1]), IM * (ph[0] - r * ph[2]), IM * (r * ph[1] - r * ph[3]), IM * (r * ph[2] - ph[4]),
                IM * ph[3]};
    Z ax, ay;
    double fz = 0;
    const double sz[5] = {2, 1, 0, -1, -2};
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      ax = ax + cph[m] * sxp[m];
      ay = ay + cph[m] * syp[m];
      fz += sz[m] * am[m];
    }
    const double fx = ax.r, fy = ay.r;
    Z fm = Z(fx, -fy);
    Z cc[6];
    cc[0] = t1 * fm - 0.4 * t2 * ph[3] * cj(ph[4]);
    cc[1] = 0.2 * t2 * ph[2] * cj(ph[4]);
    cc[2] = s15 * t1 * fm - 0.2 * t2 * ph[2] * cj(ph[3]);
    cc[3] = s15 * t1 * fm - 0.2 * t2 * ph[1] * cj(ph[2]);
    cc[4] = 0.2 * t2 * ph[0] * cj(ph[2]);
    cc[5] = t1 * fm - 0.4 * t2 * ph[0] * cj(ph[1]);
    Z th = (2.0 * ph[0] * ph[4] - 2.0 * ph[1] * ph[3] + ph[2] * ph[2]) / 2.23606797749978969641;
    c[0] += v * rho + t0 * rho * rho / 2.0 + t1 * (fx * fx + fy * fy + fz * fz) / 2.0 + t2 * ab2(th) / 2.0;
    c[1] += 2.0 * am[0] + am[1] - am[3] - 2.0 * am[4];
    c[2] += (Z((v + t0 * rho + 2.0 * t1 * fz + 0.4 * t2 * am[4]) * am[0]) + (cc[0] * ph[1] + cc[1] * ph[2]) * cph[0]).r;
    c[3] += (Z((v + t0 * rho + t1 * fz + 0.4 * t2 * am[3]) * am[1]) + (cj(cc[0]) * ph[0] + cc[2] * ph[2]) * cph[1]).r;
    c[4] += (Z((v + t0 * rho + 0.2 * t2 * am[2]) * am[2]) +
             (cj(cc[1]) * ph[0] + cj(cc[2]) * ph[1] + cc[3] * ph[3] + cc[4] * ph[4]) * cph[2]).r;
    c[5] += (Z((v + t0 * rho - t1 * fz + 0.4 * t2 * am[1]) * am[3]) + (cj(cc[3]) * ph[2] + cc[5] * ph[4]) * cph[3]).r;
    c[6] += (Z((v + t0 * rho - 2.0 * t1 * fz + 0.4 * t2 * am[0]) * am[4]) +
             (cj(cc[4]) * ph[2] + cj(cc[5]) * ph[3]) * cph[4]).r;
  }
#pragma unroll
  for (int m = 0; m < MS; ++m) {
    c[MS + 2 + m] += r2 * am[m];
    c[2 * MS + 2 + m] += am[m];
  }
}

// Lines per tile of an observables line pass: all MS components of L lines
// of N points (padded to N + 2) in at most 48 KB of LDS (three blocks per
// CU; 256^3 spin-2 FP32: 4 lines, 32 B rows for the y and z passes), L a
// power of two, at most 64 (at least 1).
template <typename T, int SPIN, int N> __host__ __device__ constexpr int obs_lines() {
  constexpr long per = (long)(2 * SPIN + 1) * (N + 2) * sizeof(C_t<T>);
  int l = 1;
  while (l < 64 && (long)(2 * l) * per <= 49152) l *= 2;
  return l;
}
// blocks of an observables line pass: each loops over tiles, so the partial
// rows (one per block) stay few
constexpr long OBS_MAX_BLOCKS = 1024;
template <typename T, int SPIN, int N> __host__ __device__ constexpr size_t obs_line_lds() {
  return (size_t)(2 * SPIN + 1) * obs_lines<T, SPIN, N>() * (N + 2) * sizeof(C_t<T>);
}
// threads per observables block
constexpr int OBS_TPB = 256;

// Sum of nparts rows of NC partials into out[NC] in a fixed order (one block
// of 1024 threads: thread t takes rows t, t + 1024, ...; then a fixed tree),
// so the observables are bit-reproducible and only NC values leave the device.
template <int NC>
__global__ __launch_bounds__(1024) void k_sum_rows(const double* __restrict__ partials, long nparts,
                                                   double* __restrict__ out) {
  __shared__ double red[16 * NC];
  double a[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) a[c] = 0.0;
  for (long r = threadIdx.x; r < nparts; r += 1024)
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] += partials[r * NC + c];
  wave_sum<NC>(a);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) red[wid * NC + c] = a[c];
  __syncthreads();
  if (threadIdx.x < NC) {
    double s = 0;
    for (int w = 0; w < 16; ++w) s += red[w * NC + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// One axis of the observables (2D/3D): tiles of L lines along AXIS (x: L
// consecutive lines of the (y, z) grid; y, z: L consecutive x columns at one
// z or y), all components loaded once into LDS; each block takes tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... of the ntiles, each thread adds
// the Simpson-weighted columns of its points to FP64 sums, and a block stores
// its 3 MS + 2 sums in partials[blockIdx.x].  Lines past the grid (a last
// partial tile) are skipped.
template <typename T, int SPIN, int DIM, int AXIS, int N>
__global__ __launch_bounds__(OBS_TPB) void k_obs_line(const C_t<T>* __restrict__ phi, Geom g, ObsArgs a, long ntiles,
                                                      double* __restrict__ partials) {
  using C = C_t<T>;
  constexpr int MS = 2 * SPIN + 1;
  constexpr int NC = 3 * MS + 2;
  constexpr int L = obs_lines<T, SPIN, N>();
  constexpr int NS = N + 2;   // padded line stride
  constexpr int PTS = L * N;  // points of a tile
  constexpr int PER = (PTS + OBS_TPB - 1) / OBS_TPB;
  // with 3 * 2^a points per line the last round of a tile is partial
  auto in_tile = [](int e) { return PTS % OBS_TPB == 0 || e < PTS; };
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);  // [MS][L][NS]
  static_assert(16 * NC * sizeof(double) <= (size_t)MS * L * NS * sizeof(C), "reduction must fit the tile");
  const long nlines = (long)g.ny * g.nz;
  const int nbx = (g.nx + L - 1) / L;
  const double d = AXIS == 0 ? a.dx : AXIS == 1 ? a.dy : a.dz;
  const double gam = AXIS == 0 ? a.gx : AXIS == 1 ? a.gy : a.gz;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  // y and z tiles of L columns narrower than a 128 B line: the tiles that
  // share a line go to blocks 8 apart (one XCD, dispatched together), so the
  // line is fetched into one L2 once instead of once per tile
  constexpr int SHARE = (AXIS != 0 && L * (int)sizeof(C) < 128) ? 128 / (L * (int)sizeof(C)) : 1;
  const long full = ntiles / (8 * SHARE) * (8 * SHARE);
  for (long bb = blockIdx.x; bb < ntiles; bb += gridDim.x) {
    long b = bb;
    if (SHARE > 1 && bb < full) {
      const long r = bb % 8, q = (bb / 8) % SHARE;
      b = bb - bb % (8 * SHARE) + r * SHARE + q;
    }
    // line l of the tile: its (y, z) (x lines) or x (y, z lines) and validity
    long line0 = 0;
    int ox = 0, oy = 0, oz = 0;
    if constexpr (AXIS == 0) {
      line0 = b * L;
    } else if constexpr (AXIS == 1) {
      ox = (int)(b % nbx) * L;
      oz = (int)(b / nbx);
    } else {
      ox = (int)(b % nbx) * L;
      oy = (int)(b / nbx);
    }
    auto valid = [&](int l) { return AXIS == 0 ? line0 + l < nlines : ox + l < g.nx; };
    auto gidx = [&](int l, int i) -> long {
      if constexpr (AXIS == 0) return (line0 + l) * g.nx + i;
      else if constexpr (AXIS == 1) return ((long)oz * g.ny + i) * g.nx + ox + l;
      else return ((long)i * g.ny + oy) * g.nx + ox + l;
    };
    __syncthreads();  // the previous tile's readers are done with the LDS
    // load: x lines contiguous (i fastest); y/z lines with the L columns
    // fastest; every load of a component is issued before its LDS stores
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      const C* pm = phi + m * g.npts;
      C tmp[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + k * OBS_TPB;
        const int l = AXIS == 0 ? e / N : e % L, i = AXIS == 0 ? e % N : e / L;
        tmp[k] = in_tile(e) && valid(l) ? pm[gidx(l, i)] : mk<C>((T)0, (T)0);
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + k * OBS_TPB;
        const int l = AXIS == 0 ? e / N : e % L, i = AXIS == 0 ? e % N : e / L;
        if (in_tile(e)) lds[(m * L + l) * NS + i] = tmp[k];
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < PTS; e += OBS_TPB) {
      const int l = e / N, i = e % N;
      if (!valid(l)) continue;
      int x, y, z;
      if constexpr (AXIS == 0) {
        x = i;
        y = (int)((line0 + l) % g.ny);
        z = (int)((line0 + l) / g.ny);
      } else if constexpr (AXIS == 1) {
        x = ox + l;
        y = i;
        z = oz;
      } else {
        x = ox + l;
        y = oy;
        z = i;
      }
      Z ph[MS], dp[MS], u[MS];
#pragma unroll
      for (int m = 0; m < MS; ++m) {
        const C* line = lds + (m * L + l) * NS;
        const C c0 = line[i];
        ph[m] = Z((double)c0.x, (double)c0.y);
        dp[m] = diff_at(
            [&](int t) {
              const C v = line[t];
              return Z((double)v.x, (double)v.y);
            },
            i, N, d);
      }
      double c[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) c[k] = 0.0;
      two_s<SPIN, AXIS>(dp, u);
      double so = 0.0;
#pragma unroll
      for (int m = 0; m < MS; ++m) {
        const double k2 = ab2(dp[m]);
        const double sm = gam * (ph[m].r * u[m].i - ph[m].i * u[m].r);  // gamma Im(phi_m^* u_m)
        so += sm;
        c[0] += k2;
        c[2 + m] += k2 + sm;
      }
      c[0] += so;
      if constexpr (AXIS == 0) {
        const double v = a.vxd[x] + (DIM >= 2 ? a.vyd[y] : 0.0) + (DIM >= 3 ? a.vzd[z] : 0.0);
        const double r2 = a.x2[x] + (DIM >= 2 ? a.y2[y] : 0.0) + (DIM >= 3 ? a.z2[z] : 0.0);
        obs_pointwise<SPIN>(ph, v, r2, a, c);
      }
      const double w = a.wx[x] * (DIM >= 2 ? a.wy[y] : 1.0) * (DIM >= 3 ? a.wz[z] : 1.0);
#pragma unroll
      for (int k = 0; k < NC; ++k) acc[k] += w * c[k];
    }
  }
  // the wave sums reuse the tile's LDS once every thread is done with it
  // (block_sum_store synchronises before writing)
  block_sum_store<NC>(acc, reinterpret_cast<double*>(smem_raw), partials, (long)blockIdx.x);
}

// NORMC partials over real-space points (separable Simpson weights)
template <typename T, int MS>
__global__ __launch_bounds__(256) void k_norm_reduce(const C_t<T>* __restrict__ phi, Geom g, const double* wx,
                                                     const double* wy, const double* wz, double* partials) {
  __shared__ double red[16 * MS];
  double acc[MS];
#pragma unroll
  for (int m = 0; m < MS; ++m) acc[m] = 0.0;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < g.npts; q += (long)gridDim.x * blockDim.x) {
    const int i = (int)(q % g.nx);
    const int jy = (int)((q / g.nx) % g.ny);
    const int k = (int)(q / ((long)g.nx * g.ny));
    const double w = wx[i] * wy[jy] * wz[k];
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      C_t<T> a = phi[m * g.npts + q];
      acc[m] += w * ((double)a.x * (double)a.x + (double)a.y * (double)a.y);
    }
  }
  block_sum_store<MS>(acc, red, partials);
}

// max |phi - phi_old| per block (SRC/cgpe3d.f90:503)
template <typename T>
__global__ __launch_bounds__(256) void k_conv_max(const C_t<T>* __restrict__ a, const C_t<T>* __restrict__ b, long n,
                                                  double* partials) {
  __shared__ double red[16];
  double mx = 0.0;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    C_t<T> u = a[q], v = b[q];
    double dr = (double)u.x - (double)v.x, di = (double)u.y - (double)v.y;
    mx = fmax(mx, sqrt(dr * dr + di * di));
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, red[w]);
    partials[blockIdx.x] = r;
  }
}

template <typename T>
__global__ void k_apply_scale(C_t<T>* __restrict__ phi, long npts, int ms, const C_t<T>* __restrict__ cscale) {
  const long n = npts * ms;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const C_t<T> c = cscale[q / npts];
    phi[q] = cmul(phi[q], c);
  }
}

template <typename T> __global__ void k_set_scale_one(C_t<T>* cscale, int ms) {
  if (threadIdx.x < ms) cscale[threadIdx.x] = mk<C_t<T>>((T)1, (T)0);
}

template <typename T> __global__ void k_from64(const double2* __restrict__ src, C_t<T>* __restrict__ dst, long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    double2 a = src[q];
    dst[q] = mk<C_t<T>>((T)a.x, (T)a.y);
  }
}
template <typename T> __global__ void k_to64(const C_t<T>* __restrict__ src, double2* __restrict__ dst, long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    C_t<T> a = src[q];
    dst[q] = make_double2((double)a.x, (double)a.y);
  }
}

struct GuessArgs {
  int dim, spin, fpc;
  double ax, ay, az, t0, t1, t2, mag;
  const double *x2, *y2, *z2;
};

// initial guess of INITIALIZE (SRC/cgpe3d.f90:753-833; 2D SRC/cgpe2d.f90:697-771;
// 1D SRC/cgpe1d.f90:707-786), quirks included (see the oracle's restatement)
template <typename T> __global__ void k_guess(C_t<T>* __restrict__ phi, Geom g, GuessArgs a) {
  const double PI = 3.141592653589793238462643383279502884;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < g.npts; q += (long)gridDim.x * blockDim.x) {
    const int i = (int)(q % g.nx);
    const int jy = (int)((q / g.nx) % g.ny);
    const int k = (int)(q / ((long)g.nx * g.ny));
    const double x2 = a.x2[i], y2 = a.dim >= 2 ? a.y2[jy] : 0.0, z2 = a.dim >= 3 ? a.z2[k] : 0.0;
    const double tmp = a.dim == 3 ? a.ax * x2 + a.ay * y2 + a.az * z2 : a.dim == 2 ? a.ax * x2 + a.ay * y2 : a.ax * x2;
    const double vv = a.dim == 3 ? a.ax * a.ax * x2 + a.ay * a.ay * y2 + a.az * a.az * z2
                                 : a.dim == 2 ? a.ax * a.ax * x2 + a.ay * a.ay * y2 : a.ax * a.ax * x2;
    const double normg = a.dim == 3 ? pow(PI, 0.75) : a.dim == 2 ? sqrt(PI) : sqrt(sqrt(PI));
    const double gsn = exp(-tmp / 2.0) / normg;
    const double t0 = a.t0, t1 = a.t1, t2 = a.spin == 1 ? 0.0 : a.t2, mag = a.mag;
    double c[5] = {0, 0, 0, 0, 0};
    if (a.spin == 1) {
      if (a.fpc == 1) {
        double tf, thr, thr2;
        if (a.dim == 3) {
          thr = pow(15.0 * (t0 + t1) / (8.0 * PI), 0.4);
          thr2 = pow(15.0 * (t0 + t1) / (8.0 * PI), 0.2);
          tf = (thr - vv >= 0.0) ? (thr - vv) / (t0 + t1) : 0.0;
        } else if (a.dim == 2) {
          double cond = pow(4.0 * (t0 + t1) / (2.0 * PI), 8.0);
          thr = pow(4.0 * (t0 + t1) / (2.0 * PI), 0.5);
          thr2 = pow(4.0 * (t0 + t1) / (2.0 * PI), 0.25);
          tf = (cond - vv >= 0.0) ? (thr - vv) / (t0 + t1) : 0.0;
        } else {
          thr = pow(3.0 * (t0 + t2) / 4.0, 0.67);
          thr2 = pow(3.0 * (t0 + t2) / 4.0, 0.33);
          tf = (thr - vv >= 0.0) ? (thr - vv) / (t0 + t2) : 0.0;
        }
        double b = thr2 >= 10.0 ? sqrt(tf) : gsn;
        c[0] = ((1.0 + mag) / 2.0) * b;
        c[1] = sqrt((1.0 - mag * mag) / 2.0) * b;
        c[2] = ((1.0 - mag) / 2.0) * b;
      } else if (a.fpc == 2) {
        c[0] = sqrt((1.0 + mag) / 2.0) * gsn;
        c[2] = sqrt((1.0 - mag) / 2.0) * gsn;
      } else {
        double gg = a.dim == 3 ? exp(-(x2 + y2 + z2) / 2.0) / pow(PI, 0.75) : gsn;
        c[0] = c[1] = c[2] = gg;
      }
    } else {
      if (a.fpc == 1) {
        double tt = t0 + 4.0 * t1, thr, thr2;
        if (a.dim == 3) {
          thr = pow(15.0 * tt / (8.0 * PI), 0.4);
          thr2 = pow(4.0 * tt / (2.0 * PI), 0.25);
        } else if (a.dim == 2) {
          thr = pow(4.0 * tt / (2.0 * PI), 0.5);
          thr2 = pow(4.0 * tt / (2.0 * PI), 0.25);
        } else {
          thr = pow(3.0 * tt / 4.0, 0.67);
          thr2 = 3.0 * pow(t0 + 4.0 * t1 / 4.0, 0.33);
        }
        double tf = (thr - vv >= 0.0) ? (thr - vv) / tt : 0.0;
        double b = thr2 >= 10.0 ? sqrt(tf) : gsn;
        c[0] = ((2.0 + mag) * (2.0 + mag) / 16.0) * b;
        c[1] = (sqrt(4.0 - mag * mag) * (2.0 + mag) / 8.0) * b;
        c[2] = (sqrt(1.5) * (4.0 - mag * mag) / 8.0) * b;
        c[3] = (sqrt(4.0 - mag * mag) * (2.0 - mag) / 8.0) * b;
        c[4] = ((2.0 - mag) * (2.0 - mag) / 16.0) * b;
      } else if (a.fpc == 2) {
        c[0] = (sqrt(2.0 + mag) / 2.0) * gsn;
        c[4] = (sqrt(2.0 - mag) / 2.0) * gsn;
      } else if (a.fpc == 3) {
        c[0] = sqrt((1.0 + mag) / 3.0) * gsn;
        c[3] = sqrt((2.0 - mag) / 3.0) * gsn;
      } else {
        const double f = a.dim == 3 ? 1.0 : sqrt(0.2);
        for (int m = 0; m < 5; ++m) c[m] = f * gsn;
      }
    }
    const int ms = 2 * a.spin + 1;
    for (int m = 0; m < ms; ++m) phi[m * g.npts + q] = mk<C_t<T>>((T)c[m], (T)0);
  }
}

}  // namespace fgpe

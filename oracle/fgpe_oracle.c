/*
 * fgpe_oracle.c -- FP64 CPU restatement of FORTRESS's time-splitting spectral
 * propagator (SRC/cgpe1d.f90, SRC/cgpe2d.f90, SRC/cgpe3d.f90, SRC/diff.f90,
 * SRC/simpson.f90).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libfgpe and the host-CPU baseline.
 * The product path never links or loads this file.
 *
 * Arithmetic is FP64 complex throughout, as in the reference
 * (COMPLEX(KIND=DBL), SRC/cgpe3d.f90:82).  Two deliberate differences, both at
 * rounding level, are stated where they occur:
 *   - the FFT is a self-written mixed-radix (2/3) transform instead of FFTW;
 *     both compute the same unnormalised DFT (SRC/cgpe3d.f90:852-853);
 *   - the per-point 5x5 Hermitian eigenproblem of the spin-2 SE step is solved
 *     by cyclic complex Jacobi instead of LAPACK ZHEEV
 *     (SRC/cgpe3d.f90:1725-1734); V diag(exp(-i CDT w)) V^H is the unique
 *     matrix exponential, so both give exp(-i CDT Hbar) phi to rounding.
 *   - |phi|^2 is formed as re^2 + im^2 where the reference writes ABS(PHI)**2.
 *
 * Layout: phi[m][k][j][i] = PHI(i+1,j+1,k+1,m+1), x fastest.
 */
#include "fgpe_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* SRC/cgpe3d.f90:37-41,72 */
static const double PI_ = 3.141592653589793238462643383279502884;
static const double AMU = 1.66e-27, HBAR = 1.054560653e-34, AU = 0.529177208e-10;
static const double EPS = 1.0e-40;
static const double TOL_NR = 1.0e-6;
static const int NTRIAL = 1000;

struct or_state {
  fgpe_params p;
  int nx, ny, nz, ms, spin, dim;
  long npts;
  cplx* phi;  /* [ms][npts] */
  cplx* phif; /* [ms][npts] */
  double *x, *y, *z, *x2, *y2, *z2, *kx, *ky, *kz;
  double* v;
  double* r2;
  cplx cdt;
  double norm[5];
  double sig[2];
  double lx, ly, lz;
};

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---------------------------------------------------------------- SIMPSON
 * SRC/simpson.f90:4-22.  Composite 1/3 rule on points 1..N-3 plus the 3/8 rule
 * on N-3..N (N even). */
double or_simpson(const double* f, int n, double dx) {
  /* 1-based F(i) = f[i-1] */
  double f1 = f[1] + f[n - 5];
  double f2 = f[2];
  for (int i = 4; i <= n - 6; i += 2) {
    f1 += f[i - 1];
    f2 += f[i];
  }
  return dx * (f[0] + 4.0 * f1 + 2.0 * f2 + f[n - 4]) / 3.0 +
         3.0 * dx * (f[n - 4] + 3.0 * (f[n - 3] + f[n - 2]) + f[n - 1]) / 8.0;
}

/* ------------------------------------------------------------------- DIFF
 * SRC/diff.f90:4-33.  Nine-point first derivative in the interior, five-point
 * at i = 3,4,N-3,N-2 and periodically wrapped five-point at the two ends. */
void or_diff(const double* p, int n, double dx, double* dp) {
#define P(i) p[(i)-1]
  dp[0] = (P(n - 1) - 8.0 * P(n) + 8.0 * P(2) - P(3)) / (12.0 * dx);
  dp[1] = (P(n) - 8.0 * P(1) + 8.0 * P(3) - P(4)) / (12.0 * dx);
  for (int i = 3; i <= 4; i++) dp[i - 1] = (P(i - 2) - 8.0 * P(i - 1) + 8.0 * P(i + 1) - P(i + 2)) / (12.0 * dx);
  for (int i = n - 3; i <= n - 2; i++)
    dp[i - 1] = (P(i - 2) - 8.0 * P(i - 1) + 8.0 * P(i + 1) - P(i + 2)) / (12.0 * dx);
  for (int i = 5; i <= n - 4; i++)
    dp[i - 1] = (3.0 * P(i - 4) - 32.0 * P(i - 3) + 168.0 * P(i - 2) - 672.0 * P(i - 1) + 672.0 * P(i + 1) -
                 168.0 * P(i + 2) + 32.0 * P(i + 3) - 3.0 * P(i + 4)) /
                (840.0 * dx);
  dp[n - 2] = (P(n - 3) - 8.0 * P(n - 2) + 8.0 * P(n) - P(1)) / (12.0 * dx);
  dp[n - 1] = (P(n - 2) - 8.0 * P(n - 1) + 8.0 * P(1) - P(2)) / (12.0 * dx);
#undef P
}

/* -------------------------------------------------------------------- FFT
 * Unnormalised DFT X_k = sum_j x_j exp(sign*2*pi*i*j*k/n), the transform of
 * fftw_execute_dft with FFTW_FORWARD (sign -1) / FFTW_BACKWARD (+1)
 * (SRC/cgpe3d.f90:852-853,1760,1781).  Recursive decimation in time over the
 * factors 2 and 3; other prime factors fall back to a direct sum. */
static void dft_rec(cplx* out, const cplx* in, int n, int stride, const cplx* tw, int twstride, cplx* tmp) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int r = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : n;
  int m = n / r;
  if (r == n) { /* direct DFT */
    for (int k = 0; k < n; k++) {
      cplx s = 0;
      for (int j = 0; j < n; j++) s += in[(long)j * stride] * tw[((long)j * k % n) * twstride];
      tmp[k] = s;
    }
    memcpy(out, tmp, sizeof(cplx) * n);
    return;
  }
  for (int q = 0; q < r; q++) dft_rec(out + q * m, in + (long)q * stride, m, stride * r, tw, twstride * r, tmp);
  for (int k = 0; k < m; k++) {
    cplx v[3];
    for (int q = 0; q < r; q++) v[q] = out[q * m + k] * tw[((long)q * k) * twstride];
    for (int s = 0; s < r; s++) {
      cplx acc = 0;
      for (int q = 0; q < r; q++) acc += v[q] * tw[(((long)q * s * m) % n) * twstride];
      tmp[s * m + k] = acc;
    }
  }
  memcpy(out, tmp, sizeof(cplx) * n);
}

static cplx* make_twiddles(int n, int sign) {
  cplx* tw = (cplx*)malloc(sizeof(cplx) * n);
  for (int k = 0; k < n; k++) {
    double a = 2.0 * PI_ * (double)k / (double)n;
    tw[k] = cos(a) + sign * I * sin(a);
  }
  return tw;
}

/* in-place DFT of n interleaved complex doubles */
void or_dft(double* z, int n, int sign) {
  cplx* tw = make_twiddles(n, sign);
  cplx* a = (cplx*)malloc(sizeof(cplx) * n);
  cplx* o = (cplx*)malloc(sizeof(cplx) * n);
  cplx* t = (cplx*)malloc(sizeof(cplx) * n);
  for (int i = 0; i < n; i++) a[i] = z[2 * i] + I * z[2 * i + 1];
  dft_rec(o, a, n, 1, tw, 1, t);
  for (int i = 0; i < n; i++) {
    z[2 * i] = creal(o[i]);
    z[2 * i + 1] = cimag(o[i]);
  }
  free(tw);
  free(a);
  free(o);
  free(t);
}

/* transform every line of one axis of a [nz][ny][nx] complex block */
static void fft_axis(cplx* a, int nx, int ny, int nz, int axis, int sign) {
  int n = axis == 0 ? nx : axis == 1 ? ny : nz;
  if (n == 1) return;
  long stride = axis == 0 ? 1 : axis == 1 ? nx : (long)nx * ny;
  long nlines = (long)nx * ny * nz / n;
  cplx* tw = make_twiddles(n, sign);
#pragma omp parallel
  {
    cplx* line = (cplx*)malloc(sizeof(cplx) * n);
    cplx* out = (cplx*)malloc(sizeof(cplx) * n);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
    for (long l = 0; l < nlines; l++) {
      long base;
      if (axis == 0)
        base = l * nx;
      else if (axis == 1)
        base = (l % nx) + (l / nx) * (long)nx * ny;
      else
        base = l;
      for (int q = 0; q < n; q++) line[q] = a[base + q * stride];
      dft_rec(out, line, n, 1, tw, 1, tmp);
      for (int q = 0; q < n; q++) a[base + q * stride] = out[q];
    }
    free(line);
    free(out);
    free(tmp);
  }
  free(tw);
}

/* FFT: PHIF(:,:,:,L) = DFT(PHI(:,:,:,L)) (SRC/cgpe3d.f90:1748-1763) */
void or_fft_forward(or_state* s) {
  memcpy(s->phif, s->phi, sizeof(cplx) * s->ms * s->npts);
  for (int m = 0; m < s->ms; m++)
    for (int ax = 0; ax < s->dim; ax++) fft_axis(s->phif + m * s->npts, s->nx, s->ny, s->nz, ax, -1);
}

/* BFT: PHI(:,:,:,L) = IDFT(PHIF(:,:,:,L)) / (NX*NY*NZ) (SRC/cgpe3d.f90:1768-1784) */
void or_fft_backward(or_state* s) {
  memcpy(s->phi, s->phif, sizeof(cplx) * s->ms * s->npts);
  double nn = (double)s->npts;
  for (int m = 0; m < s->ms; m++) {
    cplx* a = s->phi + m * s->npts;
    for (int ax = 0; ax < s->dim; ax++) fft_axis(a, s->nx, s->ny, s->nz, ax, +1);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < s->npts; q++) a[q] = a[q] / nn;
  }
}

/* --------------------------------------------------------------- derive
 * Driver setup: SWITCH_SOC (SRC/cgpe3d.f90:256), ALPHA (277-279), AOSC
 * (334-335) and TAU per spin (337-345); 2D TAU carries sqrt(ALPHAZ/(2 pi))
 * (SRC/cgpe2d.f90:326-336), 1D carries sqrt(ALPHAY*ALPHAZ)/(2 pi)
 * (SRC/cgpe1d.f90:324-332).  Spin-1 converts NATOMS with REAL() (default
 * real), spin-2 with DBLE(). */
int or_derive(const fgpe_input* in, fgpe_params* o, double* aosc_out) {
  memset(o, 0, sizeof(*o));
  o->dim = in->dim;
  o->spin = in->spin;
  for (int d = 0; d < 3; d++) {
    o->n[d] = d < in->dim ? in->n[d] : 1;
    o->d[d] = d < in->dim ? in->d[d] : 1.0;
    o->gamma[d] = d < in->dim ? in->gamma[d] : 0.0;
  }
  o->dt = in->dt;
  o->imag = in->switch_im;
  o->precision = FGPE_FP64;
  o->mag = in->mag;
  int soc = 0;
  for (int d = 0; d < in->dim && d < 3; d++)
    if (fabs(in->gamma[d]) > EPS) soc = 1;
  o->soc = soc;
  o->option_fpc = soc ? 4 : in->option_fpc;
  double nux = in->nu[0];
  o->alpha[0] = in->nu[0] / nux;
  o->alpha[1] = in->nu[1] / nux;
  o->alpha[2] = in->nu[2] / nux;
  double omegam = 2.0 * PI_ * nux;
  double aosc = sqrt(HBAR / (in->mass * AMU * omegam));
  if (aosc_out) *aosc_out = aosc;
  double a0 = in->a[0], a2 = in->a[1], a4 = in->a[2];
  double f;
  if (in->dim == 3)
    f = 1.0;
  else if (in->dim == 2)
    f = sqrt(o->alpha[2] / (2.0 * PI_));
  else
    f = sqrt(o->alpha[1] * o->alpha[2]) / (2.0 * PI_);
  if (in->spin == 1) {
    double na = (double)(float)in->natoms;
    o->tau[0] = 2.0 * f * (4.0 * PI_ * na * (a0 * AU + 2.0 * a2 * AU) / (3.0 * aosc));
    o->tau[1] = 2.0 * f * (4.0 * PI_ * na * (a2 * AU - a0 * AU) / (3.0 * aosc));
    o->tau[2] = 0.0;
  } else {
    double na = (double)in->natoms;
    o->tau[0] = 2.0 * (f * 4.0 * PI_ * na * (4.0 * a2 * AU + 3.0 * a4 * AU) / (7.0 * aosc));
    o->tau[1] = 2.0 * (f * 4.0 * PI_ * na * (a4 * AU - a2 * AU) / (7.0 * aosc));
    o->tau[2] = 2.0 * (f * 4.0 * PI_ * na * (7.0 * a0 * AU - 10.0 * a2 * AU + 3.0 * a4 * AU) / (7.0 * aosc));
  }
  return 0;
}

/* ---------------------------------------------------------------- create
 * Grids and potential of INITIALIZE (SRC/cgpe3d.f90:702-753; 2D
 * SRC/cgpe2d.f90:669-695; 1D SRC/cgpe1d.f90:692-711). */
static void make_axis(int n, double d, double* x, double* x2, double* k, double* lout) {
  double l = n * d;
  *lout = l;
  for (int i = 1; i <= 1 + n / 2; i++) k[i - 1] = (i - 1.0) * 2.0 * PI_ / l;
  for (int i = 1; i <= n / 2 - 1; i++) k[i + n / 2] = -k[n / 2 - i];
  for (int i = 1; i <= n; i++) {
    x[i - 1] = (-1.0 + 2.0 * (i - 1.0) / n) * l / 2.0;
    x2[i - 1] = x[i - 1] * x[i - 1];
  }
}

or_state* or_create(const fgpe_params* p) {
  or_state* s = (or_state*)calloc(1, sizeof(or_state));
  s->p = *p;
  s->dim = p->dim;
  s->spin = p->spin;
  s->ms = 2 * p->spin + 1;
  s->nx = p->n[0];
  s->ny = p->dim >= 2 ? p->n[1] : 1;
  s->nz = p->dim >= 3 ? p->n[2] : 1;
  s->npts = (long)s->nx * s->ny * s->nz;
  s->phi = (cplx*)calloc((size_t)s->ms * s->npts, sizeof(cplx));
  s->phif = (cplx*)calloc((size_t)s->ms * s->npts, sizeof(cplx));
  s->x = calloc(s->nx, sizeof(double));
  s->x2 = calloc(s->nx, sizeof(double));
  s->kx = calloc(s->nx, sizeof(double));
  s->y = calloc(s->ny, sizeof(double));
  s->y2 = calloc(s->ny, sizeof(double));
  s->ky = calloc(s->ny, sizeof(double));
  s->z = calloc(s->nz, sizeof(double));
  s->z2 = calloc(s->nz, sizeof(double));
  s->kz = calloc(s->nz, sizeof(double));
  make_axis(s->nx, p->d[0], s->x, s->x2, s->kx, &s->lx);
  if (s->dim >= 2) make_axis(s->ny, p->d[1], s->y, s->y2, s->ky, &s->ly);
  if (s->dim >= 3) make_axis(s->nz, p->d[2], s->z, s->z2, s->kz, &s->lz);
  s->v = calloc(s->npts, sizeof(double));
  s->r2 = calloc(s->npts, sizeof(double));
  const double ax = p->alpha[0], ay = p->alpha[1], az = p->alpha[2];
  for (int k = 0; k < s->nz; k++)
    for (int j = 0; j < s->ny; j++)
      for (int i = 0; i < s->nx; i++) {
        long q = i + (long)s->nx * (j + (long)s->ny * k);
        if (s->dim == 3) {
          s->r2[q] = s->x2[i] + s->y2[j] + s->z2[k];
          s->v[q] = ax * ax * s->x2[i] + ay * ay * s->y2[j] + az * az * s->z2[k];
        } else if (s->dim == 2) {
          s->r2[q] = s->x2[i] + s->y2[j];
          s->v[q] = ax * ax * s->x2[i] + ay * ay * s->y2[j];
        } else {
          s->r2[q] = s->x2[i];
          s->v[q] = ax * ax * s->x2[i];
        }
      }
  /* CDT (SRC/cgpe3d.f90:305,319) */
  s->cdt = p->imag ? -I * p->dt : p->dt + 0.0 * I;
  /* Newton start (SRC/cgpe3d.f90:443-444) */
  s->sig[0] = 20.0;
  s->sig[1] = 20.0;
  return s;
}

void or_destroy(or_state* s) {
  if (!s) return;
  free(s->phi);
  free(s->phif);
  free(s->x);
  free(s->x2);
  free(s->kx);
  free(s->y);
  free(s->y2);
  free(s->ky);
  free(s->z);
  free(s->z2);
  free(s->kz);
  free(s->v);
  free(s->r2);
  free(s);
}

void or_set_phi(or_state* s, const double* phi) {
  for (long q = 0; q < s->ms * s->npts; q++) s->phi[q] = phi[2 * q] + I * phi[2 * q + 1];
}

void or_get_phi(const or_state* s, double* phi) {
  for (long q = 0; q < s->ms * s->npts; q++) {
    phi[2 * q] = creal(s->phi[q]);
    phi[2 * q + 1] = cimag(s->phi[q]);
  }
}

void or_get_phif(const or_state* s, double* phif) {
  for (long q = 0; q < s->ms * s->npts; q++) {
    phif[2 * q] = creal(s->phif[q]);
    phif[2 * q + 1] = cimag(s->phif[q]);
  }
}

void or_set_phif(or_state* s, const double* phif) {
  for (long q = 0; q < s->ms * s->npts; q++) s->phif[q] = phif[2 * q] + I * phif[2 * q + 1];
}

/* --------------------------------------------------------- initial guess
 * INITIALIZE's guess branch: 3D SRC/cgpe3d.f90:753-833, 2D
 * SRC/cgpe2d.f90:697-771, 1D SRC/cgpe1d.f90:707-786.  The quirks are kept:
 * 3D spin-1 Gaussian ignores the anisotropy, 1D/2D spin-2 Gaussians carry
 * sqrt(0.2), the 2D spin-1 ferro test raises to the 8th power and the 1D
 * spin-2 ferro test groups 3*(TAU0+4*TAU1/4)**0.33.  TAU(2) is zero for
 * spin-1 (module storage, never assigned). */
void or_initial_guess(or_state* s) {
  const fgpe_params* p = &s->p;
  const double ax = p->alpha[0], ay = p->alpha[1], az = p->alpha[2];
  const double t0 = p->tau[0], t1 = p->tau[1], t2 = p->spin == 1 ? 0.0 : p->tau[2];
  const double mag = p->mag;
  const int fpc = p->option_fpc;
  const int dim = s->dim;
  double norm_g = dim == 3 ? pow(PI_, 0.75) : dim == 2 ? sqrt(PI_) : sqrt(sqrt(PI_));
  for (int k = 0; k < s->nz; k++)
    for (int j = 0; j < s->ny; j++)
      for (int i = 0; i < s->nx; i++) {
        long q = i + (long)s->nx * (j + (long)s->ny * k);
        double x2 = s->x2[i], y2 = s->y2[j], z2 = s->z2[k];
        double tmp = dim == 3 ? ax * x2 + ay * y2 + az * z2 : dim == 2 ? ax * x2 + ay * y2 : ax * x2;
        double vv = dim == 3 ? ax * ax * x2 + ay * ay * y2 + az * az * z2
                             : dim == 2 ? ax * ax * x2 + ay * ay * y2 : ax * ax * x2;
        double g = exp(-tmp / 2.0) / norm_g;
        double c[5] = {0, 0, 0, 0, 0};
        if (s->spin == 1) {
          if (fpc == 1) {
            double tf, thr, thr2;
            if (dim == 3) {
              thr = pow(15.0 * (t0 + t1) / (8.0 * PI_), 0.4);
              thr2 = pow(15.0 * (t0 + t1) / (8.0 * PI_), 0.2);
              tf = (thr - vv >= 0.0) ? (thr - vv) / (t0 + t1) : 0.0;
            } else if (dim == 2) {
              double cond = pow(4.0 * (t0 + t1) / (2.0 * PI_), 8.0);
              thr = pow(4.0 * (t0 + t1) / (2.0 * PI_), 0.5);
              thr2 = pow(4.0 * (t0 + t1) / (2.0 * PI_), 0.25);
              tf = (cond - vv >= 0.0) ? (thr - vv) / (t0 + t1) : 0.0;
            } else {
              thr = pow(3.0 * (t0 + t2) / 4.0, 0.67);
              thr2 = pow(3.0 * (t0 + t2) / 4.0, 0.33);
              tf = (thr - vv >= 0.0) ? (thr - vv) / (t0 + t2) : 0.0;
            }
            double b = thr2 >= 10.0 ? sqrt(tf) : g;
            c[0] = ((1.0 + mag) / 2.0) * b;
            c[1] = sqrt((1.0 - mag * mag) / 2.0) * b;
            c[2] = ((1.0 - mag) / 2.0) * b;
          } else if (fpc == 2) {
            c[0] = sqrt((1.0 + mag) / 2.0) * g;
            c[1] = 0.0;
            c[2] = sqrt((1.0 - mag) / 2.0) * g;
          } else {
            double gg = dim == 3 ? exp(-(x2 + y2 + z2) / 2.0) / pow(PI_, 0.75) : g;
            c[0] = c[1] = c[2] = gg;
          }
        } else {
          if (fpc == 1) {
            double tf, thr, thr2;
            double tt = t0 + 4.0 * t1;
            if (dim == 3) {
              thr = pow(15.0 * tt / (8.0 * PI_), 0.4);
              thr2 = pow(4.0 * tt / (2.0 * PI_), 0.25);
            } else if (dim == 2) {
              thr = pow(4.0 * tt / (2.0 * PI_), 0.5);
              thr2 = pow(4.0 * tt / (2.0 * PI_), 0.25);
            } else {
              thr = pow(3.0 * tt / 4.0, 0.67);
              thr2 = 3.0 * pow(t0 + 4.0 * t1 / 4.0, 0.33);
            }
            tf = (thr - vv >= 0.0) ? (thr - vv) / tt : 0.0;
            double b = thr2 >= 10.0 ? sqrt(tf) : g;
            c[0] = ((2.0 + mag) * (2.0 + mag) / 16.0) * b;
            c[1] = (sqrt(4.0 - mag * mag) * (2.0 + mag) / 8.0) * b;
            c[2] = (sqrt(1.5) * (4.0 - mag * mag) / 8.0) * b;
            c[3] = (sqrt(4.0 - mag * mag) * (2.0 - mag) / 8.0) * b;
            c[4] = ((2.0 - mag) * (2.0 - mag) / 16.0) * b;
          } else if (fpc == 2) {
            c[0] = (sqrt(2.0 + mag) / 2.0) * g;
            c[4] = (sqrt(2.0 - mag) / 2.0) * g;
          } else if (fpc == 3) {
            c[0] = sqrt((1.0 + mag) / 3.0) * g;
            c[3] = sqrt((2.0 - mag) / 3.0) * g;
          } else {
            double f = dim == 3 ? 1.0 : sqrt(0.2);
            for (int m = 0; m < 5; m++) c[m] = f * g;
          }
        }
        for (int m = 0; m < s->ms; m++) s->phi[m * s->npts + q] = c[m];
      }
}

/* --------------------------------------------------------------------- KE
 * 3D: exp(-i CDT (kx^2+ky^2+kz^2 - 2 L GAMMAZ kz)) per component L = -S..S
 * (SRC/cgpe3d.f90:875-899); 2D exp(-i CDT (kx^2+ky^2)) (SRC/cgpe2d.f90:812-829);
 * 1D exp(-i CDT kx^2) (SRC/cgpe1d.f90:826-842). */
void or_ke(or_state* s) {
  const int S = s->spin;
  const double gz = s->p.gamma[2];
  for (int L = -S; L <= S; L++) {
    cplx* a = s->phif + (L + S) * s->npts;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s->nz; k++)
      for (int j = 0; j < s->ny; j++)
        for (int i = 0; i < s->nx; i++) {
          long q = i + (long)s->nx * (j + (long)s->ny * k);
          double arg;
          if (s->dim == 3)
            arg = s->kx[i] * s->kx[i] + s->ky[j] * s->ky[j] + s->kz[k] * s->kz[k] - 2.0 * L * gz * s->kz[k];
          else if (s->dim == 2)
            arg = s->kx[i] * s->kx[i] + s->ky[j] * s->ky[j];
          else
            arg = s->kx[i] * s->kx[i];
          a[q] = cexp(-I * s->cdt * arg) * a[q];
        }
  }
}

/* SOC eigenbasis A(beta) of SRC/cgpe3d.f90:930-940 (spin-1) and 963-991
 * (spin-2); the 2D code builds the same matrices (SRC/cgpe2d.f90:1578-1638).
 * a[r*ms+c] = A(r+1,c+1). */
static void soc_a_beta(int spin, double beta, cplx* a) {
  cplx e1 = cexp(-I * beta), e2 = cexp(-2.0 * I * beta);
  if (spin == 1) {
#define A3(r, c) a[((r)-1) * 3 + (c)-1]
    A3(1, 1) = 0.5 * e2;
    A3(2, 1) = -sqrt(0.5) * e1;
    A3(3, 1) = 0.5;
    A3(1, 2) = -sqrt(0.5) * e2;
    A3(2, 2) = 0.0;
    A3(3, 2) = sqrt(0.5);
    A3(1, 3) = A3(1, 1);
    A3(2, 3) = -A3(2, 1);
    A3(3, 3) = A3(3, 1);
#undef A3
  } else {
#define A5(r, c) a[((r)-1) * 5 + (c)-1]
    A5(1, 1) = 0.25 * e2;
    A5(2, 1) = -0.5 * e1;
    A5(3, 1) = sqrt(0.375);
    A5(4, 1) = conj(A5(2, 1));
    A5(5, 1) = conj(A5(1, 1));
    A5(1, 2) = -0.5 * e2;
    A5(2, 2) = 0.5 * e1;
    A5(3, 2) = 0.0;
    A5(4, 2) = -conj(A5(2, 2));
    A5(5, 2) = -conj(A5(1, 2));
    A5(1, 3) = sqrt(0.375) * e2;
    A5(2, 3) = A5(3, 2);
    A5(3, 3) = -0.5;
    A5(4, 3) = A5(3, 2);
    A5(5, 3) = conj(A5(1, 3));
    A5(1, 4) = -0.5 * e2;
    A5(2, 4) = -0.5 * e1;
    A5(3, 4) = A5(3, 2);
    A5(4, 4) = -conj(A5(2, 4));
    A5(5, 4) = -conj(A5(1, 4));
    A5(1, 5) = 0.25 * e2;
    A5(2, 5) = 0.5 * e1;
    A5(3, 5) = sqrt(0.375);
    A5(4, 5) = conj(A5(2, 5));
    A5(5, 5) = conj(A5(1, 5));
#undef A5
  }
}

/* 1D constant eigenbasis (SRC/cgpe1d.f90:627-675), real. */
static void soc_a_1d(int spin, double* a) {
  if (spin == 1) {
    double s5 = sqrt(0.5);
    double v[9] = {0.5, -s5, 0.5, /* row 1: A(1,1..3) */
                   -s5, 0.0, s5,  /* row 2 */
                   0.5, s5,  0.5};
    memcpy(a, v, sizeof(v));
  } else {
    double r = sqrt(0.375);
    double v[25] = {0.25, -0.5, r,   -0.5, 0.25, /* row 1: A(1,1..5) */
                    -0.5, 0.5,  0.0, -0.5, 0.5,  /* row 2 */
                    r,    0.0,  -0.5, 0.0, r,    /* row 3 */
                    -0.5, -0.5, 0.0, 0.5,  0.5,  /* row 4 */
                    0.25, 0.5,  r,   0.5,  0.25};
    memcpy(a, v, sizeof(v));
  }
}

/* SOC factor at one k-point, for unit checks: 1D (dim == 1) uses the
 * constant basis and exponent 2 L GAMMAX kx (SRC/cgpe1d.f90:1516-1520),
 * 2D/3D the A(beta) basis (SRC/cgpe3d.f90:956-998). v, out: ms interleaved
 * complex. */
void or_soc_point(int spin, int dim, double kx, double ky, double gx, double gy, double cdt_re, double cdt_im,
                  const double* v, double* out) {
  const int S = spin, ms = 2 * spin + 1;
  cplx cdt = cdt_re + I * cdt_im;
  cplx x[5], w[5];
  for (int m = 0; m < ms; m++) x[m] = v[2 * m] + I * v[2 * m + 1];
  if (dim == 1) {
    double a[25];
    soc_a_1d(S, a);
    for (int r = 0; r < ms; r++) {
      w[r] = 0;
      for (int c = 0; c < ms; c++) w[r] += a[c * ms + r] * x[c];
    }
    for (int L = -S; L <= S; L++) w[L + S] = cexp(-I * cdt * (2.0 * L * gx * kx)) * w[L + S];
    for (int r = 0; r < ms; r++) {
      x[r] = 0;
      for (int c = 0; c < ms; c++) x[r] += a[r * ms + c] * w[c];
    }
  } else {
    double alpha = 2.0 * sqrt(gx * gx * kx * kx + gy * gy * ky * ky);
    double beta = alpha <= EPS ? 0.0 : atan2(gy * ky, gx * kx);
    cplx a[25];
    soc_a_beta(S, beta, a);
    for (int r = 0; r < ms; r++) {
      w[r] = 0;
      for (int c = 0; c < ms; c++) w[r] += conj(a[c * ms + r]) * x[c];
    }
    for (int L = -S; L <= S; L++) w[L + S] *= cexp(-I * cdt * (L * alpha));
    for (int r = 0; r < ms; r++) {
      x[r] = 0;
      for (int c = 0; c < ms; c++) x[r] += a[r * ms + c] * w[c];
    }
  }
  for (int m = 0; m < ms; m++) {
    out[2 * m] = creal(x[m]);
    out[2 * m + 1] = cimag(x[m]);
  }
}

/* -------------------------------------------------------------------- SOC
 * 2D/3D: per (kx,ky), ALPHA = 2 sqrt(gx^2 kx^2 + gy^2 ky^2), BETA = atan2 (0
 * when ALPHA <= EPS), PHIF <- A diag(exp(-i CDT L ALPHA)) A^H PHIF
 * (SRC/cgpe3d.f90:904-1004, SRC/cgpe2d.f90:1554-1650).
 * 1D: PHIF <- A diag(exp(-i CDT 2 L GAMMAX kx)) A^T PHIF with the constant A
 * (SRC/cgpe1d.f90:1504-1523). */
void or_soc(or_state* s) {
  const int S = s->spin, ms = s->ms;
  const double gx = s->p.gamma[0], gy = s->p.gamma[1];
  const cplx cdt = s->cdt;
  if (s->dim == 1) {
    double a[25];
    soc_a_1d(S, a);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->nx; i++) {
      cplx v[5], w[5];
      for (int m = 0; m < ms; m++) v[m] = s->phif[m * s->npts + i];
      for (int r = 0; r < ms; r++) { /* A^T v */
        w[r] = 0;
        for (int c = 0; c < ms; c++) w[r] += a[c * ms + r] * v[c];
      }
      for (int L = -S; L <= S; L++) w[L + S] = cexp(-I * cdt * (2.0 * L * gx * s->kx[i])) * w[L + S];
      for (int r = 0; r < ms; r++) {
        v[r] = 0;
        for (int c = 0; c < ms; c++) v[r] += a[r * ms + c] * w[c];
      }
      for (int m = 0; m < ms; m++) s->phif[m * s->npts + i] = v[m];
    }
    return;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < s->ny; j++)
    for (int i = 0; i < s->nx; i++) {
      double alpha = 2.0 * sqrt(gx * gx * s->kx[i] * s->kx[i] + gy * gy * s->ky[j] * s->ky[j]);
      double beta = alpha <= EPS ? 0.0 : atan2(gy * s->ky[j], gx * s->kx[i]);
      cplx a[25];
      soc_a_beta(S, beta, a);
      cplx ph[5];
      for (int L = -S; L <= S; L++) ph[L + S] = cexp(-I * cdt * (L * alpha));
      for (int k = 0; k < s->nz; k++) {
        long q = i + (long)s->nx * (j + (long)s->ny * k);
        cplx v[5], w[5];
        for (int m = 0; m < ms; m++) v[m] = s->phif[m * s->npts + q];
        for (int r = 0; r < ms; r++) { /* A^H v */
          w[r] = 0;
          for (int c = 0; c < ms; c++) w[r] += conj(a[c * ms + r]) * v[c];
        }
        for (int m = 0; m < ms; m++) w[m] *= ph[m];
        for (int r = 0; r < ms; r++) {
          v[r] = 0;
          for (int c = 0; c < ms; c++) v[r] += a[r * ms + c] * w[c];
        }
        for (int m = 0; m < ms; m++) s->phif[m * s->npts + q] = v[m];
      }
    }
}

/* ------------------------------------------------------ FXYZ and MAT_C
 * F_x = Re(phi^H SX phi), F_y = Re(phi^H SY phi), F_z = sum SZ |phi|^2 with the
 * reference's SX, SY, SZ (SRC/cgpe3d.f90:676-697,1070-1093); the spin-2
 * off-diagonal couplings C12..C45 (SRC/cgpe3d.f90:1099-1124). */
static void fxyz(int spin, const cplx* ph, double* fx, double* fy, double* fz) {
  if (spin == 1) {
    double s5 = sqrt(0.5);
    /* SX = sqrt(.5)[[0,1,0],[1,0,1],[0,1,0]]; SY = (i/sqrt2)*[[0,1,0],[-1,0,1],[0,-1,0]] */
    cplx sxp[3] = {s5 * ph[1], s5 * (ph[0] + ph[2]), s5 * ph[1]};
    cplx syp[3] = {(I / sqrt(2.0)) * ph[1], (I / sqrt(2.0)) * (-ph[0] + ph[2]), (I / sqrt(2.0)) * (-ph[1])};
    cplx ax = 0, ay = 0;
    for (int m = 0; m < 3; m++) {
      ax += conj(ph[m]) * sxp[m];
      ay += conj(ph[m]) * syp[m];
    }
    *fx = creal(ax);
    *fy = creal(ay);
    double a1 = creal(ph[0]) * creal(ph[0]) + cimag(ph[0]) * cimag(ph[0]);
    double a3 = creal(ph[2]) * creal(ph[2]) + cimag(ph[2]) * cimag(ph[2]);
    *fz = a1 - a3;
  } else {
    double r = sqrt(1.5);
    /* SX rows: [0,1,0,0,0],[1,0,r,0,0],[0,r,0,r,0],[0,0,r,0,1],[0,0,0,1,0]
       SY = i*[[0,-1,0,0,0],[1,0,-r,0,0],[0,r,0,-r,0],[0,0,r,0,-1],[0,0,0,1,0]] */
    cplx sxp[5] = {ph[1], ph[0] + r * ph[2], r * (ph[1] + ph[3]), r * ph[2] + ph[4], ph[3]};
    cplx syp[5] = {I * (-ph[1]), I * (ph[0] - r * ph[2]), I * (r * ph[1] - r * ph[3]), I * (r * ph[2] - ph[4]),
                   I * ph[3]};
    cplx ax = 0, ay = 0;
    double az = 0;
    const double sz[5] = {2, 1, 0, -1, -2};
    for (int m = 0; m < 5; m++) {
      ax += conj(ph[m]) * sxp[m];
      ay += conj(ph[m]) * syp[m];
      az += sz[m] * (creal(ph[m]) * creal(ph[m]) + cimag(ph[m]) * cimag(ph[m]));
    }
    *fx = creal(ax);
    *fy = creal(ay);
    *fz = az;
  }
}

static void mat_c(const cplx* ph, cplx fm, const double* tau, cplx* c) {
  /* c = {C12, C13, C23, C34, C35, C45} */
  c[0] = tau[1] * fm - 0.4 * tau[2] * ph[3] * conj(ph[4]);
  c[1] = 0.2 * tau[2] * ph[2] * conj(ph[4]);
  c[2] = sqrt(1.5) * tau[1] * fm - 0.2 * tau[2] * ph[2] * conj(ph[3]);
  c[3] = sqrt(1.5) * tau[1] * fm - 0.2 * tau[2] * ph[1] * conj(ph[2]);
  c[4] = 0.2 * tau[2] * ph[0] * conj(ph[2]);
  c[5] = tau[1] * fm - 0.4 * tau[2] * ph[0] * conj(ph[1]);
}

/* Hbar phi with Hbar(1,2)=C12, (1,3)=C13, (2,3)=C23, (3,4)=C34, (3,5)=C35,
 * (4,5)=C45 and the conjugates below the diagonal. */
static void hmul(const cplx* c, const cplx* v, cplx* o) {
  o[0] = c[0] * v[1] + c[1] * v[2];
  o[1] = conj(c[0]) * v[0] + c[2] * v[2];
  o[2] = conj(c[1]) * v[0] + conj(c[2]) * v[1] + c[3] * v[3] + c[4] * v[4];
  o[3] = conj(c[3]) * v[2] + c[5] * v[4];
  o[4] = conj(c[4]) * v[2] + conj(c[5]) * v[3];
}

/* Cyclic complex Jacobi eigen-decomposition of a Hermitian n x n matrix
 * (row-major a, overwritten); eigenvectors in the columns of vec. */
static void jacobi_herm(int n, cplx* a, cplx* vec, double* w) {
  for (int i = 0; i < n * n; i++) vec[i] = 0;
  for (int i = 0; i < n; i++) vec[i * n + i] = 1;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0, diag = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double v = cabs(a[i * n + j]);
        if (i != j)
          off += v * v;
        else
          diag += v * v;
      }
    if (off <= 1e-34 * (diag + off) || off < 1e-300) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        cplx g = a[p * n + q];
        double ag = cabs(g);
        if (ag < 1e-300) continue;
        cplx e = g / ag;
        double theta = (creal(a[q * n + q]) - creal(a[p * n + p])) / (2.0 * ag);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        /* J = U P with U_qq = conj(e):  J[:,p] = c e_p - sn conj(e) e_q,
           J[:,q] = sn e_p + c conj(e) e_q;  A <- J^H A J, V <- V J */
        cplx jpp = c, jqp = -sn * conj(e), jpq = sn, jqq = c * conj(e);
        for (int k = 0; k < n; k++) {
          cplx akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = akp * jpp + akq * jqp;
          a[k * n + q] = akp * jpq + akq * jqq;
          cplx vkp = vec[k * n + p], vkq = vec[k * n + q];
          vec[k * n + p] = vkp * jpp + vkq * jqp;
          vec[k * n + q] = vkp * jpq + vkq * jqq;
        }
        for (int k = 0; k < n; k++) {
          cplx apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = conj(jpp) * apk + conj(jqp) * aqk;
          a[q * n + k] = conj(jpq) * apk + conj(jqq) * aqk;
        }
        a[p * n + q] = 0;
        a[q * n + p] = 0;
      }
  }
  for (int i = 0; i < n; i++) w[i] = creal(a[i * n + i]);
}

/* exp(-i CDT H) v for a 5x5 Hermitian H given row-major as interleaved complex
 * (h[2*(r*5+c)]), through its eigen-decomposition, as SE does with ZHEEV
 * (SRC/cgpe3d.f90:1725-1734). */
void or_expm_herm5(const double* h, double cdt_re, double cdt_im, const double* v, double* out) {
  cplx a[25], vec[25], x[5], t[5];
  double w[5];
  for (int i = 0; i < 25; i++) a[i] = h[2 * i] + I * h[2 * i + 1];
  for (int i = 0; i < 5; i++) x[i] = v[2 * i] + I * v[2 * i + 1];
  jacobi_herm(5, a, vec, w);
  cplx cdt = cdt_re + I * cdt_im;
  for (int r = 0; r < 5; r++) {
    t[r] = 0;
    for (int c = 0; c < 5; c++) t[r] += conj(vec[c * 5 + r]) * x[c];
    t[r] *= cexp(-I * cdt * w[r]);
  }
  for (int r = 0; r < 5; r++) {
    cplx s = 0;
    for (int c = 0; c < 5; c++) s += vec[r * 5 + c] * t[c];
    out[2 * r] = creal(s);
    out[2 * r + 1] = cimag(s);
  }
}

/* --------------------------------------------------------------------- SE
 * spin-1: closed-form exp(-i CDT H) (SRC/cgpe3d.f90:1637-1672; 2D
 * SRC/cgpe2d.f90:1448-1479 with .GT.EPS; 1D SRC/cgpe1d.f90:1406-1435 with
 * .GE.EPS).
 * spin-2: FXYZ, MAT_C, predictor with CDT/2 in 1D (SRC/cgpe1d.f90:1443-1447)
 * and CDT in 2D/3D (SRC/cgpe3d.f90:1681-1686), averaged couplings, then
 * V diag(exp(-i CDT w)) V^H per point (SRC/cgpe3d.f90:1689-1741). */
void or_se(or_state* s) {
  const cplx cdt = s->cdt;
  const double* tau = s->p.tau;
  const long np = s->npts;
  if (s->spin == 1) {
    const int ge = s->dim == 1;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < np; q++) {
      cplx p1 = s->phi[q], p2 = s->phi[np + q], p3 = s->phi[2 * np + q];
      cplx a1 = tau[1] * p2 * conj(p3);
      cplx b1 = tau[1] * p2 * conj(p1);
      double ra = cabs(a1), rb = cabs(b1);
      cplx om = cdt * sqrt(ra * ra + rb * rb);
      cplx delta, kappa;
      int big = ge ? (cabs(om) >= EPS) : (cabs(om) > EPS);
      if (big) {
        delta = (ccos(om) - 1.0) / (om * om);
        kappa = -I * csin(om) / om;
      } else {
        delta = -0.5;
        kappa = -I;
      }
      cplx t1 = (delta * cdt * cdt) * (a1 * conj(a1) * p1 + a1 * conj(b1) * p3) + kappa * cdt * a1 * p2;
      cplx t2 = (delta * cdt * cdt) * (a1 * conj(a1) + b1 * conj(b1)) * p2 + kappa * cdt * (conj(a1) * p1 + conj(b1) * p3);
      cplx t3 = (delta * cdt * cdt) * (conj(a1) * b1 * p1 + b1 * conj(b1) * p3) + kappa * cdt * b1 * p2;
      s->phi[q] = p1 + t1;
      s->phi[np + q] = p2 + t2;
      s->phi[2 * np + q] = p3 + t3;
    }
    return;
  }
  const double cpred = s->dim == 1 ? 0.5 : 1.0;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < np; q++) {
    cplx ph[5], p1[5], c0[6], c1[6], hv[5];
    for (int m = 0; m < 5; m++) ph[m] = s->phi[m * np + q];
    double fx, fy, fz;
    fxyz(2, ph, &fx, &fy, &fz);
    mat_c(ph, fx - I * fy, tau, c0);
    hmul(c0, ph, hv);
    for (int m = 0; m < 5; m++) p1[m] = ph[m] - (I * cdt * cpred) * hv[m];
    fxyz(2, p1, &fx, &fy, &fz);
    mat_c(p1, fx - I * fy, tau, c1);
    for (int i = 0; i < 6; i++) c0[i] = (c0[i] + c1[i]) / 2.0;
    cplx a[25] = {0}, vec[25], t[5];
    double w[5];
    a[0 * 5 + 1] = c0[0];
    a[0 * 5 + 2] = c0[1];
    a[1 * 5 + 0] = conj(c0[0]);
    a[1 * 5 + 2] = c0[2];
    a[2 * 5 + 0] = conj(c0[1]);
    a[2 * 5 + 1] = conj(c0[2]);
    a[2 * 5 + 3] = c0[3];
    a[2 * 5 + 4] = c0[4];
    a[3 * 5 + 2] = conj(c0[3]);
    a[3 * 5 + 4] = c0[5];
    a[4 * 5 + 2] = conj(c0[4]);
    a[4 * 5 + 3] = conj(c0[5]);
    jacobi_herm(5, a, vec, w);
    for (int r = 0; r < 5; r++) {
      t[r] = 0;
      for (int c = 0; c < 5; c++) t[r] += conj(vec[c * 5 + r]) * ph[c];
      t[r] *= cexp(-I * cdt * w[r]);
    }
    for (int r = 0; r < 5; r++) {
      cplx acc = 0;
      for (int c = 0; c < 5; c++) acc += vec[r * 5 + c] * t[c];
      s->phi[r * np + q] = acc;
    }
  }
}

/* --------------------------------------------------------------------- SP
 * Diagonal trap + density phases, spin-1 SRC/cgpe3d.f90:1023-1040, spin-2
 * SRC/cgpe3d.f90:1041-1063 (same in 2D SRC/cgpe2d.f90:834-886 and 1D
 * SRC/cgpe1d.f90:847-896). */
void or_sp(or_state* s) {
  const cplx cdt = s->cdt;
  const double* tau = s->p.tau;
  const long np = s->npts;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < np; q++) {
    double a[5], e[5];
    cplx ph[5];
    double rho = 0;
    for (int m = 0; m < s->ms; m++) {
      ph[m] = s->phi[m * np + q];
      a[m] = creal(ph[m]) * creal(ph[m]) + cimag(ph[m]) * cimag(ph[m]);
      rho += a[m];
    }
    double v = s->v[q];
    if (s->spin == 1) {
      e[0] = v + tau[0] * rho + tau[1] * (a[0] + a[1] - a[2]);
      e[1] = v + tau[0] * rho + tau[1] * (a[0] + a[2]);
      e[2] = v + tau[0] * rho + tau[1] * (a[2] + a[1] - a[0]);
    } else {
      double fz = 2.0 * a[0] + a[1] - a[3] - 2.0 * a[4];
      e[0] = v + tau[0] * rho + 2.0 * tau[1] * fz + 0.4 * tau[2] * a[4];
      e[1] = v + tau[0] * rho + tau[1] * fz + 0.4 * tau[2] * a[3];
      e[2] = v + tau[0] * rho + 0.2 * tau[2] * a[2];
      e[3] = v + tau[0] * rho - tau[1] * fz + 0.4 * tau[2] * a[1];
      e[4] = v + tau[0] * rho - 2.0 * tau[1] * fz + 0.4 * tau[2] * a[0];
    }
    for (int m = 0; m < s->ms; m++) s->phi[m * np + q] = ph[m] * cexp((-I * cdt) * e[m]);
  }
}

/* ---------------------------------------------------- nested Simpson helpers
 * The reference integrates x innermost, then y, then z
 * (SRC/cgpe3d.f90:1156-1172). f is [nz][ny][nx]. */
static double nested_simpson(const or_state* s, const double* f) {
  const double dx = s->p.d[0], dy = s->p.d[1], dz = s->p.d[2];
  if (s->dim == 1) return or_simpson(f, s->nx, dx);
  double* t1 = (double*)malloc(sizeof(double) * s->nz);
#pragma omp parallel
  {
    double* t2 = (double*)malloc(sizeof(double) * s->ny);
#pragma omp for schedule(static)
    for (int k = 0; k < s->nz; k++) {
      for (int j = 0; j < s->ny; j++) t2[j] = or_simpson(f + (long)s->nx * (j + (long)s->ny * k), s->nx, dx);
      t1[k] = or_simpson(t2, s->ny, dy);
    }
    free(t2);
  }
  double r = s->dim == 2 ? t1[0] : or_simpson(t1, s->nz, dz);
  free(t1);
  return r;
}

/* NORMC (SRC/cgpe3d.f90:1129-1173) */
void or_normc(const or_state* s, double* norm) {
  double* f = (double*)malloc(sizeof(double) * s->npts);
  for (int m = 0; m < s->ms; m++) {
    const cplx* a = s->phi + m * s->npts;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < s->npts; q++) f[q] = creal(a[q]) * creal(a[q]) + cimag(a[q]) * cimag(a[q]);
    norm[m] = nested_simpson(s, f);
  }
  free(f);
}

/* NORMT (SRC/cgpe3d.f90:1178-1219) */
double or_normt(or_state* s) {
  double* f = (double*)malloc(sizeof(double) * s->npts);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < s->npts; q++) {
    double acc = 0;
    for (int m = 0; m < s->ms; m++) {
      cplx z = s->phi[m * s->npts + q];
      acc += creal(z) * creal(z) + cimag(z) * cimag(z);
    }
    f[q] = acc;
  }
  double tn = sqrt(nested_simpson(s, f));
  free(f);
  for (long q = 0; q < s->ms * s->npts; q++) s->phi[q] = s->phi[q] / tn;
  return tn;
}

/* RAD (SRC/cgpe3d.f90:1224-1268) */
void or_rad(const or_state* s, double* rms) {
  double* f = (double*)malloc(sizeof(double) * s->npts);
  for (int m = 0; m < s->ms; m++) {
    const cplx* a = s->phi + m * s->npts;
#pragma omp parallel for schedule(static)
    for (long q = 0; q < s->npts; q++) f[q] = s->r2[q] * (creal(a[q]) * creal(a[q]) + cimag(a[q]) * cimag(a[q]));
    rms[m] = sqrt(nested_simpson(s, f));
  }
  free(f);
}

/* ------------------------------------------------------------ projection
 * NEWTON/USRFUN (SRC/cgpe3d.f90:1516-1588) */
static void usrfun(const double* x, double* fv, double* fj, const double* n, double mag, double dt) {
  double e1 = exp(2.0 * (x[0] + 2.0 * x[1]) * dt), e2 = exp(2.0 * (x[0] + x[1]) * dt), e3 = exp(2.0 * x[0] * dt),
         e4 = exp(2.0 * (x[0] - x[1]) * dt), e5 = exp(2.0 * (x[0] - 2.0 * x[1]) * dt);
  fv[0] = e1 * n[0] + e2 * n[1] + e3 * n[2] + e4 * n[3] + e5 * n[4] - 1.0;
  fv[1] = 2.0 * e1 * n[0] + e2 * n[1] - e4 * n[3] - 2.0 * e5 * n[4] - mag;
  /* fj[0]=FJAC(1,1), fj[1]=FJAC(1,2), fj[2]=FJAC(2,1), fj[3]=FJAC(2,2) */
  fj[0] = (e1 * n[0] + e2 * n[1] + e3 * n[2] + e4 * n[3] + e5 * n[4]) * 2.0 * dt;
  fj[1] = e1 * n[0] * 4.0 * dt + e2 * n[1] * 2.0 * dt + e4 * n[3] * (-2.0 * dt) + e5 * n[4] * (-4.0 * dt);
  fj[2] = (2.0 * e1 * n[0] + e2 * n[1] - e4 * n[3] - 2.0 * e5 * n[4]) * 2.0 * dt;
  fj[3] = 2.0 * e1 * n[0] * 4.0 * dt + e2 * n[1] * 2.0 * dt - e4 * n[3] * (-2.0 * dt) -
          2.0 * e5 * n[4] * (-4.0 * dt);
}

static int newton(double* x, const double* n, double mag, double dt) {
  double fv[2], fj[4], p[2] = {0, 0};
  for (int k = 0; k < NTRIAL; k++) {
    usrfun(x, fv, fj, n, mag, dt);
    p[1] = (fv[1] * fj[0] - fv[0] * fj[2]) / (fj[2] * fj[1] - fj[0] * fj[3]);
    p[0] = (-fv[0] - fj[1] * p[1]) / fj[0];
    x[0] += p[0];
    x[1] += p[1];
    if (fmax(fabs(p[0]), fabs(p[1])) <= TOL_NR) break;
  }
  /* a NaN step fails too (the reference's .GT. test would let it through) */
  return !(fabs(p[0]) <= TOL_NR && fabs(p[1]) <= TOL_NR) ? FGPE_ENEWTON : 0;
}

/* ------------------------------------------------------------------ step
 * One pass of the loop body SRC/cgpe3d.f90:452-495 (2D SRC/cgpe2d.f90:440-486;
 * 1D SRC/cgpe1d.f90:433-474). */
int or_run(or_state* s, long nsteps) {
  for (long it = 0; it < nsteps; it++) {
    or_fft_forward(s);
    or_ke(s);
    or_soc(s);
    or_fft_backward(s);
    or_se(s);
    or_sp(s);
    or_normc(s, s->norm);
    if (s->p.imag) {
      if (!s->p.soc) {
        double sigma[5];
        const double mag = s->p.mag;
        const double* n = s->norm;
        if (s->spin == 1) {
          if (s->dim == 3)
            sigma[1] = sqrt(1.0 - mag * mag) /
                       sqrt(n[1] + sqrt(4.0 * (1.0 - mag * mag) * n[0] * n[2] + mag * mag * n[1] * n[1]) + EPS);
          else
            sigma[1] = sqrt(1.0 - mag * mag) /
                       (sqrt(n[1] + sqrt(4.0 * (1.0 - mag * mag) * n[0] * n[2] + mag * mag * n[1] * n[1])) + EPS);
          sigma[0] = sqrt(1.0 + mag - sigma[1] * sigma[1] * n[1]) / (sqrt(2.0 * n[0]) + EPS);
          sigma[2] = sqrt(1.0 - mag - sigma[1] * sigma[1] * n[1]) / (sqrt(2.0 * n[2]) + EPS);
        } else {
          if (newton(s->sig, n, mag, s->p.dt)) return FGPE_ENEWTON;
          const double dt = s->p.dt;
          sigma[0] = exp((s->sig[0] + 2.0 * s->sig[1]) * dt);
          sigma[1] = exp((s->sig[0] + s->sig[1]) * dt);
          sigma[2] = exp(s->sig[0] * dt);
          sigma[3] = exp((s->sig[0] - s->sig[1]) * dt);
          sigma[4] = exp((s->sig[0] - 2.0 * s->sig[1]) * dt);
        }
        for (int m = 0; m < s->ms; m++)
          for (long q = 0; q < s->npts; q++) s->phi[m * s->npts + q] *= sigma[m];
      }
      or_normt(s);
    }
  }
  return 0;
}

void or_last_norms(const or_state* s, double* norm) {
  for (int m = 0; m < s->ms; m++) norm[m] = s->norm[m];
}

/* ----------------------------------------------------------------- ENERGY
 * EN, MU, MZ with nine-point gradients and nested Simpson:
 * 3D SRC/cgpe3d.f90:1273-1511, 2D SRC/cgpe2d.f90:1076-1281,
 * 1D SRC/cgpe1d.f90:1077-1240. */
static void diff_axis(const or_state* s, const cplx* a, int axis, double d, cplx* out) {
  int n = axis == 0 ? s->nx : axis == 1 ? s->ny : s->nz;
  long stride = axis == 0 ? 1 : axis == 1 ? s->nx : (long)s->nx * s->ny;
  long nlines = s->npts / n;
#pragma omp parallel
  {
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
    double* dre = (double*)malloc(sizeof(double) * n);
    double* dim_ = (double*)malloc(sizeof(double) * n);
#pragma omp for schedule(static)
    for (long l = 0; l < nlines; l++) {
      long base;
      if (axis == 0)
        base = l * s->nx;
      else if (axis == 1)
        base = (l % s->nx) + (l / s->nx) * (long)s->nx * s->ny;
      else
        base = l;
      for (int q = 0; q < n; q++) {
        re[q] = creal(a[base + q * stride]);
        im[q] = cimag(a[base + q * stride]);
      }
      or_diff(re, n, d, dre);
      or_diff(im, n, d, dim_);
      for (int q = 0; q < n; q++) out[base + q * stride] = dre[q] + I * dim_[q];
    }
    free(re);
    free(im);
    free(dre);
    free(dim_);
  }
}

static inline double a2_(cplx z) { return creal(z) * creal(z) + cimag(z) * cimag(z); }

void or_energy(const or_state* s, double* mu, double* en, double* mz) {
  const long np = s->npts;
  const int ms = s->ms, dim = s->dim;
  const double* tau = s->p.tau;
  const double gx = s->p.gamma[0], gy = s->p.gamma[1], gz = s->p.gamma[2];
  double norm[5];
  or_normc(s, norm);
  cplx* dpx = (cplx*)calloc((size_t)ms * np, sizeof(cplx));
  cplx* dpy = dim >= 2 ? (cplx*)calloc((size_t)ms * np, sizeof(cplx)) : NULL;
  cplx* dpz = dim >= 3 ? (cplx*)calloc((size_t)ms * np, sizeof(cplx)) : NULL;
  for (int m = 0; m < ms; m++) {
    diff_axis(s, s->phi + m * np, 0, s->p.d[0], dpx + m * np);
    if (dim >= 2) diff_axis(s, s->phi + m * np, 1, s->p.d[1], dpy + m * np);
    if (dim >= 3) diff_axis(s, s->phi + m * np, 2, s->p.d[2], dpz + m * np);
  }
  const int ncol = ms + 2;
  double* col = (double*)calloc((size_t)ncol * np, sizeof(double));
  const double s2 = sqrt(2.0), s6 = sqrt(6.0), s15 = sqrt(1.5);
#pragma omp parallel for schedule(static)
  for (long q = 0; q < np; q++) {
    cplx ph[5], px[5] = {0}, py[5] = {0}, pz[5] = {0}, cph[5];
    double dp2[5], a[5];
    double rho = 0;
    for (int m = 0; m < ms; m++) {
      ph[m] = s->phi[m * np + q];
      cph[m] = conj(ph[m]);
      a[m] = a2_(ph[m]);
      rho += a[m];
      px[m] = dpx[m * np + q];
      dp2[m] = a2_(px[m]);
      if (dim >= 2) {
        py[m] = dpy[m * np + q];
        dp2[m] += a2_(py[m]);
      }
      if (dim >= 3) {
        pz[m] = dpz[m * np + q];
        dp2[m] += a2_(pz[m]);
      }
    }
    double v = s->v[q];
    double fx, fy, fz;
    fxyz(s->spin, ph, &fx, &fy, &fz);
    double c[7];
    if (s->spin == 1) {
      double base1 = v * rho + tau[0] * rho * rho / 2.0 + tau[1] * (fx * fx + fy * fy + fz * fz) / 2.0 + dp2[0] +
                     dp2[1] + dp2[2];
      cplx so1, so3, so4, so5;
      if (dim == 1) {
        so1 = s2 * gx * (-I * cph[0] * px[1] - I * cph[2] * px[1] - I * cph[1] * (px[0] + px[2]));
        so3 = s2 * gx * (-I * cph[0] * px[1]);
        so4 = s2 * gx * (-I * cph[1] * (px[0] + px[2]));
        so5 = s2 * gx * (-I * cph[2] * px[1]);
      } else {
        cplx xy = cph[0] * (-I * gx * px[1] - gy * py[1]) + cph[1] * (-I * gx * px[2] - gy * py[2]) +
                  cph[1] * (-I * gx * px[0] + gy * py[0]) + cph[2] * (-I * gx * px[1] + gy * py[1]);
        if (dim == 3) {
          so1 = 2.0 * sqrt(0.5) * xy + 2.0 * (cph[0] * (-I * gz * pz[0]) + cph[2] * (I * gz * pz[2]));
          so3 = s2 * (cph[0] * (-I * gx * px[1] - gy * py[1] - I * s2 * gz * pz[0]));
          so5 = s2 * (cph[2] * (-I * gx * px[1] + gy * py[1] + I * gz * s2 * pz[2]));
        } else {
          so1 = s2 * xy;
          so3 = s2 * (cph[0] * (-I * gx * px[1] - gy * py[1]));
          so5 = s2 * (cph[2] * (-I * gx * px[1] + gy * py[1]));
        }
        so4 = s2 * (cph[1] * (-I * gx * px[2] - gy * py[2]) + cph[1] * (-I * gx * px[0] + gy * py[0]));
      }
      c[0] = creal(base1 + so1);
      c[1] = a[0] - a[2];
      c[2] = creal(v * a[0] + dp2[0] + tau[0] * (a[0] + a[1] + a[2]) * a[0] + tau[1] * (a[0] + a[1] - a[2]) * a[0] +
                   tau[1] * cph[2] * ph[1] * ph[1] * cph[0] + so3);
      c[3] = creal(v * a[1] + dp2[1] + tau[0] * (a[0] + a[1] + a[2]) * a[1] + tau[1] * (a[0] + a[2]) * a[1] +
                   2.0 * tau[1] * ph[0] * ph[2] * cph[1] * cph[1] + so4);
      c[4] = creal(v * a[2] + dp2[2] + tau[0] * (a[0] + a[1] + a[2]) * a[2] + tau[1] * (a[1] + a[2] - a[0]) * a[2] +
                   tau[1] * cph[0] * ph[1] * ph[1] * cph[2] + so5);
    } else {
      cplx cc[6];
      mat_c(ph, fx - I * fy, tau, cc);
      cplx th = (2.0 * ph[0] * ph[4] - 2.0 * ph[1] * ph[3] + ph[2] * ph[2]) / sqrt(5.0);
      double base1 = v * rho + dp2[0] + dp2[1] + dp2[2] + dp2[3] + dp2[4] + tau[0] * rho * rho / 2.0 +
                     tau[1] * (fx * fx + fy * fy + fz * fz) / 2.0 + tau[2] * creal(th * conj(th)) / 2.0;
      cplx so1, so3, so4, so5, so6, so7;
      if (dim == 1) {
        so1 = gx * (-I * 2.0 * (cph[0] * px[1]) - I * 2.0 * (cph[1] * px[0]) -
                    I * s6 * (cph[1] * px[2] + cph[2] * px[1] + cph[2] * px[3] + cph[3] * px[2]) -
                    I * 2.0 * (cph[3] * px[4] + cph[4] * px[3]));
        so3 = -I * 2.0 * gx * (cph[0] * px[1]);
        so4 = gx * (-I * 2.0 * (cph[1] * px[0]) - I * s6 * cph[1] * px[2]);
        so5 = -I * gx * s6 * (cph[2] * px[1] + cph[2] * px[3]);
        so6 = gx * (-I * s6 * cph[3] * px[2] - I * 2.0 * cph[3] * px[4]);
        so7 = -I * 2.0 * gx * cph[4] * px[3];
      } else if (dim == 2) {
        so1 = -I * 2.0 *
              (cph[0] * gx * px[1] - I * cph[0] * gy * py[1] + cph[1] * gx * px[0] + I * cph[1] * gy * py[0] +
               s15 * (cph[1] * gx * px[2] - I * cph[1] * gy * py[2] + cph[2] * gx * px[1] + I * cph[2] * gy * py[1] +
                      cph[2] * gx * px[3] - I * cph[2] * gy * py[3] + cph[3] * gx * px[2] + I * cph[3] * gy * py[2]) +
               cph[3] * gx * px[4] - I * cph[3] * gy * py[4] + cph[4] * gx * px[3] + I * cph[4] * gy * py[3]);
        so3 = -I * 2.0 * (cph[0] * gx * px[1] - I * cph[0] * gy * py[1]);
        so4 = -I * 2.0 *
              (cph[1] * gx * px[0] + I * cph[1] * gy * py[0] + s15 * (cph[1] * gx * px[2] - I * cph[1] * gy * py[2]));
        so5 = -I * 2.0 *
              (s15 * (cph[2] * gx * px[1] + I * cph[2] * gy * py[1] + cph[2] * gx * px[3] - I * cph[2] * gy * py[3]));
        so6 = -I * 2.0 *
              (cph[3] * gx * px[4] - I * cph[3] * gy * py[4] + s15 * (cph[3] * gx * px[2] + I * cph[3] * gy * py[2]));
        so7 = -I * 2.0 * (cph[4] * gx * px[3] + I * cph[4] * gy * py[3]);
      } else {
        so1 = -I * cph[0] * (2.0 * gx * px[1] - I * 2.0 * gy * py[1] + 4.0 * gz * pz[0]) -
              I * cph[1] * (gx * (2.0 * px[0] + s6 * px[2]) + I * gy * (2.0 * py[0] - s6 * py[2]) + 2.0 * gz * pz[1]) -
              I * cph[2] * (gx * s6 * (px[1] + px[3]) + I * gy * s6 * (py[1] - py[3])) -
              I * cph[3] * (gx * (s6 * px[2] + 2.0 * px[4]) + I * gy * (s6 * py[2] - 2.0 * py[4]) - 2.0 * gz * pz[3]) -
              I * cph[4] * (2.0 * gx * px[3] + I * 2.0 * gy * py[3] - 4.0 * gz * pz[4]);
        so3 = -I * 2.0 *
              (cph[0] * gx * px[1] - I * cph[0] * gy * py[1] + 2.0 * cph[0] * gz * pz[0]);
        so4 = -I * 2.0 *
              (cph[1] * gx * px[0] + I * cph[1] * gy * py[0] +
               s15 * (cph[1] * gx * px[2] - I * cph[1] * gy * py[2]) + cph[1] * gz * pz[1]);
        so5 = -I * 2.0 *
              (s15 * (cph[2] * gx * px[1] + I * cph[2] * gy * py[1] + cph[2] * gx * px[3] - I * cph[2] * gy * py[3]));
        so6 = -I * 2.0 *
              (cph[3] * gx * px[4] - I * cph[3] * gy * py[4] +
               s15 * (cph[3] * gx * px[2] + I * cph[3] * gy * py[2]) - cph[3] * gz * pz[3]);
        so7 = -I * 2.0 *
              (cph[4] * gx * px[3] + I * cph[4] * gy * py[3] - 2.0 * cph[4] * gz * pz[4]);
      }
      c[0] = creal(base1 + so1);
      c[1] = 2.0 * a[0] + a[1] - a[3] - 2.0 * a[4];
      c[2] = creal((v + tau[0] * rho + 2.0 * tau[1] * fz + 0.4 * tau[2] * a[4]) * a[0] +
                   (cc[0] * ph[1] + cc[1] * ph[2]) * cph[0] + so3 + dp2[0]);
      c[3] = creal((v + tau[0] * rho + tau[1] * fz + 0.4 * tau[2] * a[3]) * a[1] +
                   (conj(cc[0]) * ph[0] + cc[2] * ph[2]) * cph[1] + so4 + dp2[1]);
      c[4] = creal((v + tau[0] * rho + 0.2 * tau[2] * a[2]) * a[2] +
                   (conj(cc[1]) * ph[0] + conj(cc[2]) * ph[1] + cc[3] * ph[3] + cc[4] * ph[4]) * cph[2] + so5 +
                   dp2[2]);
      c[5] = creal((v + tau[0] * rho - tau[1] * fz + 0.4 * tau[2] * a[1]) * a[3] +
                   (conj(cc[3]) * ph[2] + cc[5] * ph[4]) * cph[3] + so6 + dp2[3]);
      c[6] = creal((v + tau[0] * rho - 2.0 * tau[1] * fz + 0.4 * tau[2] * a[0]) * a[4] +
                   (conj(cc[4]) * ph[2] + conj(cc[5]) * ph[3]) * cph[4] + so7 + dp2[4]);
    }
    for (int cidx = 0; cidx < ncol; cidx++) col[cidx * np + q] = c[cidx];
  }
  double integ[7];
  for (int cidx = 0; cidx < ncol; cidx++) integ[cidx] = nested_simpson(s, col + cidx * np);
  double sumn = 0;
  for (int m = 0; m < ms; m++) sumn += norm[m];
  *en = integ[0] / sumn;
  *mz = integ[1];
  for (int m = 0; m < ms; m++) mu[m] = integ[m + 2] / (norm[m] + EPS);
  free(col);
  free(dpx);
  free(dpy);
  free(dpz);
}

/*
 * refcost.c -- host-CPU cost of the two third-party calls that dominate one
 * 3D spin-2 step of the reference, on the machine it runs on.
 *
 * CPU-BASELINE INFRASTRUCTURE ONLY (run by bench.py's cpu_baseline leg with
 * --refcost); the product path never links or runs it.
 *
 * The reference itself cannot be built in this image (its sources INCLUDE
 * 'fftw3.f03', which the image lacks).  Its step (SRC/cgpe3d.f90:452-457)
 * spends its time in library calls that can be timed without it:
 *   FFT/BFT (SRC/cgpe3d.f90:1748-1784): per component a copy of PHI into the
 *     work array, an out-of-place FFTW 3D execute (plans from
 *     fftw_plan_dft_3d(NZ,NY,NX,...,FFTW_ESTIMATE) on FFTW_THREADS threads,
 *     SRC/cgpe3d.f90:843-854) and a copy back (÷N in BFT);
 *   SE spin-2 (SRC/cgpe3d.f90:1700-1738): per grid point a 5x5 Hermitian
 *     matrix with the twelve C12..C45 entries, ZHEEV('V','L') twice (workspace
 *     query, then the solve), two 5x5 MATMULs and five complex exponentials,
 *     in an OpenMP loop over the grid.
 * Both are timed here through the same library interfaces (FFTW's API and
 * LAPACK ZHEEV, as exported by the image's MKL 2021.4, libmkl_rt.so, loaded
 * at run time with GNU OpenMP threading).  Their sum is a lower bound on the
 * reference's step time: KE, SOC (per-point MATMULs), SP, NORMC, FXYZ/MAT_C
 * and the full-grid temporaries (SRC/cgpe3d.f90:875-1124, 1593-1700) are not
 * counted.  SE is timed on a sample of points and scaled to the grid.
 *
 * Usage: refcost N THREADS SE_SAMPLE_POINTS   (prints one JSON line)
 */
#include <complex.h>
#include <dlfcn.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef double fftw_complex_t[2];
typedef void* fftw_plan_t;
typedef int (*init_threads_f)(void);
typedef void (*plan_with_nthreads_f)(int);
typedef fftw_plan_t (*plan_dft_3d_f)(int, int, int, fftw_complex_t*, fftw_complex_t*, int, unsigned);
typedef void (*execute_dft_f)(const fftw_plan_t, fftw_complex_t*, fftw_complex_t*);
typedef void (*destroy_plan_f)(fftw_plan_t);
typedef void (*zheev_f)(const char*, const char*, const int*, double complex*, const int*, double*, double complex*,
                        const int*, double*, int*);

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void* sym(void* lib, const char* name) {
  void* f = dlsym(lib, name);
  if (!f) {
    fprintf(stderr, "refcost: %s not found in the MKL library\n", name);
    exit(2);
  }
  return f;
}

/* splitmix64 -> uniform in [-1, 1) */
static double urand(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: refcost N THREADS SE_SAMPLE_POINTS\n");
    return 2;
  }
  const int n = atoi(argv[1]), threads = atoi(argv[2]);
  long sample = atol(argv[3]);
  const int ms = 5;
  if (n < 8 || threads < 1 || sample < threads) {
    fprintf(stderr, "refcost: bad arguments\n");
    return 2;
  }
  setenv("MKL_THREADING_LAYER", "GNU", 1);
  setenv("MKL_INTERFACE_LAYER", "LP64", 1);
  const char* path = getenv("FGPE_MKL_RT");
  if (!path) path = "/opt/conda/lib/libmkl_rt.so";
  void* lib = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
  if (!lib) {
    fprintf(stderr, "refcost: cannot load %s: %s\n", path, dlerror());
    return 3;
  }
  init_threads_f init_threads = (init_threads_f)sym(lib, "fftw_init_threads");
  plan_with_nthreads_f plan_with_nthreads = (plan_with_nthreads_f)sym(lib, "fftw_plan_with_nthreads");
  plan_dft_3d_f plan_dft_3d = (plan_dft_3d_f)sym(lib, "fftw_plan_dft_3d");
  execute_dft_f execute_dft = (execute_dft_f)sym(lib, "fftw_execute_dft");
  destroy_plan_f destroy_plan = (destroy_plan_f)sym(lib, "fftw_destroy_plan");
  zheev_f zheev = (zheev_f)sym(lib, "zheev_");
  void (*mkl_threads)(int) = (void (*)(int))sym(lib, "MKL_Set_Num_Threads");
  omp_set_num_threads(threads);

  /* ---- FFT + BFT of one step: MS components each way */
  const long npts = (long)n * n * n;
  if (sample > npts) sample = npts; /* each sampled point updates its own entries */
  double complex* phi = malloc(sizeof(double complex) * npts * ms);
  double complex* phif = malloc(sizeof(double complex) * npts * ms);
  fftw_complex_t* wf = malloc(sizeof(fftw_complex_t) * npts);
  fftw_complex_t* wb = malloc(sizeof(fftw_complex_t) * npts);
  if (!phi || !phif || !wf || !wb) {
    fprintf(stderr, "refcost: out of memory\n");
    return 3;
  }
  uint64_t seed = 12345;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < npts * ms; ++q) {
    uint64_t s = seed + (uint64_t)q * 2654435761ull;
    phi[q] = urand(&s) + I * urand(&s);
  }
  init_threads();
  plan_with_nthreads(threads);
  fftw_plan_t pf = plan_dft_3d(n, n, n, wf, wb, -1, 64u);
  fftw_plan_t pb = plan_dft_3d(n, n, n, wb, wf, +1, 64u);
  const double inv = 1.0 / (double)npts;
  double t_fft = 1e300;
  for (int rep = 0; rep < 2; ++rep) { /* the first repetition pages everything in */
    const double t0 = now();
    for (int m = 0; m < ms; ++m) {
      const double complex* src = phi + (long)m * npts;
      double complex* dst = phif + (long)m * npts;
#pragma omp parallel for schedule(static)
      for (long q = 0; q < npts; ++q) memcpy(&wf[q], &src[q], sizeof(double complex));
      execute_dft(pf, wf, wb);
#pragma omp parallel for schedule(static)
      for (long q = 0; q < npts; ++q) memcpy(&dst[q], &wb[q], sizeof(double complex));
    }
    for (int m = 0; m < ms; ++m) {
      const double complex* src = phif + (long)m * npts;
      double complex* dst = phi + (long)m * npts;
#pragma omp parallel for schedule(static)
      for (long q = 0; q < npts; ++q) memcpy(&wb[q], &src[q], sizeof(double complex));
      execute_dft(pb, wb, wf);
#pragma omp parallel for schedule(static)
      for (long q = 0; q < npts; ++q) {
        double complex v;
        memcpy(&v, &wf[q], sizeof v);
        dst[q] = v * inv;
      }
    }
    const double t = now() - t0;
    if (t < t_fft) t_fft = t;
  }
  destroy_plan(pf);
  destroy_plan(pb);
  free(wf);
  free(wb);
  free(phif);

  /* ---- SE spin-2 on a sample of points: the loop body of
   *      SRC/cgpe3d.f90:1707-1737 with random couplings */
  enum { LWMAX = 1000 };
  double t_se = 0.0;
  long done = 0;
  int bad = 0;
  /* one ZHEEV per thread before the clock starts (the library picks its code
   * path on the first call), each call single-threaded as in the reference's
   * OpenMP loop */
  mkl_threads(1);
#pragma omp parallel
  {
    double complex a[25] = {0}, work[64];
    double w[5], rwork[13];
    const int nn = 5, lda = 5, lwork = 64;
    int info;
    for (int i = 0; i < 5; ++i) a[6 * i] = 1.0 + i;
    zheev("V", "L", &nn, a, &lda, w, work, &lwork, rwork, &info);
  }
  {
    const double t0 = now();
#pragma omp parallel reduction(+ : done, bad)
    {
      double complex a[25], work[LWMAX], v[5], t[5];
      double w[5], rwork[3 * 5 - 2];
      const int nn = 5, lda = 5;
      uint64_t s = 777 + 1000003ull * (uint64_t)omp_get_thread_num();
      const double cdt = 5e-4;
#pragma omp for schedule(static)
      for (long q = 0; q < sample; ++q) {
        double complex* p = phi + q; /* the point's five components */
        for (int i = 0; i < 25; ++i) a[i] = 0.0;
        const double complex c12 = urand(&s) + I * urand(&s), c13 = urand(&s) + I * urand(&s),
                             c23 = urand(&s) + I * urand(&s), c34 = urand(&s) + I * urand(&s),
                             c35 = urand(&s) + I * urand(&s), c45 = urand(&s) + I * urand(&s);
        /* column-major A(i,j) = a[(j-1)*5 + (i-1)] */
        a[5 * 1 + 0] = c12;
        a[5 * 2 + 0] = c13;
        a[5 * 0 + 1] = conj(c12);
        a[5 * 2 + 1] = c23;
        a[5 * 0 + 2] = conj(c13);
        a[5 * 1 + 2] = conj(c23);
        a[5 * 3 + 2] = c34;
        a[5 * 4 + 2] = c35;
        a[5 * 2 + 3] = conj(c34);
        a[5 * 4 + 3] = c45;
        a[5 * 2 + 4] = conj(c35);
        a[5 * 3 + 4] = conj(c45);
        int lwork = -1, info = 0;
        zheev("V", "L", &nn, a, &lda, w, work, &lwork, rwork, &info);
        lwork = (int)creal(work[0]);
        if (lwork > LWMAX) lwork = LWMAX;
        zheev("V", "L", &nn, a, &lda, w, work, &lwork, rwork, &info);
        bad += info != 0;
        for (int m = 0; m < 5; ++m) v[m] = p[(long)m * npts];
        for (int i = 0; i < 5; ++i) { /* conj(A)^T v */
          double complex acc = 0.0;
          for (int k = 0; k < 5; ++k) acc += conj(a[5 * i + k]) * v[k];
          t[i] = acc * cexp(-I * cdt * w[i]);
        }
        for (int i = 0; i < 5; ++i) { /* A t */
          double complex acc = 0.0;
          for (int k = 0; k < 5; ++k) acc += a[5 * k + i] * t[k];
          v[i] = acc;
        }
        for (int m = 0; m < 5; ++m) p[(long)m * npts] = v[m];
        ++done;
      }
    }
    t_se = (now() - t0) * (double)npts / (double)done;
  }
  free(phi);
  printf("{\"n\": %d, \"ms\": %d, \"threads\": %d, \"fft_bft_s\": %.4f, \"se_zheev_s\": %.4f, "
         "\"se_sample_points\": %ld, \"zheev_failures\": %d, \"step_lower_bound_s\": %.4f, "
         "\"library\": \"MKL 2021.4 libmkl_rt (FFTW3 API, LAPACK ZHEEV), GNU OpenMP\"}\n",
         n, ms, threads, t_fft, t_se, done, bad, t_fft + t_se);
  return 0;
}

/*
 * fgpe_oracle.h -- FP64 CPU restatement of FORTRESS's split-step propagator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product path (libfgpe, cgpe.x) never does.
 *
 * Every function follows the reference line by line in arithmetic (FP64
 * complex), step order and quirks; each definition in fgpe_oracle.c cites the
 * reference file:line it restates.
 */
#ifndef FGPE_ORACLE_H
#define FGPE_ORACLE_H

#include "../include/fgpe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_state or_state;

int or_derive(const fgpe_input* in, fgpe_params* out, double* aosc);

or_state* or_create(const fgpe_params* p);
void or_destroy(or_state* s);
void or_initial_guess(or_state* s);
void or_set_phi(or_state* s, const double* phi);
void or_get_phi(const or_state* s, double* phi);
/* the transform buffer PHIF (k space between FFT and BFT), same layout as phi */
void or_get_phif(const or_state* s, double* phif);
void or_set_phif(or_state* s, const double* phif);
double or_normt(or_state* s);
int or_run(or_state* s, long nsteps);
void or_last_norms(const or_state* s, double* norm);
void or_normc(const or_state* s, double* norm);
void or_energy(const or_state* s, double* mu, double* en, double* mz);
void or_rad(const or_state* s, double* rms);
void or_set_threads(int n);

/* building blocks, exposed for closed-form checks */
double or_simpson(const double* f, int n, double dx);
void or_diff(const double* p, int n, double dx, double* dp);
void or_dft(double* z, int n, int sign);
void or_fft_forward(or_state* s);
void or_fft_backward(or_state* s);
void or_ke(or_state* s);
void or_soc(or_state* s);
void or_se(or_state* s);
void or_sp(or_state* s);
void or_soc_point(int spin, int dim, double kx, double ky, double gx, double gy, double cdt_re, double cdt_im,
                  const double* v, double* out);
void or_expm_herm5(const double* h, double cdt_re, double cdt_im, const double* v, double* out);

#ifdef __cplusplus
}
#endif
#endif

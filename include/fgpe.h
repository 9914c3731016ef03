/*
 * fgpe.h -- C ABI of libfgpe, the MI355X (gfx950) split-step Fourier propagator
 * for the coupled Gross-Pitaevskii equations of spin-orbit-coupled spin-1 and
 * spin-2 condensates in q1D, q2D and 3D.
 *
 * The reference (FORTRESS) has no library interface: its boundary is the
 * executable cgpe.x, whose time loop calls argument-less subroutines acting on
 * module globals (SRC/cgpe3d.f90:123-204 declares them; the loop body is
 * SRC/cgpe3d.f90:451-567).  Each entry point below replaces one group of those
 * calls; the comment above each declaration cites the reference code it
 * stands for.  The project's Fortran host (fortress_amd/host/cgpe.f90) binds
 * these through ISO_C_BINDING and keeps cgpe.x's INPUT/OUTPUT contract.
 *
 * Conventions
 *  - Plain C types only.  Every call returns 0 on success or a negative
 *    FGPE_E* code; fgpe_last_error() then gives a message.  Nothing throws.
 *  - Wave functions cross the boundary in the reference's Fortran layout
 *    PHI(NX,NY,NZ,MS) (SRC/cgpe3d.f90:82,637): x fastest, component slowest,
 *    each element an interleaved (re, im) pair of doubles.  Unused axes have
 *    extent 1.
 *  - A handle owns its device buffers and its HIP stream; calls on one handle
 *    are not re-entrant.  Different handles are independent.
 *  - Step semantics are the reference's: first-order Lie splitting
 *    FFT, KE, SOC, BFT, SE, SP, then NORMC and (imaginary time) the
 *    projection/NORMT renormalisation (SRC/cgpe3d.f90:452-495).
 */
#ifndef FGPE_H
#define FGPE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define FGPE_OK 0
#define FGPE_EINVAL (-1)   /* bad argument or unsupported configuration   */
#define FGPE_EDEVICE (-2)  /* HIP runtime error (allocation, launch, ...) */
#define FGPE_ENEWTON (-3)  /* spin-2 SOC=0 projection Newton did not converge */
#define FGPE_ESTATE (-4)   /* call not valid in the handle's current state */

/* precision of the propagation (reductions are always FP64) */
#define FGPE_FP64 0
#define FGPE_FP32 1

/* Physical input, one field per record of INPUT/input{1,2,3}D.dat
 * (SRC/cgpe3d.f90:239-266; 2D SRC/cgpe2d.f90:234-261; 1D SRC/cgpe1d.f90:233-260). */
typedef struct fgpe_input {
  int32_t dim;        /* 1, 2 or 3 */
  int32_t spin;       /* SPIN: 1 or 2 */
  int32_t niter;      /* NITER */
  int32_t stp;        /* STP */
  int32_t nstp;       /* NSTP */
  int32_t threads;    /* NUMBER OF THREADS (echoed only) */
  int32_t n[3];       /* NX, NY, NZ (unused axes 1) */
  double d[3];        /* DX, DY, DZ */
  double dt;          /* DT */
  double mass;        /* mass in amu */
  double a[3];        /* A0, A2, A4 in Bohr radii */
  int32_t natoms;     /* NATOMS */
  double nu[3];       /* NUX, NUY, NUZ in Hz */
  double tol;         /* TOL */
  int32_t switch_im;  /* SWITCH_IM: 1 imaginary, 0 real time */
  double gamma[3];    /* GAMMAX, GAMMAY, GAMMAZ */
  int32_t option_fpc; /* OPTION_FPC (4 = Gaussian) */
  double mag;         /* MAG */
} fgpe_input;

/* Derived, dimensionless solver parameters (what the propagator needs). */
typedef struct fgpe_params {
  int32_t dim;        /* 1, 2 or 3 */
  int32_t spin;       /* 1 or 2 */
  int32_t n[3];       /* grid points per axis (unused axes 1): 2^a or 3*2^a,
                         32..512 in 2D/3D, 64..4096 in 1D */
  int32_t imag;       /* 1 imaginary time (CDT = -i DT), 0 real time (CDT = DT) */
  int32_t precision;  /* FGPE_FP64 or FGPE_FP32 */
  int32_t option_fpc; /* initial-guess selector used by fgpe_initial_guess */
  int32_t soc;        /* SWITCH_SOC: 1 if any |gamma| > 1e-40 */
  double d[3];        /* DX, DY, DZ */
  double dt;          /* DT */
  double tau[3];      /* TAU(0..2), the reference's scaled couplings */
  double alpha[3];    /* ALPHAX, ALPHAY, ALPHAZ = NU/NUX */
  double gamma[3];    /* GAMMAX, GAMMAY, GAMMAZ */
  double mag;         /* MAG (used by the SOC = 0 projection) */
} fgpe_params;

typedef struct fgpe_handle fgpe_handle;

/* Setup arithmetic of the driver: ALPHA = NU/NUX, AOSC, TAU(0..2) per spin and
 * dimension, SWITCH_SOC.  Replaces SRC/cgpe3d.f90:256,277-283,334-345
 * (2D SRC/cgpe2d.f90:326-336; 1D SRC/cgpe1d.f90:321-332).
 * precision selects FGPE_FP64 or FGPE_FP32 for the propagation. */
int fgpe_derive(const fgpe_input* in, int32_t precision, fgpe_params* out, double* aosc);

/* Allocate device state and tables for one run.  Replaces ALLOCATE_MEM and the
 * grid/k/potential part of INITIALIZE plus CREATE_PLANS
 * (SRC/cgpe3d.f90:622-643, 666-754, 843-854). */
int fgpe_create(const fgpe_params* p, fgpe_handle** h);

/* Release everything.  Replaces DESTROY_PLANS + DEALLOCATE_MEM
 * (SRC/cgpe3d.f90:859-869, 648-660). */
void fgpe_destroy(fgpe_handle* h);

/* The initial-guess part of INITIALIZE (Gaussian default and the OPTION_FPC
 * 1-3 guesses), evaluated on the device (SRC/cgpe3d.f90:753-837;
 * 2D SRC/cgpe2d.f90:697-774; 1D SRC/cgpe1d.f90:707-789). */
int fgpe_initial_guess(fgpe_handle* h);

/* Host -> device / device -> host copies of PHI in the Fortran layout
 * (interleaved complex double).  Upload replaces the initial_sol read
 * (SRC/cgpe3d.f90:391-409); download serves the dumps (534-565, 596-602). */
int fgpe_upload(fgpe_handle* h, const double* phi);
int fgpe_download(fgpe_handle* h, double* phi);

/* NORMT: normalise the total density to 1 and return TNORM
 * (SRC/cgpe3d.f90:1178-1219; called at 412 before the loop). */
int fgpe_normalize(fgpe_handle* h, double* tnorm);

/* Advance nsteps full steps of the loop body SRC/cgpe3d.f90:452-495:
 * FFT, KE, SOC, BFT, SE, SP, NORMC and, in imaginary time, the SOC = 0
 * projection (spin-1 sigma, spin-2 NEWTON) and NORMT.  After the call the
 * wave function is in real space and normalised as after line 495. */
int fgpe_run(fgpe_handle* h, int64_t nsteps);

/* Pencil schedule of a 3D handle (no reference counterpart: FFTW's plans,
 * SRC/cgpe3d.f90:843-854, choose only how each transform is split).  A step
 * runs in two phases, each a pipeline over chunks: the x-line pass (inverse-x,
 * SE, SP, NORMC, forward-x) and the forward-y pass by z-chunks of zchunk
 * planes, then the z-line pass (KE) and the inverse-y pass (SOC, inverse-y)
 * by x-chunks of xchunk columns.  0 (or a value that does not split the axis into whole
 * blocks) means the whole grid in one launch per pass.  Chunks go round-robin
 * to nstreams streams (1-3) of the handle.  Every schedule gives bit-identical
 * results.  Other dimensions ignore it.  fgpe_get_schedule returns the chunk
 * sizes in effect and the stream count. */
int fgpe_set_schedule(fgpe_handle* h, int32_t zchunk, int32_t xchunk, int32_t nstreams);
int fgpe_get_schedule(fgpe_handle* h, int32_t* zchunk, int32_t* xchunk, int32_t* nstreams);

/* Schedule autotuner of a 3D handle: times trial_steps steps of the handle's
 * own state under each of FGPE_NSCHED_CAND candidate schedules (whole grid;
 * eighths/quarters of NZ and NX on two or three streams), restores the state
 * (and the norms and Newton state a run updates) after each, and keeps the
 * fastest.  ms_out[FGPE_NSCHED_CAND] receives milliseconds per step of each
 * candidate, chosen[3] the (zchunk, xchunk, nstreams) in effect.  It
 * allocates a temporary copy of the state.  A handle starts with eighths of
 * each axis on two streams.  Other dimensions: no-op. */
#define FGPE_NSCHED_CAND 8
int fgpe_tune_schedule(fgpe_handle* h, int32_t trial_steps, double* ms_out, int32_t* chosen);

/* fgpe_run without any device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  fgpe_run
 * allocates its NORMC block partials on the first call; this variant runs the
 * same kernels with those partials in a fixed 8 MiB device array that the
 * library holds on each GPU until the process ends, and calls no hipMalloc or
 * hipFree.  Results are bit-identical to fgpe_run.  Limit: 2^20 doubles of
 * partials (MS per pass-A block), which every grid fgpe_create accepts fits;
 * larger needs are refused with FGPE_EINVAL before anything runs.  All calls
 * on one GPU share that array: calls that overlap in time on one GPU (other
 * handles, other streams, other threads) race and must be serialised by the
 * caller. */
int fgpe_run_noalloc(fgpe_handle* h, int64_t nsteps);

/* NORMC of the most recent step (taken before renormalisation, as the
 * reference prints it: SRC/cgpe3d.f90:459,449,527). */
int fgpe_last_norms(fgpe_handle* h, double* norm);

/* NORMC of the current state (SRC/cgpe3d.f90:1129-1173). */
int fgpe_normc(fgpe_handle* h, double* norm);

/* ENERGY and RAD of the current state: EN, MU(1..MS), MZ, RMS(1..MS)
 * (SRC/cgpe3d.f90:1273-1511, 1224-1268).  Energies are returned in the
 * reference's scaled units (the files print EN/2 and MU/2). */
int fgpe_observables(fgpe_handle* h, double* en, double* mu, double* mz, double* rms);

/* PHI_OLD handling for the convergence test: keep a copy of the current state
 * (SRC/cgpe3d.f90:530) and return max |PHI - PHI_OLD| over all points and
 * components (SRC/cgpe3d.f90:503, before the division by STP*2*DT). */
int fgpe_keep_old(fgpe_handle* h);
int fgpe_conv_max(fgpe_handle* h, double* maxdiff);

/* |PHI(NX/2+1,NY/2+1,NZ/2+1,m)| for the file1 rows (SRC/cgpe3d.f90:430,533,579). */
int fgpe_center_abs(fgpe_handle* h, double* absphi);

/* Wait for all work queued on the handle's stream. */
int fgpe_sync(fgpe_handle* h);

/* Binary checkpoint and resume (no counterpart in the reference, which can
 * restart only real time, from the 10-decimal text file initial_sol:
 * SRC/cgpe3d.f90:383-410, SRC/cgpe2d.f90:378-400, SRC/cgpe1d.f90:372-383).
 * A checkpoint holds the run parameters, the caller's loop state, the device
 * scalars carried from step to step (last NORMC, spin-2 Newton state) and the
 * wave function as the device holds it (FP32 or FP64), followed by a 64-bit
 * FNV-1a checksum; it is written to path.tmp and renamed.  Loading it into a
 * handle created with the same parameters (bit-equal DT, DX, GAMMA, TAU,
 * ALPHA, MAG; same grid, spin, precision and time mode, else FGPE_EINVAL)
 * and advancing gives bit-identical results to a run never stopped.
 * fgpe_checkpoint_info reads and verifies a file without a device. */
typedef struct fgpe_loop_state {
  int64_t ii;       /* steps done (II) */
  double en_old;    /* EN of the previous STP event, for CONV_EN */
  double user[6];   /* the caller's own values, stored and returned as is */
} fgpe_loop_state;
typedef struct fgpe_checkpoint_header {
  int32_t version, dim, spin, precision, imag;
  int32_t n[3];
  double dt;
  int64_t ii;
  double en_old;
  int64_t payload_bytes;
} fgpe_checkpoint_header;
int fgpe_checkpoint_save(fgpe_handle* h, const char* path, const fgpe_loop_state* st);
int fgpe_checkpoint_load(fgpe_handle* h, const char* path, fgpe_loop_state* st);
int fgpe_checkpoint_info(const char* path, fgpe_checkpoint_header* info);

/* The field dumps of cgpe.x, formatted on `threads` host threads (0: all) and
 * written in order; the bytes are those of the reference's formatted WRITE
 * loops: records in the loop order i, j, k (k fastest), density |PHI|**2 and
 * phase ATAN2(AIMAG, REAL) per component, Fw.d fields (exact value rounded to
 * d decimals, ties to even; w asterisks on overflow), an empty record after
 * each x where the reference writes one.  phi in the Fortran layout
 * (interleaved complex double); x, y, z the grids the dump prints.
 *   FGPE_DUMP_1D_TMP  (F11.5,10F14.6)                SRC/cgpe1d.f90:513-516
 *   FGPE_DUMP_1D_SOL  (F11.5,10F18.10)               SRC/cgpe1d.f90:547-549
 *   FGPE_DUMP_2D      (2F8.2,10F10.5) + empty record SRC/cgpe2d.f90:521-528,560-565
 *   FGPE_DUMP_3D      (10F16.10)                     SRC/cgpe3d.f90:534-540,596-602
 *   FGPE_DUMP_3D_XY   (2F10.4,10F14.8) at k = NZ/2+1 SRC/cgpe3d.f90:541-551
 *   FGPE_DUMP_3D_XZ   (2F10.4,10F14.8) at j = NY/2+1 SRC/cgpe3d.f90:553-563
 * fgpe_format_f writes one Fw.d field (w <= 40) and a NUL into out. */
#define FGPE_DUMP_1D_TMP 1
#define FGPE_DUMP_1D_SOL 2
#define FGPE_DUMP_2D 3
#define FGPE_DUMP_3D 4
#define FGPE_DUMP_3D_XY 5
#define FGPE_DUMP_3D_XZ 6
int fgpe_write_dump(const char* path, int32_t kind, const int32_t* n, int32_t ms, const double* phi,
                    const double* x, const double* y, const double* z, int32_t threads);
int fgpe_format_f(double v, int32_t w, int32_t d, char* out, size_t cap);

/* Reads ./INPUT/initial_sol_spinS.dat as cgpe.x's list-directed READ loop
 * does (SRC/cgpe3d.f90:391-409, SRC/cgpe2d.f90:382-397, SRC/cgpe1d.f90:
 * 377-388): one record per grid point in the order i, j, k (empty records
 * skipped), x (1D) or x, y (2D) then density and phase per component;
 * PHI = SQRT(density) EXP(i phase) into phi (Fortran layout).  x, y (may be
 * null) receive the coordinates as the READ leaves them (the last record of
 * each index).  Records are parsed on `threads` host threads (0: all).  Takes
 * plain numbers separated by blanks or commas, D or E exponents; a record
 * with fewer values, repeat counts (r*c), '/' or text is refused with
 * FGPE_EINVAL. */
int fgpe_read_initial_sol(const char* path, int32_t dim, const int32_t* n, int32_t ms, double* phi, double* x,
                          double* y, int32_t threads);

/* Device time in milliseconds of the last fgpe_run call, measured with HIP
 * events on the handle's stream (no reference counterpart; the reference
 * only times the whole program, SRC/cgpe3d.f90:218,606-614). */
int fgpe_last_run_ms(fgpe_handle* h, double* ms);

/* Per-kernel device times of a timed run: advances nsteps steps like
 * fgpe_run, with whole-grid launches on the handle's stream (one launch per
 * pass and step, whatever the schedule), while recording HIP events around
 * every kernel launch, and returns in ms_out[FGPE_NPASS] the average duration
 * in milliseconds of each pass kind (0 where a kind did not run) and in
 * count_out[FGPE_NPASS] how many launches it averaged.  No reference
 * counterpart; it backs the per-pass bandwidth report. */
#define FGPE_PASS_X 0     /* inverse-x + SE + SP + NORMC (+ forward-x) */
#define FGPE_PASS_X0 1    /* forward-x only (first step of a run) */
#define FGPE_PASS_Y 2     /* 3D y lines, one component: forward-y, kx/ky KE */
#define FGPE_PASS_K 3     /* 2D k-space lines (KE + SOC); 3D y lines with SOC + inverse-y */
#define FGPE_PASS_FIN 4   /* NORMC finalisation / renormalisation scale */
#define FGPE_PASS_1D 5    /* q1D whole-step kernel */
#define FGPE_PASS_Z 6     /* 3D z lines, one component: forward-z, KE, inverse-z */
#define FGPE_NPASS 7
int fgpe_pass_times(fgpe_handle* h, int64_t nsteps, double* ms_out, int32_t* count_out);

/* Message of the last failing call on this thread. */
int fgpe_last_error(char* buf, size_t len);

/* ------------------------------------------------------------------------
 * Batches: many independent q2D instances advanced together.
 *
 * No reference counterpart (the reference runs one cgpe.x per parameter
 * set); each instance follows exactly the single-instance step of
 * fgpe_run (SRC/cgpe2d.f90:440-485) and the convergence test of cgpe.x
 * (CONV = max|PHI - PHI_OLD|/(STP*2*DT), CONV_EN = |EN - EN_OLD|/(STP*2*DT),
 * converged when either is <= TOL; SRC/cgpe2d.f90:488-495).  Instances are
 * grouped by grid size and every pass of a step is one launch per size over
 * all of that size's unconverged instances (blocks map to instance x lines),
 * so a step of any number of instances costs a fixed handful of launches.
 * Instances: dim 2, spin 1 or 2, 64, 128, 192 or 256 points along each axis
 * (square or not); one spin, one precision and one time mode (real or
 * imaginary) for the batch.  Instances are grouped by grid shape.
 */
typedef struct fgpe_batch fgpe_batch;

typedef struct fgpe_batch_result {
  int64_t steps;    /* steps advanced so far */
  int32_t done;     /* 1 once CONV <= TOL or CONV_EN <= TOL */
  int32_t err;      /* nonzero if the SOC = 0 projection failed */
  double en;        /* EN of the last event (scaled units, as fgpe_observables) */
  double mz;        /* MZ */
  double conv;      /* CONV of the last event */
  double conv_en;   /* CONV_EN of the last event */
  double maxdiff;   /* max |PHI - PHI_OLD| of the last event */
  double mu[5];     /* MU(1..MS) */
  double rms[5];    /* RMS(1..MS) */
  double norm[5];   /* NORMC of the last step, before renormalisation */
} fgpe_batch_result;

/* Schedules (points per thread of the x pass / the k pass), two per line
 * length: s = 2*c + v for c = 0..3 (64, 128, 192, 256).  An instance's x pass
 * follows the schedule of its x length, its k pass that of its y length.  FP32: 64 (4/8, 8/4),
 * 128 (4/8, 4/4), 192 (12/12, 6/6), 256 (4/8, 4/4); FP64: 64 (8/4, 4/8),
 * 128 (8/4, 4/8), 192 (6/6, 12/6), 256 (8/4, 4/8). */
#define FGPE_BATCH_NSCHED 8

/* Allocate a batch of count instances described by p[0..count-1] (each as
 * fgpe_create would take it). */
int fgpe_batch_create(const fgpe_params* p, int32_t count, fgpe_batch** b);
void fgpe_batch_destroy(fgpe_batch* b);

/* Gaussian initial guess and NORMT for every instance (as fgpe_initial_guess
 * followed by fgpe_normalize), then the observables of the start state. */
int fgpe_batch_initial_guess(fgpe_batch* b);

/* NORMT of every instance (as fgpe_normalize) and the observables of that
 * state, resetting the results: the start of a solve from uploaded states
 * (SRC/cgpe2d.f90:401-407, the guess of INITIALIZE replaced by the
 * uploaded states). */
int fgpe_batch_normalize(fgpe_batch* b);

/* Copy instance i in or out (layout of fgpe_upload/fgpe_download). */
int fgpe_batch_upload(fgpe_batch* b, int32_t i, const double* phi);
int fgpe_batch_download(fgpe_batch* b, int32_t i, double* phi);

/* Advance every instance (converged or not) by nsteps steps; afterwards each
 * result holds NORMC of the last step, the observables and maxdiff (PHI_OLD
 * = the state one step before the end). */
int fgpe_batch_run(fgpe_batch* b, int64_t nsteps);

/* Run to convergence: repeated rounds of stp steps over the unconverged
 * instances until all have CONV <= tol or CONV_EN <= tol, or max_steps.
 * *converged receives the number of converged instances. */
int fgpe_batch_solve(fgpe_batch* b, int32_t stp, double tol, int64_t max_steps, int32_t* converged);

/* Result of instance i. */
int fgpe_batch_result_get(fgpe_batch* b, int32_t i, fgpe_batch_result* r);

/* Autotuner: for every line length present, time trial_steps steps of the
 * instances of that square shape (or, with none, of every instance with lines
 * of that length) under each of its schedules (state restored afterwards)
 * and keep the fastest.  ms_out[FGPE_BATCH_NSCHED] receives the
 * measured milliseconds per schedule (0 where not timed), chosen[4] the
 * schedule kept for 64, 128, 192 and 256 points (-1 where absent). */
int fgpe_batch_tune(fgpe_batch* b, int32_t trial_steps, double* ms_out, int32_t* chosen);

/* Device time in ms of the last fgpe_batch_run / fgpe_batch_solve. */
int fgpe_batch_last_ms(fgpe_batch* b, double* ms);

/* ------------------------------------------------------------------------
 * Initial-guess surrogate: a fully connected network y = act(x W + b)
 * (ReLU between layers, none after the last), W stored [K][N] row major,
 * trained with mean-squared error and Adam.  Products run on the matrix
 * cores in BF16 with FP32 accumulation; master weights, biases and the loss
 * stay FP32.  No reference counterpart: the reference starts every run from
 * its Gaussian or Thomas-Fermi guess (SRC/cgpe2d.f90:697-771).  Parameters
 * are packed layer by layer as W (K*N floats) then b (N floats).
 */
typedef struct fgpe_mlp fgpe_mlp;

/* dims[0..nlayers]: input width, hidden widths, output width. */
int fgpe_mlp_create(const int32_t* dims, int32_t nlayers, int32_t max_batch, uint64_t seed, fgpe_mlp** m);
void fgpe_mlp_destroy(fgpe_mlp* m);
int fgpe_mlp_param_count(fgpe_mlp* m, int64_t* n);
int fgpe_mlp_get_params(fgpe_mlp* m, float* params);
int fgpe_mlp_set_params(fgpe_mlp* m, const float* params);
/* y[batch][out] = network(x[batch][in]) (y may be NULL). */
int fgpe_mlp_forward(fgpe_mlp* m, const float* x, int32_t batch, float* y);
/* Gradients of the mean-squared error against t[batch][out], packed like the
 * parameters (grad may be NULL), and the loss. */
int fgpe_mlp_gradients(fgpe_mlp* m, const float* x, const float* t, int32_t batch, float* grad, double* loss);
/* Keep a training set on the device: x[rows][in], y[rows][out]. */
int fgpe_mlp_set_dataset(fgpe_mlp* m, const float* x, const float* y, int32_t rows);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on dataset rows idx[0..batch). */
int fgpe_mlp_train_step(fgpe_mlp* m, const int32_t* idx, int32_t batch, float lr, double* loss);
/* The matrix-core product alone: C[M][N] = op(A) op(B) for BF16 bit patterns
 * (ta: A given as [K][M]; tb: B given as [N][K]). */
int fgpe_mlp_gemm_test(int32_t ta, int32_t tb, const uint16_t* a, const uint16_t* b, float* c, int32_t m, int32_t n,
                       int32_t k);
/* Device time of that product on resident operands (filled on the device),
 * averaged over reps launches after one untimed launch, in *ms; 2 m n k / ms
 * is its FLOP rate. */
int fgpe_mlp_gemm_time(int32_t ta, int32_t tb, int32_t m, int32_t n, int32_t k, int32_t reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* FGPE_H */

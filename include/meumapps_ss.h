/*
 * meumapps_ss.h -- C ABI of the MI355X semi-implicit spectral step of MEUMAPPS-SS.
 *
 * The reference has no plugin API: the time step is inline in the main
 * program (Multi_MEUMAPPS.f:548-1106) and its only library seam is the
 * P3DFFT module called from f_trans / inv_trans (ModFunc.f:15-40).  This
 * header is the seam that replaces both: the Fortran driver keeps input
 * parsing, set-up, initial conditions and dumps, hands the post-set-up
 * ModGparams state over once (mm_params), uploads the fields once, and then
 * calls mm_step in place of the loop body.  The whole periodic grid stays in
 * the GPU's HBM between calls.
 *
 * Conventions
 *   - Plain C types only; every host array is caller-owned and laid out the
 *     way the reference's Fortran arrays are (column-major, component index
 *     fastest for phi(V,Nx,Ny,Nz) and Cons(E,Nx,Ny,Nz)).
 *   - Every function returns 0 on success or one of the MM_ERR_* codes; the
 *     positive codes reuse the reference's err_flag values (ModFunc.f:44-72,
 *     Manual p.9) so a driver can print the reference's STOP messages.
 *   - One context = one device; a context is not reentrant.  It owns its HIP
 *     stream and runs the step on it; with MM_CH_STREAM=1 in the environment
 *     at mm_create an eager step on the fused path runs the Cahn-Hilliard
 *     update on a second stream of its own, joined within the step.
 */
#ifndef MEUMAPPS_SS_H
#define MEUMAPPS_SS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_MAX_VAR 16 /* order parameters (vartot)                       */
#define MM_MAX_PH 4   /* phases including the matrix (numph)             */
#define MM_MAX_EL 4   /* independent compositions (numel after the -1)   */
#define MM_GP_STRIDE (2 * MM_MAX_EL + 1) /* rows of Gparams per phase    */

/* precision of the step */
#define MM_FP64 0  /* every field, FFT and pointwise kernel in binary64       */
#define MM_MIXED 1 /* FP32 FFTs and pointwise kernels; FP64 k=0 mode, mass
                      sums and semi-implicit denominators                     */

/* return codes */
#define MM_OK 0
#define MM_ERR_ZSOLVE 2      /* KKS Zintf singular   (err_flag 2, Multi_MEUMAPPS.f:690) */
#define MM_ERR_QINV 3        /* acoustic tensor      (err_flag 3, ModElastic.f:187)     */
#define MM_ERR_CINV_ITER 4   /* 6x6 compliance mix   (err_flag 4, ModElastic.f:270)     */
#define MM_ERR_CINV_ENER 6   /* 6x6 compliance mix   (err_flag 6, ModElastic.f:599)     */
#define MM_ERR_NEWFAC 7      /* elastic backtracking newfac < 1e-7 (ModElastic.f:446)    */
#define MM_ERR_ARG 20001     /* bad argument / unsupported size                         */
#define MM_ERR_NOMEM 20002   /* device allocation failed                                */
#define MM_ERR_HIP 20003     /* HIP runtime error                                       */
#define MM_ERR_NODEV 20004   /* no GPU visible                                          */

/*
 * Post-set-up contents of ModGparams (ModGparams.f:6-90) that the step reads.
 * Multi-dimensional Fortran arrays are stored column-major with a fixed
 * leading stride so the struct can be mirrored by a BIND(C) derived type:
 *   grad_coeff(i,j,v) -> grad_coeff[(v-1)*9 + (j-1)*3 + (i-1)]  (already / del_N**2)
 *   trans_ppt(m,v)    -> trans_ppt[(v-1)*6 + (m-1)]
 *   CEla(i,j,p)       -> cela[(p-1)*36 + (j-1)*6 + (i-1)]        (after eltensor)
 *   SEla(i,j,p)       -> sela[(p-1)*36 + (j-1)*6 + (i-1)]        (= CEla^-1)
 *   Gparams(j,p)      -> gparams[(p-1)*MM_GP_STRIDE + (j-1)]     (after the /1.055e-5
 *                        conversion and disc_ener, Multi_MEUMAPPS.f:218-236)
 *   Nvar(p)           -> nvar[p-1]  (nvar[0] == 1: the matrix)
 */
typedef struct mm_params {
  int32_t nx, ny, nz;      /* Nx, Ny, Nz (even, 2^a 3^b, 8..512)                 */
  int32_t vartot;          /* V                                                  */
  int32_t numel;           /* E (numel - 1 of Input.txt)                         */
  int32_t numph;           /* phases including the matrix                       */
  int32_t nvar[MM_MAX_PH]; /* Nvar(1..numph)                                     */
  int32_t step;            /* current value of the reference's `step`            */
  int32_t elastep0;        /* elasticity runs when step >= elastep0              */
  int32_t noise_step;      /* Langevin noise is added while step <= noise_step   */
  int32_t pad0;
  double t_step;           /* dt                                                 */
  double b_mob;            /* Allen-Cahn mobility L                              */
  double cmobs[MM_MAX_EL]; /* Cahn-Hilliard mobility per composition           */
  double grad_coeff[MM_MAX_VAR * 9];
  double trans_ppt[MM_MAX_VAR * 6];
  double cela[MM_MAX_PH * 36];
  double sela[MM_MAX_PH * 36];
  double gparams[MM_MAX_PH * MM_GP_STRIDE];
  double wellph[MM_MAX_PH];
  double e_mean[6];        /* imposed mean strain (Voigt, tensor shear)          */
} mm_params;

/*
 * Per-step diagnostics: the fort.98 row (Multi_MEUMAPPS.f:905-906) and the
 * fort.99 row (:1064) of one step.  glob_min, glob_max, phi_max and
 * solid_tot lag by one step exactly as in the reference, because the
 * reference writes fort.98 before its check loop (:1008-1065).
 */
typedef struct mm_diag {
  int32_t step;        /* value of `step` when the row was written          */
  int32_t n_elastic;   /* elastic fixed-point iterations run in the step    */
  double fint, fw, fch, elen, etot;
  double glob_min, glob_max, phi_max, solid_tot;
  double caviter[MM_MAX_EL];
  double cfac;         /* final relaxation factor of the elastic iteration  */
} mm_diag;

typedef struct mm_ctx mm_ctx;
typedef struct mm_fft mm_fft;

/* ---- whole-step interface: replaces the loop body Multi_MEUMAPPS.f:548-1106 ---- */

/* Allocate device state for one grid.  precision = MM_FP64 or MM_MIXED. */
int mm_create(const mm_params *p, int precision, mm_ctx **out);
void mm_destroy(mm_ctx *ctx);

/* Upload the restart record set of data_save.NNNN (Multi_MEUMAPPS.f:473-480):
 *   phi(V,Nx,Ny,Nz), cons(E,Nx,Ny,Nz), el_en(Nx,Ny,Nz) (may be NULL),
 *   u = u1|u2|u3 (3 fields), gradu = grad_1_u1, grad_1_u2, grad_1_u3,
 *   grad_2_u1, ..., grad_3_u3 (9 fields, record order of :478-480).  */
int mm_upload(mm_ctx *ctx, const double *phi, const double *cons, const double *el_en,
              const double *u, const double *gradu);

/* Download the same record set (the dump block Multi_MEUMAPPS.f:1087-1098).
 * Any pointer may be NULL to skip that record. */
int mm_download(mm_ctx *ctx, double *phi, double *cons, double *el_en, double *u,
                double *gradu);

/* Advance nsteps steps.  diag (may be NULL) receives nsteps rows.  The
 * step counter in the context advances like the reference's `step`. */
int mm_step(mm_ctx *ctx, int32_t nsteps, mm_diag *diag);

int mm_get_step(const mm_ctx *ctx);
int mm_set_step(mm_ctx *ctx, int32_t step);

/* ---- Langevin noise (Multi_MEUMAPPS.f:552-571, 963-978) ----
 * Enable the noise of the steps with step <= noise_step: every such step
 * adds one langevin(iseed, sd) draw per point and variant (mm_langevin
 * below), in the reference's order (z, y, x outer to inner, variants
 * innermost), to the accumulated field phi_noise, whose transform enters the
 * Allen-Cahn update (+ t_step * F(phi_noise), :973); later steps reset
 * phi_noise to 0 (:569).  (idum, state) is the ran_2 generator at the first
 * noise step, as for mm_ran2: for a restart run idum = iseed - myid (Input.txt
 * `iseed`, :140, :229) and state = NULL (program-start state).  The draws are
 * generated on the host (the stream is sequential).  Without this call a
 * step with step <= noise_step returns MM_ERR_ARG.  Needs b_mob != 0.
 * mm_get_noise_state returns the generator after the last noise step. */
int mm_set_noise(mm_ctx *ctx, double sd, int32_t idum, const int32_t *state);
int mm_get_noise_state(const mm_ctx *ctx, int32_t *idum, int32_t *state);

/* Graph mode (no reference counterpart): the regular step (step > elastep0)
 * is recorded once into a HIP graph and replayed, with the elastic
 * convergence control of ModElastic.f:435-481 running on the device over
 * `iters` recorded iterations (0 keeps the current count, default 4); a step
 * whose elastic loop needs more iterations finishes eagerly after the
 * replay.  Results are identical to eager mode. */
int mm_set_graph(mm_ctx *ctx, int32_t on, int32_t iters);

/* Adjoint of the step (no reference counterpart): reverse-mode derivative of
 * the regular step (step > elastep0) for FP64 contexts with the same elastic
 * moduli in every phase.  State of the derivative: lam_phi (V fields), lam_cons
 * (E), lam_strain (6 engineering strain components), host arrays in the
 * mm_download field order (structure of arrays, x fastest per field).
 *
 *   mm_adj_enable      allocate the tape and the adjoint buffers (the fused x
 *                      path and graph mode are switched off on this context)
 *   mm_state_bytes / mm_state_save / mm_state_load
 *                      device-memory record of the dynamic state (phi, c, strain,
 *                      u_hat, step, lagged check-loop values and, in MM_MIXED, the
 *                      FP64 mean composition the stored FP32 deviation is taken
 *                      from): checkpoints; a record loads into any context of the
 *                      same model and precision (MM_ERR_ARG for another E)
 *   mm_adj_tape_step   one forward step from the current state, recording what
 *                      its adjoint needs (state advances as in mm_step)
 *   mm_adj_step        pull the adjoint state back through the taped step and
 *                      add the parameter derivatives into *grads
 *   mm_adj_set / mm_adj_get / mm_adj_zero
 *                      adjoint state from / to host memory
 *
 * grads holds dL/d of the mm_params arrays of the same names (grad_coeff
 * entries (i,j,v) at [v*9 + j*3 + i], trans_ppt (m,v) at [v*6 + m]). */
typedef struct mm_adj_grads {
  double grad_coeff[MM_MAX_VAR * 9];
  double wellph[MM_MAX_PH];
  double cmobs[MM_MAX_EL];
  double b_mob;
  double trans_ppt[MM_MAX_VAR * 6];
} mm_adj_grads;
int mm_adj_enable(mm_ctx *ctx);
int mm_adj_zero(mm_ctx *ctx);
int mm_adj_set(mm_ctx *ctx, const double *lam_phi, const double *lam_cons, const double *lam_strain);
int mm_adj_get(mm_ctx *ctx, double *lam_phi, double *lam_cons, double *lam_strain);
int mm_state_bytes(const mm_ctx *ctx, int64_t *bytes);
int mm_state_save(mm_ctx *ctx, void *dev_dst);
int mm_state_load(mm_ctx *ctx, const void *dev_src);
int mm_adj_tape_step(mm_ctx *ctx, mm_diag *diag);
int mm_adj_step(mm_ctx *ctx, mm_adj_grads *grads);

/* Branch tape (no reference counterpart; for derivative checks).  The step
 * takes two data-dependent branches per point and variant: the phi > 0 tests
 * of the f_diff correction (Multi_MEUMAPPS.f:829-831) and the clamp of the
 * check loop (:1034-1035).  Along a run the map from inputs to outputs is
 * smooth only while those decisions stay fixed; the tape records them, counts
 * how many a perturbed run changes, or replays them, so that finite
 * differences can be taken on the frozen-branch map whose derivative the
 * adjoint computes.  dev_buf is a caller-owned device buffer of nsteps *
 * mm_branch_bytes() bytes (one byte per point and variant per step); each
 * eager step uses the next slice.  mode: 0 off, 1 record, 2 count (the step
 * decides as usual; decisions that differ from the tape are counted), 3
 * replay (the tape's decisions are taken).  While a tape is set, steps run
 * eagerly with the unfused clamp.  A step past nsteps returns MM_ERR_ARG.
 * mm_branch_flips returns the count of mode 2 since mm_branch_tape. */
#define MM_BRANCH_OFF 0
#define MM_BRANCH_RECORD 1
#define MM_BRANCH_COUNT 2
#define MM_BRANCH_REPLAY 3
int mm_branch_bytes(const mm_ctx *ctx, int64_t *bytes_per_step);
int mm_branch_tape(mm_ctx *ctx, int32_t mode, void *dev_buf, int32_t nsteps);
int mm_branch_flips(mm_ctx *ctx, int64_t *flips);

/* Ensembles (no reference counterpart; the reference runs one simulation per
 * MPI job, Multi_MEUMAPPS.f:548-1106).  n independent members, each with its
 * own parameters and grid (params[0..n-1]) in one precision.  Fields go in
 * and out through the member contexts (mm_batch_member + mm_upload /
 * mm_download / mm_field / mm_set_step); those contexts belong to the batch
 * and must not be destroyed separately.  mm_batch_step advances every member
 * by nsteps; diag, when not NULL, holds nsteps * n rows, row k * n + i for
 * member i at step k.  Steps where every member is past elastep0 run as one
 * graph launch (one branch per member, graph_iters elastic iterations each as
 * set on the member contexts before the first such step); other steps run the
 * members one after another.  Each member's results equal its single run. */
typedef struct mm_batch mm_batch;
int mm_batch_create(int32_t n, const mm_params *params, int32_t precision, mm_batch **out);
int mm_batch_member(mm_batch *batch, int32_t i, mm_ctx **ctx);
int mm_batch_step(mm_batch *batch, int32_t nsteps, mm_diag *diag);
void mm_batch_destroy(mm_batch *batch);

/* Device-resident field access for in-process callers (torch, tests):
 * returns the device address of one SoA field and its element size (8 or 4).
 * In MM_MIXED the device composition field (which = 1) holds the FP32
 * deviation c - cbar from the FP64 mean cbar of the last upload;
 * mm_field_to_host adds cbar back.
 * which: 0 phi, 1 cons, 2 el_en, 3 f_diff (last step), 4 strain (6 fields:
 * e11, e22, e33, 2e23, 2e13, 2e12 of the displacement u; e_mean is not
 * included), 5 grad_phi (3V fields),
 * 6 k_del_sq (V fields), 7 el_grad (V fields).  With the fused x path
 * (homogeneous moduli) the strain is kept as x half-spectra between steps:
 * which = 4 forms the real fields on this call, valid until the next step. */
int mm_field(mm_ctx *ctx, int32_t which, void **dev_ptr, int32_t *elem_size,
             int32_t *nfields);

/* Copy one SoA field group to the host as binary64 (converts in MIXED). */
int mm_field_to_host(mm_ctx *ctx, int32_t which, double *host);

/* Live kernel timing (no reference counterpart; the reference only times
 * eldis with MPI_Wtime, Multi_MEUMAPPS.f:861-871).  mm_probe selects one
 * kernel class; every launch of it inside mm_step is bracketed by HIP events
 * on the context's stream.  mm_probe_read returns the accumulated kernel
 * time, launch count and algorithmic bytes (reads + writes) of those launches
 * and resets the counters.  kind = -1 switches timing off. */
#define MM_PROBE_X_R2C 0
#define MM_PROBE_Y 1
#define MM_PROBE_Z 2
#define MM_PROBE_X_C2R 3
#define MM_PROBE_CHEM 4
#define MM_PROBE_POLAR 5
#define MM_PROBE_ELENER 6
#define MM_PROBE_KS_ELASTIC 7
#define MM_PROBE_ZF_ELASTIC 8 /* fused z pass of the elastic iteration */
#define MM_PROBE_ZF_OTHER 9   /* fused z passes of gradients, CH, AC */
#define MM_PROBE_X_POLAR 10   /* fused x pass + polarization stress of the elastic iterations */
int mm_probe(mm_ctx *ctx, int32_t kind);
int mm_probe_read(mm_ctx *ctx, double *ms_total, int64_t *launches, double *bytes_total);

/* FFT plan autotuner (no reference counterpart): at mm_create / mm_fft_create
 * every radix plan of the x_r2c, x_c2r, y and z passes and of the y pass of
 * the fused path (forward + inverse with the change to the column-tiled
 * layout) (default, lean, 4-point) is timed on the grid's own size and the
 * fastest kept; set the environment variable MM_AUTOTUNE=0 to keep the
 * defaults.  The choice is made once per grid size, precision and device in a
 * process and reused by every later context of that shape (so contexts of one
 * process, ensemble members and single runs alike, round the same way).  variant[5] gets the chosen plan per pass, ms[15] the measured
 * kernel times (ms, 4 fields). */
int mm_get_tuning(const mm_ctx *ctx, int32_t *variant, float *ms);

/* ---- FFT interface: replaces f_trans / inv_trans over P3DFFT (ModFunc.f:15-40) ----
 * Spectrum layout (device): element (kx, ky, kz), kx in [0, Nx/2],
 * at index ((kz * (Nx/2+1)) + kx) * Ny + ky, interleaved complex.           */
int mm_fft_create(int32_t nx, int32_t ny, int32_t nz, int32_t precision, mm_fft **out);
void mm_fft_destroy(mm_fft *f);
/* f_trans: X = (1/(Nx Ny Nz)) * sum x exp(-2 pi i k.n / N)  (ModFunc.f:22-23) */
int mm_fft_forward(mm_fft *f, const void *real_dev, void *spec_dev, void *hip_stream);
/* inv_trans: unnormalised c2r; the imaginary parts of the kx = 0 and
 * kx = Nx/2 entries are ignored after the y and z passes (FFTW c2r
 * semantics, ModFunc.f:37).  The spectrum is overwritten (used as scratch). */
int mm_fft_inverse(mm_fft *f, void *spec_dev, void *real_dev, void *hip_stream);
size_t mm_fft_spectrum_elems(const mm_fft *f);

/* mm_fft_forward without a plan and without any allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The grid is given by (nx, ny, nz, precision) with the same
 * limits as mm_fft_create: every axis 8..512 of the form 2^a 3^b, Nx/2 one of
 * the x-pass lengths; other sizes return MM_ERR_ARG.  It runs the default
 * plan of every pass (the plans mm_fft_create keeps with MM_FFT_VARIANT=0),
 * so its spectra equal those of such a plan bit for bit.  The twiddle tables
 * of every supported length live in a file-scope device array of 64 KiB per
 * GPU (held until the process ends), filled once at the first call on each
 * device (a synchronous host-to-device copy) and only read afterwards, so
 * concurrent calls on one GPU do not interfere through it. */
int mm_fft_forward_noalloc(int32_t nx, int32_t ny, int32_t nz, int32_t precision, const void *real_dev,
                           void *spec_dev, void *hip_stream);

/* ---- host helpers (host-side set-up; no device work) ---- */
/* Numerical-Recipes ran2 as in ran_2.f:1-42, including the float32-folded
 * constants; state is kept in *state (35 int32) instead of SAVE storage. */
double mm_ran2(int32_t *idum, int32_t *state);
/* n consecutive ran2 draws into out (the initial-field loops of
 * ModInitials.f:415-690 draw one per point). */
int mm_ran2_fill(int32_t *idum, int32_t *state, int64_t n, double *out);
/* langevin(iseed, sd, harvest) (ModFunc.f:77-100): polar Box-Muller on two
 * ran2 draws per attempt; returns the Gaussian number times sd. */
double mm_langevin(int32_t *idum, int32_t *state, double sd);
/* The noise loop of one step for one process (Multi_MEUMAPPS.f:554-562):
 * phi_noise[v*n3 + i] += mm_langevin(...) for the points i in linear
 * (z, y, x) order with the variants innermost; the draws and the generator
 * left behind are those of the sequential calls (threads are used where the
 * stream allows it). */
int mm_noise_accumulate(int32_t *idum, int32_t *state, double sd, int32_t nvar, int64_t n3,
                        double *phi_noise);
/* Uniform initial fields c = cav + a*(2r-1), phi_v = a*(2r-1) drawn from one
 * ran2 stream in Fortran order (x fastest; per cell: c_1..c_E then phi_1..V). */
int mm_ic_uniform(int32_t nx, int32_t ny, int32_t nz, int32_t nvar, int32_t nel,
                  int32_t iseed, double amp, const double *cav, double *phi,
                  double *cons);

const char *mm_errstr(int code);

#ifdef __cplusplus
}
#endif
#endif

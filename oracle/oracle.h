/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU binary64 restatement of the MEUMAPPS-SS semi-implicit spectral step
 * (Multi_MEUMAPPS.f:548-1106, ModElastic.f:11-691, ModFunc.f:15-72,
 * ModInitials.f:348-409).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Fields are structure-of-arrays, each (Nx,Ny,Nz) with x fastest (the
 * reference's Fortran order without the leading component index).
 * Spectra use [kz][ky][kx] with kx in [0, Nx/2] fastest.
 */
#ifndef MM_ORACLE_H
#define MM_ORACLE_H
#include <stdint.h>
#include "../include/meumapps_ss.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

oracle_t *oracle_create(const mm_params *p, int nthreads);
void oracle_destroy(oracle_t *o);
int oracle_set_threads(oracle_t *o, int nthreads);

/* SoA state access: phi (V*N3), cons (E*N3), u (3*N3), gradu (9*N3 in the
 * record order grad_1_u1, grad_1_u2, grad_1_u3, grad_2_u1, ...), el_en (N3). */
int oracle_set_state(oracle_t *o, const double *phi, const double *cons, const double *u,
                     const double *gradu, const double *el_en);
int oracle_get_state(oracle_t *o, double *phi, double *cons, double *u, double *gradu,
                     double *el_en);
/* intermediate fields of the last step: which = 0 grad_x|grad_y|grad_z (3V),
 * 1 k_del_sq (V), 2 f_grad (V), 3 f_diff (V, incl. el_grad), 4 el_grad (V),
 * 5 Fconx|Fcony|Fconz (3E), 6 P_phi (V) */
int oracle_get_work(oracle_t *o, int which, double *out);
int oracle_get_step(oracle_t *o);
/* check-loop values of the last step (glob_min, glob_max, phi_max, solid
 * sum): the lagged columns of the next row; get / set for checkpoints */
void oracle_get_lagged(oracle_t *o, double *v4);
void oracle_set_lagged(oracle_t *o, const double *v4);
void oracle_set_step(oracle_t *o, int step);

/* one full step of the loop body; returns 0 or an MM_ERR_* code */
int oracle_step(oracle_t *o, mm_diag *d);

/* f_trans / inv_trans restated (ModFunc.f:15-40) */
int oracle_f_trans(oracle_t *o, const double *in, double *spec /* 2*Nz*Ny*H */);
int oracle_inv_trans(oracle_t *o, const double *spec, double *out);

/* LINPACK-style inverse (ModFunc.f:44-72): returns 0 or flag when singular */
int oracle_invert(double *a /* n*n column-major */, int n, int flag);

/* k-vector tables for one axis of length n (ModInitials.f:355-401):
 * kf_im[m] = imaginary part of kf, kq[m] = kf_sq (real, negative) */
void oracle_kaxis(int n, int count, double *kf_im, double *kq);

/* ran_2.f:1-43 and langevin (ModFunc.f:77-100); st holds 35 int32 (idum2,
 * iy, iv(32), init flag), all zero at program start */
double oracle_ran2(int32_t *idum, int32_t *st);
double oracle_langevin(int32_t *idum, int32_t *st, double sd);
/* the noise loop of one step, sequentially: acc[v*n3 + i] += langevin */
void oracle_noise_accumulate(int32_t *idum, int32_t *st, double sd, int nvar, long n3, double *acc);
/* noise of the steps step <= noise_step (Multi_MEUMAPPS.f:552-571, 963-978)
 * from the generator (idum, state; state NULL = program start) */
int oracle_set_noise(oracle_t *o, double sd, int32_t idum, const int32_t *state);
/* generator after the last noise step and the accumulated phi_noise (V*N3);
 * any pointer may be NULL */
int oracle_get_noise(oracle_t *o, int32_t *idum, int32_t *state, double *phi_noise);

/* elastic solve alone (eldis + elener) on the current state, given P_phi
 * fields computed from phi; returns 0 or error and the iteration count */
int oracle_eldis(oracle_t *o, int *n_iter, double *cfac);
/* P_phi = p(phi) of the current phi alone */
int oracle_set_pphi(oracle_t *o);
/* elener alone at the current phi, P_phi and grad u (el_en: oracle_get_state,
 * el_grad: oracle_get_work(4)) */
int oracle_elener(oracle_t *o);
/* the zeroth-order guess of step == elastep0 alone (ModElastic.f:105-236)
 * from given P_phi fields (V*N3): u (3*N3) and grad u (9*N3, d_d u_i at
 * d*3+i); either output may be NULL */
int oracle_elastic_zeroth(oracle_t *o, const double *pphi, double *u, double *gradu);

/* Tangent-linear step (forward-mode derivative along a direction).  Once a
 * direction is set, every oracle_step also advances the tangent of
 * (phi, c, grad u), taking the primal's branch decisions (phi > 0 in the
 * f_diff correction, the clamp, the elastic iteration count and cfac
 * sequence).  dpar is a direction in the parameter arrays grad_coeff,
 * wellph, cmobs, b_mob and trans_ppt (mm_adj_grads layout); NULL pointers
 * are zero directions.  A noise step or step == elastep0 returns MM_ERR_ARG
 * while a tangent is set. */
int oracle_tlm_set(oracle_t *o, const double *dphi, const double *dcons, const double *dgradu,
                   const mm_adj_grads *dpar);
int oracle_tlm_get(oracle_t *o, double *dphi, double *dcons, double *dgradu);
void oracle_tlm_off(oracle_t *o);

#ifdef __cplusplus
}
#endif
#endif

// step.hip -- device state, kernels and host orchestration of one
// MEUMAPPS-SS semi-implicit spectral step (Multi_MEUMAPPS.f:548-1106), and
// the mm_* whole-step entry points of include/meumapps_ss.h.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fft3d.h"
#include "mm_common.h"
#include "step_kernels.h"
#include "zfused.h"

namespace mm {

template <typename T, class Op>
int launch_zfused(const Fft3& f, const Op& op, int ngroups, hipStream_t s, int probe_kind, double bytes,
                  int* nblk);
template <typename T>
int launch_zf_elastic(const Fft3& f, const OpElastic<T>& op, int v, hipStream_t s, int probe_kind, double bytes,
                      int* nblk);
template <typename T>
int launch_y_ch(const Fft3& f, const DevParams* P, const void* in, void* out, int ne, int e0, hipStream_t s);
template <typename T>
int launch_y_grad(const Fft3& f, const DevParams* P, const void* in, void* out, int nv, hipStream_t s);

// ============================================================== kernels

// One workgroup of kFinBlock threads; each thread keeps four independent
// partial results (fixed assignment of partials to accumulators), so the
// loads of a slot are in flight together; the combination order is fixed.
constexpr int kFinBlock = 1024;
template <int OP> __device__ __forceinline__ double fin_slot(const double* __restrict__ p, int nblk, double* sh) {
  const double id = OP == 0 ? 0.0 : (OP == 1 ? HUGE_VAL : -HUGE_VAL);
  double a0 = id, a1 = id, a2 = id, a3 = id;
  int b = threadIdx.x;
  for (; b + 3 * kFinBlock < nblk; b += 4 * kFinBlock) {
    const double x0 = p[b], x1 = p[b + kFinBlock], x2 = p[b + 2 * kFinBlock], x3 = p[b + 3 * kFinBlock];
    a0 = rop<OP>(a0, x0);
    a1 = rop<OP>(a1, x1);
    a2 = rop<OP>(a2, x2);
    a3 = rop<OP>(a3, x3);
  }
  for (; b < nblk; b += kFinBlock) a0 = rop<OP>(a0, p[b]);
  return block_reduce<OP, kFinBlock / 64>(rop<OP>(rop<OP>(a0, a1), rop<OP>(a2, a3)), sh);
}
__global__ void __launch_bounds__(kFinBlock) k_finalize(const double* __restrict__ part, int nblk, int K,
                                                        FinalizeOps ops, double* __restrict__ out,
                                                        const int* __restrict__ gate) {
  if (gated_off(gate)) return;
  __shared__ double sh[kFinBlock / 64];
  for (int k = 0; k < K; ++k) {
    const double* p = part + (size_t)k * nblk;
    const int op = ops.op[k];
    const double r = op == 0 ? fin_slot<0>(p, nblk, sh) : (op == 1 ? fin_slot<1>(p, nblk, sh) : fin_slot<2>(p, nblk, sh));
    if (threadIdx.x == 0) out[k] = r;
  }
}

// result slots in dres
enum { R_FINT = 0, R_FCH, R_FW, R_EL, R_ERR0 /* 6 */, R_CLAMP = 10 /* E+4 */, R_N = 24 };

// graph-mode helpers: fill, and the one-thread elastic control
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill(T* __restrict__ p, size_t n, T v, const int* __restrict__ gate) {
  if (gated_off(gate)) return;
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) p[i] = v;
}
// Allen-Cahn source with Langevin noise (Multi_MEUMAPPS.f:963-973):
// F(phi) - dt b F(f_diff) + dt F(phi_noise) = F(phi) - dt b F(f_diff - phi_noise / b),
// so the noise joins f_diff before the one forward transform
// a[i] += b[i] (f_diff + el_grad formed on demand)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_add_to(T* __restrict__ a, const T* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    a[i] = (T)((double)a[i] + (double)b[i]);
}
template <typename T>
__global__ void __launch_bounds__(kBlock) k_noise_fold(const T* __restrict__ fdiff, const T* __restrict__ noise,
                                                        T* __restrict__ out, size_t n, double b_mob) {
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    out[i] = (T)((double)fdiff[i] - (double)noise[i] / b_mob);
}
__global__ void k_ctl_start(Ctl* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl_start(*c);
}
__global__ void k_ctl_iter(Ctl* c, const double* __restrict__ err) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl_after_iter(*c, err);
}

// Stage layout after a replay (doubles): [0, 12) the Ctl block, [16, 16 + R_N)
// the result slots, [kStageErr] the error flag.
constexpr int kStage = 64, kStageRes = 16, kStageErr = 48;
static_assert(sizeof(Ctl) == 12 * sizeof(double), "Ctl block is 12 doubles");
static_assert(kStageRes + R_N <= kStageErr, "stage layout");
__global__ void k_gather(const Ctl* __restrict__ ctl, const double* __restrict__ dres, const int* __restrict__ derr,
                         double* __restrict__ out) {
  const int t = threadIdx.x;
  if (t < 12) out[t] = reinterpret_cast<const double*>(ctl)[t];
  if (t < R_N) out[kStageRes + t] = dres[t];
  if (t == 0) out[kStageErr] = (double)*derr;
}

#define GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < (n); i += (size_t)gridDim.x * kBlock)

// The acoustic tensor Q(k) of every non-zero wave vector must be invertible:
// eldis inverts it (invert(Q,3,3), ModElastic.f:187-189) and stops with
// err_flag 3 when the reciprocal 1-norm condition number is <= 1e-16.  Q
// depends only on CEla(:,:,2) and k, so a context checks it once, before its
// first step that runs elasticity; the inverse used for the condition number
// is the adjugate one of the step's acoustic solve.
__global__ void k_qcheck(const DevParams* __restrict__ P, size_t nc, int* __restrict__ err) {
  bool bad = false;
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) continue;
    double Q[3][3];
    acoustic_tensor(P, k, Q);
    const double c00 = Q[1][1] * Q[2][2] - Q[1][2] * Q[2][1];
    const double c01 = Q[1][2] * Q[2][0] - Q[1][0] * Q[2][2];
    const double c02 = Q[1][0] * Q[2][1] - Q[1][1] * Q[2][0];
    const double det = Q[0][0] * c00 + Q[0][1] * c01 + Q[0][2] * c02;
    if (det == 0.0) { bad = true; continue; }
    double A[3][3];
    A[0][0] = c00; A[1][0] = c01; A[2][0] = c02;
    A[0][1] = Q[0][2] * Q[2][1] - Q[0][1] * Q[2][2];
    A[1][1] = Q[0][0] * Q[2][2] - Q[0][2] * Q[2][0];
    A[2][1] = Q[0][1] * Q[2][0] - Q[0][0] * Q[2][1];
    A[0][2] = Q[0][1] * Q[1][2] - Q[0][2] * Q[1][1];
    A[1][2] = Q[0][2] * Q[1][0] - Q[0][0] * Q[1][2];
    A[2][2] = Q[0][0] * Q[1][1] - Q[0][1] * Q[1][0];
    double anorm = 0.0, inorm = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double a = fabs(Q[0][j]) + fabs(Q[1][j]) + fabs(Q[2][j]);
      const double b = (fabs(A[0][j]) + fabs(A[1][j]) + fabs(A[2][j])) / fabs(det);
      anorm = a > anorm ? a : anorm;
      inorm = b > inorm ? b : inorm;
    }
    const double rcond = (anorm > 0.0 && inorm > 0.0) ? 1.0 / (anorm * inorm) : 0.0;
    if (!(rcond > 1.0e-16)) bad = true;
  }
  if (bad) *err = MM_ERR_QINV;
}

// -------- Gaussian elimination, partial pivoting, N compile-time (KKS Zintf)
template <int N>
__device__ __forceinline__ bool solve_pivot(double (&A)[N][N], double (&b)[N]) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int p = k;
    double amax = fabs(A[k][k]);
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double a = fabs(A[i][k]);
      if (a > amax) { amax = a; p = i; }
    }
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      if (i == p) {
#pragma unroll
        for (int j = 0; j < N; ++j) { double t = A[k][j]; A[k][j] = A[i][j]; A[i][j] = t; }
        double t = b[k]; b[k] = b[i]; b[i] = t;
      }
    }
    if (!(amax > 0.0)) ok = false;
    const double inv = 1.0 / A[k][k];
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double l = A[i][k] * inv;
#pragma unroll
      for (int j = k + 1; j < N; ++j) A[i][j] -= l * A[k][j];
      b[i] -= l * b[k];
    }
  }
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    double s = b[k];
#pragma unroll
    for (int j = k + 1; j < N; ++j) s -= A[k][j] * b[j];
    b[k] = s / A[k][k];
  }
  return ok;
}

// -------- 6x6 solve without pivoting: the guarded mixture of the elastic
// iteration (ModElastic.f:262-268) is a nonnegative combination of SPD
// compliances once phi is in [0,1]
__device__ __forceinline__ bool solve6(double (&A)[6][6], double (&b)[6]) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (!(A[k][k] > 0.0)) ok = false;
    const double inv = 1.0 / A[k][k];
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const double l = A[i][k] * inv;
#pragma unroll
      for (int j = k + 1; j < 6; ++j) A[i][j] -= l * A[k][j];
      b[i] -= l * b[k];
    }
  }
#pragma unroll
  for (int k = 5; k >= 0; --k) {
    double s = b[k];
#pragma unroll
    for (int j = k + 1; j < 6; ++j) s -= A[k][j] * b[j];
    b[k] = s / A[k][k];
  }
  return ok;
}

__device__ __forceinline__ double pfun(double f) {  // p(phi), Multi_MEUMAPPS.f:653
  const double fs = f * f;
  return fs * f * (6.0 * fs - 15.0 * f + 10.0);
}
__device__ __forceinline__ double ppr(double f) {  // p'(phi), :656
  const double fs = f * f;
  return 30.0 * fs * (fs - 2.0 * f + 1.0);
}

// Branch tape of one step (mm_branch_tape): one byte per point and variant,
// bit 0 the phi > 0 test of the f_diff correction (Multi_MEUMAPPS.f:829-831),
// bits 1-2 the clamp of the check loop (:1034-1035; 0 in [0, 1], 1 below 0,
// 2 above 1).  mode 0 off, 1 record, 2 count (decide as usual, count the
// decisions that differ from the tape), 3 replay (take the tape's decisions).
struct BranchTape {
  uint8_t* buf;
  unsigned long long* flips;
  int mode;
};

// ---------------------------------------------------------------- chemistry
// KKS chemistry at one point (Multi_MEUMAPPS.f:634-803): p, p', Pph, Gph, the
// interface compositions (Zintf x = Zr, :677-710), Gibbs energies, chemical
// potential; returns the driving forces fg (before the f_diff correction),
// the flux coefficients cf[v][e] = p'(phi_v)(x_{e,1} - x_{e,p(v)}), so that
// F_e = sum_v cf[v][e] grad phi_v in variant order, and the energy densities
// fch, fw.  false: Zintf singular.
template <int V, int NPH, int E>
__device__ __forceinline__ bool chem_point(const DevParams* __restrict__ P, const double (&f)[V], const double (&c)[E],
                                           double (&fg)[V], double (&cf)[V][E], double& fch_out, double& fw_out) {
  constexpr int Z = NPH * E;
  double Ppr[V], Pph[NPH + 1], Gph[NPH + 1];
#pragma unroll
  for (int q = 0; q <= NPH; ++q) { Pph[q] = 0.0; Gph[q] = 0.0; }
#pragma unroll
  for (int q = 2; q <= NPH; ++q)
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (P->phase_of[v] == q) {
        const double fi = f[v], fs = fi * fi;
        const double pl = fs * fi * (6.0 * fs - 15.0 * fi + 10.0);
        Ppr[v] = 30.0 * fs * (fs - 2.0 * fi + 1.0);
        Pph[1] += pl;
        Pph[q] += pl;
        Gph[q] += fs * ((1.0 - fi) * (1.0 - fi));
#pragma unroll
        for (int jv = 0; jv < V; ++jv)
          if (jv != v) Gph[1] += fs * (f[jv] * f[jv]);
      }
  Pph[1] = 1.0 - Pph[1];
  // Zintf x = Zr (constant rows :426-448, mixture rows :677-688)
  double A[Z][Z], b[Z];
#pragma unroll
  for (int i = 0; i < Z; ++i) {
#pragma unroll
    for (int j = 0; j < Z; ++j) A[i][j] = P->zc[i * 16 + j];
    b[i] = P->zrc[i];
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i0 = Z - E + e;
#pragma unroll
    for (int ph = 0; ph < NPH; ++ph) A[i0][NPH * e + ph] = Pph[ph + 1];
    b[i0] = c[e];
  }
  const bool ok = solve_pivot<Z>(A, b);
  // b[NPH*e + ph] = Xintf(e+1, ph+1)
  const double* gp = P->gp;
#define GP(j, ph) gp[((ph) - 1) * MM_GP_STRIDE + ((j) - 1)]
  double Gf[NPH + 1], Mu[E];
#pragma unroll
  for (int ph = 1; ph <= NPH; ++ph) {
    double g = 0.0;
#pragma unroll
    for (int e = 1; e <= E; ++e) {
      const double dx = b[NPH * (e - 1) + ph - 1] - GP(e + E + 1, ph);
      g += GP(e, ph) * (dx * dx);
    }
    Gf[ph] = g + GP(E + 1, ph);
  }
  Gf[0] = 0.0;
#pragma unroll
  for (int e = 1; e <= E; ++e) Mu[e - 1] = 2.0 * GP(e, 1) * (b[NPH * (e - 1)] - GP(e + E + 1, 1));
  double fch = Pph[1] * Gf[1];
  double fw = P->well[0] * Gph[1];
#pragma unroll
  for (int q = 2; q <= NPH; ++q) {
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (P->phase_of[v] == q) {
        const double fi = f[v];
        double termch = Gf[q] - Gf[1];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double xd = b[NPH * e] - b[NPH * e + q - 1];
          termch = termch + Mu[e] * xd;
          cf[v][e] = Ppr[v] * xd;
        }
        double wg = 2.0 * P->well[q - 1] * fi * (1.0 - fi) * (1.0 - 2.0 * fi);
#pragma unroll
        for (int jv = 0; jv < V; ++jv)
          if (jv != v) wg = wg + 4.0 * P->well[0] * fi * (f[jv] * f[jv]);
        fg[v] = termch * Ppr[v] + wg;
      }
    fch = fch + Pph[q] * Gf[q];
    fw = fw + P->well[q - 1] * Gph[q];
  }
#undef GP
  fch_out = fch;
  fw_out = fw;
  return ok;
}

// gradient energy density 1/2 grad phi . G grad phi of variant v (:768-775)
__device__ __forceinline__ double grad_term(const DevParams* __restrict__ P, int v, double gx, double gy, double gz) {
  const double* g = P->gc + v * 9;
  return g[0] * (gx * gx) + g[4] * (gy * gy) + g[8] * (gz * gz) + (g[3] + g[1]) * gx * gy +
         (g[6] + g[2]) * gx * gz + (g[7] + g[5]) * gy * gz;
}

// f_diff correction (:824-841) from the driving forces, del.G del phi and the
// phi > 0 decisions
template <int V>
__device__ __forceinline__ double fdiff_corr(int jv, const double (&fg)[V], const double (&kd)[V], const bool (&pos)[V]) {
  double t11 = 0.0;
  int nt = 1;
  if (pos[jv]) {
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (v != jv && pos[v]) { t11 = t11 + fg[v] - kd[v]; nt += 1; }
  }
  return fg[jv] - t11 / (double)nt;
}

// Multi_MEUMAPPS.f:634-845: p, p', Pph, Gph, KKS interface compositions,
// Gibbs energies, chemical potential, driving forces f_grad, fluxes F and
// the f_diff correction.  Partial sums: f_int, fch, fw.
template <typename T, int V, int NPH, int E>
__global__ void __launch_bounds__(kBlock)
k_chem(const DevParams* __restrict__ P, const T* __restrict__ phi, const T* __restrict__ gphi,
       const T* __restrict__ kdel, const T* __restrict__ cons, T* __restrict__ fdiff,
       T* __restrict__ flux, double* __restrict__ part, int* __restrict__ err, size_t n3, BranchTape bt) {
  __shared__ double sh[kWaves];
  double s_fint = 0.0, s_fch = 0.0, s_fw = 0.0;
  bool bad = false;
  unsigned long long nflip = 0;
  GRID_STRIDE(pt, n3) {
    double f[V], gx[V], gy[V], gz[V], kd[V], c[E];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      f[v] = ld(phi, v * n3 + pt);
      gx[v] = ld(gphi, (3 * v + 0) * n3 + pt);
      gy[v] = ld(gphi, (3 * v + 1) * n3 + pt);
      gz[v] = ld(gphi, (3 * v + 2) * n3 + pt);
      kd[v] = ld(kdel, v * n3 + pt);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) c[e] = P->cbar[e] + ld(cons, e * n3 + pt);
    double fg[V], cf[V][E], fch, fw;
    if (!chem_point<V, NPH, E>(P, f, c, fg, cf, fch, fw)) bad = true;
    double Fx[E], Fy[E], Fz[E];
#pragma unroll
    for (int e = 0; e < E; ++e) { Fx[e] = 0.0; Fy[e] = 0.0; Fz[e] = 0.0; }
#pragma unroll
    for (int v = 0; v < V; ++v) {
      s_fint += 0.5 * grad_term(P, v, gx[v], gy[v], gz[v]);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        Fx[e] += cf[v][e] * gx[v];
        Fy[e] += cf[v][e] * gy[v];
        Fz[e] += cf[v][e] * gz[v];
      }
    }
    s_fch += fch;
    s_fw += fw;
    // f_diff correction (:824-841)
    bool pos[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pos[v] = f[v] > 0.0;
    if (bt.mode) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const size_t i = v * n3 + pt;
        if (bt.mode == 1) {
          bt.buf[i] = pos[v] ? 1 : 0;
        } else {
          const bool r = (bt.buf[i] & 1) != 0;
          if (bt.mode == 2) nflip += (r != pos[v]);
          else pos[v] = r;
        }
      }
    }
#pragma unroll
    for (int jv = 0; jv < V; ++jv) st(fdiff, jv * n3 + pt, fdiff_corr<V>(jv, fg, kd, pos));
#pragma unroll
    for (int e = 0; e < E; ++e) {
      st(flux, (3 * e + 0) * n3 + pt, Fx[e]);
      st(flux, (3 * e + 1) * n3 + pt, Fy[e]);
      st(flux, (3 * e + 2) * n3 + pt, Fz[e]);
    }
  }
  block_store<0>(s_fint, part, 0, sh);
  block_store<0>(s_fch, part, 1, sh);
  block_store<0>(s_fw, part, 2, sh);
  if (bad) *err = MM_ERR_ZSOLVE;
  if (nflip) atomicAdd(bt.flips, nflip);
}

// p(phi) fields for the zeroth-order elastic solve (ModElastic.f:105-112)
template <typename T, int V>
__global__ void __launch_bounds__(kBlock) k_ppoly(const T* __restrict__ phi, T* __restrict__ out, size_t n3) {
  GRID_STRIDE(pt, n3) {
#pragma unroll
    for (int v = 0; v < V; ++v) st(out, v * n3 + pt, pfun(ld(phi, v * n3 + pt)));
  }
}

// compliance mixture  sum_v P_v S_ph(v) [+ (1 - sum P) S_1]  (ModElastic.f:252-268, :583-597)
template <int V, int NPH>
__device__ __forceinline__ void compliance_mix(const DevParams* __restrict__ P, const double (&Pv)[V],
                                               double ps, bool matrix, double (&M)[6][6]) {
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int c = 0; c < 6; ++c) M[a][c] = 0.0;
#pragma unroll
  for (int q = 2; q <= NPH; ++q)
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (P->phase_of[v] == q) {
        const double* S = P->sela + (q - 1) * 36;
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int c = 0; c < 6; ++c) M[a][c] = M[a][c] + Pv[v] * S[c * 6 + a];
      }
  if (matrix) {
    const double w = 1.0 - ps;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int c = 0; c < 6; ++c) M[a][c] = M[a][c] + w * P->sela[c * 6 + a];
  }
}

// ------------------------------------------------- stress polarisation x_ij
// ModElastic.f:248-327: C(x) = (sum P S)^-1, C' = CEla(:,:,2) - C(x),
// x = cfac C' (eps + e_mean) + C(x) sum_v P_v eps0_v  with engineering
// shears, evaluated as  cfac CEla2 b + C(x) (a - cfac b)  by one 6x6 solve.
// Polarization stress at one point when every phase has the same compliance
// S (ModElastic.f:262-323): the guarded mixture is S when sum P < 1 and
// (sum P) S otherwise, so C(x) = sc C with sc = 1 / max(1, sum P),
// C' = (1 - sc) C, and x = C (cfac (1 - sc) b + sc a), with b = strain +
// mean strain and a = sum_v P_v eps0_v (engineering shears).
// SYM: CEla is symmetric (checked bitwise on the host), so C(n,m) is read as
// C(min, max): 21 distinct values instead of 36, the same products and sums.
// In the fused x pass every thread evaluates this at 2E points with the
// moduli, eigenstrains and mean strain as wave-uniform scalars; with 36
// moduli the set exceeds the scalar registers and spills into vector lanes.
// SP: CEla has the block form of the cubic, tetragonal and hexagonal classes
// (no coupling between normal and shear components, shear block diagonal;
// checked exactly at context creation), so the terms that are exact zeros
// are left out: the same sums, 12 products instead of 36.
template <int V, bool SYM = false, bool SP = false>
__device__ __forceinline__ void polar_homog(const DevParams* __restrict__ P, const double (&f)[V],
                                            const double (&e)[6], double cfac, double (&x)[6]) {
  double Pv[V], ps = 0.0;
#pragma unroll
  for (int v = 0; v < V; ++v) { Pv[v] = pfun(f[v]); ps = ps + Pv[v]; }
  const double sc = ps < 1.0 ? 1.0 : 1.0 / ps;
  const double cb = cfac * (1.0 - sc);
  double r[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const double w = m < 3 ? 1.0 : 2.0;
    const double bv = e[m] + w * P->e_mean[m];
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < V; ++v) s += Pv[v] * P->tp[v * 6 + m];
    r[m] = cb * bv + sc * (w * s);
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    double s = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      if (SP && (m < 3 ? n >= 3 : n != m)) continue;
      s += P->cela[SYM ? (n < m ? n * 6 + m : m * 6 + n) : n * 6 + m] * r[n];
    }
    x[m] = s;
  }
}

template <typename T, int V, int NPH, bool HOMOG>
__global__ void __launch_bounds__(kBlock)
k_polar(const DevParams* __restrict__ P, const T* __restrict__ phi, const T* __restrict__ strain,
        double cfac_arg, const double* __restrict__ cfac_dev, T* __restrict__ xpol, int* __restrict__ err,
        size_t n3, const int* __restrict__ gate) {
  if (gated_off(gate)) return;
  const double cfac = cfac_dev ? *cfac_dev : cfac_arg;
  bool bad = false;
  GRID_STRIDE(pt, n3) {
    if constexpr (HOMOG) {
      double f[V], e[6], x[6];
#pragma unroll
      for (int v = 0; v < V; ++v) f[v] = ld(phi, v * n3 + pt);
#pragma unroll
      for (int m = 0; m < 6; ++m) e[m] = ld(strain, m * n3 + pt);
      polar_homog<V>(P, f, e, cfac, x);
#pragma unroll
      for (int m = 0; m < 6; ++m) st(xpol, m * n3 + pt, x[m]);
    } else {
      double Pv[V], ps = 0.0;
#pragma unroll
      for (int v = 0; v < V; ++v) { Pv[v] = pfun(ld(phi, v * n3 + pt)); ps = ps + Pv[v]; }
      double bv[6], a[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const double w = m < 3 ? 1.0 : 2.0;
        bv[m] = ld(strain, m * n3 + pt) + w * P->e_mean[m];
        double s = 0.0;
#pragma unroll
        for (int v = 0; v < V; ++v) s += Pv[v] * P->tp[v * 6 + m];
        a[m] = w * s;
      }
      double M[6][6];
      compliance_mix<V, NPH>(P, Pv, ps, ps < 1.0, M);
      double r[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) r[m] = a[m] - cfac * bv[m];
      if (!solve6(M, r)) bad = true;
      const double* C2m = P->cela + 36;
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 6; ++n) s += C2m[n * 6 + m] * bv[n];
        st(xpol, m * n3 + pt, cfac * s + r[m]);
      }
    }
  }
  if (bad) *err = MM_ERR_CINV_ITER;
}

// ------------------------------------------- fused x-pass kernels: shared parts
// One workgroup owns B rows (z, y0 .. y0+B-1) of x half-spectra; each thread
// keeps the points x = 2n, 2n+1 (n = tl + TPL q) of its row through every
// phase, so pointwise work between the x transforms needs no exchange.
template <typename T, class P> struct XRows {
  static constexpr int L = P::L, TPL = P::TPL, E = P::E, B = TileLM<T, P>::B, LP = line_pitch<L>();
  int z, y0, t, bl, tl, yl, bg, r, yg;
  __device__ XRows(int ny) {
    z = blockIdx.y; y0 = blockIdx.x * B; t = threadIdx.x;
    bl = t / TPL; tl = t % TPL; yl = y0 + bl;  // line view (transforms, pointwise)
    bg = t % B; r = t / B; yg = y0 + bg;       // gather view (transposed loads and stores)
    (void)ny;
  }
};

// two register sets in turn: component M+1 loads (into xb) while M (in xa)
// is transformed
template <int NC, int M, class LD, class CV, class X>
__device__ __forceinline__ void x_c2r_pipe(LD& load, CV& c2r, X& xa, X& xb) {
  if constexpr (M + 1 < NC) load(M + 1, xb);
  else load(-1, xb);  // the caller's own loads, if any, go out with the last component
  c2r(M, xa);
  if constexpr (M + 1 < NC) x_c2r_pipe<NC, M + 1>(load, c2r, xb, xa);
}

// c2r of NC x half-spectra (in + m*nc, layout S after the inverse y pass) into
// s[m][q] = (x[2n], x[2n+1]); imaginary parts of X_0 and X_L are ignored.
// The loads of component m+1 are issued before component m is transformed;
// they are unconditional (indices clamped into the array; rows past ny are
// computed but never stored), so that no load waits behind a branch.
// Untangle factors exp(-2 pi i k / Nx) of the x c2r, k = tl + TPL q: read from
// the table at every use (UntTable) or held in registers (UntRegs).
template <typename T> struct UntTable {
  const C2<T>* p;
  __device__ __forceinline__ C2<T> get(int, int k) const { return p[k]; }
};
template <typename T, int E> struct UntRegs {
  C2<T> w[E];
  __device__ __forceinline__ C2<T> get(int q, int) const { return w[q]; }
};

// twl: the stage twiddles, the LDS table or a TwRegs set; twn: the untangle
// factors, a table pointer or an UntRegs set.  NBUF = 2: lds0 holds two row
// buffers and component m uses buffer m mod 2; the previous user of that
// buffer, component m - 2, finished before every thread passed the barrier
// after component m - 1's gather, so no barrier is needed before the gather.
// last_loads (optional): called with no argument once the loads of the last
// component are in flight, before it is transformed (loads the caller needs
// after the transforms, issued early)
struct NoLoads {
  __device__ __forceinline__ void operator()() const {}
};
template <typename T, class P, int NC, class TWS, class UN, int NBUF = 1, class LL = NoLoads>
__device__ __forceinline__ void x_c2r_n(const XRows<T, P>& w, const C2<T>* in, size_t nc, int ny, C2<T>* lds0,
                                        const TWS& twl, const UN& twn, C2<T> (&s)[NC][P::E],
                                        const LL& last_loads = LL{}) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, LP = line_pitch<L>();
  const IdxLineMajor idx{w.bl * LP};
  const size_t base = (size_t)w.z * (L + 1) * ny + (size_t)(w.yg < ny ? w.yg : ny - 1);
  // point i < E is kx = r + TPL i (< L); point E is kx = L, kept by r == 0 only
  // (component-wise copies: a whole-struct copy of double2 into the register
  // sets becomes a memcpy that keeps them in scratch memory)
  auto load = [&](int m, C2<T>(&x)[E + 1]) {
    if (m < 0) {
      last_loads();
      return;
    }
    const C2<T>* src = in + (size_t)m * nc;
#pragma unroll
    for (int i = 0; i <= E; ++i) {
      const C2<T>* p = src + base + (size_t)(i < E ? w.r + TPL * i : L) * ny;
      x[i] = mk<T>(p->x, p->y);
    }
  };
  auto c2r = [&](int m, const C2<T>(&x)[E + 1]) {
    C2<T>* lds = lds0 + (NBUF == 2 ? (m & 1) * TileLM<T, P>::LDS : 0);
    if (NBUF == 1 && m > 0) __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int kx = w.r + TPL * i;
      C2<T>* d = lds + w.bg * LP + kx + (kx >> 4);
      d->x = x[i].x;
      d->y = x[i].y;
    }
    if (w.r == 0) {
      C2<T>* d = lds + w.bg * LP + L + (L >> 4);
      d->x = x[E].x;
      d->y = x[E].y;
    }
    __syncthreads();
    C2<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int k = w.tl + TPL * q;
      C2<T> xk = lds[idx(k)], xc = lds[idx(L - k)];
      if (k == 0) { xk.y = 0; xc.y = 0; }
      C2<T> ev = mk<T>(xk.x + xc.x, xk.y - xc.y);
      C2<T> dd = mk<T>(xk.x - xc.x, xk.y + xc.y);
      C2<T> tw;
      if constexpr (std::is_pointer<UN>::value || std::is_array<UN>::value) tw = twn[k];
      else tw = twn.get(q, k);
      tw.y = -tw.y;
      C2<T> od = cmul<T>(dd, tw);
      v[q] = mk<T>(ev.x - od.y, ev.y + od.x);
    }
    line_sync<wave_lines<P>()>();  // the reads above touch this line only
    line_fft<T, +1, P, IdxLineMajor, wave_lines<P>()>(v, w.tl, lds, idx, twl);
#pragma unroll
    for (int q = 0; q < E; ++q) s[m][q] = v[q];
  };
  C2<T> xa[E + 1], xb[E + 1];
  load(0, xa);
  x_c2r_pipe<NC, 0>(load, c2r, xa, xb);
}

// ------------------------------------ fused x pass with the polarization stress
// The elastic iterations after the first (ModElastic.f:241-323 between the
// inverse transforms of the strain and the forward transforms of x): one
// workgroup owns B rows (z, y0 .. y0+B-1) of the six strain spectra left by
// the inverse y pass, runs their x c2r transforms into registers, forms the
// polarization stress at every point of its rows (homogeneous moduli), runs
// the six x r2c transforms and writes the spectra back in place.  The real
// strain and stress never reach HBM.  in and work may be the same buffer
// (each workgroup reads its rows of every component before it writes).
template <typename T, class P, int V, bool SP>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS), 2)
k_x_polar(const DevParams* __restrict__ Pp, const T* __restrict__ phi, const C2<T>* in, C2<T>* work, size_t nc,
          size_t n3, const C2<T>* __restrict__ twL, const C2<T>* __restrict__ twN, int ny, T scale, double cfac_arg,
          const double* __restrict__ cfac_dev, const int* __restrict__ gate) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, LP = line_pitch<L>();
  // two row buffers: component m of a phase works in buffer m mod 2
  __shared__ C2<T> ldsb[2 * TileLM<T, P>::LDS];
  if (gated_off(gate)) return;
  const XRows<T, P> w(ny);
  const int z = w.z, t = w.t, tl = w.tl, yl = w.yl, bg = w.bg, r = w.r, yg = w.yg;
  (void)t;
  const IdxLineMajor idx{w.bl * LP};
  // the twiddles of this thread (stages, and the untangle factors of the
  // c2r and r2c), in registers for all twelve transforms
  TwRegs<T, P> trw;
  trw.load(twL, tl);
  UntRegs<T, E> ur;
#pragma unroll
  for (int q = 0; q < E; ++q) ur.w[q] = twN[tl + TPL * q];
  C2<T> s[6][E];
  // 1. c2r of the six strain components (in; may be the same buffer as work);
  // phi of this thread's points is loaded while the last one is transformed
  C2<T> ph[E][V];
  auto load_phi = [&]() {
    const size_t row = ((size_t)z * ny + (yl < ny ? yl : ny - 1)) * (size_t)(2 * L);
#pragma unroll
    for (int q = 0; q < E; ++q)
#pragma unroll
      for (int v = 0; v < V; ++v) ph[q][v] = ((const C2<T>*)(phi + (size_t)v * n3 + row))[tl + TPL * q];
  };
  x_c2r_n<T, P, 6, TwRegs<T, P>, UntRegs<T, E>, 2>(w, in, nc, ny, ldsb, trw, ur, s, load_phi);
  // 2. polarization stress at the 2E points of this thread
  if (yl < ny) {
    const double cfac = cfac_dev ? *cfac_dev : cfac_arg;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      double f[V], e[6], xo[6];
#pragma unroll
      for (int v = 0; v < V; ++v) f[v] = (double)ph[q][v].x;
#pragma unroll
      for (int m = 0; m < 6; ++m) e[m] = (double)s[m][q].x;
      polar_homog<V, true, SP>(Pp, f, e, cfac, xo);
#pragma unroll
      for (int m = 0; m < 6; ++m) s[m][q].x = (T)xo[m];
#pragma unroll
      for (int v = 0; v < V; ++v) f[v] = (double)ph[q][v].y;
#pragma unroll
      for (int m = 0; m < 6; ++m) e[m] = (double)s[m][q].y;
      polar_homog<V, true, SP>(Pp, f, e, cfac, xo);
#pragma unroll
      for (int m = 0; m < 6; ++m) s[m][q].y = (T)xo[m];
    }
  }
  // 3. r2c of the six stress components, scaled, written back in place; the
  // untangle factors of this thread's kx (gather view) in registers
  const T hs = (T)0.5 * scale;
  C2<T> wn[E + 1];
#pragma unroll
  for (int i = 0; i <= E; ++i) {
    const int kx = r + TPL * i;
    wn[i] = twN[kx <= L ? kx : L];
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    // buffer m mod 2: its previous user (component m - 2, or the strain
    // transforms) finished before the barrier of component m - 1
    C2<T>* lds = ldsb + (m & 1) * TileLM<T, P>::LDS;
    C2<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = s[m][q];
    line_fft<T, -1, P, IdxLineMajor, wave_lines<P>()>(v, tl, lds, idx, trw);
    if constexpr (P::NS > 1) line_sync<wave_lines<P>()>();
#pragma unroll
    for (int q = 0; q < E; ++q) lds[idx(tl + TPL * q)] = v[q];
    __syncthreads();
    if (yg < ny) {
      C2<T>* out = work + (size_t)m * nc;
#pragma unroll
      for (int i = 0; i <= E; ++i) {
        const int kx = r + TPL * i;
        if (kx > L) break;
        const int k1 = kx == L ? 0 : kx, k2 = kx == 0 ? 0 : L - kx;
        C2<T> zk = lds[bg * LP + k1 + (k1 >> 4)];
        C2<T> zc = lds[bg * LP + k2 + (k2 >> 4)];
        C2<T> ev = mk<T>(zk.x + zc.x, zk.y - zc.y);
        C2<T> od = mk<T>(zk.y + zc.y, -(zk.x - zc.x));
        C2<T> xx = cadd<T>(ev, cmul<T>(wn[i], od));
        out[((size_t)z * (L + 1) + kx) * ny + yg] = cscale<T>(xx, hs);
      }
    }
  }
}

// x r2c of one component held as v[q] = (x[2n], x[2n+1]), n = tl + TPL q,
// scaled by `scale`, into the S-layout half-spectrum out (rows past ny are
// transformed but not stored).  Barriers inside: every thread of the
// workgroup calls it.
template <typename T, class P>
__device__ __forceinline__ void x_r2c_store(const XRows<T, P>& w, C2<T> (&v)[P::E], C2<T>* lds, const C2<T>* twl,
                                            const C2<T>* twn, C2<T>* __restrict__ out, int ny, T scale) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, LP = line_pitch<L>();
  const IdxLineMajor idx{w.bl * LP};
  const T hs = (T)0.5 * scale;
  __syncthreads();
  line_fft<T, -1, P, IdxLineMajor, wave_lines<P>()>(v, w.tl, lds, idx, twl);
  if constexpr (P::NS > 1) line_sync<wave_lines<P>()>();
#pragma unroll
  for (int q = 0; q < E; ++q) lds[idx(w.tl + TPL * q)] = v[q];
  __syncthreads();
  if (w.yg < ny) {
#pragma unroll
    for (int i = 0; i <= E; ++i) {
      const int kx = w.r + TPL * i;
      if (kx > L) break;
      const int k1 = kx == L ? 0 : kx, k2 = kx == 0 ? 0 : L - kx;
      C2<T> zk = lds[w.bg * LP + k1 + (k1 >> 4)];
      C2<T> zc = lds[w.bg * LP + k2 + (k2 >> 4)];
      C2<T> ev = mk<T>(zk.x + zc.x, zk.y - zc.y);
      C2<T> od = mk<T>(zk.y + zc.y, -(zk.x - zc.x));
      C2<T> tw = twn[kx];
      C2<T> xx = cadd<T>(ev, cmul<T>(tw, od));
      out[((size_t)w.z * (L + 1) + kx) * ny + w.yg] = cscale<T>(xx, hs);
    }
  }
}

// Elastic energy density and driving force at one point when every phase has
// the same moduli (ModElastic.f:493-691): sig = C eps with
// eps = strain + w e_mean - w sum_v p_v eps0_v, every phase's energy density
// equals the matrix one, and el_grad_v = p'(phi_v)(el_mat - el_ppt) -
// p'(phi_v) sig:eps0_v where sum_m trans_ppt(m,v) > 1e-9.  Returns el_en.
// SYM: CEla read as C(min, max), as in polar_homog.
template <int V, int NPH, bool SYM = false, bool SP = false>
__device__ __forceinline__ double elener_homog(const DevParams* __restrict__ P, const double (&f)[V],
                                               const double (&strain)[6], double (&eg)[V]) {
  double Pv[V], ps = 0.0, te[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int v = 0; v < V; ++v) {
    eg[v] = 0.0;
    Pv[v] = pfun(f[v]);
    ps = ps + Pv[v];
#pragma unroll
    for (int m = 0; m < 6; ++m) te[m] = te[m] + Pv[v] * P->tp[v * 6 + m];
  }
  double eps[6], sig[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const double w = m < 3 ? 1.0 : 2.0;
    eps[m] = strain[m] + w * P->e_mean[m] - w * te[m];
  }
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      if (SP && (a < 3 ? c >= 3 : c != a)) continue;
      s += P->cela[SYM ? (c < a ? c * 6 + a : a * 6 + c) : c * 6 + a] * eps[c];
    }
    sig[a] = s;
  }
  double e2 = 0.0;
#pragma unroll
  for (int a = 0; a < 6; ++a) e2 += sig[a] * eps[a];
  const double el_mat = 0.5 * e2, el_ppt = el_mat;
  double een = (1.0 - ps) * el_mat;
#pragma unroll
  for (int q = 2; q <= NPH; ++q) {
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (P->phase_of[v] == q) {
        eg[v] = 0.0;
        if (P->tsum_ok[v]) {
          een = een + el_ppt * Pv[v];
          const double* t = P->tp + v * 6;
          const double d = (sig[0] * t[0] + sig[1] * t[1] + sig[2] * t[2] +
                            2.0 * (sig[3] * t[3] + sig[4] * t[4] + sig[5] * t[5]));
          const double pp = ppr(f[v]);
          eg[v] = pp * (el_mat - el_ppt) - d * pp;
        }
      }
  }
  return een;
}

// ------------------------------------------ fused x c2r of the strain + elener
// The strain is kept between steps as x half-spectra (the layout S after the
// inverse y pass); this kernel runs their x c2r transforms into registers and
// evaluates elener (homogeneous moduli) at every point of its rows: el_en,
// el_grad, f_diff += el_grad, and the partial sums of el_en per workgroup
// (part[blockIdx.y * gridDim.x + blockIdx.x]).  The real strain is not written.
// XS: f_diff + el_grad is not stored as a real field but run through the x r2c
// transform (scaled by `scale`) into the half-spectra sfd + v nc, the input of
// the Allen-Cahn update's y pass; fdiff is then only read.
template <typename T, class P, int V, int NPH, bool XS, bool SP>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS), 2)
k_x_elener(const DevParams* __restrict__ Pp, const T* __restrict__ phi, const C2<T>* __restrict__ sstrain, size_t nc,
           size_t n3, const C2<T>* __restrict__ twL, const C2<T>* __restrict__ twN, int ny, T* __restrict__ el_en,
           T* __restrict__ elgrad, T* __restrict__ fdiff, C2<T>* __restrict__ sfd, T scale,
           double* __restrict__ part, const int* __restrict__ gate) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, NT = TileLM<T, P>::THREADS, NW = (NT + 63) / 64;
  // row buffers: two (used in turn by the strain components) unless XS
  __shared__ C2<T> ldsb[(XS ? 1 : 2) * TileLM<T, P>::LDS];
  C2<T>* const lds = ldsb;
  __shared__ C2<T> twl[L], twn[2 * L];
  __shared__ double sh[NW];
  // XS: el_grad of this thread's points, kept in LDS (slot (v E + q) NT + t)
  // until the r2c transforms add it to f_diff: registers hold the strain
  __shared__ C2<T> fdl[XS ? V * E * NT : 1];
  if (gated_off(gate)) return;
  const XRows<T, P> w(ny);
  for (int e = w.t; e < L; e += NT) twl[e] = twL[e];
  for (int e = w.t; e < 2 * L; e += NT) twn[e] = twN[e];
  __syncthreads();
  C2<T> s[6][E];
  if constexpr (XS) x_c2r_n<T, P, 6>(w, sstrain, nc, ny, lds, twl, twn, s);
  else x_c2r_n<T, P, 6, C2<T>[L], C2<T>[2 * L], 2>(w, sstrain, nc, ny, ldsb, twl, twn, s);
  // rows past ny read row ny-1 and store nothing
  const size_t row = ((size_t)w.z * ny + (w.yl < ny ? w.yl : ny - 1)) * (size_t)(2 * L);
  double s_el = 0.0;
  if (w.yl < ny) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int n = w.tl + TPL * q;
      C2<T> eo, go[V], ph[V], fd[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        ph[v] = ((const C2<T>*)(phi + (size_t)v * n3 + row))[n];
        // XS adds the chemistry's f_diff at the r2c (binary64 only: the same sum)
        fd[v] = XS ? mk<T>(0, 0) : ((const C2<T>*)(fdiff + (size_t)v * n3 + row))[n];
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double f[V], e[6], eg[V];
#pragma unroll
        for (int v = 0; v < V; ++v) f[v] = (double)(h ? ph[v].y : ph[v].x);
#pragma unroll
        for (int m = 0; m < 6; ++m) e[m] = (double)(h ? s[m][q].y : s[m][q].x);
        const double een = elener_homog<V, NPH, true, SP>(Pp, f, e, eg);
        s_el += een;
        if (h) eo.y = (T)een; else eo.x = (T)een;
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const T g = (T)eg[v];
          const T nf = (T)((double)(h ? fd[v].y : fd[v].x) + eg[v]);
          if (h) { go[v].y = g; fd[v].y = nf; } else { go[v].x = g; fd[v].x = nf; }
        }
      }
      ((C2<T>*)(el_en + row))[n] = eo;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        ((C2<T>*)(elgrad + (size_t)v * n3 + row))[n] = go[v];
        if constexpr (XS) fdl[(v * E + q) * NT + w.t] = go[v];
        else ((C2<T>*)(fdiff + (size_t)v * n3 + row))[n] = fd[v];
      }
    }
  }
  if constexpr (XS) {
    // one variant at a time (not unrolled: the loads of the next would be
    // hoisted into the transform of this one)
#pragma unroll 1
    for (int v = 0; v < V; ++v) {
      C2<T> x[E];
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const C2<T> fc = ((const C2<T>*)(fdiff + (size_t)v * n3 + row))[w.tl + TPL * q];
        const C2<T> g = fdl[(v * E + q) * NT + w.t];
        x[q] = mk<T>((T)((double)fc.x + (double)g.x), (T)((double)fc.y + (double)g.y));
      }
      x_r2c_store<T, P>(w, x, lds, twl, twn, sfd + (size_t)v * nc, ny, scale);
    }
  }
  const double rsum = block_reduce<0, NW>(s_el, sh);
  if (w.t == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = rsum;
}

// ------------------------------------------ fused x pass of the chemistry
// Multi_MEUMAPPS.f:575-845 after the gradient block's z and inverse y passes,
// whose 4V x half-spectra (grad phi_v x, y, z and del.G del phi_v at
// sgrad + (4v + d) nc) this kernel takes over: one workgroup owns B rows
// (z, y0 .. y0+B-1).  It transforms del.G del phi of every variant to real
// rows in registers, evaluates the KKS chemistry at every point of its rows
// (phi and c read once), writes f_diff with the phi > 0 correction, then
// transforms the gradient of one variant at a time and accumulates the
// fluxes F_e = sum_v cf[v][e] grad phi_v and the gradient energy, and last
// runs the x r2c transforms of c and F into sch + (4e + m) nc (m = 0: c,
// 1-3: F_x, F_y, F_z), the input of the Cahn-Hilliard update's y pass.  The
// gradients and fluxes never reach HBM as real fields.  Partial sums (f_int,
// fch, fw) go to part[slot * nblk + b], b = blockIdx.y * gridDim.x + blockIdx.x.
template <typename T, class P, int V, int NPH, int E>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS), 2)
k_x_chem(const DevParams* __restrict__ Pp, const T* __restrict__ phi, const T* __restrict__ cons,
         const C2<T>* __restrict__ sgrad, size_t nc, size_t n3, const C2<T>* __restrict__ twL,
         const C2<T>* __restrict__ twN, int ny, T scale, T* __restrict__ fdiff, C2<T>* __restrict__ sch,
         double* __restrict__ part, int* __restrict__ err, const int* __restrict__ gate) {
  constexpr int L = P::L, TPL = P::TPL, EP = P::E, NT = TileLM<T, P>::THREADS, NW = (NT + 63) / 64;
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  __shared__ C2<T> twl[L], twn[2 * L];
  __shared__ double sh[NW];
  // the flux coefficients of this thread's points, kept in LDS between the
  // chemistry and the gradient passes (slot ((q*2 + h)*V + v)*E + e, thread
  // fastest): registers hold the transforms
  __shared__ double cfl[EP * 2 * V * E * NT];
  if (gated_off(gate)) return;
  const XRows<T, P> w(ny);
  for (int e = w.t; e < L; e += NT) twl[e] = twL[e];
  for (int e = w.t; e < 2 * L; e += NT) twn[e] = twN[e];
  __syncthreads();
  const bool live = w.yl < ny;
  // rows past ny read row ny-1 and store nothing
  const size_t row = ((size_t)w.z * ny + (live ? w.yl : ny - 1)) * (size_t)(2 * L);
  auto cfs = [&](int q, int h, int v, int e) -> double& { return cfl[(((q * 2 + h) * V + v) * E + e) * NT + w.t]; };
  double s_fint = 0.0, s_fch = 0.0, s_fw = 0.0;
  bool bad = false;
  {
    // 1. del.G del phi of every variant into registers
    C2<T> kd[V][EP];
    x_c2r_n<T, P, V>(w, sgrad + 3 * nc, 4 * nc, ny, lds, twl, twn, kd);
    // 2. chemistry and f_diff at the 2 EP points of this thread
#pragma unroll
    for (int q = 0; q < EP; ++q) {
      const int n = w.tl + TPL * q;
      C2<T> ph[V], cr[E], fo[V];
#pragma unroll
      for (int v = 0; v < V; ++v) ph[v] = ((const C2<T>*)(phi + (size_t)v * n3 + row))[n];
#pragma unroll
      for (int e = 0; e < E; ++e) cr[e] = ((const C2<T>*)(cons + (size_t)e * n3 + row))[n];
      // the two points of a pair one after the other (not unrolled: the
      // chemistry's temporaries of both would be live at once)
#pragma unroll 1
      for (int h = 0; h < 2; ++h) {
        double f[V], c[E], fg[V], kdv[V], cf[V][E], fch, fw;
        bool pos[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          f[v] = (double)(h ? ph[v].y : ph[v].x);
          kdv[v] = (double)(h ? kd[v][q].y : kd[v][q].x);
          pos[v] = f[v] > 0.0;
        }
#pragma unroll
        for (int e = 0; e < E; ++e) c[e] = Pp->cbar[e] + (double)(h ? cr[e].y : cr[e].x);
        const bool ok = chem_point<V, NPH, E>(Pp, f, c, fg, cf, fch, fw);
        if (live) {
          bad = bad || !ok;
          s_fch += fch;
          s_fw += fw;
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
#pragma unroll
          for (int e = 0; e < E; ++e) cfs(q, h, v, e) = cf[v][e];
          const T d = (T)fdiff_corr<V>(v, fg, kdv, pos);
          if (h) fo[v].y = d; else fo[v].x = d;
        }
      }
      if (live) {
#pragma unroll
        for (int v = 0; v < V; ++v) ((C2<T>*)(fdiff + (size_t)v * n3 + row))[n] = fo[v];
      }
      // one point pair at a time: keeps the loads of the next pair from
      // being hoisted over this one's chemistry (register pressure)
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // 3. gradients of one variant at a time: fluxes and gradient energy
  double F[EP][2][3 * E];
#pragma unroll
  for (int q = 0; q < EP; ++q)
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int m = 0; m < 3 * E; ++m) F[q][h][m] = 0.0;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    C2<T> g[3][EP];
    __syncthreads();
    x_c2r_n<T, P, 3>(w, sgrad + (size_t)(4 * v) * nc, nc, ny, lds, twl, twn, g);
#pragma unroll
    for (int q = 0; q < EP; ++q)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double gx = (double)(h ? g[0][q].y : g[0][q].x);
        const double gy = (double)(h ? g[1][q].y : g[1][q].x);
        const double gz = (double)(h ? g[2][q].y : g[2][q].x);
        if (live) s_fint += 0.5 * grad_term(Pp, v, gx, gy, gz);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double cf = cfs(q, h, v, e);
          F[q][h][3 * e + 0] += cf * gx;
          F[q][h][3 * e + 1] += cf * gy;
          F[q][h][3 * e + 2] += cf * gz;
        }
      }
  }
  // 4. x r2c of c and the fluxes (the values the unfused path would store)
#pragma unroll
  for (int e = 0; e < E; ++e) {
    C2<T> cr[EP];
#pragma unroll
    for (int q = 0; q < EP; ++q) cr[q] = ((const C2<T>*)(cons + (size_t)e * n3 + row))[w.tl + TPL * q];
    x_r2c_store<T, P>(w, cr, lds, twl, twn, sch + (size_t)(4 * e) * nc, ny, scale);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      C2<T> v[EP];
#pragma unroll
      for (int q = 0; q < EP; ++q) v[q] = mk<T>((T)F[q][0][3 * e + d], (T)F[q][1][3 * e + d]);
      x_r2c_store<T, P>(w, v, lds, twl, twn, sch + (size_t)(4 * e + 1 + d) * nc, ny, scale);
    }
  }
  // 5. partial sums
  const size_t nb = (size_t)gridDim.x * gridDim.y, b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  const double r0 = block_reduce<0, NW>(s_fint, sh);
  const double r1 = block_reduce<0, NW>(s_fch, sh);
  const double r2 = block_reduce<0, NW>(s_fw, sh);
  if (w.t == 0) {
    part[b] = r0;
    part[nb + b] = r1;
    part[2 * nb + b] = r2;
  }
  if (bad) *err = MM_ERR_ZSOLVE;
}

// ------------------------------- Allen-Cahn inverse x pass fused with the clamp
// Multi_MEUMAPPS.f:1001-1043: the V phi half-spectra left by the inverse y
// pass are transformed to real rows in registers, clamped to [0, 1] and
// written once; the check-loop partial sums of k_clamp (slots 0..E-1 sum of
// the stored compositions, E min c, E+1 max c, E+2 max phi, E+3 sum of phi > 0)
// go to part[slot * nblk + b], b = blockIdx.y * gridDim.x + blockIdx.x.
template <typename T, class P, int V, int E>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS), 2)
k_x_clamp(const DevParams* __restrict__ Pp, const C2<T>* __restrict__ in, size_t nc, T* __restrict__ phi,
          const T* __restrict__ cons, size_t n3, const C2<T>* __restrict__ twL, const C2<T>* __restrict__ twN, int ny,
          double* __restrict__ part, const int* __restrict__ gate) {
  constexpr int L = P::L, TPL = P::TPL, E_ = P::E, NW = (TileLM<T, P>::THREADS + 63) / 64;
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  __shared__ C2<T> twl[L], twn[2 * L];
  __shared__ double sh[NW];
  if (gated_off(gate)) return;
  const XRows<T, P> w(ny);
  for (int e = w.t; e < L; e += TileLM<T, P>::THREADS) twl[e] = twL[e];
  for (int e = w.t; e < 2 * L; e += TileLM<T, P>::THREADS) twn[e] = twN[e];
  __syncthreads();
  C2<T> s[V][E_];
  x_c2r_n<T, P, V>(w, in, nc, ny, lds, twl, twn, s);
  double cs[E], cmin = HUGE_VAL, cmax = -HUGE_VAL, pmax = -HUGE_VAL, sf = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) cs[e] = 0.0;
  if (w.yl < ny) {
    const size_t row = ((size_t)w.z * ny + w.yl) * (size_t)(2 * L);
    C2<T> cv[E][E_];
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
      for (int q = 0; q < E_; ++q) {
        const C2<T>* pc = (const C2<T>*)(cons + (size_t)e * n3 + row) + (w.tl + TPL * q);
        cv[e][q] = mk<T>(pc->x, pc->y);
      }
#pragma unroll
    for (int q = 0; q < E_; ++q) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const double raw = (double)(h ? cv[e][q].y : cv[e][q].x);
          const double c = Pp->cbar[e] + raw;
          if (c < cmin) cmin = c;
          if (c > cmax) cmax = c;
          cs[e] += raw;
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
          double f = (double)(h ? s[v][q].y : s[v][q].x);
          if (f < 0.0) f = 0.0;
          if (f > 1.0) f = 1.0;
          const T ft = (T)f;
          if (h) s[v][q].y = ft; else s[v][q].x = ft;
          f = (double)ft;
          if (f > pmax) pmax = f;
          if (f > 0.0) sf = sf + f;
        }
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        C2<T>* pp = (C2<T>*)(phi + (size_t)v * n3 + row) + (w.tl + TPL * q);
        pp->x = s[v][q].x;
        pp->y = s[v][q].y;
      }
    }
  }
  const size_t nblk = (size_t)gridDim.x * gridDim.y, b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  double r[E + 4];
#pragma unroll
  for (int e = 0; e < E; ++e) r[e] = block_reduce<0, NW>(cs[e], sh);
  r[E] = block_reduce<1, NW>(cmin, sh);
  r[E + 1] = block_reduce<2, NW>(cmax, sh);
  r[E + 2] = block_reduce<2, NW>(pmax, sh);
  r[E + 3] = block_reduce<0, NW>(sf, sh);
  if (w.t == 0) {
#pragma unroll
    for (int k = 0; k < E + 4; ++k) part[(size_t)k * nblk + b] = r[k];
  }
}

// --------------------------------------------------------------- elener
// ModElastic.f:493-691 fused with f_diff += el_grad (Multi_MEUMAPPS.f:880).
template <typename T, int V, int NPH, bool HOMOG>
__global__ void __launch_bounds__(kBlock)
k_elener(const DevParams* __restrict__ P, const T* __restrict__ phi, const T* __restrict__ strain,
         T* __restrict__ el_en, T* __restrict__ elgrad, T* __restrict__ fdiff,
         double* __restrict__ part, int* __restrict__ err, size_t n3, const int* __restrict__ gate) {
  if (gated_off(gate)) return;
  __shared__ double sh[kWaves];
  double s_el = 0.0;
  bool bad = false;
  GRID_STRIDE(pt, n3) {
    if constexpr (HOMOG) {
      double f[V], e[6], eg[V];
#pragma unroll
      for (int v = 0; v < V; ++v) f[v] = ld(phi, v * n3 + pt);
#pragma unroll
      for (int m = 0; m < 6; ++m) e[m] = ld(strain, m * n3 + pt);
      const double een = elener_homog<V, NPH>(P, f, e, eg);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        st(elgrad, v * n3 + pt, eg[v]);
        st(fdiff, v * n3 + pt, ld(fdiff, v * n3 + pt) + eg[v]);
      }
      st(el_en, pt, een);
      s_el += een;
      continue;
    }
    double f[V], Pv[V], ps = 0.0, te[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < V; ++v) {
      f[v] = ld(phi, v * n3 + pt);
      Pv[v] = pfun(f[v]);
      ps = ps + Pv[v];
#pragma unroll
      for (int m = 0; m < 6; ++m) te[m] = te[m] + Pv[v] * P->tp[v * 6 + m];
    }
    double eps[6], sig[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double w = m < 3 ? 1.0 : 2.0;
      eps[m] = ld(strain, m * n3 + pt) + w * P->e_mean[m] - w * te[m];
    }
    double el_mat, el_ppt[NPH + 1];
    {
      // the matrix term is unguarded here (ModElastic.f:597): with sum P > 1
      // the mixture can be indefinite, so this solve pivots
      double M[6][6];
      compliance_mix<V, NPH>(P, Pv, ps, true, M);
#pragma unroll
      for (int m = 0; m < 6; ++m) sig[m] = eps[m];
      if (!solve_pivot<6>(M, sig)) bad = true;  // sig = C(x) eps
      // el_mat = 1/2 sig . S_1 sig
      el_mat = 0.0;
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        double em = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) em = em + P->sela[c * 6 + a] * sig[c];
        el_mat = el_mat + sig[a] * em;
      }
      el_mat = 0.5 * el_mat;
#pragma unroll
      for (int q = 2; q <= NPH; ++q) {
        double e = 0.0;
        const double* S = P->sela + (q - 1) * 36;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
          double ep = 0.0;
#pragma unroll
          for (int c = 0; c < 6; ++c) ep = ep + S[c * 6 + a] * sig[c];
          e = e + sig[a] * ep;
        }
        el_ppt[q] = 0.5 * e;
      }
    }
    double een = (1.0 - ps) * el_mat;
#pragma unroll
    for (int q = 2; q <= NPH; ++q) {
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (P->phase_of[v] == q) {
          double eg = 0.0;
          if (P->tsum_ok[v]) {
            een = een + el_ppt[q] * Pv[v];
            const double* t = P->tp + v * 6;
            const double d = (sig[0] * t[0] + sig[1] * t[1] + sig[2] * t[2] +
                              2.0 * (sig[3] * t[3] + sig[4] * t[4] + sig[5] * t[5]));
            const double pp = ppr(f[v]);
            eg = pp * (el_mat - el_ppt[q]) - d * pp;
          }
          st(elgrad, v * n3 + pt, eg);
          st(fdiff, v * n3 + pt, ld(fdiff, v * n3 + pt) + eg);
        }
    }
    st(el_en, pt, een);
    s_el += een;
  }
  block_store<0>(s_el, part, 0, sh);
  if (bad) *err = MM_ERR_CINV_ENER;
}

// ------------------------------------------------- clamp + check loop sums
// Multi_MEUMAPPS.f:1016-1043: clamp phi to [0,1]; sum c, min/max c, max phi,
// sum of phi > 0.  Slots: 0..E-1 sum c (stored values), E min, E+1 max,
// E+2 max phi, E+3 solid sum.
template <typename T, int V, int E>
__global__ void __launch_bounds__(kBlock)
k_clamp(const DevParams* __restrict__ P, T* __restrict__ phi, const T* __restrict__ cons,
        double* __restrict__ part, size_t n3, const int* __restrict__ gate, BranchTape bt) {
  if (gated_off(gate)) return;
  __shared__ double sh[kWaves];
  double cs[E], cmin = HUGE_VAL, cmax = -HUGE_VAL, pmax = -HUGE_VAL, sf = 0.0;
  unsigned long long nflip = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) cs[e] = 0.0;
  GRID_STRIDE(pt, n3) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const double raw = ld(cons, e * n3 + pt);
      const double c = P->cbar[e] + raw;
      if (c < cmin) cmin = c;
      if (c > cmax) cmax = c;
      cs[e] += raw;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double f = ld(phi, v * n3 + pt);
      int cl = f < 0.0 ? 1 : (f > 1.0 ? 2 : 0);
      if (bt.mode) {
        const size_t i = v * n3 + pt;
        const uint8_t b = bt.buf[i];
        if (bt.mode == 1) {
          bt.buf[i] = (uint8_t)((b & 1) | (cl << 1));
        } else {
          const int r = (b >> 1) & 3;
          if (bt.mode == 2) nflip += (r != cl);
          else cl = r;
        }
      }
      if (cl == 1) f = 0.0;
      if (cl == 2) f = 1.0;
      st(phi, v * n3 + pt, f);
      f = (double)(T)f;
      if (f > pmax) pmax = f;
      if (f > 0.0) sf = sf + f;
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) block_store<0>(cs[e], part, e, sh);
  block_store<1>(cmin, part, E, sh);
  block_store<2>(cmax, part, E + 1, sh);
  block_store<2>(pmax, part, E + 2, sh);
  block_store<0>(sf, part, E + 3, sh);
  if (nflip) atomicAdd(bt.flips, nflip);
}

// ============================================================ k-space kernels

// gradient block (Multi_MEUMAPPS.f:581-605): out[4j+d] = phi_hat * i kf_d (d<3),
// out[4j+3] = phi_hat * sum G kk.  blockIdx.y = j, variant v0 + j.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ks_grad(const DevParams* __restrict__ P, const C2<T>* __restrict__ sphi, C2<T>* __restrict__ work,
          int v0, size_t nc) {
  const int j = blockIdx.y, v = v0 + j;
  const C2<T>* in = sphi + (size_t)v * nc;
  C2<T>* o = work + (size_t)4 * j * nc;
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    const C2<T> s = in[i];
    const double sr = (double)s.x, si = (double)s.y;
#pragma unroll
    for (int d = 0; d < 3; ++d) o[(size_t)d * nc + i] = mk<T>((T)(-si * k.kf[d]), (T)(sr * k.kf[d]));
    const double g = kgk(P, v, k);
    o[(size_t)3 * nc + i] = mk<T>((T)(sr * g), (T)(si * g));
  }
}

// zeroth-order displacement at step == elastep0 (ModElastic.f:115-236):
// R_i = i sum_j kf_j sum_v Re(P_hat_v) sig0_v(ij) (sum_0 is real*8, so only the
// real part of P_hat survives), u = Q^-1 R, u(k=0) = 0.
template <typename T, int V>
__global__ void __launch_bounds__(kBlock)
k_ks_zeroth(const DevParams* __restrict__ P, const C2<T>* __restrict__ phat, C2<T>* __restrict__ uhat,
            C2<T>* __restrict__ work, size_t nc) {
  C2<T>* outs[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) outs[m] = work + (size_t)m * nc;
  const int vm[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    Cplx3 u;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) { u.re[c] = 0.0; u.im[c] = 0.0; }
    } else {
      double pre[V];
#pragma unroll
      for (int v = 0; v < V; ++v) pre[v] = (double)phat[(size_t)v * nc + i].x;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double ri = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          double s0 = 0.0;
#pragma unroll
          for (int v = 0; v < V; ++v) s0 += pre[v] * P->sig0[v * 6 + vm[a][b]];
          ri += s0 * k.kf[b];
        }
        u.re[a] = 0.0;
        u.im[a] = ri;
      }
      acoustic_solve(P, k, u);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) uhat[(size_t)c * nc + i] = mk<T>((T)u.re[c], (T)u.im[c]);
    store_strain<T>(outs, i, k, u);
  }
}

// elastic iteration k-space step (ModElastic.f:340-394): R = kf . x_hat,
// u = Q^-1 R (u(0) = 0), strain spectra, and Parseval forms of the
// convergence sums of :413-430 over the half spectrum.  work[0..5] holds
// x_hat on entry and the strain spectra on exit; uhat holds u_old / u.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ks_elastic(const DevParams* __restrict__ P, C2<T>* __restrict__ work, C2<T>* __restrict__ uhat,
             double* __restrict__ part, size_t nc) {
  __shared__ double sh[kWaves];
  C2<T>* outs[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) outs[m] = work + (size_t)m * nc;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    double xr[6], xi[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const C2<T> x = outs[m][i];
      xr[m] = (double)x.x;
      xi[m] = (double)x.y;
    }
    Cplx3 u;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) { u.re[c] = 0.0; u.im[c] = 0.0; }
    } else {
      // R(1) = x1 kf1 + x6 kf2 + x5 kf3 ; R(2) = x6 kf1 + x2 kf2 + x4 kf3 ; R(3) = x5 kf1 + x4 kf2 + x3 kf3
      const int vm[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double rr = 0.0, ri = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          rr += -xi[vm[a][b]] * k.kf[b];
          ri += xr[vm[a][b]] * k.kf[b];
        }
        u.re[a] = rr;
        u.im[a] = ri;
      }
      acoustic_solve(P, k, u);
    }
    const double w = (k.kx == 0 || 2 * k.kx == P->nx) ? 1.0 : 2.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const C2<T> uo = uhat[(size_t)c * nc + i];
      const double dr = u.re[c] - (double)uo.x, di = u.im[c] - (double)uo.y;
      acc[2 * c] += w * (dr * dr + di * di);
      acc[2 * c + 1] += w * ((double)uo.x * (double)uo.x + (double)uo.y * (double)uo.y);
      uhat[(size_t)c * nc + i] = mk<T>((T)u.re[c], (T)u.im[c]);
    }
    store_strain<T>(outs, i, k, u);
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) block_store<0>(acc[m], part, m, sh);
}

// Cahn-Hilliard update (Multi_MEUMAPPS.f:924-943); blockIdx.y = e,
// work[4e] = c_hat (in/out), work[4e+1..3] = F_hat.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ks_ch(const DevParams* __restrict__ P, C2<T>* __restrict__ work, size_t nc) {
  const int e = blockIdx.y;
  C2<T>* ch = work + (size_t)(4 * e) * nc;
  const C2<T>* fx = ch + nc;
  const C2<T>* fy = ch + 2 * nc;
  const C2<T>* fz = ch + 3 * nc;
  const double dtm = P->dt * P->cmobs[e];
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    const double ksq = k.kq[0] + k.kq[1] + k.kq[2];
    const C2<T> a = fx[i], b = fy[i], c = fz[i];
    // term1 = i (kf1 Fx + kf2 Fy + kf3 Fz)
    const double tr = -((double)a.y * k.kf[0] + (double)b.y * k.kf[1] + (double)c.y * k.kf[2]);
    const double ti = (double)a.x * k.kf[0] + (double)b.x * k.kf[1] + (double)c.x * k.kf[2];
    const double den = 1.0 - dtm * ksq;
    const C2<T> s = ch[i];
    ch[i] = mk<T>((T)(((double)s.x + dtm * tr) / den), (T)(((double)s.y + dtm * ti) / den));
    // MIXED stores c - cbar: its mean is zero, and the FP32 rounding the
    // forward transform leaves in the k = 0 mode must not accumulate
    if (sizeof(T) == 4 && i == 0) ch[i] = mk<T>(0, 0);
  }
}

// Allen-Cahn update (Multi_MEUMAPPS.f:966-991), noise term zero;
// blockIdx.y = v, sphi[v] <- (phi_hat - dt L f_hat) / (1 - dt L sum G kk).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ks_ac(const DevParams* __restrict__ P, C2<T>* __restrict__ sphi, const C2<T>* __restrict__ fhat, size_t nc) {
  const int v = blockIdx.y;
  C2<T>* ph = sphi + (size_t)v * nc;
  const C2<T>* fh = fhat + (size_t)v * nc;
  const double dtl = P->dt * P->bmob;
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    const double den = 1.0 - dtl * kgk(P, v, k);
    const C2<T> s = ph[i], f = fh[i];
    ph[i] = mk<T>((T)(((double)s.x - dtl * (double)f.x) / den), (T)(((double)s.y - dtl * (double)f.y) / den));
  }
}

// out = i kf_d in   (displacement gradients for the dump records)
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ks_deriv(const DevParams* __restrict__ P, const C2<T>* __restrict__ in, C2<T>* __restrict__ out, int d,
           size_t nc) {
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    const C2<T> s = in[i];
    out[i] = mk<T>((T)(-(double)s.y * k.kf[d]), (T)((double)s.x * k.kf[d]));
  }
}

// ============================================================ dispatch table
template <typename T> struct Kern {
  void (*chem)(dim3, hipStream_t, const DevParams*, const T*, const T*, const T*, const T*, T*, T*, double*, int*, size_t,
               BranchTape) = nullptr;
  void (*ppoly)(dim3, hipStream_t, const T*, T*, size_t) = nullptr;
  void (*polar)(dim3, hipStream_t, const DevParams*, const T*, const T*, double, const double*, T*, int*, size_t,
                const int*) = nullptr;
  void (*elener)(dim3, hipStream_t, const DevParams*, const T*, const T*, T*, T*, T*, double*, int*, size_t,
                 const int*) = nullptr;
  void (*clamp)(dim3, hipStream_t, const DevParams*, T*, const T*, double*, size_t, const int*, BranchTape) = nullptr;
  void (*zeroth)(dim3, hipStream_t, const DevParams*, const C2<T>*, C2<T>*, C2<T>*, size_t) = nullptr;
  // fused x pass + polarization stress (homogeneous moduli, supported x lengths only):
  // returns false when it has no kernel for this x length
  bool (*xpolar)(bool launch, hipStream_t, const DevParams*, const T*, const C2<T>*, C2<T>*, size_t, size_t,
                 const void*, const void*, int, int, int, T, double, const double*, const int*, bool) = nullptr;
  // fused x c2r of the strain + elener (homogeneous moduli): nblk gets the workgroup count
  // (sfd non-null: f_diff goes out as x half-spectra into sfd, scaled by scale)
  // the last argument of both: moduli of the block form (see polar_homog)
  bool (*xelener)(bool launch, hipStream_t, const DevParams*, const T*, const C2<T>*, size_t, size_t, const void*,
                  const void*, int, int, int, T*, T*, T*, C2<T>*, T, double*, const int*, int*, bool) = nullptr;
  // Allen-Cahn inverse x pass + clamp + check-loop sums: nblk gets the workgroup count
  bool (*xclamp)(bool launch, hipStream_t, const DevParams*, const C2<T>*, size_t, T*, const T*, size_t, const void*,
                 const void*, int, int, int, double*, const int*, int*) = nullptr;
  // fused x pass of the chemistry (gradient c2r, KKS, f_diff, flux r2c): nblk gets the workgroup count
  bool (*xchem)(bool launch, hipStream_t, const DevParams*, const T*, const T*, const C2<T>*, size_t, size_t,
                const void*, const void*, int, int, int, T, T*, C2<T>*, double*, int*, const int*, int*) = nullptr;
};

// plan of the fused x kernel: six components of E points stay in registers
// (binary32: the 8-point plans, and 6 points for the radix-3 lengths whose
// lean plan holds 12)
template <typename T, int L>
using XPlan = std::conditional_t<sizeof(T) == 8 || (PlanLean<L>::E > 8), PlanLean4<L>, PlanLean<L>>;
template <typename T, int L> constexpr bool xplan_ok() { return XPlan<T, L>::E <= (sizeof(T) == 8 ? 4 : 8); }

template <typename T, int L, int V>
static bool go_xpolar(bool launch, hipStream_t s, const DevParams* P, const T* phi, const C2<T>* in, C2<T>* work,
                      size_t nc, size_t n3, const void* twL, const void* twN, int ny, int nz, T scale, double cfac,
                      const double* cfd, const int* gate, bool sp) {
  if constexpr (!xplan_ok<T, L>()) {
    return false;
  } else {
    using XP = XPlan<T, L>;
    using TL = TileLM<T, XP>;
    if (launch) {
      dim3 grid((ny + TL::B - 1) / TL::B, nz);
      if (sp)
        hipLaunchKernelGGL((k_x_polar<T, XP, V, true>), grid, dim3(TL::THREADS), 0, s, P, phi, in, work, nc, n3,
                           (const C2<T>*)twL, (const C2<T>*)twN, ny, scale, cfac, cfd, gate);
      else
        hipLaunchKernelGGL((k_x_polar<T, XP, V, false>), grid, dim3(TL::THREADS), 0, s, P, phi, in, work, nc, n3,
                           (const C2<T>*)twL, (const C2<T>*)twN, ny, scale, cfac, cfd, gate);
    }
    return true;
  }
}

// x lengths (Nx/2) whose binary32 elener hand-over runs on the 4-point plan
constexpr bool xs32_rows(int L) { return L == 8 || L == 16 || L == 32 || L == 64 || L == 128 || L == 256; }

template <typename T, int L, int V, int NPH>
static bool go_xelener(bool launch, hipStream_t s, const DevParams* P, const T* phi, const C2<T>* sstrain, size_t nc,
                       size_t n3, const void* twL, const void* twN, int ny, int nz, T* el_en, T* elgrad, T* fdiff,
                       C2<T>* sfd, T scale, double* part, const int* gate, int* nblk, bool sp) {
  if constexpr (!xplan_ok<T, L>()) {
    return false;
  } else {
    using XP = XPlan<T, L>;
    using TL = TileLM<T, XP>;
    dim3 grid((ny + TL::B - 1) / TL::B, nz);
    if (nblk) *nblk = (int)(grid.x * grid.y);
#define MM_XEL(XS_, SP_)                                                                                         \
  hipLaunchKernelGGL((k_x_elener<T, XP, V, NPH, XS_, SP_>), grid, dim3(TL::THREADS), 0, s, P, phi, sstrain, nc, n3, \
                     (const C2<T>*)twL, (const C2<T>*)twN, ny, el_en, elgrad, fdiff, sfd, scale, part, gate)
    if constexpr (sizeof(T) == 8) {
      if (launch && sfd) {
        if (sp) MM_XEL(true, true);
        else MM_XEL(true, false);
        return true;
      }
    } else if constexpr (PlanLean4<L>::E == 4) {
      // binary32 hand-over: the 4-point plan (its own row tiling and grid)
      if (launch && sfd) {
        using XP4 = PlanLean4<L>;
        using TL4 = TileLM<T, XP4>;
        dim3 g4((ny + TL4::B - 1) / TL4::B, nz);
        if (nblk) *nblk = (int)(g4.x * g4.y);
#define MM_XEL4(SP_)                                                                                           \
  hipLaunchKernelGGL((k_x_elener<T, XP4, V, NPH, true, SP_>), g4, dim3(TL4::THREADS), 0, s, P, phi, sstrain, nc, n3, \
                     (const C2<T>*)twL, (const C2<T>*)twN, ny, el_en, elgrad, fdiff, sfd, scale, part, gate)
        if (sp) MM_XEL4(true);
        else MM_XEL4(false);
#undef MM_XEL4
        return true;
      }
    }
    if (launch) {
      if (sp) MM_XEL(false, true);
      else MM_XEL(false, false);
    }
#undef MM_XEL
    return true;
  }
}

template <typename T, int L, int V, int E>
static bool go_xclamp(bool launch, hipStream_t s, const DevParams* P, const C2<T>* in, size_t nc, T* phi,
                      const T* cons, size_t n3, const void* twL, const void* twN, int ny, int nz, double* part,
                      const int* gate, int* nblk) {
  if constexpr (!xplan_ok<T, L>()) {
    return false;
  } else {
    using XP = XPlan<T, L>;
    using TL = TileLM<T, XP>;
    dim3 grid((ny + TL::B - 1) / TL::B, nz);
    if (nblk) *nblk = (int)(grid.x * grid.y);
    if (launch)
      hipLaunchKernelGGL((k_x_clamp<T, XP, V, E>), grid, dim3(TL::THREADS), 0, s, P, in, nc, phi, cons, n3,
                         (const C2<T>*)twL, (const C2<T>*)twN, ny, part, gate);
    return true;
  }
}

// the chemistry pass keeps the flux coefficients, fluxes and three gradient
// components of 2E points per thread in registers: 4 complex points per
// thread in both precisions
template <int L> constexpr bool xchem_ok() { return PlanLean4<L>::E <= 4; }

template <typename T, int L, int V, int NPH, int E>
static bool go_xchem(bool launch, hipStream_t s, const DevParams* P, const T* phi, const T* cons,
                     const C2<T>* sgrad, size_t nc, size_t n3, const void* twL, const void* twN, int ny, int nz,
                     T scale, T* fdiff, C2<T>* sch, double* part, int* err, const int* gate, int* nblk) {
  if constexpr (!xchem_ok<L>()) {
    return false;
  } else {
    using XP = PlanLean4<L>;
    using TL = TileLM<T, XP>;
    dim3 grid((ny + TL::B - 1) / TL::B, nz);
    if (nblk) *nblk = (int)(grid.x * grid.y);
    if (launch)
      hipLaunchKernelGGL((k_x_chem<T, XP, V, NPH, E>), grid, dim3(TL::THREADS), 0, s, P, phi, cons, sgrad, nc, n3,
                         (const C2<T>*)twL, (const C2<T>*)twN, ny, scale, fdiff, sch, part, err, gate);
    return true;
  }
}

template <typename T, int V, int NPH, int E> struct KernImpl {
  static void chem(dim3 g, hipStream_t s, const DevParams* P, const T* a, const T* b, const T* c, const T* d,
                   T* e, T* f, double* p, int* err, size_t n, BranchTape bt) {
    hipLaunchKernelGGL((k_chem<T, V, NPH, E>), g, dim3(kBlock), 0, s, P, a, b, c, d, e, f, p, err, n, bt);
  }
  static void ppoly(dim3 g, hipStream_t s, const T* a, T* b, size_t n) {
    hipLaunchKernelGGL((k_ppoly<T, V>), g, dim3(kBlock), 0, s, a, b, n);
  }
  template <bool H>
  static void polar(dim3 g, hipStream_t s, const DevParams* P, const T* a, const T* b, double cf,
                    const double* cfd, T* x, int* err, size_t n, const int* gate) {
    hipLaunchKernelGGL((k_polar<T, V, NPH, H>), g, dim3(kBlock), 0, s, P, a, b, cf, cfd, x, err, n, gate);
  }
  template <bool H>
  static void elener(dim3 g, hipStream_t s, const DevParams* P, const T* a, const T* b, T* c, T* d, T* e,
                     double* p, int* err, size_t n, const int* gate) {
    hipLaunchKernelGGL((k_elener<T, V, NPH, H>), g, dim3(kBlock), 0, s, P, a, b, c, d, e, p, err, n, gate);
  }
  static void clamp(dim3 g, hipStream_t s, const DevParams* P, T* a, const T* b, double* p, size_t n,
                    const int* gate, BranchTape bt) {
    hipLaunchKernelGGL((k_clamp<T, V, E>), g, dim3(kBlock), 0, s, P, a, b, p, n, gate, bt);
  }
  static void zeroth(dim3 g, hipStream_t s, const DevParams* P, const C2<T>* a, C2<T>* b, C2<T>* c, size_t n) {
    hipLaunchKernelGGL((k_ks_zeroth<T, V>), g, dim3(kBlock), 0, s, P, a, b, c, n);
  }
  static bool xpolar(bool launch, hipStream_t s, const DevParams* P, const T* phi, const C2<T>* in, C2<T>* work,
                     size_t nc, size_t n3, const void* twL, const void* twN, int nx, int ny, int nz, T scale,
                     double cfac, const double* cfd, const int* gate, bool sp) {
    switch (nx / 2) {
#define X(L_) \
  case L_: return go_xpolar<T, L_, V>(launch, s, P, phi, in, work, nc, n3, twL, twN, ny, nz, scale, cfac, cfd, gate, sp);
      X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
#undef X
      default: return false;
    }
  }
  static bool xelener(bool launch, hipStream_t s, const DevParams* P, const T* phi, const C2<T>* sstrain, size_t nc,
                      size_t n3, const void* twL, const void* twN, int nx, int ny, int nz, T* el_en, T* elgrad,
                      T* fdiff, C2<T>* sfd, T scale, double* part, const int* gate, int* nblk, bool sp) {
    switch (nx / 2) {
#define X(L_)                                                                                                  \
  case L_:                                                                                                     \
    return go_xelener<T, L_, V, NPH>(launch, s, P, phi, sstrain, nc, n3, twL, twN, ny, nz, el_en, elgrad, fdiff, \
                                     sfd, scale, part, gate, nblk, sp);
      X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
#undef X
      default: return false;
    }
  }
  static bool xclamp(bool launch, hipStream_t s, const DevParams* P, const C2<T>* in, size_t nc, T* phi,
                     const T* cons, size_t n3, const void* twL, const void* twN, int nx, int ny, int nz, double* part,
                     const int* gate, int* nblk) {
    switch (nx / 2) {
#define X(L_)                                                                                                  \
  case L_:                                                                                                     \
    return go_xclamp<T, L_, V, E>(launch, s, P, in, nc, phi, cons, n3, twL, twN, ny, nz, part, gate, nblk);
      X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
#undef X
      default: return false;
    }
  }
  static bool xchem(bool launch, hipStream_t s, const DevParams* P, const T* phi, const T* cons, const C2<T>* sgrad,
                    size_t nc, size_t n3, const void* twL, const void* twN, int nx, int ny, int nz, T scale, T* fdiff,
                    C2<T>* sch, double* part, int* err, const int* gate, int* nblk) {
    switch (nx / 2) {
#define X(L_)                                                                                                     \
  case L_:                                                                                                        \
    return go_xchem<T, L_, V, NPH, E>(launch, s, P, phi, cons, sgrad, nc, n3, twL, twN, ny, nz, scale, fdiff, sch, \
                                      part, err, gate, nblk);
      X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
#undef X
      default: return false;
    }
  }
  static Kern<T> table(bool homog) {
    Kern<T> k;
    k.chem = chem; k.ppoly = ppoly; k.clamp = clamp; k.zeroth = zeroth;
    // the fused x kernels keep V fields per point in registers: up to 4
    if constexpr (V <= 4) {
      k.xclamp = xclamp;
      if constexpr (E == 1) k.xchem = xchem;
      k.xpolar = homog ? xpolar : nullptr;
      k.xelener = homog ? xelener : nullptr;
    }
    k.polar = homog ? polar<true> : polar<false>;
    k.elener = homog ? elener<true> : elener<false>;
    return k;
  }
};

// supported (V, NPH, E)
#define MM_COMBOS(X) \
  X(1, 2, 1) X(2, 2, 1) X(3, 2, 1) X(4, 2, 1) X(2, 3, 1) X(3, 3, 1) X(4, 3, 1) X(1, 2, 2) X(3, 2, 2) \
  X(12, 2, 1)

template <typename T> static bool kern_table(int V, int NPH, int E, bool homog, Kern<T>& out) {
#define X(v_, p_, e_) \
  if (V == v_ && NPH == p_ && E == e_) { out = KernImpl<T, v_, p_, e_>::table(homog); return true; }
  MM_COMBOS(X)
#undef X
  return false;
}

}  // namespace mm

// ================================================================= context
using namespace mm;

struct mm_adj;  // adjoint tape and state (adjoint.h)
void adj_release(mm_adj* a);

struct mm_ctx {
  mm_params prm{};
  int prec = 0;
  int V = 0, E = 0, NPH = 0;
  size_t n3 = 0, nc = 0;
  int step = 0;
  Fft3 fft;
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  DevParams hp{};
  DevParams* dp = nullptr;
  size_t esz = 8;
  void *phi = nullptr, *cons = nullptr, *gphi = nullptr, *kdel = nullptr, *fdiff = nullptr, *flux = nullptr,
       *strain = nullptr, *xpol = nullptr, *el_en = nullptr, *elgrad = nullptr;
  void *sphi = nullptr, *suhat = nullptr, *swork = nullptr;
  // fused path: spectra between the x and y passes (S layout), kStageN of them;
  // the y pass moves them to and from the column-tiled layout of the z kernels
  void* sstage = nullptr;
  // eager steps on the fused path run the Cahn-Hilliard update (which needs only
  // c and the fluxes of the chemistry) on a second stream, beside the elastic
  // solve, with its own work spectra (4E) and stage (4), when MM_CH_STREAM=1
  bool ch_async = false;
  bool ch_launched = false;
  hipStream_t s2 = nullptr;
  hipEvent_t ev_head = nullptr, ev_ch = nullptr;
  void* sch = nullptr;
  // with the fused x path: the strain between steps, as x half-spectra (S
  // layout after the inverse y pass, 6 components); c->strain is then filled
  // from it only on demand (field access, checkpoints)
  void* sstrain = nullptr;
  int nwork = 0, nxpol = 0;
  double* part = nullptr;
  double* dres = nullptr;
  double* hres = nullptr;  // pinned
  int* derr = nullptr;
  int nblk_pw = 0, nblk_k = 0;
  double glob_min = 1.0, glob_max = 0.0, phi_max = 0.0, solid_frac = 0.0;
  Kern<double> k64;
  Kern<float> k32;
  Probe probe;
  bool zfused = true;  // z passes with the k-space operations fused in (MM_ZFUSED=0: separate kernels)
  bool autotune = true;  // time the FFT plan variants at creation (MM_AUTOTUNE=0: defaults)
  int zf_el_var = 0;     // launch-shape variant of the fused elastic z kernel
  // elastic iterations after the first run the fused x pass + polarization
  // stress (needs zfused, homogeneous moduli and a fused kernel for Nx;
  // MM_XFUSED=0 switches it off)
  bool xpolar = false;
  bool csparse = false;  // CEla of the block form: the fused kernels skip its zero terms
  // the Allen-Cahn inverse x pass runs fused with the clamp (needs zfused and a
  // fused kernel for Nx; MM_XFUSED=0 switches it off)
  bool xclamp = false;
  // Langevin noise (steps with step <= noise_step): generator state, the
  // accumulated phi_noise (host, binary64, SoA), its device copy and a pinned
  // staging buffer in the storage type
  bool noise_on = false;
  bool noise_live = false;  // phi_noise is nonzero (a noise step ran since the last reset)
  double noise_sd = 0.0;
  int32_t rng_idum = 0;      // ran_2's idum and its saved state (mm_ran2)
  int32_t rng_state[35] = {};
  std::vector<double> noise_acc;
  void* noise_dev = nullptr;
  void* noise_pin = nullptr;
  int last_niter = 0;
  bool qchecked = false;  // the acoustic tensor passed k_qcheck
  // fused chemistry pass (k_x_chem): the gradient block's inverse y pass
  // writes its 4V spectra to sgrad (S layout), k_x_chem forms f_diff and the
  // Cahn-Hilliard x spectra, and the Cahn-Hilliard update runs at the end of
  // the head (ch_done); grad phi and del.G del phi are formed on demand only
  bool xchem = false;
  bool ch_done = false;
  bool step_ran = false;  // a step has formed phi_hat since the last upload
  bool xs_tail = false;   // the last tail run or recorded handed f_diff over as x half-spectra
  bool xs_graph = false;  // the same for the recorded step of graph mode
  bool fd_split = false;  // fdiff holds the chemistry's f_diff without el_grad
  bool rec_ch_head = false;  // the recorded graph's head runs the Cahn-Hilliard update
  void* sgrad = nullptr;
  // branch tape (mm_branch_tape): caller's device buffer of br_cap steps,
  // br_stride bytes each; bt_now is the tape of the step being run (mode 0
  // outside run_step_eager)
  BranchTape br{nullptr, nullptr, 0};
  BranchTape bt_now{nullptr, nullptr, 0};
  int64_t br_cap = 0, br_cursor = 0;
  unsigned long long* dflips = nullptr;
  double last_cfac = 1.0;
  // graph mode: the regular step (step > elastep0) recorded once and replayed
  Ctl* ctl = nullptr;       // device control of the elastic iteration
  double* dstage = nullptr;  // kStage doubles: Ctl, the result sums and the error flag after a replay
  double* hstage = nullptr;  // pinned host copy
  const int* gate = nullptr;  // gate applied to launches while recording
  bool graph = false;       // mm_set_graph
  int graph_iters = 4;      // elastic iterations recorded per step
  hipGraph_t g = nullptr;
  hipGraphExec_t gx = nullptr;
  mm_adj* adj = nullptr;    // mm_adj_enable
};

namespace {

constexpr int kMaxPartSlots = 8;
constexpr int kStageN = 6;
constexpr int kMaxBlocks = 2048;      // grid of the pointwise and k-space kernels
constexpr int kMaxPartBlocks = 65536; // largest grid that writes partial sums

int blocks_for(size_t n) { return (int)std::min<size_t>((n + kBlock - 1) / kBlock, (size_t)kMaxBlocks); }

void host_params(const mm_params& p, int prec, DevParams& d) {
  std::memset(&d, 0, sizeof(d));
  d.nx = p.nx; d.ny = p.ny; d.nz = p.nz; d.H = p.nx / 2 + 1;
  d.V = p.vartot; d.E = p.numel; d.NPH = p.numph; d.zdim = p.numph * p.numel;
  int ivn = 0;
  for (int ph = 2; ph <= p.numph; ++ph) {
    ivn += p.nvar[ph - 2];
    for (int v = ivn; v <= ivn + p.nvar[ph - 1] - 1; ++v) d.phase_of[v - 1] = ph;
  }
  for (int v = 0; v < p.vartot; ++v) {
    double s = 0.0;
    for (int m = 0; m < 6; ++m) s += p.trans_ppt[v * 6 + m];
    d.tsum_ok[v] = s > (double)1.e-9f;
  }
  const double pi = std::atan(1.0) * 4.0;
  d.unit[0] = (2.0 * pi) / (double)(float)p.nx;
  d.unit[1] = (2.0 * pi) / (double)(float)p.ny;
  d.unit[2] = (2.0 * pi) / (double)(float)p.nz;
  d.dt = p.t_step; d.bmob = p.b_mob;
  for (int e = 0; e < MM_MAX_EL; ++e) d.cmobs[e] = p.cmobs[e];
  std::memcpy(d.gc, p.grad_coeff, sizeof(d.gc));
  std::memcpy(d.tp, p.trans_ppt, sizeof(d.tp));
  std::memcpy(d.cela, p.cela, sizeof(d.cela));
  std::memcpy(d.sela, p.sela, sizeof(d.sela));
  std::memcpy(d.gp, p.gparams, sizeof(d.gp));
  std::memcpy(d.well, p.wellph, sizeof(d.well));
  std::memcpy(d.e_mean, p.e_mean, sizeof(d.e_mean));
  // sig0_v(ij) = sum_m CEla(ij,m,ph(v)) w_m trans_ppt(m,v)
  for (int v = 0; v < p.vartot; ++v) {
    const int ph = d.phase_of[v];
    for (int a = 0; a < 6; ++a) {
      double s = 0.0;
      for (int m = 0; m < 6; ++m) s += p.cela[(ph - 1) * 36 + m * 6 + a] * (m < 3 ? 1.0 : 2.0) * p.trans_ppt[v * 6 + m];
      d.sig0[v * 6 + a] = s;
    }
  }
  // constant rows of Zintf / Zr (Multi_MEUMAPPS.f:430-448)
  const int E = p.numel, NPH = p.numph;
  auto GP = [&](int j, int ph) { return p.gparams[(ph - 1) * MM_GP_STRIDE + (j - 1)]; };
  for (int e = 1; e <= E; ++e) {
    const int i1 = e + E + 1;
    for (int ph = 1; ph <= NPH - 1; ++ph) {
      const int i0 = (NPH - 1) * (e - 1) + ph, j0 = NPH * (e - 1) + ph;
      d.zc[(i0 - 1) * 16 + (j0 - 1)] = GP(e, ph);
      d.zc[(i0 - 1) * 16 + j0] = -GP(e, ph + 1);
      d.zrc[i0 - 1] = GP(e, ph) * GP(i1, ph) - GP(e, ph + 1) * GP(i1, ph + 1);
    }
  }
  d.homog = 1;
  for (int ph = 1; ph < p.numph; ++ph)
    if (std::memcmp(p.cela + ph * 36, p.cela, 36 * sizeof(double)) != 0 ||
        std::memcmp(p.sela + ph * 36, p.sela, 36 * sizeof(double)) != 0)
      d.homog = 0;
  (void)prec;
}

int dalloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return MM_ERR_NOMEM; }
  return 0;
}

template <typename T> T* F(void* base, size_t n, int k) { return (T*)base + (size_t)k * n; }
template <typename T> C2<T>* S(void* base, size_t nc, int k) { return (C2<T>*)base + (size_t)k * nc; }

int finalize(mm_ctx* c, int nblk, int K, std::initializer_list<int> ops, int slot0) {
  FinalizeOps fo;
  int i = 0;
  for (int o : ops) fo.op[i++] = o;
  for (; i < 8; ++i) fo.op[i] = 0;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, c->s, c->part, nblk, K, fo, c->dres + slot0, c->gate);
  MM_HIP(hipGetLastError());
  return 0;
}

void set_gate(mm_ctx* c, const int* g) {
  c->gate = g;
  c->fft.gate = g;
}

template <typename T> const Kern<T>& kern(const mm_ctx* c);
template <> const Kern<double>& kern<double>(const mm_ctx* c) { return c->k64; }
template <> const Kern<float>& kern<float>(const mm_ctx* c) { return c->k32; }

// batched forward FFT of nf real fields (stride n3) into spectra (stride nc)
template <typename T> int fwd(mm_ctx* c, T* in, size_t in_stride, C2<T>* out, int nf) {
  const T scale = (T)(1.0 / (double)(float)c->n3);
  for (int f0 = 0; f0 < nf; f0 += kMaxBatch) {
    const int g = std::min(kMaxBatch, nf - f0);
    const T* ip[kMaxBatch];
    void* op[kMaxBatch];
    for (int j = 0; j < g; ++j) { ip[j] = in + (size_t)(f0 + j) * in_stride; op[j] = out + (size_t)(f0 + j) * c->nc; }
    MM_TRY(c->fft.forward<T>(ip, op, g, c->s, scale));
  }
  return 0;
}
template <typename T> int inv(mm_ctx* c, C2<T>* const* sp, T* const* out, int nf) {
  for (int f0 = 0; f0 < nf; f0 += kMaxBatch) {
    const int g = std::min(kMaxBatch, nf - f0);
    MM_TRY(c->fft.inverse<T>((void* const*)(sp + f0), out + f0, g, c->s));
  }
  return 0;
}

// x and y passes only (the z pass is fused with a k-space operation): real
// fields -> S spectra in the stage -> column-tiled spectra out[]
// (on stream s, default c->s, through a stage of nst spectra, default the
// context's kStageN)
template <typename T>
int fwd_xy(mm_ctx* c, const T* const* in, C2<T>* const* out, int nf, hipStream_t s = nullptr,
           C2<T>* stg = nullptr, int nst = kStageN) {
  const T scale = (T)(1.0 / (double)(float)c->n3);
  if (nst < 1 || nst > kStageN) return MM_ERR_ARG;
  if (!s) s = c->s;
  if (!stg) stg = (C2<T>*)c->sstage;
  void* sp[kStageN];
  for (int j = 0; j < nst; ++j) sp[j] = stg + (size_t)j * c->nc;
  for (int f0 = 0; f0 < nf; f0 += nst) {
    const int g = std::min(nst, nf - f0);
    MM_TRY(c->fft.x_r2c<T>(in + f0, sp, g, s, scale));
    MM_TRY(c->fft.y_pass_t<T>(sp, (void* const*)(out + f0), g, -1, 1, s));
  }
  return 0;
}
// column-tiled spectra sp[] -> S spectra in the stage -> real fields out[]
template <typename T>
int inv_yx(mm_ctx* c, C2<T>* const* sp, T* const* out, int nf, hipStream_t s = nullptr, C2<T>* stg = nullptr,
           int nst = kStageN) {
  if (nst < 1 || nst > kStageN) return MM_ERR_ARG;
  if (!s) s = c->s;
  if (!stg) stg = (C2<T>*)c->sstage;
  void* st[kStageN];
  for (int j = 0; j < nst; ++j) st[j] = stg + (size_t)j * c->nc;
  for (int f0 = 0; f0 < nf; f0 += nst) {
    const int g = std::min(nst, nf - f0);
    MM_TRY(c->fft.y_pass_t<T>((void* const*)(sp + f0), st, g, +1, 2, s));
    MM_TRY(c->fft.x_c2r<T>(st, out + f0, g, s));
  }
  return 0;
}
// u_hat between the full-transform S layout and the tiled layout of the fused path
template <typename T> int uhat_s_to_z(mm_ctx* c, C2<T>* scratch3) {
  if (!c->zfused) return 0;
  MM_TRY(c->fft.perm_sz<T>(c->suhat, scratch3, 3, 1, c->s));
  MM_HIP(hipMemcpyAsync(c->suhat, scratch3, 3 * c->nc * sizeof(C2<T>), hipMemcpyDeviceToDevice, c->s));
  return 0;
}

template <typename T> int strain_from_real(mm_ctx* c);

// ------------------------------------------------------------ elasticity
// zeroth-order displacement at step == elastep0 (ModElastic.f:105-236)
template <typename T> int elastic_zeroth(mm_ctx* c) {
  const size_t n3 = c->n3, nc = c->nc;
  const Kern<T>& K = kern<T>(c);
  C2<T>* work = (C2<T>*)c->swork;
  C2<T>* wsp[6];
  T* wre[6];
  for (int m = 0; m < 6; ++m) { wsp[m] = work + (size_t)m * nc; wre[m] = (T*)c->strain + (size_t)m * n3; }
  K.ppoly(dim3(c->nblk_pw), c->s, (T*)c->phi, (T*)c->xpol, n3);
  MM_HIP(hipGetLastError());
  C2<T>* phat = work + (size_t)6 * nc;
  MM_TRY(fwd<T>(c, (T*)c->xpol, n3, phat, c->V));
  K.zeroth(dim3(c->nblk_k), c->s, c->dp, phat, (C2<T>*)c->suhat, work, nc);
  MM_HIP(hipGetLastError());
  MM_TRY(inv<T>(c, wsp, wre, 6));
  MM_TRY(uhat_s_to_z<T>(c, work));
  if (c->xpolar) MM_TRY(strain_from_real<T>(c));
  return 0;
}

// One fixed-point iteration (ModElastic.f:241-433) up to the six convergence
// sums in dres[R_ERR0..]; the strain inverse transforms follow separately.
// cfac comes from the argument (eager) or from the device control (graph).
// spectral: the strain of the previous iteration is held as six spectra
// after the inverse y pass (work[0..5]) and the fused x kernel takes it from
// there; otherwise it is read from the real strain fields.
template <typename T> int elastic_iter(mm_ctx* c, double cfac, const double* cfac_dev, bool spectral) {
  const size_t n3 = c->n3, nc = c->nc;
  const Kern<T>& K = kern<T>(c);
  T* xpol = (T*)c->xpol;
  C2<T>* work = (C2<T>*)c->swork;
  C2<T>* uhat = (C2<T>*)c->suhat;
  C2<T>* wsp[6];
  for (int m = 0; m < 6; ++m) wsp[m] = work + (size_t)m * nc;
  if (spectral) {
    const T scale = (T)(1.0 / (double)(float)n3);
    c->probe.begin(MM_PROBE_X_POLAR, c->s);
    // in place: the strain half-spectra are replaced by the polarization ones
    // (the inverse pass of this iteration writes the new strain back)
    C2<T>* stg = (C2<T>*)c->sstrain;
    K.xpolar(true, c->s, c->dp, (const T*)c->phi, (const C2<T>*)c->sstrain, stg, nc, n3, c->fft.twLx, c->fft.twNx,
             c->fft.nx, c->fft.ny, c->fft.nz, scale, cfac, cfac_dev, c->gate, c->csparse);
    c->probe.end(MM_PROBE_X_POLAR, c->s, (double)sizeof(T) * (24.0 * nc + (double)c->V * n3));
    MM_HIP(hipGetLastError());
    void* sp[6];
    for (int m = 0; m < 6; ++m) sp[m] = stg + (size_t)m * nc;
    MM_TRY(c->fft.y_pass_t<T>(sp, (void* const*)wsp, 6, -1, 1, c->s));
  } else {
    c->probe.begin(MM_PROBE_POLAR, c->s);
    K.polar(dim3(c->nblk_pw), c->s, c->dp, (T*)c->phi, (T*)c->strain, cfac, cfac_dev, xpol, c->derr, n3, c->gate);
    c->probe.end(MM_PROBE_POLAR, c->s, (double)n3 * sizeof(T) * (c->V + 12));
    MM_HIP(hipGetLastError());
    if (c->zfused) {
      const T* xin[6];
      for (int m = 0; m < 6; ++m) xin[m] = xpol + (size_t)m * n3;
      MM_TRY(fwd_xy<T>(c, xin, wsp, 6));
    } else {
      MM_TRY(fwd<T>(c, xpol, n3, work, 6));
    }
  }
  if (c->zfused) {
    OpElastic<T> op{c->dp, work, uhat, c->part, nc};
    op.gate = c->gate;
    int nblk = 0;
    MM_TRY(launch_zf_elastic<T>(c->fft, op, c->zf_el_var, c->s, MM_PROBE_ZF_ELASTIC, (double)nc * 2 * sizeof(T) * 18,
                                &nblk));
    MM_TRY(finalize(c, nblk, 6, {0, 0, 0, 0, 0, 0}, R_ERR0));
  } else {
    c->probe.begin(MM_PROBE_KS_ELASTIC, c->s);
    hipLaunchKernelGGL(k_ks_elastic<T>, dim3(c->nblk_k), dim3(kBlock), 0, c->s, c->dp, work, uhat, c->part, nc);
    c->probe.end(MM_PROBE_KS_ELASTIC, c->s, (double)nc * 2 * sizeof(T) * 18);
    MM_HIP(hipGetLastError());
    MM_TRY(finalize(c, c->nblk_k, 6, {0, 0, 0, 0, 0, 0}, R_ERR0));
  }
  return 0;
}

// Inverse transforms of the new strain spectra: with the fused x kernel only
// the y pass (the strain stays spectral, returns true), else back to real.
template <typename T> int elastic_strain_inverse(mm_ctx* c, bool* spectral) {
  const size_t n3 = c->n3, nc = c->nc;
  C2<T>* work = (C2<T>*)c->swork;
  C2<T>* wsp[6];
  T* wre[6];
  for (int m = 0; m < 6; ++m) { wsp[m] = work + (size_t)m * nc; wre[m] = (T*)c->strain + (size_t)m * n3; }
  if (c->xpolar) {
    // the strain stays as x half-spectra for the fused x kernels
    *spectral = true;
    C2<T>* ss = (C2<T>*)c->sstrain;
    void* sp[6];
    for (int m = 0; m < 6; ++m) sp[m] = ss + (size_t)m * nc;
    return c->fft.y_pass_t<T>((void* const*)wsp, sp, 6, +1, 2, c->s);
  }
  *spectral = false;
  if (c->zfused) return inv_yx<T>(c, wsp, wre, 6);
  return inv<T>(c, wsp, wre, 6);
}

// the real strain from its x half-spectra, and back (scale 1/Nx: the x c2r
// is unnormalised)
template <typename T> int strain_to_real(mm_ctx* c) {
  const size_t n3 = c->n3, nc = c->nc;
  C2<T>* ss = (C2<T>*)c->sstrain;
  void* sp[6];
  T* re[6];
  for (int m = 0; m < 6; ++m) { sp[m] = ss + (size_t)m * nc; re[m] = (T*)c->strain + (size_t)m * n3; }
  return c->fft.x_c2r<T>(sp, re, 6, c->s);
}
template <typename T> int strain_from_real(mm_ctx* c) {
  const size_t n3 = c->n3, nc = c->nc;
  C2<T>* ss = (C2<T>*)c->sstrain;
  void* sp[6];
  const T* re[6];
  for (int m = 0; m < 6; ++m) { sp[m] = ss + (size_t)m * nc; re[m] = (const T*)c->strain + (size_t)m * n3; }
  return c->fft.x_r2c<T>(re, sp, 6, c->s, (T)(1.0 / (double)c->fft.nx));
}
int strain_to_real_any(mm_ctx* c) {
  if (!c->xpolar) return 0;
  return c->prec == MM_FP64 ? strain_to_real<double>(c) : strain_to_real<float>(c);
}
int strain_from_real_any(mm_ctx* c) {
  if (!c->xpolar) return 0;
  return c->prec == MM_FP64 ? strain_from_real<double>(c) : strain_from_real<float>(c);
}

// eager iterations with host control, starting from the control state h
template <typename T> int elastic_loop_eager(mm_ctx* c, Ctl& h, bool spectral) {
  while (h.active) {
    MM_TRY(elastic_iter<T>(c, h.cfac, nullptr, spectral));
    MM_HIP(hipMemcpyAsync(c->hres + R_ERR0, c->dres + R_ERR0, 6 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    MM_HIP(hipEventRecord(c->ev, c->s));
    MM_TRY(elastic_strain_inverse<T>(c, &spectral));
    MM_HIP(hipEventSynchronize(c->ev));
    ctl_after_iter(h, c->hres + R_ERR0);
  }
  c->last_niter = h.total;
  c->last_cfac = h.cfac;
  return h.err;
}

template <typename T> int ch_update(mm_ctx* c, hipStream_t s, C2<T>* work, C2<T>* stg, int nst,
                                    bool staged = false);

// the gradient block's spectra (all V variants at once) into sgrad: z pass
// with the gradient multipliers, then the inverse y pass to the S layout
template <typename T> int grad_spectra(mm_ctx* c) {
  const int V = c->V;
  const size_t nc = c->nc;
  C2<T>* work = (C2<T>*)c->swork;
  // z pass: phi_hat, phi_hat i kf_z and phi_hat G_zz kq_z per variant; the
  // inverse y pass applies the kx and ky factors and forms the 4V spectra
  OpGrad3<T> op{c->dp, (C2<T>*)c->sphi, work, nc};
  op.gate = c->gate;
  MM_TRY(launch_zfused<T>(c->fft, op, V, c->s, MM_PROBE_ZF_OTHER, (double)nc * 2 * sizeof(T) * 5 * V, nullptr));
  MM_HIP(hipGetLastError());
  return launch_y_grad<T>(c->fft, c->dp, work, c->sgrad, V, c->s);
}

// head with the fused chemistry pass (phi_hat already formed): gradient
// spectra, k_x_chem (f_diff and the Cahn-Hilliard x spectra in the stage),
// then the rest of the Cahn-Hilliard update (:910-949), which needs nothing
// from the elastic solve
template <typename T> int head_xchem(mm_ctx* c) {
  const size_t n3 = c->n3, nc = c->nc;
  const Kern<T>& K = kern<T>(c);
  MM_TRY(grad_spectra<T>(c));
  C2<T>* stg = (C2<T>*)c->sstage;
  int nblk = 0;
  c->probe.begin(MM_PROBE_CHEM, c->s);
  K.xchem(true, c->s, c->dp, (const T*)c->phi, (const T*)c->cons, (const C2<T>*)c->sgrad, nc, n3, c->fft.twLx,
          c->fft.twNx, c->fft.nx, c->fft.ny, c->fft.nz, (T)(1.0 / (double)(float)n3), (T*)c->fdiff, stg, c->part,
          c->derr, c->gate, &nblk);
  c->probe.end(MM_PROBE_CHEM, c->s,
               (double)sizeof(T) * ((double)(c->V + c->E + c->V) * n3 + 2.0 * (4 * c->V + 4 * c->E) * nc));
  MM_HIP(hipGetLastError());
  MM_TRY(finalize(c, nblk, 3, {0, 0, 0}, R_FINT));
  MM_TRY(ch_update<T>(c, c->s, (C2<T>*)c->swork, stg, 4 * c->E, true));
  c->ch_done = true;
  c->step_ran = true;
  return 0;
}

// grad phi and del.G del phi of the last step for mm_field (5, 6) on the
// fused chemistry path, where the step does not form them: recomputed from
// phi_hat, which the head's gradient z pass left as the full transform in
// the column-tiled layout (back to S, k-space multipliers, full inverses)
template <typename T> int grad_fields(mm_ctx* c) {
  const int V = c->V;
  const size_t nc = c->nc, n3 = c->n3;
  C2<T>* sph = (C2<T>*)c->sgrad;
  C2<T>* work = (C2<T>*)c->swork;
  MM_TRY(c->fft.perm_sz<T>(c->sphi, sph, V, 2, c->s));
  hipLaunchKernelGGL(k_ks_grad<T>, dim3(c->nblk_k, V), dim3(kBlock), 0, c->s, c->dp, (const C2<T>*)sph, work, 0, nc);
  MM_HIP(hipGetLastError());
  for (int v = 0; v < V; ++v) {
    void* sp[4];
    T* rp[4];
    for (int d = 0; d < 4; ++d) sp[d] = work + (size_t)(4 * v + d) * nc;
    for (int d = 0; d < 3; ++d) rp[d] = (T*)c->gphi + (size_t)(3 * v + d) * n3;
    rp[3] = (T*)c->kdel + (size_t)v * n3;
    MM_TRY(c->fft.inverse<T>(sp, rp, 4, c->s));
  }
  MM_HIP(hipStreamSynchronize(c->s));
  return 0;
}

// -------------------------------------------------------------- one step
// head: phi_hat, gradient block, chemistry (Multi_MEUMAPPS.f:575-845)
template <typename T> int step_head(mm_ctx* c) {
  const size_t n3 = c->n3, nc = c->nc;
  c->ch_launched = false;  // set by the eager step once the update is on the second stream
  const int V = c->V, E = c->E;
  const Kern<T>& K = kern<T>(c);
  T* phi = (T*)c->phi;
  C2<T>* sphi = (C2<T>*)c->sphi;
  C2<T>* work = (C2<T>*)c->swork;
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(kBlock), 0, c->s, c->derr, (size_t)1, 0, c->gate);
  MM_HIP(hipGetLastError());
  // 1. phi_hat (:577-578); kept for the Allen-Cahn update
  if (c->zfused) {
    const T* pin[MM_MAX_VAR];
    C2<T>* pout[MM_MAX_VAR];
    for (int v = 0; v < V; ++v) { pin[v] = phi + (size_t)v * n3; pout[v] = sphi + (size_t)v * nc; }
    MM_TRY(fwd_xy<T>(c, pin, pout, V));
  } else {
    MM_TRY(fwd<T>(c, phi, n3, sphi, V));
  }
  if (c->xchem && !c->bt_now.mode) return head_xchem<T>(c);
  // 2. gradient block (:581-613) in groups of up to 4 variants
  const int G = std::max(1, std::min({V, c->nwork / 4, 4}));
  for (int v0 = 0; v0 < V; v0 += G) {
    const int g = std::min(G, V - v0);
    if (c->zfused) {
      OpGrad<T> op{c->dp, sphi, work, nc, v0};
      op.gate = c->gate;
      MM_TRY(launch_zfused<T>(c->fft, op, g, c->s, MM_PROBE_ZF_OTHER, (double)nc * 2 * sizeof(T) * 6 * g, nullptr));
    } else {
      hipLaunchKernelGGL(k_ks_grad<T>, dim3(c->nblk_k, g), dim3(kBlock), 0, c->s, c->dp, sphi, work, v0, nc);
    }
    MM_HIP(hipGetLastError());
    C2<T>* sp[16];
    T* rp[16];
    for (int j = 0; j < g; ++j) {
      const int v = v0 + j;
      for (int dd = 0; dd < 3; ++dd) { sp[4 * j + dd] = work + (size_t)(4 * j + dd) * nc; rp[4 * j + dd] = (T*)c->gphi + (size_t)(3 * v + dd) * n3; }
      sp[4 * j + 3] = work + (size_t)(4 * j + 3) * nc;
      rp[4 * j + 3] = (T*)c->kdel + (size_t)v * n3;
    }
    if (c->zfused) {
      MM_TRY(inv_yx<T>(c, sp, rp, 4 * g));
    } else {
      MM_TRY(inv<T>(c, sp, rp, 4 * g));
    }
  }
  // 3. chemistry (:634-845)
  c->probe.begin(MM_PROBE_CHEM, c->s);
  K.chem(dim3(c->nblk_pw), c->s, c->dp, phi, (T*)c->gphi, (T*)c->kdel, (T*)c->cons, (T*)c->fdiff, (T*)c->flux,
         c->part, c->derr, n3, c->bt_now);
  c->probe.end(MM_PROBE_CHEM, c->s, (double)n3 * sizeof(T) * (6 * V + 4 * E));
  MM_HIP(hipGetLastError());
  MM_TRY(finalize(c, c->nblk_pw, 3, {0, 0, 0}, R_FINT));
  return 0;
}

// check-loop sums of the clamp from nblk partials per slot
int finalize_clamp(mm_ctx* c, int nblk) {
  const int E = c->E;
  FinalizeOps fo;
  for (int e = 0; e < E; ++e) fo.op[e] = 0;
  fo.op[E] = 1; fo.op[E + 1] = 2; fo.op[E + 2] = 2; fo.op[E + 3] = 0;
  for (int i = E + 4; i < 8; ++i) fo.op[i] = 0;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinBlock), 0, c->s, c->part, nblk, E + 4, fo, c->dres + R_CLAMP,
                     c->gate);
  MM_HIP(hipGetLastError());
  return 0;
}

// Cahn-Hilliard update (Multi_MEUMAPPS.f:910-949) on stream s: work[4e] = c_hat,
// work[4e+1..3] = F_hat (4E spectra), the fused passes staging through stg
// (nst spectra; null = the context's stage); writes c
template <typename T> int ch_update(mm_ctx* c, hipStream_t s, C2<T>* work, C2<T>* stg, int nst, bool staged) {
  const size_t n3 = c->n3, nc = c->nc;
  const int E = c->E;
  const T scale = (T)(1.0 / (double)(float)n3);
  if (!stg) stg = (C2<T>*)c->sstage;
  if (c->zfused) {
    // x pass of c and the fluxes into the stage (already done by k_x_chem when
    // staged), then the y pass with the x and y flux terms folded in: work[2e]
    // = c_hat + dt M (i kf_x F_x + i kf_y F_y), work[2e+1] = F_z
    if (staged) {
      MM_TRY(launch_y_ch<T>(c->fft, c->dp, stg, work, E, 0, s));
    } else {
      if (nst < 4) return MM_ERR_ARG;
      for (int e = 0; e < E; ++e) {
        const T* ip[4] = {(T*)c->cons + (size_t)e * n3, (T*)c->flux + (size_t)(3 * e) * n3,
                          (T*)c->flux + (size_t)(3 * e + 1) * n3, (T*)c->flux + (size_t)(3 * e + 2) * n3};
        void* sp4[4];
        for (int m = 0; m < 4; ++m) sp4[m] = stg + (size_t)m * nc;
        MM_TRY(c->fft.x_r2c<T>(ip, sp4, 4, s, scale));
        MM_TRY(launch_y_ch<T>(c->fft, c->dp, stg, work + (size_t)(2 * e) * nc, 1, e, s));
      }
    }
    OpCH2<T> opch{c->dp, work, nc};
    opch.gate = c->gate;
    MM_TRY(launch_zfused<T>(c->fft, opch, E, s, MM_PROBE_ZF_OTHER, (double)nc * 2 * sizeof(T) * 3 * E, nullptr));
  } else {
    for (int e = 0; e < E; ++e) {
      const T* ip[4] = {(T*)c->cons + (size_t)e * n3, (T*)c->flux + (size_t)(3 * e) * n3,
                        (T*)c->flux + (size_t)(3 * e + 1) * n3, (T*)c->flux + (size_t)(3 * e + 2) * n3};
      void* op[4] = {work + (size_t)(4 * e) * nc, work + (size_t)(4 * e + 1) * nc, work + (size_t)(4 * e + 2) * nc,
                     work + (size_t)(4 * e + 3) * nc};
      MM_TRY(c->fft.forward<T>(ip, op, 4, s, scale));
    }
    hipLaunchKernelGGL(k_ks_ch<T>, dim3(c->nblk_k, E), dim3(kBlock), 0, s, c->dp, work, nc);
  }
  MM_HIP(hipGetLastError());
  C2<T>* sp[MM_MAX_EL];
  T* rp[MM_MAX_EL];
  const int cs = c->zfused ? 2 : 4;  // spacing of the new c_hat in work
  for (int e = 0; e < E; ++e) { sp[e] = work + (size_t)(cs * e) * nc; rp[e] = (T*)c->cons + (size_t)e * n3; }
  if (c->zfused) {
    MM_TRY(inv_yx<T>(c, sp, rp, E, s, stg, nst));
  } else {
    MM_TRY(inv<T>(c, sp, rp, E));
  }
  return 0;
}

// tail: elener (when the elastic solve ran), Cahn-Hilliard, Allen-Cahn, clamp
// rec_fd / rec_phipre (FP64 tape): copies of f_diff after elener and of phi
// before the clamp
template <typename T> int step_tail(mm_ctx* c, bool elastic, double* rec_fd = nullptr, double* rec_phipre = nullptr) {
  const size_t n3 = c->n3, nc = c->nc;
  const int V = c->V, E = c->E;
  const Kern<T>& K = kern<T>(c);
  T* phi = (T*)c->phi;
  C2<T>* sphi = (C2<T>*)c->sphi;
  C2<T>* work = (C2<T>*)c->swork;
  const dim3 gp(c->nblk_pw);
  // With the fused x kernels and the Cahn-Hilliard update already done, elener
  // hands f_diff to the Allen-Cahn update as x half-spectra in the stage (the
  // real f_diff + el_grad is not formed; fdiff keeps the chemistry's part and
  // mm_field adds el_grad on demand).  Noise steps and recorded steps need the
  // real f_diff.  Rows of up to 256 points only (the binary64 kernel on
  // 512-point rows spills: 5.4 -> 7.0 ms at 512^3); binary32 runs it on the
  // 4-point plan (8 points per thread spill with the three extra transforms)
  // and so only for the power-of-two rows that plan covers.
  const bool xs = (sizeof(T) == 8 || xs32_rows(c->fft.nx / 2)) && c->fft.nx <= 256 && elastic && c->xpolar &&
                  c->zfused && c->ch_done && c->step > c->prm.noise_step && !rec_fd;
  c->xs_tail = xs;
  c->fd_split = xs;
  if (elastic) {
    c->probe.begin(MM_PROBE_ELENER, c->s);
    int nblk = c->nblk_pw;
    if (c->xpolar) {
      // the strain's x c2r fused with elener (the real strain is not formed)
      K.xelener(true, c->s, c->dp, phi, (const C2<T>*)c->sstrain, nc, n3, c->fft.twLx, c->fft.twNx, c->fft.nx,
                c->fft.ny, c->fft.nz, (T*)c->el_en, (T*)c->elgrad, (T*)c->fdiff,
                xs ? (C2<T>*)c->sstage : nullptr, (T)(1.0 / (double)(float)n3), c->part, c->gate, &nblk,
                c->csparse);
    } else {
      K.elener(gp, c->s, c->dp, phi, (T*)c->strain, (T*)c->el_en, (T*)c->elgrad, (T*)c->fdiff, c->part, c->derr,
               n3, c->gate);
    }
    c->probe.end(MM_PROBE_ELENER, c->s, (double)n3 * sizeof(T) * (4 * c->V + 7));
    MM_HIP(hipGetLastError());
    MM_TRY(finalize(c, nblk, 1, {0}, R_EL));
    if (rec_fd) MM_HIP(hipMemcpyAsync(rec_fd, c->fdiff, (size_t)V * n3 * sizeof(T), hipMemcpyDeviceToDevice, c->s));
  } else {
    hipLaunchKernelGGL(k_fill<T>, gp, dim3(kBlock), 0, c->s, (T*)c->el_en, n3, (T)0, c->gate);
    hipLaunchKernelGGL(k_fill<T>, gp, dim3(kBlock), 0, c->s, (T*)c->elgrad, (size_t)V * n3, (T)0, c->gate);
    hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(kBlock), 0, c->s, c->dres + R_EL, (size_t)1, 0.0, c->gate);
    MM_HIP(hipGetLastError());
  }
  // Cahn-Hilliard (:910-949), unless it already ran beside the elastic solve
  // or at the end of the head (fused chemistry pass)
  if (c->ch_done) {
    c->ch_done = false;
  } else if (c->ch_launched) {
    MM_HIP(hipStreamWaitEvent(c->s, c->ev_ch, 0));
    c->ch_launched = false;
  } else {
    MM_TRY(ch_update<T>(c, c->s, work, nullptr, kStageN));
  }
  // Allen-Cahn (:955-1001)
  {
    C2<T>* sp[MM_MAX_VAR];
    T* rp[MM_MAX_VAR];
    T* fsrc = (T*)c->fdiff;
    if (c->step <= c->prm.noise_step) {
      // a noise step: phi_noise (already on the device) joins f_diff
      hipLaunchKernelGGL(k_noise_fold<T>, gp, dim3(kBlock), 0, c->s, (const T*)c->fdiff, (const T*)c->noise_dev,
                         (T*)c->xpol, (size_t)V * n3, c->prm.b_mob);
      MM_HIP(hipGetLastError());
      fsrc = (T*)c->xpol;
    }
    if (c->zfused) {
      const T* fin[MM_MAX_VAR];
      C2<T>* fout[MM_MAX_VAR];
      for (int v = 0; v < V; ++v) { fin[v] = fsrc + (size_t)v * n3; fout[v] = work + (size_t)v * nc; }
      if (xs) {
        // the x half-spectra of f_diff are in the stage already
        void* st[MM_MAX_VAR];
        for (int v = 0; v < V; ++v) st[v] = (C2<T>*)c->sstage + (size_t)v * nc;
        MM_TRY(c->fft.y_pass_t<T>(st, (void* const*)fout, V, -1, 1, c->s));
      } else {
        MM_TRY(fwd_xy<T>(c, fin, fout, V));
      }
      OpAC<T> opac{c->dp, work, sphi, nc};
      opac.gate = c->gate;
      MM_TRY(launch_zfused<T>(c->fft, opac, V, c->s, MM_PROBE_ZF_OTHER, (double)nc * 2 * sizeof(T) * 3 * V, nullptr));
      for (int v = 0; v < V; ++v) { sp[v] = work + (size_t)v * nc; rp[v] = phi + (size_t)v * n3; }
      if (c->xclamp && !rec_phipre && !c->bt_now.mode) {
        // inverse y pass into the stage; the x pass, the clamp and the
        // check-loop sums in one kernel (phi written once)
        C2<T>* stg = (C2<T>*)c->sstage;
        void* st[kStageN];
        for (int v = 0; v < V; ++v) st[v] = stg + (size_t)v * nc;
        MM_TRY(c->fft.y_pass_t<T>((void* const*)sp, st, V, +1, 2, c->s));
        int nblk = 0;
        K.xclamp(true, c->s, c->dp, stg, nc, phi, (const T*)c->cons, n3, c->fft.twLx, c->fft.twNx, c->fft.nx,
                 c->fft.ny, c->fft.nz, c->part, c->gate, &nblk);
        MM_HIP(hipGetLastError());
        return finalize_clamp(c, nblk);
      }
      MM_TRY(inv_yx<T>(c, sp, rp, V));
    } else {
      MM_TRY(fwd<T>(c, fsrc, n3, work, V));
      hipLaunchKernelGGL(k_ks_ac<T>, dim3(c->nblk_k, V), dim3(kBlock), 0, c->s, c->dp, sphi, work, nc);
      MM_HIP(hipGetLastError());
      for (int v = 0; v < V; ++v) { sp[v] = sphi + (size_t)v * nc; rp[v] = phi + (size_t)v * n3; }
      MM_TRY(inv<T>(c, sp, rp, V));
    }
  }
  // clamp and check-loop sums (:1008-1065)
  if (rec_phipre) MM_HIP(hipMemcpyAsync(rec_phipre, phi, (size_t)V * n3 * sizeof(T), hipMemcpyDeviceToDevice, c->s));
  K.clamp(gp, c->s, c->dp, phi, (T*)c->cons, c->part, n3, c->gate, c->bt_now);
  MM_HIP(hipGetLastError());
  return finalize_clamp(c, c->nblk_pw);
}

// fill the fort.98/99 row from the result slots r, advance `step`
int step_account(mm_ctx* c, const double* r, int herr, mm_diag* d) {
  const int E = c->E;
  const size_t n3 = c->n3;
  MM_TRY(c->probe.collect());
  if (herr) return herr;
  if (d) {
    std::memset(d, 0, sizeof(*d));
    d->step = c->step;
    d->n_elastic = c->last_niter;
    d->cfac = c->last_cfac;
    d->fint = r[R_FINT];
    d->fch = r[R_FCH];
    d->fw = r[R_FW];
    d->elen = r[R_EL];
    d->etot = d->fint + d->fch + d->elen + d->fw;
    d->glob_min = c->glob_min;
    d->glob_max = c->glob_max;
    d->phi_max = c->phi_max;
    d->solid_tot = c->solid_frac / (double)(float)n3;
    for (int e = 0; e < E; ++e) d->caviter[e] = r[R_CLAMP + e] * (1.0 / (double)(float)n3) + c->hp.cbar[e];
  }
  // lagged check-loop values for the next row (con_min starts at 1, con_max and max_phi at 0)
  c->glob_min = std::min(1.0, r[R_CLAMP + E]);
  c->glob_max = std::max(0.0, r[R_CLAMP + E + 1]);
  c->phi_max = std::max(0.0, r[R_CLAMP + E + 2]);
  c->solid_frac = r[R_CLAMP + E + 3];
  c->step += 1;
  return 0;
}

// copy the step's results to the host, then account them
int step_collect(mm_ctx* c, mm_diag* d) {
  MM_HIP(hipMemcpyAsync(c->hres, c->dres, R_N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  MM_HIP(hipMemcpyAsync(c->hres + R_N, c->derr, sizeof(int), hipMemcpyDeviceToHost, c->s));
  MM_HIP(hipStreamSynchronize(c->s));
  int herr = 0;
  std::memcpy(&herr, c->hres + R_N, sizeof(int));
  return step_account(c, c->hres, herr, d);
}

// eager step: host-driven elastic control
// Multi_MEUMAPPS.f:552-571: this step's draws added to phi_noise on the host
// (binary64, the reference's order), then copied to the device
template <typename T> int noise_advance(mm_ctx* c) {
  const size_t n3 = c->n3;
  const int V = c->V;
  double* acc = c->noise_acc.data();
  noise_add(&c->rng_idum, c->rng_state, c->noise_sd, V, n3, acc);
  T* pin = (T*)c->noise_pin;
  for (size_t i = 0; i < (size_t)V * n3; ++i) pin[i] = (T)acc[i];
  MM_HIP(hipMemcpyAsync(c->noise_dev, pin, (size_t)V * n3 * sizeof(T), hipMemcpyHostToDevice, c->s));
  c->noise_live = true;
  return 0;
}

template <typename T> int run_step_eager_body(mm_ctx* c, mm_diag* d);
template <typename T> int run_step_eager(mm_ctx* c, mm_diag* d) {
  if (!c->br.mode) return run_step_eager_body<T>(c, d);
  if (c->br_cursor >= c->br_cap) return MM_ERR_ARG;  // the tape is full
  const size_t stride = (size_t)c->V * c->n3;
  c->bt_now = BranchTape{c->br.buf + (size_t)c->br_cursor * stride, c->dflips, c->br.mode};
  const int rc = run_step_eager_body<T>(c, d);
  c->bt_now = BranchTape{nullptr, nullptr, 0};
  if (!rc) c->br_cursor += 1;
  return rc;
}

template <typename T> int run_step_eager_body(mm_ctx* c, mm_diag* d) {
  const bool noisy = c->step <= c->prm.noise_step;
  if (noisy && !c->noise_on) return MM_ERR_ARG;  // mm_set_noise first
  if (!noisy && c->noise_live) {
    // steps past noise_step: phi_noise = 0 (:569)
    std::fill(c->noise_acc.begin(), c->noise_acc.end(), 0.0);
    c->noise_live = false;
  }
  set_gate(c, nullptr);
  MM_TRY(step_head<T>(c));
  // the Cahn-Hilliard update needs only c and the fluxes: on the second stream
  // it runs beside the elastic solve (joined before the Allen-Cahn update).
  // While a probe times a kernel class the update stays on the main stream,
  // so that the probed kernels run alone.
  if (c->ch_async && c->probe.kind < 0 && !c->ch_done) {
    C2<T>* w = (C2<T>*)c->sch;
    MM_HIP(hipEventRecord(c->ev_head, c->s));
    MM_HIP(hipStreamWaitEvent(c->s2, c->ev_head, 0));
    MM_TRY(ch_update<T>(c, c->s2, w, w + (size_t)4 * c->E * c->nc, 4));
    MM_HIP(hipEventRecord(c->ev_ch, c->s2));
    c->ch_launched = true;
  }
  // the host draws overlap the head on the device
  if (noisy) MM_TRY(noise_advance<T>(c));
  const bool elastic = c->step >= c->prm.elastep0;
  c->last_niter = 0;
  c->last_cfac = 1.0;
  if (elastic) {
    if (c->step == c->prm.elastep0) MM_TRY(elastic_zeroth<T>(c));
    Ctl h;
    ctl_start(h);
    MM_TRY(elastic_loop_eager<T>(c, h, c->xpolar));
  }
  MM_TRY(step_tail<T>(c, elastic));
  return step_collect(c, d);
}

// Record the regular step (step > elastep0) on c->s while it is capturing:
// head, graph_iters gated iterations with device control, gated tail.
template <typename T> int record_step(mm_ctx* c, double* stage) {
  set_gate(c, nullptr);
  MM_TRY(step_head<T>(c));
  c->rec_ch_head = c->ch_done;  // the recorded head runs the Cahn-Hilliard update
  hipLaunchKernelGGL(k_ctl_start, dim3(1), dim3(1), 0, c->s, c->ctl);
  MM_HIP(hipGetLastError());
  bool spectral = c->xpolar;
  for (int i = 0; i < c->graph_iters; ++i) {
    set_gate(c, &c->ctl->active);
    MM_TRY(elastic_iter<T>(c, 0.0, &c->ctl->cfac, spectral));
    set_gate(c, nullptr);
    hipLaunchKernelGGL(k_ctl_iter, dim3(1), dim3(1), 0, c->s, c->ctl, (const double*)(c->dres + R_ERR0));
    MM_HIP(hipGetLastError());
    set_gate(c, &c->ctl->ran);
    MM_TRY(elastic_strain_inverse<T>(c, &spectral));
  }
  set_gate(c, &c->ctl->done);
  MM_TRY(step_tail<T>(c, true));
  c->xs_graph = c->xs_tail;
  set_gate(c, nullptr);
  hipLaunchKernelGGL(k_gather, dim3(1), dim3(64), 0, c->s, (const Ctl*)c->ctl, (const double*)c->dres,
                     (const int*)c->derr, stage);
  MM_HIP(hipGetLastError());
  return 0;
}

// After a replay, from the staged control block and sums st: if the recorded
// iterations did not finish the elastic loop (backtracking, slow
// convergence), continue eagerly from the device state and run the tail.
template <typename T> int finish_replay(mm_ctx* c, const double* st, mm_diag* d) {
  Ctl h;
  std::memcpy(&h, st, sizeof(Ctl));
  c->fd_split = c->xs_graph;  // the replayed tail's hand-over of f_diff
  if (!h.done) {
    set_gate(c, nullptr);
    c->ch_done = c->rec_ch_head;  // the replayed head already updated c
    MM_TRY(elastic_loop_eager<T>(c, h, c->xpolar));
    MM_TRY(step_tail<T>(c, true));
    return step_collect(c, d);
  }
  if (h.err) return h.err;
  c->last_niter = h.total;
  c->last_cfac = h.cfac;
  return step_account(c, st + kStageRes, (int)st[kStageErr], d);
}

int graph_build(mm_ctx* c, hipStream_t cap);

// Once per context, before the first step that runs elasticity (step >=
// elastep0): the acoustic tensor check of k_qcheck (err_flag 3).
int qcheck_once(mm_ctx* c) {
  if (c->qchecked || c->step < c->prm.elastep0) return 0;
  hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(kBlock), 0, c->s, c->derr, (size_t)1, 0, (const int*)nullptr);
  hipLaunchKernelGGL(k_qcheck, dim3(c->nblk_k), dim3(kBlock), 0, c->s, c->dp, c->nc, c->derr);
  MM_HIP(hipGetLastError());
  int e = 0;
  MM_HIP(hipMemcpyAsync(&e, c->derr, sizeof(int), hipMemcpyDeviceToHost, c->s));
  MM_HIP(hipStreamSynchronize(c->s));
  if (e) return e;
  c->qchecked = true;
  return 0;
}

template <typename T> int run_step(mm_ctx* c, mm_diag* d) {
  MM_TRY(qcheck_once(c));
  const bool regular = c->step > c->prm.elastep0 && c->step > c->prm.noise_step;
  if (!(c->graph && c->zfused && regular) || c->br.mode) return run_step_eager<T>(c, d);
  if (!c->gx) MM_TRY(graph_build(c, c->s));
  MM_HIP(hipGraphLaunch(c->gx, c->s));
  MM_HIP(hipMemcpyAsync(c->hstage, c->dstage, kStage * sizeof(double), hipMemcpyDeviceToHost, c->s));
  MM_HIP(hipStreamSynchronize(c->s));
  return finish_replay<T>(c, c->hstage, d);
}


int graph_build(mm_ctx* c, hipStream_t cap) {
  const int kind = c->probe.kind;
  c->probe.kind = -1;  // no event records inside the graph
  MM_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  int rc = c->prec == MM_FP64 ? record_step<double>(c, c->dstage) : record_step<float>(c, c->dstage);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(cap, &g);
  c->probe.kind = kind;
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  c->g = g;
  MM_HIP(hipGraphInstantiate(&c->gx, c->g, nullptr, nullptr, 0));
  return 0;
}

void free_all(mm_ctx* c) {
  adj_release(c->adj);
  c->adj = nullptr;
  if (c->gx) (void)hipGraphExecDestroy(c->gx);
  if (c->g) (void)hipGraphDestroy(c->g);
  c->gx = nullptr;
  c->g = nullptr;
  void* bufs[] = {c->phi, c->cons, c->gphi, c->kdel, c->fdiff, c->flux, c->strain, c->xpol, c->el_en,
                  c->elgrad, c->sphi, c->suhat, c->swork, c->part, c->dres, c->derr, c->dp, c->ctl,
                  c->dstage, c->sstage, c->sstrain, c->dflips, c->sgrad};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->hres) (void)hipHostFree(c->hres);
  if (c->noise_dev) (void)hipFree(c->noise_dev);
  if (c->noise_pin) (void)hipHostFree(c->noise_pin);
  if (c->hstage) (void)hipHostFree(c->hstage);
  if (c->ev) (void)hipEventDestroy(c->ev);
  if (c->ev_head) (void)hipEventDestroy(c->ev_head);
  if (c->ev_ch) (void)hipEventDestroy(c->ev_ch);
  if (c->sch) (void)hipFree(c->sch);
  if (c->s2) (void)hipStreamDestroy(c->s2);
  c->probe.destroy();
  if (c->s) (void)hipStreamDestroy(c->s);
  c->fft.destroy();
}

// upload helpers ---------------------------------------------------------
template <typename T> int put_field(mm_ctx*, void* dst, const std::vector<double>& h) {
  if (sizeof(T) == 8) {
    MM_HIP(hipMemcpy(dst, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(h.size());
    for (size_t i = 0; i < h.size(); ++i) f[i] = (float)h[i];
    MM_HIP(hipMemcpy(dst, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  }
  return 0;
}
template <typename T> int get_field(mm_ctx*, const void* src, std::vector<double>& h) {
  if (sizeof(T) == 8) {
    MM_HIP(hipMemcpy(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(h.size());
    MM_HIP(hipMemcpy(f.data(), src, f.size() * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h.size(); ++i) h[i] = (double)f[i];
  }
  return 0;
}

template <typename T>
int upload(mm_ctx* c, const double* phi, const double* cons, const double* el_en, const double* u,
           const double* gradu) {
  const size_t n3 = c->n3;
  const int V = c->V, E = c->E;
  if (phi) {
    std::vector<double> h((size_t)V * n3);
    for (size_t i = 0; i < n3; ++i)
      for (int v = 0; v < V; ++v) h[(size_t)v * n3 + i] = phi[i * V + v];
    MM_TRY(put_field<T>(c, c->phi, h));
  }
  if (cons) {
    std::vector<double> h((size_t)E * n3);
    for (int e = 0; e < E; ++e) {
      double mean = 0.0;
      if (c->prec == MM_MIXED) {
        for (size_t i = 0; i < n3; ++i) mean += cons[i * E + e];
        mean /= (double)n3;
      }
      c->hp.cbar[e] = mean;
      for (size_t i = 0; i < n3; ++i) h[(size_t)e * n3 + i] = cons[i * E + e] - mean;
    }
    MM_TRY(put_field<T>(c, c->cons, h));
    MM_HIP(hipMemcpy(c->dp, &c->hp, sizeof(DevParams), hipMemcpyHostToDevice));
  }
  if (el_en) {
    std::vector<double> h(el_en, el_en + n3);
    MM_TRY(put_field<T>(c, c->el_en, h));
  }
  if (gradu) {
    // engineering strain from the nine gradients (record order d_1 u_1, d_1 u_2, d_1 u_3, d_2 u_1, ...)
    std::vector<double> h(6 * n3);
    auto g = [&](int dd, int i, size_t p) { return gradu[(size_t)(dd * 3 + i) * n3 + p]; };
    for (size_t p = 0; p < n3; ++p) {
      h[0 * n3 + p] = g(0, 0, p);
      h[1 * n3 + p] = g(1, 1, p);
      h[2 * n3 + p] = g(2, 2, p);
      h[3 * n3 + p] = g(1, 2, p) + g(2, 1, p);
      h[4 * n3 + p] = g(0, 2, p) + g(2, 0, p);
      h[5 * n3 + p] = g(1, 0, p) + g(0, 1, p);
    }
    MM_TRY(put_field<T>(c, c->strain, h));
    if (c->xpolar) {
      MM_TRY(strain_from_real<T>(c));
      MM_HIP(hipStreamSynchronize(c->s));
    }
  }
  if (u) {
    std::vector<double> h(u, u + 3 * n3);
    MM_TRY(put_field<T>(c, c->xpol, h));
    MM_TRY(fwd<T>(c, (T*)c->xpol, n3, (C2<T>*)c->suhat, 3));
    MM_TRY(uhat_s_to_z<T>(c, (C2<T>*)c->swork));
    MM_HIP(hipStreamSynchronize(c->s));
  }
  return 0;
}

template <typename T>
int download(mm_ctx* c, double* phi, double* cons, double* el_en, double* u, double* gradu) {
  const size_t n3 = c->n3, nc = c->nc;
  const int V = c->V, E = c->E;
  MM_HIP(hipStreamSynchronize(c->s));
  if (phi) {
    std::vector<double> h((size_t)V * n3);
    MM_TRY(get_field<T>(c, c->phi, h));
    for (size_t i = 0; i < n3; ++i)
      for (int v = 0; v < V; ++v) phi[i * V + v] = h[(size_t)v * n3 + i];
  }
  if (cons) {
    std::vector<double> h((size_t)E * n3);
    MM_TRY(get_field<T>(c, c->cons, h));
    for (size_t i = 0; i < n3; ++i)
      for (int e = 0; e < E; ++e) cons[i * E + e] = c->hp.cbar[e] + h[(size_t)e * n3 + i];
  }
  if (el_en) {
    std::vector<double> h(n3);
    MM_TRY(get_field<T>(c, c->el_en, h));
    std::memcpy(el_en, h.data(), n3 * 8);
  }
  C2<T>* work = (C2<T>*)c->swork;
  C2<T>* uh = (C2<T>*)c->suhat;
  T* scratch = (T*)c->xpol;
  if (c->zfused && (u || gradu)) {
    // S-layout copy of u_hat in work[3..5]
    uh = work + 3 * nc;
    MM_TRY(c->fft.perm_sz<T>(c->suhat, uh, 3, 2, c->s));
  }
  if (u) {
    MM_HIP(hipMemcpyAsync(work, uh, 3 * nc * sizeof(C2<T>), hipMemcpyDeviceToDevice, c->s));
    C2<T>* sp[3] = {work, work + nc, work + 2 * nc};
    T* rp[3] = {scratch, scratch + n3, scratch + 2 * n3};
    MM_TRY(inv<T>(c, sp, rp, 3));
    MM_HIP(hipStreamSynchronize(c->s));
    std::vector<double> h(3 * n3);
    MM_TRY(get_field<T>(c, scratch, h));
    std::memcpy(u, h.data(), 3 * n3 * 8);
  }
  if (gradu) {
    for (int dd = 0; dd < 3; ++dd) {
      for (int i = 0; i < 3; ++i)
        hipLaunchKernelGGL(k_ks_deriv<T>, dim3(c->nblk_k), dim3(kBlock), 0, c->s, c->dp, uh + (size_t)i * nc,
                           work + (size_t)i * nc, dd, nc);
      MM_HIP(hipGetLastError());
      C2<T>* sp[3] = {work, work + nc, work + 2 * nc};
      T* rp[3] = {scratch, scratch + n3, scratch + 2 * n3};
      MM_TRY(inv<T>(c, sp, rp, 3));
      MM_HIP(hipStreamSynchronize(c->s));
      std::vector<double> h(3 * n3);
      MM_TRY(get_field<T>(c, scratch, h));
      std::memcpy(gradu + (size_t)dd * 3 * n3, h.data(), 3 * n3 * 8);
    }
  }
  return 0;
}

}  // namespace

// ================================================================== C ABI
extern "C" int mm_create(const mm_params* p, int precision, mm_ctx** out) {
  if (!p || !out) return MM_ERR_ARG;
  *out = nullptr;
  if (precision != MM_FP64 && precision != MM_MIXED) return MM_ERR_ARG;
  if (p->vartot < 1 || p->vartot > MM_MAX_VAR || p->numel < 1 || p->numel > MM_MAX_EL || p->numph < 2 ||
      p->numph > MM_MAX_PH)
    return MM_ERR_ARG;
  int s = 0;
  for (int q = 0; q < p->numph; ++q) s += p->nvar[q];
  if (s != p->vartot + 1 || p->nvar[0] != 1) return MM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MM_ERR_NODEV;
  mm_ctx* c = new mm_ctx();
  c->prm = *p;
  c->prec = precision;
  c->V = p->vartot; c->E = p->numel; c->NPH = p->numph;
  c->step = p->step;
  host_params(*p, precision, c->hp);
  bool ok = precision == MM_FP64 ? kern_table<double>(c->V, c->NPH, c->E, c->hp.homog != 0, c->k64)
                                 : kern_table<float>(c->V, c->NPH, c->E, c->hp.homog != 0, c->k32);
  int rc = ok ? c->fft.init(p->nx, p->ny, p->nz, precision) : MM_ERR_ARG;
  c->fft.probe = &c->probe;
  {
    const char* zf = std::getenv("MM_ZFUSED");
    c->zfused = !(zf && zf[0] == '0');
  }
  c->autotune = true;
  {
    const char* at = std::getenv("MM_AUTOTUNE");
    if (at && at[0] == '0') c->autotune = false;
    const char* xf = std::getenv("MM_XFUSED");
    const bool xf_on = !(xf && xf[0] == '0');
    if (!rc && xf_on && c->zfused) {
      // both fused x kernels exist for this x length, and the elener grid
      // fits the partial-sum buffer
      int nb = 0;
      if (precision == MM_FP64)
        c->xpolar = c->k64.xpolar && c->k64.xelener &&
                    c->k64.xpolar(false, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, p->nx,
                                  p->ny, p->nz, 0.0, 0.0, nullptr, nullptr, false) &&
                    c->k64.xelener(false, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, p->nx, p->ny,
                                   p->nz, nullptr, nullptr, nullptr, nullptr, 0.0, nullptr, nullptr, &nb, false);
      else
        c->xpolar = c->k32.xpolar && c->k32.xelener &&
                    c->k32.xpolar(false, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, p->nx,
                                  p->ny, p->nz, 0.0f, 0.0, nullptr, nullptr, false) &&
                    c->k32.xelener(false, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, p->nx, p->ny,
                                   p->nz, nullptr, nullptr, nullptr, nullptr, 0.0f, nullptr, nullptr, &nb, false);
      if (nb > kMaxPartBlocks) c->xpolar = false;
      // the fused x kernels read CEla as a symmetric matrix (every phase has
      // the same moduli here); a non-symmetric one runs the unfused passes
      for (int i = 0; i < 6 && c->xpolar; ++i)
        for (int j = i + 1; j < 6; ++j)
          if (std::memcmp(&c->hp.cela[i * 6 + j], &c->hp.cela[j * 6 + i], sizeof(double)) != 0) c->xpolar = false;
      // block form (cubic, tetragonal, hexagonal): normal-shear coupling and
      // off-diagonal shear terms exactly zero
      c->csparse = c->xpolar;
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
          if ((i < 3 ? j >= 3 : j != i) && c->hp.cela[i * 6 + j] != 0.0) c->csparse = false;
      nb = 0;
      if (precision == MM_FP64)
        c->xclamp = c->k64.xclamp &&
                    c->k64.xclamp(false, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr, p->nx,
                                  p->ny, p->nz, nullptr, nullptr, &nb);
      else
        c->xclamp = c->k32.xclamp &&
                    c->k32.xclamp(false, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr, p->nx,
                                  p->ny, p->nz, nullptr, nullptr, &nb);
      if (nb > kMaxPartBlocks || c->V > kStageN) c->xclamp = false;
      // the chemistry pass: measured 4.6% (256^3) and 2.4% (512^3) faster FP64
      // steps, 2% faster mixed at 256^3 and 1% slower mixed at 512^3 (its
      // FP64 chemistry at 2 waves per SIMD is latency bound), so on for FP64
      // and for mixed up to 256^3 points; MM_XCHEM=0 / 1 forces it off / on
      const char* xc = std::getenv("MM_XCHEM");
      const bool xc_want = xc && xc[0] ? xc[0] != '0'
                                       : (precision == MM_FP64 || (size_t)p->nx * p->ny * p->nz <= ((size_t)1 << 24));
      if (xc_want) {
        nb = 0;
        if (precision == MM_FP64)
          c->xchem = c->k64.xchem &&
                     c->k64.xchem(false, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, p->nx,
                                  p->ny, p->nz, 0.0, nullptr, nullptr, nullptr, nullptr, nullptr, &nb);
        else
          c->xchem = c->k32.xchem &&
                     c->k32.xchem(false, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, p->nx,
                                  p->ny, p->nz, 0.0f, nullptr, nullptr, nullptr, nullptr, nullptr, &nb);
        if (nb > kMaxPartBlocks || 4 * c->E > kStageN || 4 * c->V > kMaxBatch) c->xchem = false;
      }
    }
    // binary32: each column of the elastic z kernel held by whole wavefronts
    // (its LDS exchanges without workgroup barriers): 0.34 -> 0.32 ms per
    // launch at 256^3; binary64 keeps the interleaved columns (0.47 -> 0.56 ms
    // with wavefront columns).  Columns of 512: 8 points per thread so that
    // one wavefront holds a column, at one wave per SIMD with the excess
    // registers in AGPRs (3.63 -> 3.01 ms per launch at 512^3; at 256^3 the
    // same shape is slower, 0.32 -> 0.37 ms)
    c->zf_el_var = precision == MM_MIXED ? (p->nz == 512 ? 3 : 2) : 0;
    const char* zv = std::getenv("MM_ZF_ELASTIC_VARIANT");
    if (zv) c->zf_el_var = std::max(0, std::min(kZfElasticVariants - 1, std::atoi(zv)));
  }
  if (rc) { free_all(c); delete c; return rc; }
  c->n3 = c->fft.nreal();
  c->nc = c->fft.ncplx();
  c->esz = precision == MM_FP64 ? 8 : 4;
  const size_t n3 = c->n3, nc = c->nc, rb = c->esz, cb = 2 * c->esz;
  const int V = c->V, E = c->E;
  c->nwork = std::max({6 + V, 4 * std::min(V, 4), 4 * E, 6});
  c->nxpol = std::max(6, V);
  c->nblk_pw = blocks_for(n3);
  c->nblk_k = blocks_for(nc);
  rc = 0;
  rc = rc ? rc : dalloc(&c->phi, V * n3 * rb);
  rc = rc ? rc : dalloc(&c->cons, E * n3 * rb);
  rc = rc ? rc : dalloc(&c->gphi, 3 * V * n3 * rb);
  rc = rc ? rc : dalloc(&c->kdel, V * n3 * rb);
  rc = rc ? rc : dalloc(&c->fdiff, V * n3 * rb);
  rc = rc ? rc : dalloc(&c->flux, 3 * E * n3 * rb);
  rc = rc ? rc : dalloc(&c->strain, 6 * n3 * rb);
  rc = rc ? rc : dalloc(&c->xpol, c->nxpol * n3 * rb);
  rc = rc ? rc : dalloc(&c->el_en, n3 * rb);
  rc = rc ? rc : dalloc(&c->elgrad, V * n3 * rb);
  rc = rc ? rc : dalloc(&c->sphi, V * nc * cb);
  rc = rc ? rc : dalloc(&c->suhat, 3 * nc * cb);
  if (c->xchem) rc = rc ? rc : dalloc(&c->sgrad, (size_t)4 * V * nc * cb);
  rc = rc ? rc : dalloc(&c->swork, c->nwork * nc * cb);
  if (c->zfused) rc = rc ? rc : dalloc(&c->sstage, (size_t)kStageN * nc * cb);
  if (c->xpolar) rc = rc ? rc : dalloc(&c->sstrain, (size_t)6 * nc * cb);
  rc = rc ? rc : dalloc((void**)&c->part, (size_t)kMaxPartSlots * kMaxPartBlocks * sizeof(double));
  rc = rc ? rc : dalloc((void**)&c->dres, 64 * sizeof(double));
  rc = rc ? rc : dalloc((void**)&c->derr, 16);
  rc = rc ? rc : dalloc((void**)&c->dp, sizeof(DevParams));
  rc = rc ? rc : dalloc((void**)&c->ctl, sizeof(Ctl));
  rc = rc ? rc : dalloc((void**)&c->dstage, kStage * sizeof(double));
  if (!rc && hipHostMalloc((void**)&c->hstage, kStage * sizeof(double), hipHostMallocDefault) != hipSuccess)
    rc = MM_ERR_NOMEM;
  if (!rc && hipHostMalloc((void**)&c->hres, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) rc = MM_ERR_NOMEM;
  if (!rc && hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) rc = MM_ERR_HIP;
  if (!rc && hipEventCreateWithFlags(&c->ev, hipEventDisableTiming) != hipSuccess) rc = MM_ERR_HIP;
  {
    // off by default: measured 1% faster at 256^3 but it slows the elastic
    // kernels it runs beside (the polarization pass by 19%), and 2% slower at
    // 512^3 (its 512-thread elastic workgroups fill a CU each and lose CUs to
    // the update); MM_CH_STREAM=1 switches it on
    const char* e = std::getenv("MM_CH_STREAM");
    const bool want = e && e[0] && e[0] != '0';
    c->ch_async = c->zfused && c->E <= MM_MAX_EL && want;
  }
  if (!rc && c->ch_async) {
    rc = dalloc(&c->sch, (size_t)(4 * E + 4) * nc * cb);
    if (!rc && hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess) rc = MM_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&c->ev_head, hipEventDisableTiming) != hipSuccess) rc = MM_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&c->ev_ch, hipEventDisableTiming) != hipSuccess) rc = MM_ERR_HIP;
  }
  if (!rc && hipMemcpy(c->dp, &c->hp, sizeof(DevParams), hipMemcpyHostToDevice) != hipSuccess) rc = MM_ERR_HIP;
  if (!rc && c->autotune) rc = c->fft.autotune(c->s);
  if (!rc) {
    // zero state so that a step without upload is well defined
    void* z[] = {c->phi, c->cons, c->strain, c->el_en, c->elgrad, c->suhat, c->sstrain};
    size_t zb[] = {V * n3 * rb, E * n3 * rb, 6 * n3 * rb, n3 * rb, V * n3 * rb, 3 * nc * cb, 6 * nc * cb};
    for (int i = 0; i < 7 && !rc; ++i)
      if (z[i] && hipMemset(z[i], 0, zb[i]) != hipSuccess) rc = MM_ERR_HIP;
  }
  if (rc) { free_all(c); delete c; return rc; }
  *out = c;
  return 0;
}

extern "C" void mm_destroy(mm_ctx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->s);
  free_all(c);
  delete c;
}

extern "C" int mm_upload(mm_ctx* c, const double* phi, const double* cons, const double* el_en, const double* u,
                         const double* gradu) {
  if (c && phi) c->step_ran = false;  // phi_hat is stale
  if (!c) return MM_ERR_ARG;
  return c->prec == MM_FP64 ? upload<double>(c, phi, cons, el_en, u, gradu) : upload<float>(c, phi, cons, el_en, u, gradu);
}

extern "C" int mm_download(mm_ctx* c, double* phi, double* cons, double* el_en, double* u, double* gradu) {
  if (!c) return MM_ERR_ARG;
  return c->prec == MM_FP64 ? download<double>(c, phi, cons, el_en, u, gradu)
                            : download<float>(c, phi, cons, el_en, u, gradu);
}

extern "C" int mm_step(mm_ctx* c, int32_t nsteps, mm_diag* diag) {
  if (!c || nsteps < 0) return MM_ERR_ARG;
  for (int i = 0; i < nsteps; ++i) {
    int rc = c->prec == MM_FP64 ? run_step<double>(c, diag ? diag + i : nullptr)
                                : run_step<float>(c, diag ? diag + i : nullptr);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int mm_branch_bytes(const mm_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return MM_ERR_ARG;
  *bytes = (int64_t)c->V * (int64_t)c->n3;
  return 0;
}

extern "C" int mm_branch_tape(mm_ctx* c, int32_t mode, void* dev_buf, int32_t nsteps) {
  if (!c || mode < 0 || mode > 3) return MM_ERR_ARG;
  if (mode && (!dev_buf || nsteps <= 0)) return MM_ERR_ARG;
  if (mode && !c->dflips) MM_TRY(dalloc((void**)&c->dflips, sizeof(unsigned long long)));
  if (c->dflips) MM_HIP(hipMemsetAsync(c->dflips, 0, sizeof(unsigned long long), c->s));
  MM_HIP(hipStreamSynchronize(c->s));
  c->br = BranchTape{(uint8_t*)dev_buf, c->dflips, mode};
  c->br_cap = mode ? nsteps : 0;
  c->br_cursor = 0;
  return 0;
}

extern "C" int mm_branch_flips(mm_ctx* c, int64_t* flips) {
  if (!c || !flips) return MM_ERR_ARG;
  unsigned long long h = 0;
  if (c->dflips) {
    MM_HIP(hipMemcpyAsync(&h, c->dflips, sizeof(h), hipMemcpyDeviceToHost, c->s));
    MM_HIP(hipStreamSynchronize(c->s));
  }
  *flips = (int64_t)h;
  return 0;
}

extern "C" int mm_get_step(const mm_ctx* c) { return c ? c->step : 0; }
extern "C" int mm_set_step(mm_ctx* c, int32_t step) {
  if (!c) return MM_ERR_ARG;
  c->step = step;
  return 0;
}

extern "C" int mm_set_noise(mm_ctx* c, double sd, int32_t idum, const int32_t* state) {
  if (!c || c->prm.b_mob == 0.0) return MM_ERR_ARG;
  const size_t bytes = (size_t)c->V * c->n3 * c->esz;
  if (!c->noise_dev) MM_TRY(dalloc(&c->noise_dev, bytes));
  if (!c->noise_pin && hipHostMalloc(&c->noise_pin, bytes, hipHostMallocDefault) != hipSuccess) {
    c->noise_pin = nullptr;
    return MM_ERR_NOMEM;
  }
  c->noise_acc.assign((size_t)c->V * c->n3, 0.0);
  c->rng_idum = idum;
  if (state) std::memcpy(c->rng_state, state, sizeof(c->rng_state));
  else std::memset(c->rng_state, 0, sizeof(c->rng_state));
  c->noise_sd = sd;
  c->noise_on = true;
  c->noise_live = false;
  return 0;
}

extern "C" int mm_get_noise_state(const mm_ctx* c, int32_t* idum, int32_t* state) {
  if (!c || !idum || !state) return MM_ERR_ARG;
  *idum = c->rng_idum;
  std::memcpy(state, c->rng_state, sizeof(c->rng_state));
  return 0;
}

extern "C" int mm_field(mm_ctx* c, int32_t which, void** dev, int32_t* esz, int32_t* nf) {
  if (!c || !dev) return MM_ERR_ARG;
  void* p = nullptr;
  int n = 0;
  switch (which) {
    case 0: p = c->phi; n = c->V; break;
    case 1: p = c->cons; n = c->E; break;
    case 2: p = c->el_en; n = 1; break;
    case 3:
      if (c->fd_split) {
        // the last step passed f_diff + el_grad on as spectra: form it now
        const size_t nn = (size_t)c->V * c->n3;
        if (c->prec == MM_FP64)
          hipLaunchKernelGGL(k_add_to<double>, dim3(blocks_for(nn)), dim3(kBlock), 0, c->s, (double*)c->fdiff,
                             (const double*)c->elgrad, nn);
        else
          hipLaunchKernelGGL(k_add_to<float>, dim3(blocks_for(nn)), dim3(kBlock), 0, c->s, (float*)c->fdiff,
                             (const float*)c->elgrad, nn);
        MM_HIP(hipGetLastError());
        MM_HIP(hipStreamSynchronize(c->s));
        c->fd_split = false;
      }
      p = c->fdiff;
      n = c->V;
      break;
    case 4:
      // with the fused x path the strain lives as x half-spectra: form the
      // real fields now (valid until the next step)
      MM_TRY(strain_to_real_any(c));
      MM_HIP(hipStreamSynchronize(c->s));
      p = c->strain;
      n = 6;
      break;
    case 5: p = c->gphi; n = 3 * c->V; break;
    case 6: p = c->kdel; n = c->V; break;
    case 7: p = c->elgrad; n = c->V; break;
    default: return MM_ERR_ARG;
  }
  if ((which == 5 || which == 6) && c->xchem && c->step_ran) {
    // the fused chemistry pass does not form them: recompute from phi_hat
    set_gate(c, nullptr);
    MM_TRY(c->prec == MM_FP64 ? grad_fields<double>(c) : grad_fields<float>(c));
  }
  *dev = p;
  if (esz) *esz = (int32_t)c->esz;
  if (nf) *nf = n;
  return 0;
}

extern "C" int mm_field_to_host(mm_ctx* c, int32_t which, double* host) {
  void* p;
  int32_t esz, n;
  MM_TRY(mm_field(c, which, &p, &esz, &n));
  if (!host) return MM_ERR_ARG;
  MM_HIP(hipStreamSynchronize(c->s));
  std::vector<double> h((size_t)n * c->n3);
  if (c->prec == MM_FP64) { MM_TRY(get_field<double>(c, p, h)); } else { MM_TRY(get_field<float>(c, p, h)); }
  if (which == 1)
    for (int e = 0; e < c->E; ++e)
      for (size_t i = 0; i < c->n3; ++i) h[(size_t)e * c->n3 + i] += c->hp.cbar[e];
  std::memcpy(host, h.data(), h.size() * 8);
  return 0;
}

extern "C" int mm_probe(mm_ctx* c, int32_t kind) {
  if (!c || kind < -1 || kind > MM_PROBE_X_POLAR) return MM_ERR_ARG;
  MM_HIP(hipStreamSynchronize(c->s));
  MM_TRY(c->probe.collect());
  c->probe.reset();
  c->probe.kind = kind;
  return 0;
}

extern "C" int mm_probe_read(mm_ctx* c, double* ms, int64_t* launches, double* bytes) {
  if (!c) return MM_ERR_ARG;
  MM_HIP(hipStreamSynchronize(c->s));
  MM_TRY(c->probe.collect());
  if (ms) *ms = c->probe.ms;
  if (launches) *launches = c->probe.launches;
  if (bytes) *bytes = c->probe.bytes;
  const int kind = c->probe.kind;
  c->probe.reset();
  c->probe.kind = kind;
  return 0;
}

extern "C" int mm_get_tuning(const mm_ctx* c, int32_t* variant, float* ms) {
  if (!c) return MM_ERR_ARG;
  for (int k = 0; k < 5; ++k) {
    if (variant) variant[k] = c->fft.var[k];
    for (int v = 0; v < 3; ++v)
      if (ms) ms[k * 3 + v] = c->fft.tune_ms[k][v];
  }
  return 0;
}

extern "C" int mm_set_graph(mm_ctx* c, int32_t on, int32_t iters) {
  if (!c || iters < 0 || iters > 64) return MM_ERR_ARG;
  if (c->gx) { (void)hipGraphExecDestroy(c->gx); c->gx = nullptr; }
  if (c->g) { (void)hipGraphDestroy(c->g); c->g = nullptr; }
  c->graph = on != 0;
  if (iters > 0) c->graph_iters = iters;
  return 0;
}

// ---------------------------------------------------------------- ensembles
// Independent members (own parameters, grid and fields) advanced together:
// the regular step of every member is recorded into one graph with one
// branch per member, launched once per step.
struct mm_batch {
  std::vector<mm_ctx*> m;
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  double* dstage = nullptr;
  double* hstage = nullptr;
  hipGraph_t g = nullptr;
  hipGraphExec_t gx = nullptr;
  int prec = MM_FP64;
};

static void batch_free(mm_batch* b) {
  if (b->s) (void)hipStreamSynchronize(b->s);
  if (b->gx) (void)hipGraphExecDestroy(b->gx);
  if (b->g) (void)hipGraphDestroy(b->g);
  for (mm_ctx* c : b->m) mm_destroy(c);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->dstage) (void)hipFree(b->dstage);
  if (b->hstage) (void)hipHostFree(b->hstage);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
}

extern "C" int mm_batch_create(int32_t n, const mm_params* params, int32_t precision, mm_batch** out) {
  if (!out || !params || n < 1 || n > 4096) return MM_ERR_ARG;
  *out = nullptr;
  mm_batch* b = new mm_batch;
  b->prec = precision;
  int rc = 0;
  for (int i = 0; i < n && !rc; ++i) {
    mm_ctx* c = nullptr;
    rc = mm_create(params + i, precision, &c);
    if (!rc) b->m.push_back(c);
  }
  if (!rc && hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess) rc = MM_ERR_HIP;
  if (!rc && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) rc = MM_ERR_HIP;
  for (int i = 0; i < n && !rc; ++i) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { rc = MM_ERR_HIP; break; }
    b->join.push_back(e);
  }
  if (!rc && hipMalloc((void**)&b->dstage, (size_t)n * kStage * sizeof(double)) != hipSuccess) rc = MM_ERR_NOMEM;
  if (!rc && hipHostMalloc((void**)&b->hstage, (size_t)n * kStage * sizeof(double), hipHostMallocDefault) != hipSuccess)
    rc = MM_ERR_NOMEM;
  if (rc) { batch_free(b); return rc; }
  *out = b;
  return 0;
}

extern "C" void mm_batch_destroy(mm_batch* b) {
  if (b) batch_free(b);
}

extern "C" int mm_batch_member(mm_batch* b, int32_t i, mm_ctx** ctx) {
  if (!b || !ctx || i < 0 || i >= (int)b->m.size()) return MM_ERR_ARG;
  *ctx = b->m[i];
  return 0;
}

static int batch_build(mm_batch* b) {
  MM_HIP(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
  int rc = hipEventRecord(b->fork, b->s) == hipSuccess ? 0 : MM_ERR_HIP;
  for (size_t i = 0; i < b->m.size() && !rc; ++i) {
    mm_ctx* c = b->m[i];
    if (hipStreamWaitEvent(c->s, b->fork, 0) != hipSuccess) { rc = MM_ERR_HIP; break; }
    const int kind = c->probe.kind;
    c->probe.kind = -1;
    double* st = b->dstage + i * kStage;
    rc = c->prec == MM_FP64 ? record_step<double>(c, st) : record_step<float>(c, st);
    c->probe.kind = kind;
    if (!rc && hipEventRecord(b->join[i], c->s) != hipSuccess) rc = MM_ERR_HIP;
    if (!rc && hipStreamWaitEvent(b->s, b->join[i], 0) != hipSuccess) rc = MM_ERR_HIP;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->s, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  b->g = g;
  MM_HIP(hipGraphInstantiate(&b->gx, b->g, nullptr, nullptr, 0));
  return 0;
}

template <typename T> static int batch_step_one(mm_batch* b, mm_diag* d) {
  const size_t n = b->m.size();
  bool regular = true;
  for (mm_ctx* c : b->m)
    regular = regular && c->zfused && c->step > c->prm.elastep0 && c->step > c->prm.noise_step;
  if (!regular) {
    for (size_t i = 0; i < n; ++i) {
      mm_ctx* c = b->m[i];
      MM_TRY(c->prec == MM_FP64 ? run_step_eager<double>(c, d ? d + i : nullptr)
                                : run_step_eager<float>(c, d ? d + i : nullptr));
    }
    return 0;
  }
  if (!b->gx) MM_TRY(batch_build(b));
  MM_HIP(hipGraphLaunch(b->gx, b->s));
  MM_HIP(hipMemcpyAsync(b->hstage, b->dstage, n * kStage * sizeof(double), hipMemcpyDeviceToHost, b->s));
  MM_HIP(hipStreamSynchronize(b->s));
  int first = 0;
  for (size_t i = 0; i < n; ++i) {
    mm_ctx* c = b->m[i];
    const int rc = c->prec == MM_FP64 ? finish_replay<double>(c, b->hstage + i * kStage, d ? d + i : nullptr)
                                      : finish_replay<float>(c, b->hstage + i * kStage, d ? d + i : nullptr);
    if (rc && !first) first = rc;
  }
  return first;
}

extern "C" int mm_batch_step(mm_batch* b, int32_t nsteps, mm_diag* diag) {
  if (!b || nsteps < 0) return MM_ERR_ARG;
  const size_t n = b->m.size();
  for (int k = 0; k < nsteps; ++k) {
    mm_diag* d = diag ? diag + (size_t)k * n : nullptr;
    for (mm_ctx* c : b->m) MM_TRY(qcheck_once(c));
    MM_TRY(b->prec == MM_FP64 ? batch_step_one<double>(b, d) : batch_step_one<float>(b, d));
  }
  return 0;
}

#include "adjoint.h"

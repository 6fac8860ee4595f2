// step_kernels.h -- real-space and k-space kernels of the semi-implicit step.
//
// Storage type T is double (FP64) or float (MIXED); all pointwise and k-space
// arithmetic is binary64 in both modes (fields are converted on load/store),
// so MIXED differs from FP64 only through FP32 storage and FP32 FFTs.
//
// Field layout: structure of arrays, each field (Nx,Ny,Nz) x fastest; spectra
// [kz][kx][ky] (see fft_pass.h).  Every pointwise kernel is a grid-stride
// loop over a grid that depends only on the problem size, and every global
// sum is a fixed-order block reduction followed by a fixed-order finalize,
// so results are bitwise reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "../../include/meumapps_ss.h"

namespace mm {

// Parameters of the step, read with wave-uniform (scalar) loads.
struct DevParams {
  int nx, ny, nz, H, V, E, NPH, zdim;
  int phase_of[MM_MAX_VAR];   // 1-based phase of each variant
  int tsum_ok[MM_MAX_VAR];    // sum(trans_ppt(:,v)) > 1e-9 (ModElastic.f:534,653)
  double unit[3];             // fl(2 pi / N) per axis (ModInitials.f:355-360)
  double dt, bmob;
  double cmobs[MM_MAX_EL];
  double gc[MM_MAX_VAR * 9];  // grad_coeff(i,j,v) at [v*9 + j*3 + i]
  double tp[MM_MAX_VAR * 6];  // trans_ppt(m,v)
  double sig0[MM_MAX_VAR * 6];  // sum_m CEla(ij,m,ph(v)) w_m trans_ppt(m,v)
  double cela[MM_MAX_PH * 36];  // column-major per phase
  double sela[MM_MAX_PH * 36];
  double gp[MM_MAX_PH * MM_GP_STRIDE];
  double well[MM_MAX_PH];
  double e_mean[6];
  double zc[16 * 16];  // constant rows of Zintf, row-major [i][j], 0-based
  double zrc[16];      // constant entries of Zr
  double cbar[MM_MAX_EL];  // MIXED: mean composition held in FP64 (0 in FP64 mode)
  int homog;               // every phase has the same CEla and SEla (bitwise)
};

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// State of the elastic fixed-point iteration of one step (ModElastic.f:100-103,
// 435-481).  The same update runs on the host (eager steps) and in a one-thread
// kernel (graph-recorded steps), so both paths take identical decisions.
struct Ctl {
  double cfac, newfac;
  double err_old[6];
  int niter, total, converged;
  int active;  // another iteration follows
  int ran;     // the iteration just finished ran (gates its inverse transforms)
  int done;    // the loop has ended (gates the rest of the step)
  int err;     // MM_ERR_NEWFAC when backtracking gave up
  int pad;
};

__host__ __device__ inline void ctl_start(Ctl& c) {
  c.cfac = 1.0;
  c.newfac = 0.5;
  for (int m = 0; m < 6; ++m) c.err_old[m] = 0.0;
  c.niter = 0;
  c.total = 0;
  c.converged = 0;
  c.active = 1;
  c.ran = 0;
  c.done = 0;
  c.err = 0;
  c.pad = 0;
}

// After one iteration with the six sums err = {sum (u1-u1_old)^2, sum u1_old^2, ...}.
__host__ __device__ inline void ctl_after_iter(Ctl& c, const double* err) {
  c.ran = c.active;
  if (!c.active) return;
  c.niter += 1;
  c.total += 1;
  if (err[0] / err[1] < 1.e-7 && err[2] / err[3] < 1.e-7 && err[4] / err[5] < 1.e-7) c.converged = 1;
  if (c.niter > 1 && !c.converged) {
    if (err[0] / err[1] > c.err_old[0] / c.err_old[1] || err[2] / err[3] > c.err_old[2] / c.err_old[3] ||
        err[4] / err[5] > c.err_old[4] / c.err_old[5]) {
      c.cfac = c.cfac * c.newfac;
      c.newfac = c.newfac * 0.5;
      if (c.newfac < 1.e-7) {
        c.err = MM_ERR_NEWFAC;
        c.active = 0;
        c.done = 1;
        return;
      }
      c.niter = 0;
      c.converged = 0;
      return;
    }
  }
  for (int m = 0; m < 6; ++m) c.err_old[m] = err[m];
  if (c.niter == 10) c.converged = 1;
  if (!c.converged || c.niter < 4) return;
  if (c.cfac < 1.0 - 1.e-6) {
    c.cfac = c.cfac + 0.1;
    if (c.cfac > 1.0) c.cfac = 1.0;
    c.niter = 0;
    c.converged = 0;
    return;
  }
  c.active = 0;
  c.done = 1;
}

template <typename T> __device__ __forceinline__ double ld(const T* p, size_t i) { return (double)p[i]; }
template <typename T> __device__ __forceinline__ void st(T* p, size_t i, double v) { p[i] = (T)v; }

// ------------------------------------------------------------ reductions
template <int OP> __device__ __forceinline__ double rop(double a, double b) {
  if constexpr (OP == 0) return a + b;
  else if constexpr (OP == 1) return b < a ? b : a;
  else return b > a ? b : a;
}

// Block-wide reduction in a fixed order over NW waves; every thread gets the result.
template <int OP, int NW = kWaves> __device__ __forceinline__ double block_reduce(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = rop<OP>(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = rop<OP>(r, sh[i]);
  __syncthreads();
  return r;
}

// part[k*nblk + blockIdx.x] = block result of slot k
template <int OP, int NW = kWaves>
__device__ __forceinline__ void block_store(double v, double* part, int k, double* sh) {
  double r = block_reduce<OP, NW>(v, sh);
  if (threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = r;
}

// reduction op per slot of k_finalize (out[k] = reduce_op[k] over
// part[k*nblk + 0..nblk)): 0 sum, 1 min, 2 max
struct FinalizeOps {
  int op[8];
};

// ---------------------------------------------------------- k-space helpers
// kf = i*a with a = m*fl(2 pi/N), m wrapped to (-N/2, N/2]; kf zeroed at the
// Nyquist index; kf_sq = -(a*a) kept there (ModInitials.f:367-401).
__device__ __forceinline__ void kaxis(int m, int n, double unit, double& kf, double& kq) {
  const int mw = (m <= n / 2) ? m : m - n;
  const double a = (double)mw * unit;
  kf = (2 * m == n) ? 0.0 : a;
  kq = -(a * a);
}

struct KVec {
  double kf[3], kq[3];
  int kx, ky, kz;
};

__device__ __forceinline__ KVec kvec3(const DevParams* __restrict__ P, int kx, int ky, int kz) {
  KVec k;
  k.kx = kx;
  k.ky = ky;
  k.kz = kz;
  kaxis(kx, P->nx, P->unit[0], k.kf[0], k.kq[0]);
  kaxis(ky, P->ny, P->unit[1], k.kf[1], k.kq[1]);
  kaxis(kz, P->nz, P->unit[2], k.kf[2], k.kq[2]);
  return k;
}

// spectrum index i -> wave vector; layout [kz][kx][ky]
__device__ __forceinline__ KVec kvec(const DevParams* __restrict__ P, size_t i) {
  const int ky = (int)(i % (size_t)P->ny);
  const size_t r = i / (size_t)P->ny;
  return kvec3(P, (int)(r % (size_t)P->H), ky, (int)(r / (size_t)P->H));
}

// sum G_ij k_i k_j with kf_sq on the diagonal (Multi_MEUMAPPS.f:596-600)
__device__ __forceinline__ double kgk(const DevParams* __restrict__ P, int v, const KVec& k) {
  const double* g = P->gc + v * 9;  // g[j*3+i] = grad_coeff(i,j)
  const double k1k2 = -(k.kf[0] * k.kf[1]), k2k3 = -(k.kf[1] * k.kf[2]), k3k1 = -(k.kf[2] * k.kf[0]);
  return g[0] * k.kq[0] + g[4] * k.kq[1] + g[8] * k.kq[2] + (g[3] + g[1]) * k1k2 +
         (g[7] + g[5]) * k2k3 + (g[2] + g[6]) * k3k1;
}

// Solve Q(k) u = r for the acoustic tensor Q_il = sum_jk CEla(ij,kl,2) kpr_jk,
// kpr_jk = kf_j kf_k (kf_sq on j = k), ModElastic.f:119-191, by Cramer's rule.
struct Cplx3 {
  double re[3], im[3];
};
__device__ __forceinline__ void acoustic_tensor(const DevParams* __restrict__ P, const KVec& k,
                                                double (&Q)[3][3]) {
  const int vm[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
  const double* C = P->cela + 36;  // phase 2, column-major: C(a,b) = C[b*6+a]
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
          const double kpr = (kk == j) ? k.kq[kk] : -(k.kf[kk] * k.kf[j]);
          s += C[vm[kk][l] * 6 + vm[i][j]] * kpr;
        }
      Q[i][l] = s;
    }
}

__device__ __forceinline__ void acoustic_solve(const DevParams* __restrict__ P, const KVec& k,
                                               Cplx3& u) {
  double Q[3][3];
  acoustic_tensor(P, k, Q);
  const double c00 = Q[1][1] * Q[2][2] - Q[1][2] * Q[2][1];
  const double c01 = Q[1][2] * Q[2][0] - Q[1][0] * Q[2][2];
  const double c02 = Q[1][0] * Q[2][1] - Q[1][1] * Q[2][0];
  const double det = Q[0][0] * c00 + Q[0][1] * c01 + Q[0][2] * c02;
  const double id = 1.0 / det;
  // inverse = adj(Q)/det, adj(Q)[i][j] = cofactor(j,i)
  double A[3][3];
  A[0][0] = c00 * id;
  A[1][0] = c01 * id;
  A[2][0] = c02 * id;
  A[0][1] = (Q[0][2] * Q[2][1] - Q[0][1] * Q[2][2]) * id;
  A[1][1] = (Q[0][0] * Q[2][2] - Q[0][2] * Q[2][0]) * id;
  A[2][1] = (Q[0][1] * Q[2][0] - Q[0][0] * Q[2][1]) * id;
  A[0][2] = (Q[0][1] * Q[1][2] - Q[0][2] * Q[1][1]) * id;
  A[1][2] = (Q[0][2] * Q[1][0] - Q[0][0] * Q[1][2]) * id;
  A[2][2] = (Q[0][0] * Q[1][1] - Q[0][1] * Q[1][0]) * id;
  Cplx3 r = u;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    u.re[i] = A[i][0] * r.re[0] + A[i][1] * r.re[1] + A[i][2] * r.re[2];
    u.im[i] = A[i][0] * r.im[0] + A[i][1] * r.im[1] + A[i][2] * r.im[2];
  }
}

// engineering-strain spectra from u: d_d u_i = i kf_d u_i
template <typename T>
__device__ __forceinline__ void store_strain(C2<T>* const* out, size_t i, const KVec& k, const Cplx3& u) {
  // e11, e22, e33, g23, g13, g12
  const int A[6] = {0, 1, 2, 1, 0, 0}, Bi[6] = {0, 1, 2, 2, 2, 1};  // term 1: d_A u_Bi
  const int A2[6] = {-1, -1, -1, 2, 2, 1}, B2[6] = {-1, -1, -1, 1, 0, 0};
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    // i*a*(x + i y) = (-a y, a x)
    double re = -k.kf[A[m]] * u.im[Bi[m]], im = k.kf[A[m]] * u.re[Bi[m]];
    if (A2[m] >= 0) {
      re += -k.kf[A2[m]] * u.im[B2[m]];
      im += k.kf[A2[m]] * u.re[B2[m]];
    }
    out[m][i] = mk<T>((T)re, (T)im);
  }
}

}  // namespace mm

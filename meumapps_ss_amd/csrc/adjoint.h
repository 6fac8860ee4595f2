// adjoint.h -- reverse-mode derivative of the regular step (FP64, homogeneous
// moduli, step > elastep0).  Included once, at the end of step.hip.
//
// The step is a fixed composition of pointwise maps and Fourier multipliers
// (Multi_MEUMAPPS.f:575-1043, ModElastic.f:241-691):
//
//   g = grad phi, kd = div G grad phi           (gradient block)
//   fd0, F = chem(phi, c, g, kd)                 (KKS chemistry + f_diff correction)
//   eps_k = L_el polar(phi, eps_{k-1}; cfac_k)   (k = 1..n elastic iterations)
//   fd = fd0 + elgrad(phi, eps_n)                (elener)
//   c'  = A_c c + dtm A_c div F                  (Cahn-Hilliard)
//   phi' = A_v phi - dtl A_v fd                  (Allen-Cahn)
//   phi_new = clamp(phi', 0, 1)
//
// with state (phi, c, eps) -> (phi_new, c', eps_n).  The iteration count and
// the cfac sequence of the elastic control are piecewise constant and are
// replayed from the tape; so are the branches of the f_diff correction and of
// the clamp.  A_v, A_c, B_v (kd), L_el are real, even Fourier multipliers
// (L_el = -K Q^-1 K^T per wave vector is symmetric), hence self-adjoint;
// d/dx_d (i kf_d, zero at Nyquist) is antisymmetric.  Parameter derivatives
// come from Parseval sums over the half spectrum (weight 1 on the kx = 0 and
// kx = Nx/2 planes, 2 elsewhere) and real-space dot products, accumulated in
// a fixed order.

namespace mm {

// ------------------------------------------------------------- accumulation
// out[idx[j]] += scale[j] * sum_b part[slot[j]*nblk + b], fixed order
struct AccumOps {
  int n;
  int slot[64];
  int idx[64];
  double scale[64];
};
__global__ void __launch_bounds__(kBlock) k_accum(const double* __restrict__ part, int nblk, AccumOps ops,
                                                  double* __restrict__ out) {
  __shared__ double sh[kWaves];
  for (int j = 0; j < ops.n; ++j) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) s += part[(size_t)ops.slot[j] * nblk + b];
    s = block_reduce<0>(s, sh);
    if (threadIdx.x == 0) out[ops.idx[j]] += ops.scale[j] * s;
  }
}

// half-spectrum Parseval weight of the wave vector
__device__ __forceinline__ double pweight(const DevParams* __restrict__ P, const KVec& k) {
  return (k.kx == 0 || 2 * k.kx == P->nx) ? 1.0 : 2.0;
}
// the six quadratic forms kk_j with kgk = sum_j Gsym_j kk_j:
// (11) kq0, (22) kq1, (33) kq2, (12) -kf0 kf1, (23) -kf1 kf2, (31) -kf2 kf0
__device__ __forceinline__ void kforms(const KVec& k, double (&kk)[6]) {
  kk[0] = k.kq[0];
  kk[1] = k.kq[1];
  kk[2] = k.kq[2];
  kk[3] = -(k.kf[0] * k.kf[1]);
  kk[4] = -(k.kf[1] * k.kf[2]);
  kk[5] = -(k.kf[2] * k.kf[0]);
}
__device__ __forceinline__ double pfun2(double f) {  // p''(phi)
  return 60.0 * f * (2.0 * f * f - 3.0 * f + 1.0);
}

// ------------------------------------------------------------------ clamp
// lam_pre = lam_new where 0 <= phi' <= 1 (the unclamped branch), else 0
__global__ void __launch_bounds__(kBlock) k_adj_mask(const double* __restrict__ phipre, double* __restrict__ lam,
                                                     size_t n) {
  GRID_STRIDE(i, n) {
    const double f = phipre[i];
    if (f < 0.0 || f > 1.0) lam[i] = 0.0;
  }
}

// lam += a * x
__global__ void __launch_bounds__(kBlock) k_axpy(double* __restrict__ lam, const double* __restrict__ x, double a,
                                                 size_t n) {
  GRID_STRIDE(i, n) lam[i] += a * x[i];
}

// ------------------------------------------------------------- Allen-Cahn
// S0[v] = D_v lam_hat (in place), D_v = 1/(1 - dtl kgk_v).  Sums per v:
// slot v: sum w Re(conj(S0) kgk_v S1), slot V + 6v + j: sum w Re(conj(S0) kk_j S1)
// with S1 = spectrum of phi' (pre-clamp).  blockIdx.y = v.
__global__ void __launch_bounds__(kBlock) k_adj_ac_ks(const DevParams* __restrict__ P, C2<double>* __restrict__ S0,
                                                      const C2<double>* __restrict__ S1, double* __restrict__ part,
                                                      int V, size_t nc) {
  __shared__ double sh[kWaves];
  const int v = blockIdx.y;
  const double dtl = P->dt * P->bmob;
  double sK = 0.0, sj[6] = {0, 0, 0, 0, 0, 0};
  C2<double>* a = S0 + (size_t)v * nc;
  const C2<double>* b = S1 + (size_t)v * nc;
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    const double K = kgk(P, v, k);
    const double D = 1.0 / (1.0 - dtl * K);
    const C2<double> l = a[i];
    const C2<double> l2 = mk<double>(l.x * D, l.y * D);
    a[i] = l2;
    const C2<double> p = b[i];
    const double w = pweight(P, k);
    const double re = l2.x * p.x + l2.y * p.y;  // Re(conj(l2) p)
    sK += w * K * re;
    double kk[6];
    kforms(k, kk);
#pragma unroll
    for (int j = 0; j < 6; ++j) sj[j] += w * kk[j] * re;
  }
  // partial layout: slot s at part[(s*gridDim.y + ...)]: one slot row per (v, quantity)
  const int nb = gridDim.x;
  double r = block_reduce<0>(sK, sh);
  if (threadIdx.x == 0) part[(size_t)v * nb + blockIdx.x] = r;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    r = block_reduce<0>(sj[j], sh);
    if (threadIdx.x == 0) part[(size_t)(V + 6 * v + j) * nb + blockIdx.x] = r;
  }
}

// lam_fd = -dtl lamt; slot v: sum lamt * fd
__global__ void __launch_bounds__(kBlock) k_adj_ac_pw(const DevParams* __restrict__ P, const double* __restrict__ lamt,
                                                      const double* __restrict__ fd, double* __restrict__ lfd,
                                                      double* __restrict__ part, int V, size_t n3) {
  __shared__ double sh[kWaves];
  const double dtl = P->dt * P->bmob;
  for (int v = 0; v < V; ++v) {
    double s = 0.0;
    GRID_STRIDE(pt, n3) {
      const size_t i = (size_t)v * n3 + pt;
      s += lamt[i] * fd[i];
      lfd[i] = -dtl * lamt[i];
    }
    s = block_reduce<0>(s, sh);
    if (threadIdx.x == 0) part[(size_t)v * gridDim.x + blockIdx.x] = s;
  }
}

// ---------------------------------------------------------- Cahn-Hilliard
// component e: S[0] = D_c lam_hat (in place), S[1+d] = i kf_d S[0];
// slot 0: sum w Re(conj(S[0]) ksq Sc), Sc = spectrum of c'
__global__ void __launch_bounds__(kBlock) k_adj_ch_ks(const DevParams* __restrict__ P, C2<double>* __restrict__ S,
                                                      const C2<double>* __restrict__ Sc, double* __restrict__ part,
                                                      int e, size_t nc) {
  __shared__ double sh[kWaves];
  const double dtm = P->dt * P->cmobs[e];
  double s = 0.0;
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    const double ksq = k.kq[0] + k.kq[1] + k.kq[2];
    const double D = 1.0 / (1.0 - dtm * ksq);
    const C2<double> l = S[i];
    const C2<double> l2 = mk<double>(l.x * D, l.y * D);
    S[i] = l2;
#pragma unroll
    for (int d = 0; d < 3; ++d) S[(size_t)(1 + d) * nc + i] = mk<double>(-l2.y * k.kf[d], l2.x * k.kf[d]);
    const C2<double> c = Sc[i];
    s += pweight(P, k) * ksq * (l2.x * c.x + l2.y * c.y);
  }
  s = block_reduce<0>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// lam_F_d = -dtm G_d (in place); slot 0: -sum_d G_d F_d
__global__ void __launch_bounds__(kBlock) k_adj_ch_pw(const DevParams* __restrict__ P, double* __restrict__ G,
                                                      const double* __restrict__ F, double* __restrict__ part, int e,
                                                      size_t n3) {
  __shared__ double sh[kWaves];
  const double dtm = P->dt * P->cmobs[e];
  double s = 0.0;
  GRID_STRIDE(pt, n3) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const size_t i = (size_t)d * n3 + pt;
      s -= G[i] * F[i];
      G[i] = -dtm * G[i];
    }
  }
  s = block_reduce<0>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------------------------------------------------------- elener
// homogeneous moduli: elgrad_v = -p'(phi_v) sum_m w_m sig_m tp_vm (where
// sum tp_v > 1e-9), sig = C (eps + w e_mean - w te), te = sum_v p_v tp_v.
// leps_out = leps_in + d/d strain; lphi += d/d phi; slot 6v+m: d/d tp_vm
template <int V>
__global__ void __launch_bounds__(kBlock)
k_adj_elener(const DevParams* __restrict__ P, const double* __restrict__ phi, const double* __restrict__ strain,
             const double* __restrict__ lfd, const double* __restrict__ leps_in, double* __restrict__ leps_out,
             double* __restrict__ lphi, double* __restrict__ part, size_t n3) {
  __shared__ double sh[kWaves];
  double stp[V][6];
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int m = 0; m < 6; ++m) stp[v][m] = 0.0;
  GRID_STRIDE(pt, n3) {
    double f[V], Pv[V], pp[V], te[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < V; ++v) {
      f[v] = phi[v * n3 + pt];
      Pv[v] = pfun(f[v]);
      pp[v] = ppr(f[v]);
#pragma unroll
      for (int m = 0; m < 6; ++m) te[m] = te[m] + Pv[v] * P->tp[v * 6 + m];
    }
    double eps[6], sig[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double w = m < 3 ? 1.0 : 2.0;
      eps[m] = strain[m * n3 + pt] + w * P->e_mean[m] - w * te[m];
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) s += P->cela[c * 6 + a] * eps[c];
      sig[a] = s;
    }
    double lsig[6] = {0, 0, 0, 0, 0, 0}, lpv[V], lpp[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      lpv[v] = 0.0;
      lpp[v] = 0.0;
      if (!P->tsum_ok[v]) continue;
      const double leg = lfd[v * n3 + pt];
      const double* t = P->tp + v * 6;
      const double d = sig[0] * t[0] + sig[1] * t[1] + sig[2] * t[2] +
                       2.0 * (sig[3] * t[3] + sig[4] * t[4] + sig[5] * t[5]);
      lpp[v] = -leg * d;
      const double ld_ = -leg * pp[v];
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const double w = m < 3 ? 1.0 : 2.0;
        lsig[m] += ld_ * w * t[m];
        stp[v][m] += ld_ * w * sig[m];
      }
    }
    double lte[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 6; ++a) s += P->cela[c * 6 + a] * lsig[a];
      const double w = c < 3 ? 1.0 : 2.0;
      leps_out[c * n3 + pt] = leps_in[c * n3 + pt] + s;
      lte[c] = -w * s;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        lpv[v] += lte[m] * P->tp[v * 6 + m];
        stp[v][m] += lte[m] * Pv[v];
      }
      lphi[v * n3 + pt] += lpv[v] * pp[v] + lpp[v] * pfun2(f[v]);
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int m = 0; m < 6; ++m) block_store<0>(stp[v][m], part, 6 * v + m, sh);
}

// --------------------------------------------------------------- elastic
// x -> eps = L_el x per wave vector (the iteration's k-space map without the
// displacement record and sums): R = i K^T x, u = Q^-1 R (u(0) = 0), eps = i K u
__global__ void __launch_bounds__(kBlock) k_adj_el_ks(const DevParams* __restrict__ P, C2<double>* __restrict__ S,
                                                      size_t nc) {
  const int vm[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
  C2<double>* outs[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) outs[m] = S + (size_t)m * nc;
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    double xr[6], xi[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const C2<double> x = outs[m][i];
      xr[m] = x.x;
      xi[m] = x.y;
    }
    Cplx3 u;
    if (k.kx == 0 && k.ky == 0 && k.kz == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) { u.re[c] = 0.0; u.im[c] = 0.0; }
    } else {
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
    store_strain<double>(outs, i, k, u);
  }
}

// polarization stress (homogeneous moduli), transposed: from lx = d/dx of
// iteration k, write leps = d/d eps_{k-1}; lphi += d/d phi; slot 6v+m: d/d tp
template <int V>
__global__ void __launch_bounds__(kBlock)
k_adj_polar(const DevParams* __restrict__ P, const double* __restrict__ phi, const double* __restrict__ strain,
            const double* __restrict__ lx, double cfac, double* __restrict__ leps, double* __restrict__ lphi,
            double* __restrict__ part, size_t n3) {
  __shared__ double sh[kWaves];
  double stp[V][6];
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int m = 0; m < 6; ++m) stp[v][m] = 0.0;
  GRID_STRIDE(pt, n3) {
    double f[V], Pv[V], ps = 0.0;
#pragma unroll
    for (int v = 0; v < V; ++v) { f[v] = phi[v * n3 + pt]; Pv[v] = pfun(f[v]); ps = ps + Pv[v]; }
    const double sc = ps < 1.0 ? 1.0 : 1.0 / ps;
    const double cb = cfac * (1.0 - sc);
    double lr[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 6; ++m) s += P->cela[n * 6 + m] * lx[m * n3 + pt];
      lr[n] = s;
    }
    double lcb = 0.0, lsc = 0.0, la[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double w = m < 3 ? 1.0 : 2.0;
      const double bv = strain[m * n3 + pt] + w * P->e_mean[m];
      double s = 0.0;
#pragma unroll
      for (int v = 0; v < V; ++v) s += Pv[v] * P->tp[v * 6 + m];
      const double a = w * s;
      leps[m * n3 + pt] = cb * lr[m];
      lcb += lr[m] * bv;
      lsc += lr[m] * a;
      la[m] = sc * lr[m];
    }
    lsc += -cfac * lcb;
    const double lps = ps < 1.0 ? 0.0 : -lsc / (ps * ps);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double lp = lps;
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const double w = m < 3 ? 1.0 : 2.0;
        lp += la[m] * w * P->tp[v * 6 + m];
        stp[v][m] += la[m] * w * Pv[v];
      }
      lphi[v * n3 + pt] += lp * ppr(f[v]);
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int m = 0; m < 6; ++m) block_store<0>(stp[v][m], part, 6 * v + m, sh);
}

// -------------------------------------------------------------- chemistry
// Transposed Jacobian of k_chem at (phi, c, g, kd): from lfd (= d/d f_diff)
// and lF (d/d flux) -> lphi +=, lc +=, lg = d/d grad phi, lkd = d/d kd;
// slot q-1: d/d well(q), q = 1..NPH.
template <int V, int NPH, int E>
__global__ void __launch_bounds__(kBlock)
k_adj_chem(const DevParams* __restrict__ P, const double* __restrict__ phi, const double* __restrict__ cons,
           const double* __restrict__ gphi, const double* __restrict__ kdel, const double* __restrict__ lfd,
           const double* __restrict__ lF, double* __restrict__ lphi, double* __restrict__ lc,
           double* __restrict__ lg, double* __restrict__ lkd, double* __restrict__ part, int* __restrict__ err,
           size_t n3) {
  constexpr int Z = NPH * E;
  __shared__ double sh[kWaves];
  double swell[NPH];
#pragma unroll
  for (int q = 0; q < NPH; ++q) swell[q] = 0.0;
  bool bad = false;
  const double* gp = P->gp;
#define GP(j, ph) gp[((ph) - 1) * MM_GP_STRIDE + ((j) - 1)]
  GRID_STRIDE(pt, n3) {
    double f[V], g[V][3], c[E];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      f[v] = phi[v * n3 + pt];
#pragma unroll
      for (int d = 0; d < 3; ++d) g[v][d] = gphi[(3 * v + d) * n3 + pt];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) c[e] = cons[e * n3 + pt];
    // ---- forward recompute
    double Ppr[V], Pph[NPH + 1];
#pragma unroll
    for (int q = 0; q <= NPH; ++q) Pph[q] = 0.0;
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
        }
    Pph[1] = 1.0 - Pph[1];
    double A[Z][Z], x[Z];
#pragma unroll
    for (int i = 0; i < Z; ++i) {
#pragma unroll
      for (int j = 0; j < Z; ++j) A[i][j] = P->zc[i * 16 + j];
      x[i] = P->zrc[i];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
#pragma unroll
      for (int ph = 0; ph < NPH; ++ph) A[Z - E + e][NPH * e + ph] = Pph[ph + 1];
      x[Z - E + e] = c[e];
    }
    double At[Z][Z];
#pragma unroll
    for (int i = 0; i < Z; ++i)
#pragma unroll
      for (int j = 0; j < Z; ++j) At[i][j] = A[j][i];
    if (!solve_pivot<Z>(A, x)) bad = true;
    // X(e, ph) = x[NPH*e + ph - 1]
    double Gf[NPH + 1], Mu[E];
#pragma unroll
    for (int ph = 1; ph <= NPH; ++ph) {
      double s = 0.0;
#pragma unroll
      for (int e = 1; e <= E; ++e) {
        const double dx = x[NPH * (e - 1) + ph - 1] - GP(e + E + 1, ph);
        s += GP(e, ph) * (dx * dx);
      }
      Gf[ph] = s + GP(E + 1, ph);
    }
#pragma unroll
    for (int e = 1; e <= E; ++e) Mu[e - 1] = 2.0 * GP(e, 1) * (x[NPH * (e - 1)] - GP(e + E + 1, 1));
    // ---- reverse
    double lfg[V], lf[V], lPpr[V];
#pragma unroll
    for (int v = 0; v < V; ++v) { lfg[v] = lfd[v * n3 + pt]; lf[v] = 0.0; lPpr[v] = 0.0; }
    // f_diff correction: fd_j = fg_j - sum_{v != j, phi_v > 0} (fg_v - kd_v) / nt_j
    double lkdl[V];
#pragma unroll
    for (int v = 0; v < V; ++v) lkdl[v] = 0.0;
#pragma unroll
    for (int jv = 0; jv < V; ++jv) {
      if (!(f[jv] > 0.0)) continue;
      int nt = 1;
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (v != jv && f[v] > 0.0) nt += 1;
      const double l = lfd[jv * n3 + pt] / (double)nt;
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (v != jv && f[v] > 0.0) { lfg[v] -= l; lkdl[v] += l; }
    }
    double lX[E][NPH + 1], lGf[NPH + 1], lMu[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      lMu[e] = 0.0;
#pragma unroll
      for (int q = 0; q <= NPH; ++q) lX[e][q] = 0.0;
    }
#pragma unroll
    for (int q = 0; q <= NPH; ++q) lGf[q] = 0.0;
#pragma unroll
    for (int q = 2; q <= NPH; ++q)
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (P->phase_of[v] == q) {
          // flux F_e = sum_v Ppr_v (X(e,1) - X(e,q)) g_v
          double termch = Gf[q] - Gf[1];
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const double xd = x[NPH * e] - x[NPH * e + q - 1];
            termch = termch + Mu[e] * xd;
            const double s = lF[(3 * e + 0) * n3 + pt] * g[v][0] + lF[(3 * e + 1) * n3 + pt] * g[v][1] +
                             lF[(3 * e + 2) * n3 + pt] * g[v][2];
            lPpr[v] += xd * s;
            const double lxd = Ppr[v] * s;
            lX[e][1] += lxd;
            lX[e][q] -= lxd;
          }
          // fg_v = termch Ppr_v + wg_v
          const double l = lfg[v];
          lPpr[v] += l * termch;
          const double lt = l * Ppr[v];
          lGf[q] += lt;
          lGf[1] -= lt;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const double xd = x[NPH * e] - x[NPH * e + q - 1];
            lMu[e] += lt * xd;
            lX[e][1] += lt * Mu[e];
            lX[e][q] -= lt * Mu[e];
          }
          const double fi = f[v];
          double s2 = 0.0;
#pragma unroll
          for (int jv = 0; jv < V; ++jv)
            if (jv != v) {
              s2 += f[jv] * f[jv];
              lf[jv] += l * 8.0 * P->well[0] * fi * f[jv];
            }
          lf[v] += l * (2.0 * P->well[q - 1] * (1.0 - 6.0 * fi + 6.0 * fi * fi) + 4.0 * P->well[0] * s2);
          swell[q - 1] += l * 2.0 * fi * (1.0 - fi) * (1.0 - 2.0 * fi);
          swell[0] += l * 4.0 * fi * s2;
        }
    // Gf, Mu
#pragma unroll
    for (int ph = 1; ph <= NPH; ++ph)
#pragma unroll
      for (int e = 1; e <= E; ++e)
        lX[e - 1][ph] += lGf[ph] * 2.0 * GP(e, ph) * (x[NPH * (e - 1) + ph - 1] - GP(e + E + 1, ph));
#pragma unroll
    for (int e = 1; e <= E; ++e) lX[e - 1][1] += lMu[e - 1] * 2.0 * GP(e, 1);
    // KKS: A x = b -> mu = A^-T lx; lc_e = mu[Z-E+e]; lPph[ph] = -sum_e mu[Z-E+e] X(e,ph)
    double mu[Z];
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
      for (int ph = 0; ph < NPH; ++ph) mu[NPH * e + ph] = lX[e][ph + 1];
    if (!solve_pivot<Z>(At, mu)) bad = true;
    double lPph[NPH + 1];
#pragma unroll
    for (int q = 0; q <= NPH; ++q) lPph[q] = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const double m = mu[Z - E + e];
      lc[e * n3 + pt] += m;
#pragma unroll
      for (int ph = 0; ph < NPH; ++ph) lPph[ph + 1] -= m * x[NPH * e + ph];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const double lpl = -lPph[1] + lPph[P->phase_of[v]];
      lf[v] += lpl * ppr(f[v]) + lPpr[v] * pfun2(f[v]);
      lphi[v * n3 + pt] += lf[v];
      lkd[v * n3 + pt] = lkdl[v];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        const int q = P->phase_of[v];
#pragma unroll
        for (int e = 0; e < E; ++e) s += lF[(3 * e + d) * n3 + pt] * (x[NPH * e] - x[NPH * e + q - 1]);
        lg[(3 * v + d) * n3 + pt] = Ppr[v] * s;
      }
    }
  }
#undef GP
#pragma unroll
  for (int q = 0; q < NPH; ++q) block_store<0>(swell[q], part, q, sh);
  if (bad) *err = MM_ERR_ZSOLVE;
}

// ------------------------------------------------------- gradient block
// variant v: out = sum_d (-i kf_d) S[d] + kgk_v S[3]; S[0..2] = spectra of lg,
// S[3] = spectrum of lkd, S[4] = phi_hat; slot j: sum w Re(conj(S[3]) kk_j phi_hat)
__global__ void __launch_bounds__(kBlock) k_adj_grad_ks(const DevParams* __restrict__ P, C2<double>* __restrict__ S,
                                                        double* __restrict__ part, int v, size_t nc) {
  __shared__ double sh[kWaves];
  double sj[6] = {0, 0, 0, 0, 0, 0};
  GRID_STRIDE(i, nc) {
    const KVec k = kvec(P, i);
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const C2<double> a = S[(size_t)d * nc + i];
      // (-i kf)(a.x + i a.y) = (kf a.y, -kf a.x)
      re += k.kf[d] * a.y;
      im += -k.kf[d] * a.x;
    }
    const C2<double> l = S[(size_t)3 * nc + i];
    const double K = kgk(P, v, k);
    re += K * l.x;
    im += K * l.y;
    const C2<double> p = S[(size_t)4 * nc + i];
    const double w = pweight(P, k);
    const double pr = l.x * p.x + l.y * p.y;
    double kk[6];
    kforms(k, kk);
#pragma unroll
    for (int j = 0; j < 6; ++j) sj[j] += w * kk[j] * pr;
    S[i] = mk<double>(re, im);
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) block_store<0>(sj[j], part, j, sh);
}

template <int V, int NPH, int E> struct AdjImpl {
  static void elener(dim3 g, hipStream_t s, const DevParams* P, const double* phi, const double* strain,
                     const double* lfd, const double* li, double* lo, double* lphi, double* part, size_t n3) {
    hipLaunchKernelGGL((k_adj_elener<V>), g, dim3(kBlock), 0, s, P, phi, strain, lfd, li, lo, lphi, part, n3);
  }
  static void polar(dim3 g, hipStream_t s, const DevParams* P, const double* phi, const double* strain,
                    const double* lx, double cfac, double* leps, double* lphi, double* part, size_t n3) {
    hipLaunchKernelGGL((k_adj_polar<V>), g, dim3(kBlock), 0, s, P, phi, strain, lx, cfac, leps, lphi, part, n3);
  }
  static void chem(dim3 g, hipStream_t s, const DevParams* P, const double* phi, const double* cons,
                   const double* gphi, const double* kdel, const double* lfd, const double* lF, double* lphi,
                   double* lc, double* lg, double* lkd, double* part, int* err, size_t n3) {
    hipLaunchKernelGGL((k_adj_chem<V, NPH, E>), g, dim3(kBlock), 0, s, P, phi, cons, gphi, kdel, lfd, lF, lphi, lc,
                       lg, lkd, part, err, n3);
  }
};

struct AdjKern {
  decltype(&AdjImpl<1, 2, 1>::elener) elener = nullptr;
  decltype(&AdjImpl<1, 2, 1>::polar) polar = nullptr;
  decltype(&AdjImpl<1, 2, 1>::chem) chem = nullptr;
};

static bool adj_table(int V, int NPH, int E, AdjKern& k) {
#define X(v_, p_, e_)                              \
  if (V == v_ && NPH == p_ && E == e_) {           \
    k.elener = AdjImpl<v_, p_, e_>::elener;        \
    k.polar = AdjImpl<v_, p_, e_>::polar;          \
    k.chem = AdjImpl<v_, p_, e_>::chem;            \
    return true;                                   \
  }
  MM_COMBOS(X)
#undef X
  return false;
}

}  // namespace mm

// ================================================================= host side
using namespace mm;

// gradient accumulator layout (doubles), mirrors mm_adj_grads
enum { AG_GC = 0, AG_WELL = AG_GC + MM_MAX_VAR * 9, AG_CMOB = AG_WELL + MM_MAX_PH, AG_BMOB = AG_CMOB + MM_MAX_EL,
       AG_TP = AG_BMOB + 1, AG_N = AG_TP + MM_MAX_VAR * 6 };
static_assert(sizeof(mm_adj_grads) == AG_N * sizeof(double), "mm_adj_grads layout");

struct mm_adj {
  AdjKern k;
  // tape of the last mm_adj_tape_step
  double *phi0 = nullptr, *cons0 = nullptr, *g = nullptr, *kd = nullptr, *F = nullptr, *fd = nullptr,
         *phipre = nullptr, *cons1 = nullptr;
  std::vector<double*> eps;  // eps[k], k = 0..n (6 fields each)
  std::vector<double> cfac;  // cfac of iteration k+1
  int n = 0;                 // iterations on the tape (-1: no tape)
  bool taped = false;
  // adjoint state and scratch
  double *lphi = nullptr, *lcons = nullptr, *leps = nullptr;
  double *acc_phi = nullptr, *lc = nullptr, *lfd = nullptr, *lF = nullptr, *lg = nullptr, *lkd = nullptr,
         *lx = nullptr, *le = nullptr, *tmp = nullptr;
  C2<double>* spec = nullptr;  // 8 spectra
  double* part = nullptr;      // 64 slots x kMaxBlocks
  double* agrad = nullptr;     // AG_N
};

void adj_release(mm_adj* a) {
  if (!a) return;
  double* bufs[] = {a->phi0, a->cons0, a->g, a->kd, a->F, a->fd, a->phipre, a->cons1, a->lphi, a->lcons, a->leps,
                    a->acc_phi, a->lc, a->lfd, a->lF, a->lg, a->lkd, a->lx, a->le, a->tmp, a->part, a->agrad};
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  for (double* p : a->eps)
    if (p) (void)hipFree(p);
  if (a->spec) (void)hipFree(a->spec);
  delete a;
}

namespace {
int accum(mm_ctx* c, int nblk, const AccumOps& ops) {
  hipLaunchKernelGGL(k_accum, dim3(1), dim3(kBlock), 0, c->s, (const double*)c->adj->part, nblk, ops,
                     c->adj->agrad);
  MM_HIP(hipGetLastError());
  return 0;
}

int fwd64(mm_ctx* c, const double* in, C2<double>* out, int nf) {
  for (int f = 0; f < nf; ++f) MM_TRY(fwd<double>(c, (double*)in + (size_t)f * c->n3, c->n3, out + (size_t)f * c->nc, 1));
  return 0;
}
int inv64(mm_ctx* c, C2<double>* in, double* out, int nf) {
  for (int f = 0; f < nf; ++f) {
    C2<double>* sp = in + (size_t)f * c->nc;
    double* rp = out + (size_t)f * c->n3;
    MM_TRY(inv<double>(c, &sp, &rp, 1));
  }
  return 0;
}
int d2d(mm_ctx* c, void* dst, const void* src, size_t bytes) {
  MM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->s));
  return 0;
}
}  // namespace

extern "C" int mm_adj_enable(mm_ctx* c) {
  if (!c) return MM_ERR_ARG;
  if (c->adj) return 0;
  if (c->prec != MM_FP64 || !c->hp.homog) return MM_ERR_ARG;
  mm_adj* a = new mm_adj;
  if (!adj_table(c->V, c->NPH, c->E, a->k)) { delete a; return MM_ERR_ARG; }
  const size_t n3 = c->n3, nc = c->nc, b = sizeof(double);
  const int V = c->V, E = c->E;
  int rc = 0;
  auto al = [&](double** p, size_t n) {
    if (!rc && hipMalloc((void**)p, n * b) != hipSuccess) { *p = nullptr; rc = MM_ERR_NOMEM; }
  };
  al(&a->phi0, V * n3); al(&a->cons0, E * n3); al(&a->g, 3 * V * n3); al(&a->kd, V * n3); al(&a->F, 3 * E * n3);
  al(&a->fd, V * n3); al(&a->phipre, V * n3); al(&a->cons1, E * n3);
  al(&a->lphi, V * n3); al(&a->lcons, E * n3); al(&a->leps, 6 * n3);
  al(&a->acc_phi, V * n3); al(&a->lc, E * n3); al(&a->lfd, V * n3); al(&a->lF, 3 * E * n3); al(&a->lg, 3 * V * n3);
  al(&a->lkd, V * n3); al(&a->lx, 6 * n3); al(&a->le, 6 * n3); al(&a->tmp, std::max<size_t>(3, V) * n3);
  al(&a->part, (size_t)64 * kMaxBlocks); al(&a->agrad, AG_N);
  if (!rc && hipMalloc((void**)&a->spec, 8 * nc * sizeof(C2<double>)) != hipSuccess) rc = MM_ERR_NOMEM;
  if (rc) { adj_release(a); return rc; }
  c->adj = a;
  // the fused x path computes the same iteration with other FFT plans; the
  // tape replays the step bitwise only through one code path (the real
  // strain is formed from its x half-spectra first)
  rc = strain_to_real_any(c);
  if (rc) { adj_release(a); return rc; }
  c->xpolar = false;
  c->xchem = false;  // the tape copies grad phi and del.G del phi after the head
  if (c->gx) { (void)hipGraphExecDestroy(c->gx); c->gx = nullptr; }
  if (c->g) { (void)hipGraphDestroy(c->g); c->g = nullptr; }
  c->graph = false;
  return mm_adj_zero(c);
}

extern "C" int mm_adj_zero(mm_ctx* c) {
  if (!c || !c->adj) return MM_ERR_ARG;
  mm_adj* a = c->adj;
  MM_HIP(hipMemsetAsync(a->lphi, 0, c->V * c->n3 * sizeof(double), c->s));
  MM_HIP(hipMemsetAsync(a->lcons, 0, c->E * c->n3 * sizeof(double), c->s));
  MM_HIP(hipMemsetAsync(a->leps, 0, 6 * c->n3 * sizeof(double), c->s));
  MM_HIP(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int mm_adj_set(mm_ctx* c, const double* lphi, const double* lcons, const double* leps) {
  if (!c || !c->adj) return MM_ERR_ARG;
  mm_adj* a = c->adj;
  const size_t n3 = c->n3;
  if (lphi) MM_HIP(hipMemcpy(a->lphi, lphi, c->V * n3 * sizeof(double), hipMemcpyHostToDevice));
  if (lcons) MM_HIP(hipMemcpy(a->lcons, lcons, c->E * n3 * sizeof(double), hipMemcpyHostToDevice));
  if (leps) MM_HIP(hipMemcpy(a->leps, leps, 6 * n3 * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int mm_adj_get(mm_ctx* c, double* lphi, double* lcons, double* leps) {
  if (!c || !c->adj) return MM_ERR_ARG;
  mm_adj* a = c->adj;
  const size_t n3 = c->n3;
  MM_HIP(hipStreamSynchronize(c->s));
  if (lphi) MM_HIP(hipMemcpy(lphi, a->lphi, c->V * n3 * sizeof(double), hipMemcpyDeviceToHost));
  if (lcons) MM_HIP(hipMemcpy(lcons, a->lcons, c->E * n3 * sizeof(double), hipMemcpyDeviceToHost));
  if (leps) MM_HIP(hipMemcpy(leps, a->leps, 6 * n3 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// ------------------------------------------------------------ state records
// [header 8 doubles: step, glob_min, glob_max, phi_max, solid_frac, 0, 0, 0]
// phi (V n3), cons (E n3), strain (6 n3), u_hat (3 nc complex)
constexpr int kStateHead = 6 + MM_MAX_EL + 2;  // doubles of the state record header

extern "C" int mm_state_bytes(const mm_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return MM_ERR_ARG;
  *bytes = (int64_t)(kStateHead * sizeof(double) + (c->V + c->E + 6) * c->n3 * c->esz + 3 * c->nc * 2 * c->esz);
  return 0;
}

extern "C" int mm_state_save(mm_ctx* c, void* dst) {
  if (!c || !dst) return MM_ERR_ARG;
  const size_t n3 = c->n3, es = c->esz;
  // header: step, the lagged check-loop values, E and the mean composition
  // cbar (MIXED stores c - cbar; 0 in FP64)
  double h[kStateHead] = {(double)c->step, c->glob_min, c->glob_max, c->phi_max, c->solid_frac, (double)c->E};
  for (int e = 0; e < c->E; ++e) h[6 + e] = c->hp.cbar[e];
  char* p = (char*)dst;
  MM_HIP(hipStreamSynchronize(c->s));
  MM_HIP(hipMemcpy(p, h, sizeof(h), hipMemcpyHostToDevice));
  p += sizeof(h);
  MM_TRY(d2d(c, p, c->phi, c->V * n3 * es)); p += c->V * n3 * es;
  MM_TRY(d2d(c, p, c->cons, c->E * n3 * es)); p += c->E * n3 * es;
  MM_TRY(strain_to_real_any(c));
  MM_TRY(d2d(c, p, c->strain, 6 * n3 * es)); p += 6 * n3 * es;
  MM_TRY(d2d(c, p, c->suhat, 3 * c->nc * 2 * es));
  MM_HIP(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int mm_state_load(mm_ctx* c, const void* src) {
  if (!c || !src) return MM_ERR_ARG;
  const size_t n3 = c->n3, es = c->esz;
  double h[kStateHead];
  const char* p = (const char*)src;
  MM_HIP(hipStreamSynchronize(c->s));
  MM_HIP(hipMemcpy(h, p, sizeof(h), hipMemcpyDeviceToHost));
  if ((int)h[5] != c->E) return MM_ERR_ARG;  // a record of another model
  p += sizeof(h);
  MM_TRY(d2d(c, c->phi, p, c->V * n3 * es)); p += c->V * n3 * es;
  MM_TRY(d2d(c, c->cons, p, c->E * n3 * es)); p += c->E * n3 * es;
  MM_TRY(d2d(c, c->strain, p, 6 * n3 * es)); p += 6 * n3 * es;
  MM_TRY(strain_from_real_any(c));
  MM_TRY(d2d(c, c->suhat, p, 3 * c->nc * 2 * es));
  MM_HIP(hipStreamSynchronize(c->s));
  c->step = (int)h[0];
  c->glob_min = h[1]; c->glob_max = h[2]; c->phi_max = h[3]; c->solid_frac = h[4];
  bool same = true;
  for (int e = 0; e < c->E; ++e) same = same && c->hp.cbar[e] == h[6 + e];
  if (!same) {
    for (int e = 0; e < c->E; ++e) c->hp.cbar[e] = h[6 + e];
    MM_HIP(hipMemcpy(c->dp, &c->hp, sizeof(DevParams), hipMemcpyHostToDevice));
  }
  if (c->adj) c->adj->taped = false;
  return 0;
}

// ------------------------------------------------------------ taped step
extern "C" int mm_adj_tape_step(mm_ctx* c, mm_diag* d) {
  if (!c || !c->adj) return MM_ERR_ARG;
  if (c->step <= c->prm.elastep0 || c->step <= c->prm.noise_step) return MM_ERR_ARG;
  MM_TRY(qcheck_once(c));
  mm_adj* a = c->adj;
  const size_t n3 = c->n3, b = sizeof(double);
  const int V = c->V, E = c->E;
  a->taped = false;
  set_gate(c, nullptr);
  MM_TRY(d2d(c, a->phi0, c->phi, V * n3 * b));
  MM_TRY(d2d(c, a->cons0, c->cons, E * n3 * b));
  MM_TRY(step_head<double>(c));
  MM_TRY(d2d(c, a->g, c->gphi, 3 * V * n3 * b));
  MM_TRY(d2d(c, a->kd, c->kdel, V * n3 * b));
  MM_TRY(d2d(c, a->F, c->flux, 3 * E * n3 * b));
  // elastic loop of run_step_eager with the strain of every iteration recorded
  c->last_niter = 0;
  c->last_cfac = 1.0;
  Ctl h;
  ctl_start(h);
  a->cfac.clear();
  int k = 0;
  bool spectral = false;
  while (h.active) {
    if ((int)a->eps.size() <= k) {
      double* p = nullptr;
      if (hipMalloc((void**)&p, 6 * n3 * b) != hipSuccess) return MM_ERR_NOMEM;
      a->eps.push_back(p);
    }
    MM_TRY(d2d(c, a->eps[k], c->strain, 6 * n3 * b));
    a->cfac.push_back(h.cfac);
    MM_TRY(elastic_iter<double>(c, h.cfac, nullptr, false));
    MM_HIP(hipMemcpyAsync(c->hres + R_ERR0, c->dres + R_ERR0, 6 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    MM_HIP(hipEventRecord(c->ev, c->s));
    MM_TRY(elastic_strain_inverse<double>(c, &spectral));
    MM_HIP(hipEventSynchronize(c->ev));
    ctl_after_iter(h, c->hres + R_ERR0);
    ++k;
  }
  if (h.err) return h.err;
  c->last_niter = h.total;
  c->last_cfac = h.cfac;
  if ((int)a->eps.size() <= k) {
    double* p = nullptr;
    if (hipMalloc((void**)&p, 6 * n3 * b) != hipSuccess) return MM_ERR_NOMEM;
    a->eps.push_back(p);
  }
  MM_TRY(d2d(c, a->eps[k], c->strain, 6 * n3 * b));
  a->n = k;
  MM_TRY(step_tail<double>(c, true, a->fd, a->phipre));
  MM_TRY(d2d(c, a->cons1, c->cons, E * n3 * b));
  MM_TRY(step_collect(c, d));
  a->taped = true;
  return 0;
}

// ------------------------------------------------------------ adjoint step
extern "C" int mm_adj_step(mm_ctx* c, mm_adj_grads* grads) {
  if (!c || !c->adj || !c->adj->taped) return MM_ERR_ARG;
  mm_adj* a = c->adj;
  const size_t n3 = c->n3, nc = c->nc, b = sizeof(double);
  const int V = c->V, E = c->E, NPH = c->NPH;
  const dim3 gp(c->nblk_pw), gk(c->nblk_k);
  const int nbp = c->nblk_pw, nbk = c->nblk_k;
  const double N3 = (double)n3, dt = c->hp.dt, dtl = c->hp.dt * c->hp.bmob;
  C2<double>* S = a->spec;
  MM_HIP(hipMemsetAsync(a->agrad, 0, AG_N * b, c->s));
  MM_HIP(hipMemsetAsync(c->derr, 0, sizeof(int), c->s));
  auto gcidx = [](int v, int j, int second) {
    // kk_j multiplies (11) g0, (22) g4, (33) g8, (12) g1+g3, (23) g5+g7, (31) g2+g6
    const int one[6] = {0, 4, 8, 1, 5, 2}, two[6] = {-1, -1, -1, 3, 7, 6};
    return second ? (two[j] < 0 ? -1 : AG_GC + 9 * v + two[j]) : AG_GC + 9 * v + one[j];
  };
  // 1. clamp
  hipLaunchKernelGGL(k_adj_mask, gp, dim3(kBlock), 0, c->s, (const double*)a->phipre, a->lphi, (size_t)V * n3);
  MM_HIP(hipGetLastError());
  // 2. Allen-Cahn: S[0..V) = D_v F(lam'), S[V..2V) = F(phi')
  MM_TRY(fwd64(c, a->lphi, S, V));
  MM_TRY(fwd64(c, a->phipre, S + (size_t)V * nc, V));
  hipLaunchKernelGGL(k_adj_ac_ks, dim3(nbk, V), dim3(kBlock), 0, c->s, (const DevParams*)c->dp, S,
                     (const C2<double>*)(S + (size_t)V * nc), a->part, V, nc);
  MM_HIP(hipGetLastError());
  {
    // k_adj_ac_ks writes slot s of variant v at part[s*nbk + block] where the
    // slot rows are indexed over (quantity) with gridDim.x = nbk
    AccumOps o{};
    for (int v = 0; v < V; ++v) {
      o.slot[o.n] = v; o.idx[o.n] = AG_BMOB; o.scale[o.n] = dt * N3; ++o.n;
      for (int j = 0; j < 6; ++j) {
        for (int s2 = 0; s2 < 2; ++s2) {
          const int id = gcidx(v, j, s2);
          if (id < 0) continue;
          o.slot[o.n] = V + 6 * v + j; o.idx[o.n] = id; o.scale[o.n] = dtl * N3; ++o.n;
        }
      }
    }
    MM_TRY(accum(c, nbk, o));
  }
  MM_TRY(inv64(c, S, a->acc_phi, V));  // lamt = A_v lam'; lphi accumulator starts there
  hipLaunchKernelGGL(k_adj_ac_pw, gp, dim3(kBlock), 0, c->s, (const DevParams*)c->dp, (const double*)a->acc_phi,
                     (const double*)a->fd, a->lfd, a->part, V, n3);
  MM_HIP(hipGetLastError());
  {
    AccumOps o{};
    for (int v = 0; v < V; ++v) { o.slot[o.n] = v; o.idx[o.n] = AG_BMOB; o.scale[o.n] = -dt; ++o.n; }
    MM_TRY(accum(c, nbp, o));
  }
  // 3. Cahn-Hilliard, per composition
  for (int e = 0; e < E; ++e) {
    MM_TRY(fwd64(c, a->lcons + (size_t)e * n3, S, 1));
    MM_TRY(fwd64(c, a->cons1 + (size_t)e * n3, S + (size_t)4 * nc, 1));
    hipLaunchKernelGGL(k_adj_ch_ks, gk, dim3(kBlock), 0, c->s, (const DevParams*)c->dp, S,
                       (const C2<double>*)(S + (size_t)4 * nc), a->part, e, nc);
    MM_HIP(hipGetLastError());
    {
      AccumOps o{};
      o.slot[0] = 0; o.idx[0] = AG_CMOB + e; o.scale[0] = dt * N3; o.n = 1;
      MM_TRY(accum(c, nbk, o));
    }
    MM_TRY(inv64(c, S, a->lc + (size_t)e * n3, 1));
    MM_TRY(inv64(c, S + nc, a->lF + (size_t)3 * e * n3, 3));
    hipLaunchKernelGGL(k_adj_ch_pw, gp, dim3(kBlock), 0, c->s, (const DevParams*)c->dp, a->lF + (size_t)3 * e * n3,
                       (const double*)(a->F + (size_t)3 * e * n3), a->part, e, n3);
    MM_HIP(hipGetLastError());
    {
      AccumOps o{};
      o.slot[0] = 0; o.idx[0] = AG_CMOB + e; o.scale[0] = dt; o.n = 1;
      MM_TRY(accum(c, nbp, o));
    }
  }
  // 4. elener (f_diff += el_grad): le = leps + d/d eps_n
  a->k.elener(gp, c->s, c->dp, a->phi0, a->eps[a->n], a->lfd, a->leps, a->le, a->acc_phi, a->part, n3);
  MM_HIP(hipGetLastError());
  AccumOps otp{};
  for (int v = 0; v < V; ++v)
    for (int m = 0; m < 6; ++m) { otp.slot[otp.n] = 6 * v + m; otp.idx[otp.n] = AG_TP + 6 * v + m; otp.scale[otp.n] = 1.0; ++otp.n; }
  MM_TRY(accum(c, nbp, otp));
  // 5. elastic iterations, last to first
  for (int k = a->n; k >= 1; --k) {
    MM_TRY(fwd64(c, a->le, S, 6));
    hipLaunchKernelGGL(k_adj_el_ks, gk, dim3(kBlock), 0, c->s, (const DevParams*)c->dp, S, nc);
    MM_HIP(hipGetLastError());
    MM_TRY(inv64(c, S, a->lx, 6));
    a->k.polar(gp, c->s, c->dp, a->phi0, a->eps[k - 1], a->lx, a->cfac[k - 1], a->le, a->acc_phi, a->part, n3);
    MM_HIP(hipGetLastError());
    MM_TRY(accum(c, nbp, otp));
  }
  MM_TRY(d2d(c, a->leps, a->le, 6 * n3 * b));
  // 6. chemistry
  a->k.chem(gp, c->s, c->dp, a->phi0, a->cons0, a->g, a->kd, a->lfd, a->lF, a->acc_phi, a->lc, a->lg, a->lkd, a->part,
            c->derr, n3);
  MM_HIP(hipGetLastError());
  {
    AccumOps o{};
    for (int q = 0; q < NPH; ++q) { o.slot[o.n] = q; o.idx[o.n] = AG_WELL + q; o.scale[o.n] = 1.0; ++o.n; }
    MM_TRY(accum(c, nbp, o));
  }
  // 7. gradient block, per variant
  for (int v = 0; v < V; ++v) {
    MM_TRY(fwd64(c, a->lg + (size_t)3 * v * n3, S, 3));
    MM_TRY(fwd64(c, a->lkd + (size_t)v * n3, S + (size_t)3 * nc, 1));
    MM_TRY(fwd64(c, a->phi0 + (size_t)v * n3, S + (size_t)4 * nc, 1));
    hipLaunchKernelGGL(k_adj_grad_ks, gk, dim3(kBlock), 0, c->s, (const DevParams*)c->dp, S, a->part, v, nc);
    MM_HIP(hipGetLastError());
    {
      AccumOps o{};
      for (int j = 0; j < 6; ++j)
        for (int s2 = 0; s2 < 2; ++s2) {
          const int id = gcidx(v, j, s2);
          if (id < 0) continue;
          o.slot[o.n] = j; o.idx[o.n] = id; o.scale[o.n] = N3; ++o.n;
        }
      MM_TRY(accum(c, nbk, o));
    }
    MM_TRY(inv64(c, S, a->tmp, 1));
    hipLaunchKernelGGL(k_axpy, gp, dim3(kBlock), 0, c->s, a->acc_phi + (size_t)v * n3, (const double*)a->tmp, 1.0, n3);
    MM_HIP(hipGetLastError());
  }
  // 8. new adjoint state
  MM_TRY(d2d(c, a->lphi, a->acc_phi, V * n3 * b));
  MM_TRY(d2d(c, a->lcons, a->lc, E * n3 * b));
  double hg[AG_N];
  int herr = 0;
  MM_HIP(hipMemcpyAsync(hg, a->agrad, AG_N * b, hipMemcpyDeviceToHost, c->s));
  MM_HIP(hipMemcpyAsync(&herr, c->derr, sizeof(int), hipMemcpyDeviceToHost, c->s));
  MM_HIP(hipStreamSynchronize(c->s));
  if (herr) return herr;
  if (grads) {
    double* out = reinterpret_cast<double*>(grads);
    for (int i = 0; i < AG_N; ++i) out[i] += hg[i];
  }
  a->taped = false;
  return 0;
}

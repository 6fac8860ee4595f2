// zfused.h -- z pass with the k-space operation of the step fused between the
// forward and the inverse transform of every column.
//
// The x and y passes leave spectra partially transformed; one workgroup owns
// B whole z-columns (kx, ky), runs the forward z transform of every input
// spectrum of its columns, applies the pointwise k-space operation with all
// inputs of a wave vector in registers, and runs the inverse z transform of
// every output.  Each input and output crosses HBM once instead of three
// times (z forward, k-space kernel, z inverse).
//
//   OpGrad     gradient block      Multi_MEUMAPPS.f:577-613   1 in  -> 4 out (+ full phi_hat)
//   OpElastic  elastic iteration   ModElastic.f:329-433       6 in  -> 6 out (+ u_hat, sums)
//   OpCH       Cahn-Hilliard       Multi_MEUMAPPS.f:912-947   4 in  -> 1 out
//   OpAC       Allen-Cahn          Multi_MEUMAPPS.f:958-997   1 in  -> 1 out (reads phi_hat)
#pragma once
#include <type_traits>

#include "fft_pass.h"
#include "step_kernels.h"

namespace mm {

// Workgroup size: up to 512 threads; plans with more than 8 points per
// thread (radix-3 lengths) use 256 so that their registers fit.
template <typename T, class P, int MAXT> struct TileZF {
  static constexpr int TPL = P::TPL;
  // at least kTW columns, so that a workgroup owns a whole tile of the layout
  static constexpr int MT0 = P::E > 8 ? cmin(MAXT, 256) : MAXT;
  static constexpr int MT = MT0 > kTW * TPL ? MT0 : kTW * TPL;
  // minimum waves per SIMD requested from the register allocator
  template <int MINW> static constexpr int minw() { return P::E <= 4 ? MINW : (MINW > 2 ? 2 : MINW); }
  static constexpr int B = pow2floor(cmin(MT / TPL, kLdsBudget / (P::L * (int)sizeof(C2<T>))));
  static constexpr int THREADS = B * TPL;
  static constexpr int LDS = B * P::L;
  static constexpr int NW = (THREADS + 63) / 64;
};

// per-thread view of one column: kx, ky fixed, kz = tl + TPL*q
struct ColK {
  double kf[2], kq[2];
  int kx, ky;
};
__device__ __forceinline__ ColK colk(const DevParams* __restrict__ P, int col) {
  ColK c;
  c.kx = col / P->ny;
  c.ky = col % P->ny;
  kaxis(c.kx, P->nx, P->unit[0], c.kf[0], c.kq[0]);
  kaxis(c.ky, P->ny, P->unit[1], c.kf[1], c.kq[1]);
  return c;
}
__device__ __forceinline__ KVec kfull(const DevParams* __restrict__ P, const ColK& c, int kz) {
  KVec k;
  k.kx = c.kx; k.ky = c.ky; k.kz = kz;
  k.kf[0] = c.kf[0]; k.kq[0] = c.kq[0];
  k.kf[1] = c.kf[1]; k.kq[1] = c.kq[1];
  kaxis(kz, P->nz, P->unit[2], k.kf[2], k.kq[2]);
  return k;
}

// plan kind of a fused op: 0 lean (radix 8), 1 lean-4 (radix 4), 2 default
template <int L, int K> struct ZPlan { using type = PlanLean<L>; };
template <int L> struct ZPlan<L, 1> { using type = PlanLean4<L>; };
template <int L> struct ZPlan<L, 2> { using type = Plan<L>; };

// Spectra are in the column-tiled layout Z (fft_pass.h): a workgroup owns B
// columns, whole tiles, so each of its loads and stores walks contiguous
// memory.  The twiddle table is staged in LDS once per workgroup: the
// per-stage twiddle reads of the 12 transforms of the elastic operation
// would otherwise be a chain of dependent global loads.
// The model parameters come in as a separate restrict pointer (Pr, the same
// block as op.P): nothing the kernel stores can alias them, so their reads
// stay scalar loads into SGPRs instead of vector reloads after every store.
// Loads of a column past ncols read column 0 (in bounds, never stored), so
// that no load sits behind a branch.
template <typename T, int L, class Op>
__global__ void __launch_bounds__((TileZF<T, typename ZPlan<L, Op::PLAN>::type, Op::MAXT>::THREADS),
                                  (TileZF<T, typename ZPlan<L, Op::PLAN>::type, Op::MAXT>::template minw<Op::MINW>()))
k_z_fused(Op op, const DevParams* __restrict__ Pr, const C2<T>* __restrict__ twL, int ncols) {
  using P = typename ZPlan<L, Op::PLAN>::type;
  using TL = TileZF<T, P, Op::MAXT>;
  op.P = Pr;
  if (gated_off(op.gate)) return;
  constexpr int TPL = P::TPL, E = P::E, B = TL::B;
  static_assert(B % kTW == 0, "a workgroup owns whole tiles");
  // WL: each column is held by whole wavefronts (thread t works on column
  // t / TPL), so the LDS exchanges of a transform need only wavefront order;
  // otherwise the columns of a tile are interleaved thread by thread
  // (element-major LDS), and every exchange is a workgroup barrier
  constexpr bool WL = Op::WL && wave_lines<P>();
  using Idx = std::conditional_t<WL, IdxLineMajor, IdxElemMajor>;
  __shared__ C2<T> lds[WL ? B * line_pitch<L>() : TL::LDS];
  __shared__ C2<T> tw[L];
  const int t = threadIdx.x, b = WL ? t / TPL : t % B, tl = WL ? t % TPL : t / B;
  const int col = blockIdx.x * B + b, g = blockIdx.y;
  for (int e = t; e < L; e += TL::THREADS) tw[e] = twL[e];
  __syncthreads();
  const bool valid = col < ncols;
  const int colc = valid ? col : 0;
  const ColK ck = colk(op.P, colc);
  const Idx idx = [&]() {
    if constexpr (WL) return IdxLineMajor{b * line_pitch<L>()};
    else return IdxElemMajor{b, B};
  }();
  typename Op::template State<E> st;
  // the column of input f+1 is loaded while input f is transformed
  C2<T> vn[E];
  {
    const C2<T>* in = op.input(g, 0);
#pragma unroll
    for (int q = 0; q < E; ++q) vn[q] = in[zcol(colc, tl + TPL * q, L)];
  }
#pragma unroll 1
  for (int f = 0; f < Op::NIN; ++f) {
    C2<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = vn[q];
    if (f + 1 < Op::NIN) {
      const C2<T>* in = op.input(g, f + 1);
#pragma unroll
      for (int q = 0; q < E; ++q) vn[q] = in[zcol(colc, tl + TPL * q, L)];
    }
    if (f > 0) line_sync<WL>();
    line_fft<T, -1, P, Idx, WL>(v, tl, lds, idx, tw);
    op.template absorb<E>(st, g, f, v, ck, tl, TPL);
  }
  op.template finish<E>(st, g, ck, colc, tl, TPL, ncols, valid);
#pragma unroll 1
  for (int o = 0; o < Op::NOUT; ++o) {
    C2<T> v[E];
    op.template emit<E>(st, g, o, v, ck, tl, TPL);
    line_sync<WL>();
    line_fft<T, +1, P, Idx, WL>(v, tl, lds, idx, tw);
    C2<T>* out = op.output(g, o);
    if (valid) {
#pragma unroll
      for (int q = 0; q < E; ++q) out[zcol(col, tl + TPL * q, L)] = v[q];
    }
  }
  op.template reduce<TL::NW>(st);
}

// ----------------------------------------------------------- gradient block
// in: phi_hat after the x and y passes (sphi[v]); the z-transformed phi_hat is
// written back in place for the Allen-Cahn update.  out[4j+d] = phi_hat * i kf_d,
// out[4j+3] = phi_hat * sum G kk.
template <typename T> struct OpGrad {
  static constexpr int NIN = 1, NOUT = 4, MAXT = 512, MINW = 4;
  static constexpr int PLAN = 0;
  static constexpr bool WL = false;
  const DevParams* P;
  C2<T>* sphi;
  C2<T>* work;
  size_t nc;
  int v0;
  const int* gate = nullptr;
  template <int E> struct State { C2<T> s[E]; };
  __device__ const C2<T>* input(int g, int) const { return sphi + (size_t)(v0 + g) * nc; }
  __device__ C2<T>* output(int g, int o) const { return work + (size_t)(4 * g + o) * nc; }
  template <int E>
  __device__ void absorb(State<E>& st, int, int, const C2<T> (&v)[E], const ColK&, int, int) const {
#pragma unroll
    for (int q = 0; q < E; ++q) st.s[q] = v[q];
  }
  template <int E>
  __device__ void finish(State<E>& st, int g, const ColK&, int col, int tl, int TPL, int ncols, bool valid) const {
    if (!valid) return;
    C2<T>* dst = sphi + (size_t)(v0 + g) * nc;
#pragma unroll
    for (int q = 0; q < E; ++q) dst[zcol(col, tl + TPL * q, P->nz)] = st.s[q];
  }
  template <int O, int E>
  __device__ void emit_o(const State<E>& st, int g, C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    const int var = v0 + g;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const KVec k = kfull(P, ck, tl + TPL * q);
      const double sr = (double)st.s[q].x, si = (double)st.s[q].y;
      if constexpr (O < 3) {
        const double a = k.kf[O];
        v[q] = mk<T>((T)(-si * a), (T)(sr * a));
      } else {
        const double gk = kgk(P, var, k);
        v[q] = mk<T>((T)(sr * gk), (T)(si * gk));
      }
    }
  }
  template <int E>
  __device__ void emit(const State<E>& st, int g, int o, C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    switch (o) {
      case 0: emit_o<0, E>(st, g, v, ck, tl, TPL); break;
      case 1: emit_o<1, E>(st, g, v, ck, tl, TPL); break;
      case 2: emit_o<2, E>(st, g, v, ck, tl, TPL); break;
      default: emit_o<3, E>(st, g, v, ck, tl, TPL); break;
    }
  }
  template <int NW, class S> __device__ void reduce(const S&) const {}
};

// The gradient block with the kx and ky factors left to the inverse y pass
// (k_y_pass_grad): in: phi_hat after the x and y passes (sphi[g]); the
// z-transformed phi_hat is written back in place for the Allen-Cahn update.
// out[3g] = phi_hat, out[3g+1] = phi_hat i kf_z, out[3g+2] = phi_hat G_zz kq_z.
template <typename T> struct OpGrad3 {
  static constexpr int NIN = 1, NOUT = 3, MAXT = 512, MINW = 4;
  static constexpr int PLAN = 0;
  static constexpr bool WL = false;
  const DevParams* P;
  C2<T>* sphi;
  C2<T>* work;
  size_t nc;
  const int* gate = nullptr;
  template <int E> struct State { C2<T> s[E]; };
  __device__ const C2<T>* input(int g, int) const { return sphi + (size_t)g * nc; }
  __device__ C2<T>* output(int g, int o) const { return work + (size_t)(3 * g + o) * nc; }
  template <int E>
  __device__ void absorb(State<E>& st, int, int, const C2<T> (&v)[E], const ColK&, int, int) const {
#pragma unroll
    for (int q = 0; q < E; ++q) st.s[q] = v[q];
  }
  template <int E>
  __device__ void finish(State<E>& st, int g, const ColK&, int col, int tl, int TPL, int ncols, bool valid) const {
    if (!valid) return;
    C2<T>* dst = sphi + (size_t)g * nc;
#pragma unroll
    for (int q = 0; q < E; ++q) dst[zcol(col, tl + TPL * q, P->nz)] = st.s[q];
  }
  template <int E>
  __device__ void emit(const State<E>& st, int g, int o, C2<T> (&v)[E], const ColK&, int tl, int TPL) const {
    const double gzz = P->gc[g * 9 + 8];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      double kf, kq;
      kaxis(tl + TPL * q, P->nz, P->unit[2], kf, kq);
      const double sr = (double)st.s[q].x, si = (double)st.s[q].y;
      if (o == 0) v[q] = st.s[q];
      else if (o == 1) v[q] = mk<T>((T)(-si * kf), (T)(sr * kf));
      else v[q] = mk<T>((T)(sr * (gzz * kq)), (T)(si * (gzz * kq)));
    }
  }
  template <int NW, class S> __device__ void reduce(const S&) const {}
};

// Inverse y pass of the gradient block (Z -> S) after OpGrad3: per line
// (kx, z) of variant g = blockIdx.z, from A = phi_hat, C = phi_hat i kf_z and
// Z = phi_hat G_zz kq_z (in + (3g + m) nc, Z layout) it forms the spectra of
// Multi_MEUMAPPS.f:581-613 with the kx and ky factors applied here,
//   grad_x = i kf_x A, grad_y = i kf_y A, grad_z = C,
//   del.G del = a A + Z + i b C,
//   a = G_xx kq_x + G_yy kq_y - (G_xy + G_yx) kf_x kf_y,
//   b = (G_yz + G_zy) kf_y + (G_xz + G_zx) kf_x,
// (so that the sum is phi_hat sum G_ij k_i k_j with kf_sq on the diagonal)
// and transforms them along y into out + (4g + d) nc (S layout, the input of
// k_x_chem).  Three spectra per variant cross HBM between the z and y passes
// instead of four.
template <typename T, class P>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS))
k_y_pass_grad(const DevParams* __restrict__ Pp, const C2<T>* __restrict__ in, C2<T>* __restrict__ out, size_t nc,
              const C2<T>* __restrict__ twL, int nz, int H, const int* __restrict__ gate) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, B = TileLM<T, P>::B, LP = line_pitch<L>();
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  __shared__ C2<T> tw[L];
  if (gated_off(gate)) return;
  const int g = blockIdx.z, t = threadIdx.x;
  for (int i = t; i < L; i += TileLM<T, P>::THREADS) tw[i] = twL[i];
  __syncthreads();
  const int b = t / TPL, tl = t % TPL, line = blockIdx.x * B + b;
  const bool ok = line < nz * H;
  const int kx = ok ? line / nz : 0, z = ok ? line % nz : 0;
  const size_t srow = ((size_t)z * H + kx) * L;
  const double* G = Pp->gc + g * 9;  // G[j*3+i] = grad_coeff(i,j)
  double kfx, kqx;
  kaxis(kx, Pp->nx, Pp->unit[0], kfx, kqx);
  C2<T> A[E], C[E], Z[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const size_t zi = zidx(z, kx, tl + TPL * q, L, nz);
    A[q] = in[(size_t)(3 * g) * nc + zi];
    C[q] = in[(size_t)(3 * g + 1) * nc + zi];
    Z[q] = in[(size_t)(3 * g + 2) * nc + zi];
  }
  IdxLineMajor idx{b * LP};
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    C2<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      double kfy, kqy;
      kaxis(tl + TPL * q, Pp->ny, Pp->unit[1], kfy, kqy);
      const double ar = (double)A[q].x, ai = (double)A[q].y;
      if (d == 0) {
        v[q] = mk<T>((T)(-ai * kfx), (T)(ar * kfx));
      } else if (d == 1) {
        v[q] = mk<T>((T)(-ai * kfy), (T)(ar * kfy));
      } else if (d == 2) {
        v[q] = C[q];
      } else {
        const double a = G[0] * kqx + G[4] * kqy + (G[3] + G[1]) * (-(kfx * kfy));
        const double bb = (G[7] + G[5]) * kfy + (G[2] + G[6]) * kfx;
        const double cr = (double)C[q].x, ci = (double)C[q].y;
        v[q] = mk<T>((T)(a * ar + (double)Z[q].x - bb * ci), (T)(a * ai + (double)Z[q].y + bb * cr));
      }
    }
    if (d > 0) line_sync<wave_lines<P>()>();
    line_fft<T, +1, P, IdxLineMajor, wave_lines<P>()>(v, tl, lds, idx, tw);
    if (ok) {
      C2<T>* dst = out + (size_t)(4 * g + d) * nc + srow;
#pragma unroll
      for (int q = 0; q < E; ++q) dst[tl + TPL * q] = v[q];
    }
  }
}

// --------------------------------------------------------- elastic iteration
// in: stress-polarisation spectra x_1..x_6 after x and y (work[0..5]);
// R = kf . x_hat, u = Q^-1 R (u(0) = 0); Parseval sums of |u - u_old|^2 and
// |u_old|^2 over the half spectrum; out: engineering-strain spectra (in place).
template <typename T> struct OpElastic {
  static constexpr int NIN = 6, NOUT = 6, MAXT = 256, MINW = 2;
  static constexpr int PLAN = 1;
  static constexpr bool WL = false;
  const DevParams* P;
  C2<T>* work;
  C2<T>* uhat;
  double* part;
  size_t nc;
  const int* gate = nullptr;
  template <int E> struct State {
    double rr[3][E], ri[3][E];  // R, then u
    double acc[6];
  };
  __device__ const C2<T>* input(int, int f) const { return work + (size_t)f * nc; }
  __device__ C2<T>* output(int, int o) const { return work + (size_t)o * nc; }
  // x_f contributes  i kf_j x_f  to R_i for every (i,j) with voigt(i,j) = f:
  // R(1) = x1 kf1 + x6 kf2 + x5 kf3; R(2) = x6 kf1 + x2 kf2 + x4 kf3; R(3) = x5 kf1 + x4 kf2 + x3 kf3
  template <int F, int E>
  __device__ void absorb_f(State<E>& st, const C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    constexpr int ia[6] = {0, 1, 2, 1, 0, 0}, ja[6] = {0, 1, 2, 2, 2, 1};
    constexpr int ib[6] = {-1, -1, -1, 2, 2, 1}, jb[6] = {-1, -1, -1, 1, 0, 0};
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if constexpr (F == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) { st.rr[i][q] = 0.0; st.ri[i][q] = 0.0; }
      }
      double kf[3], kq2;
      kf[0] = ck.kf[0];
      kf[1] = ck.kf[1];
      kaxis(tl + TPL * q, P->nz, P->unit[2], kf[2], kq2);
      const double xr = (double)v[q].x, xi = (double)v[q].y;
      // i*a*(xr + i xi) = (-a xi, a xr)
      st.rr[ia[F]][q] += -kf[ja[F]] * xi;
      st.ri[ia[F]][q] += kf[ja[F]] * xr;
      if constexpr (ib[F] >= 0) {
        st.rr[ib[F]][q] += -kf[jb[F]] * xi;
        st.ri[ib[F]][q] += kf[jb[F]] * xr;
      }
    }
  }
  template <int E>
  __device__ void absorb(State<E>& st, int, int f, const C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    switch (f) {
      case 0: absorb_f<0, E>(st, v, ck, tl, TPL); break;
      case 1: absorb_f<1, E>(st, v, ck, tl, TPL); break;
      case 2: absorb_f<2, E>(st, v, ck, tl, TPL); break;
      case 3: absorb_f<3, E>(st, v, ck, tl, TPL); break;
      case 4: absorb_f<4, E>(st, v, ck, tl, TPL); break;
      default: absorb_f<5, E>(st, v, ck, tl, TPL); break;
    }
  }
  // every solve first, then one batch of u_hat loads and one of stores
  template <int E>
  __device__ void finish(State<E>& st, int, const ColK& ck, int col, int tl, int TPL, int, bool valid) const {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int kz = tl + TPL * q;
      const KVec k = kfull(P, ck, kz);
      Cplx3 u;
#pragma unroll
      for (int i = 0; i < 3; ++i) { u.re[i] = st.rr[i][q]; u.im[i] = st.ri[i][q]; }
      if (k.kx == 0 && k.ky == 0 && kz == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) { u.re[i] = 0.0; u.im[i] = 0.0; }
      } else {
        acoustic_solve(P, k, u);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) { st.rr[i][q] = u.re[i]; st.ri[i][q] = u.im[i]; }
    }
    const int nz = P->nz;
    C2<T> uo[3][E];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int q = 0; q < E; ++q) uo[i][q] = uhat[(size_t)i * nc + zcol(col, tl + TPL * q, nz)];
#pragma unroll
    for (int m = 0; m < 6; ++m) st.acc[m] = 0.0;
    const double w = (ck.kx == 0 || 2 * ck.kx == P->nx) ? 1.0 : 2.0;
#pragma unroll
    for (int q = 0; q < E; ++q)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double ur = (double)uo[i][q].x, ui = (double)uo[i][q].y;
        const double dr = st.rr[i][q] - ur, di = st.ri[i][q] - ui;
        st.acc[2 * i] += w * (dr * dr + di * di);
        st.acc[2 * i + 1] += w * (ur * ur + ui * ui);
      }
    if (!valid) {
#pragma unroll
      for (int m = 0; m < 6; ++m) st.acc[m] = 0.0;
      return;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int q = 0; q < E; ++q)
        uhat[(size_t)i * nc + zcol(col, tl + TPL * q, nz)] = mk<T>((T)st.rr[i][q], (T)st.ri[i][q]);
  }
  // e11, e22, e33, g23, g13, g12 = i kf_A u_B (+ i kf_A2 u_B2)
  template <int O, int E>
  __device__ void emit_o(const State<E>& st, C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    constexpr int A[6] = {0, 1, 2, 1, 0, 0}, Bi[6] = {0, 1, 2, 2, 2, 1};
    constexpr int A2[6] = {-1, -1, -1, 2, 2, 1}, B2[6] = {-1, -1, -1, 1, 0, 0};
#pragma unroll
    for (int q = 0; q < E; ++q) {
      double kf[3], kq2;
      kf[0] = ck.kf[0];
      kf[1] = ck.kf[1];
      kaxis(tl + TPL * q, P->nz, P->unit[2], kf[2], kq2);
      double re = -kf[A[O]] * st.ri[Bi[O]][q], im = kf[A[O]] * st.rr[Bi[O]][q];
      if constexpr (A2[O] >= 0) {
        re += -kf[A2[O]] * st.ri[B2[O]][q];
        im += kf[A2[O]] * st.rr[B2[O]][q];
      }
      v[q] = mk<T>((T)re, (T)im);
    }
  }
  template <int E>
  __device__ void emit(const State<E>& st, int, int o, C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    switch (o) {
      case 0: emit_o<0, E>(st, v, ck, tl, TPL); break;
      case 1: emit_o<1, E>(st, v, ck, tl, TPL); break;
      case 2: emit_o<2, E>(st, v, ck, tl, TPL); break;
      case 3: emit_o<3, E>(st, v, ck, tl, TPL); break;
      case 4: emit_o<4, E>(st, v, ck, tl, TPL); break;
      default: emit_o<5, E>(st, v, ck, tl, TPL); break;
    }
  }
  template <int NW, class S> __device__ void reduce(const S& st) const {
    __shared__ double sh[NW];
#pragma unroll
    for (int m = 0; m < 6; ++m) block_store<0, NW>(st.acc[m], part, m, sh);
  }
};

// launch-shape variants of the elastic op (plan kind, workgroup size cap,
// minimum waves per SIMD, columns held by whole wavefronts), chosen at
// context creation by precision and column length
template <typename T, int PLAN_, int MAXT_, int MINW_, bool WL_ = false> struct OpElasticV : OpElastic<T> {
  static constexpr int PLAN = PLAN_, MAXT = MAXT_, MINW = MINW_;
  static constexpr bool WL = WL_;
  OpElasticV(const OpElastic<T>& b) : OpElastic<T>(b) {}
};
constexpr int kZfElasticVariants = 4;

// ------------------------------------------------------------ Cahn-Hilliard
// in: work[4e] = c, work[4e+1..3] = Fx, Fy, Fz after x and y; out: c_hat new.
template <typename T> struct OpCH {
  static constexpr int NIN = 4, NOUT = 1, MAXT = 512, MINW = 4;
  static constexpr int PLAN = 1;
  static constexpr bool WL = false;
  const DevParams* P;
  C2<T>* work;
  size_t nc;
  const int* gate = nullptr;
  template <int E> struct State { double cr[E], ci[E], tr[E], ti[E]; };
  __device__ const C2<T>* input(int g, int f) const { return work + (size_t)(4 * g + f) * nc; }
  __device__ C2<T>* output(int g, int) const { return work + (size_t)(4 * g) * nc; }
  template <int F, int E>
  __device__ void absorb_f(State<E>& st, const C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if constexpr (F == 0) {
        st.cr[q] = (double)v[q].x;
        st.ci[q] = (double)v[q].y;
        st.tr[q] = 0.0;
        st.ti[q] = 0.0;
      } else {
        double a, kq2;
        if constexpr (F < 3) a = ck.kf[F - 1];
        else kaxis(tl + TPL * q, P->nz, P->unit[2], a, kq2);
        st.tr[q] += -a * (double)v[q].y;
        st.ti[q] += a * (double)v[q].x;
      }
    }
  }
  template <int E>
  __device__ void absorb(State<E>& st, int, int f, const C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    switch (f) {
      case 0: absorb_f<0, E>(st, v, ck, tl, TPL); break;
      case 1: absorb_f<1, E>(st, v, ck, tl, TPL); break;
      case 2: absorb_f<2, E>(st, v, ck, tl, TPL); break;
      default: absorb_f<3, E>(st, v, ck, tl, TPL); break;
    }
  }
  template <int E>
  __device__ void finish(State<E>&, int, const ColK&, int, int, int, int, bool) const {}
  template <int E>
  __device__ void emit(const State<E>& st, int g, int, C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    const double dtm = P->dt * P->cmobs[g];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const KVec k = kfull(P, ck, tl + TPL * q);
      const double den = 1.0 - dtm * (k.kq[0] + k.kq[1] + k.kq[2]);
      v[q] = mk<T>((T)((st.cr[q] + dtm * st.tr[q]) / den), (T)((st.ci[q] + dtm * st.ti[q]) / den));
      // MIXED stores c - cbar: its mean is zero, and the FP32 rounding the
      // forward transform leaves in the k = 0 mode must not accumulate
      if (sizeof(T) == 4 && k.kx == 0 && k.ky == 0 && k.kz == 0) v[q] = mk<T>(0, 0);
    }
  }
  template <int NW, class S> __device__ void reduce(const S&) const {}
};

// The same update with the x and y flux terms folded into the y pass
// (k_y_pass_ch): in: work[2e] = c_hat + dt M (i kf_x F_x + i kf_y F_y),
// work[2e+1] = F_z after x and y; out (work[2e]): c_hat new.
template <typename T> struct OpCH2 {
  static constexpr int NIN = 2, NOUT = 1, MAXT = 512, MINW = 4;
  static constexpr int PLAN = 1;
  static constexpr bool WL = false;
  const DevParams* P;
  C2<T>* work;
  size_t nc;
  const int* gate = nullptr;
  template <int E> struct State { double cr[E], ci[E], tr[E], ti[E]; };
  __device__ const C2<T>* input(int g, int f) const { return work + (size_t)(2 * g + f) * nc; }
  __device__ C2<T>* output(int g, int) const { return work + (size_t)(2 * g) * nc; }
  template <int E>
  __device__ void absorb(State<E>& st, int, int f, const C2<T> (&v)[E], const ColK&, int tl, int TPL) const {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if (f == 0) {
        st.cr[q] = (double)v[q].x;
        st.ci[q] = (double)v[q].y;
      } else {
        double a, kq2;
        kaxis(tl + TPL * q, P->nz, P->unit[2], a, kq2);
        st.tr[q] = -a * (double)v[q].y;
        st.ti[q] = a * (double)v[q].x;
      }
    }
  }
  template <int E>
  __device__ void finish(State<E>&, int, const ColK&, int, int, int, int, bool) const {}
  template <int E>
  __device__ void emit(const State<E>& st, int g, int, C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    const double dtm = P->dt * P->cmobs[g];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const KVec k = kfull(P, ck, tl + TPL * q);
      const double den = 1.0 - dtm * (k.kq[0] + k.kq[1] + k.kq[2]);
      v[q] = mk<T>((T)((st.cr[q] + dtm * st.tr[q]) / den), (T)((st.ci[q] + dtm * st.ti[q]) / den));
      // MIXED stores c - cbar: its mean is zero, and the FP32 rounding the
      // forward transforms leave in the k = 0 mode must not accumulate
      if (sizeof(T) == 4 && k.kx == 0 && k.ky == 0 && k.kz == 0) v[q] = mk<T>(0, 0);
    }
  }
  template <int NW, class S> __device__ void reduce(const S&) const {}
};

// Forward y pass of the Cahn-Hilliard update (S -> Z) with the x and y flux
// terms of Multi_MEUMAPPS.f:924-943 folded in: per line (z, kx) of
// composition e = blockIdx.z the x half-spectra c, F_x, F_y, F_z
// (in + (4e + m) nc) are transformed along y, and
//   N = c_hat + dt M_{e0+e} (i kf_x F_x_hat + i kf_y F_y_hat)
// goes to out + 2e nc, F_z_hat to out + (2e + 1) nc.  OpCH2 then adds the z
// term and divides by the semi-implicit denominator: two spectra cross HBM
// between the passes instead of four.
template <typename T, class P>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS))
k_y_pass_ch(const DevParams* __restrict__ Pp, const C2<T>* __restrict__ in, C2<T>* __restrict__ out, size_t nc,
            const C2<T>* __restrict__ twL, int nz, int H, int e0, const int* __restrict__ gate) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, B = TileLM<T, P>::B, LP = line_pitch<L>();
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  __shared__ C2<T> tw[L];
  if (gated_off(gate)) return;
  // e: the buffers' composition index; e0 + e that of the parameters
  const int e = blockIdx.z, t = threadIdx.x;
  for (int i = t; i < L; i += TileLM<T, P>::THREADS) tw[i] = twL[i];
  __syncthreads();
  const int b = t / TPL, tl = t % TPL, line = blockIdx.x * B + b;
  const bool ok = line < nz * H;
  const int kx = ok ? line / nz : 0, z = ok ? line % nz : 0;
  const size_t srow = ((size_t)z * H + kx) * L;
  double kfx, kqx;
  kaxis(kx, Pp->nx, Pp->unit[0], kfx, kqx);
  const double dtm = Pp->dt * Pp->cmobs[e0 + e];
  IdxLineMajor idx{b * LP};
  double ar[E], ai[E];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const C2<T>* src = in + (size_t)(4 * e + m) * nc + srow;
    C2<T> v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = ok ? src[tl + TPL * q] : mk<T>(0, 0);
    if (m > 0) line_sync<wave_lines<P>()>();
    line_fft<T, -1, P, IdxLineMajor, wave_lines<P>()>(v, tl, lds, idx, tw);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int ky = tl + TPL * q;
      if (m == 0) {
        ar[q] = (double)v[q].x;
        ai[q] = (double)v[q].y;
      } else if (m < 3) {
        double a = kfx, kq2;
        if (m == 2) kaxis(ky, Pp->ny, Pp->unit[1], a, kq2);
        ar[q] += dtm * (-a * (double)v[q].y);
        ai[q] += dtm * (a * (double)v[q].x);
      }
      if (ok && m == 2) out[(size_t)(2 * e) * nc + zidx(z, kx, ky, L, nz)] = mk<T>((T)ar[q], (T)ai[q]);
      if (ok && m == 3) out[(size_t)(2 * e + 1) * nc + zidx(z, kx, ky, L, nz)] = v[q];
    }
  }
}

// -------------------------------------------------------------- Allen-Cahn
// in: work[v] = f_diff after x and y; reads the full phi_hat (sphi[v]);
// out (in place): (phi_hat - dt L f_hat) / (1 - dt L sum G kk).
template <typename T> struct OpAC {
  static constexpr int NIN = 1, NOUT = 1, MAXT = 512, MINW = 4;
  static constexpr int PLAN = 0;
  static constexpr bool WL = false;
  const DevParams* P;
  C2<T>* work;
  const C2<T>* sphi;
  size_t nc;
  const int* gate = nullptr;
  template <int E> struct State { double fr[E], fi[E], pr[E], pi[E]; };
  __device__ const C2<T>* input(int g, int) const { return work + (size_t)g * nc; }
  __device__ C2<T>* output(int g, int) const { return work + (size_t)g * nc; }
  template <int E>
  __device__ void absorb(State<E>& st, int, int, const C2<T> (&v)[E], const ColK&, int, int) const {
#pragma unroll
    for (int q = 0; q < E; ++q) { st.fr[q] = (double)v[q].x; st.fi[q] = (double)v[q].y; }
  }
  template <int E>
  __device__ void finish(State<E>& st, int g, const ColK&, int col, int tl, int TPL, int ncols, bool valid) const {
    const C2<T>* src = sphi + (size_t)g * nc;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const C2<T> s = src[zcol(col, tl + TPL * q, P->nz)];
      st.pr[q] = (double)s.x;
      st.pi[q] = (double)s.y;
    }
  }
  template <int E>
  __device__ void emit(const State<E>& st, int g, int, C2<T> (&v)[E], const ColK& ck, int tl, int TPL) const {
    const double dtl = P->dt * P->bmob;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const KVec k = kfull(P, ck, tl + TPL * q);
      const double den = 1.0 - dtl * kgk(P, g, k);
      v[q] = mk<T>((T)((st.pr[q] - dtl * st.fr[q]) / den), (T)((st.pi[q] - dtl * st.fi[q]) / den));
    }
  }
  template <int NW, class S> __device__ void reduce(const S&) const {}
};

}  // namespace mm

// fft_core.h -- in-register DFT codelets and the LDS-staged Stockham line
// transform used by every FFT pass kernel.
//
// A line of L complex points is owned by TPL = L / E threads; thread `tl`
// holds the E points  e = tl + TPL*q  (q = 0..E-1) in registers.  The radix
// plan L = R0*R1*...  is applied as Stockham autosort stages:
//   stage s (radix R, Ls = R0*...*R_{s-1}); butterfly j = tl + TPL*i:
//     inputs  j + m*L/R,               twiddle  w_{Ls*R}^{m*(j mod Ls)}
//     outputs (j/Ls)*Ls*R + (j mod Ls) + m*Ls
// Stage 0 reads the registers, the last stage writes the registers (its
// outputs j + m*L/R are exactly the register positions), and every stage in
// between exchanges through LDS.  Each thread's registers therefore hold the
// same element positions before and after the transform, which is what makes
// the global loads and stores of the pass kernels coalesced.
#pragma once
#include <hip/hip_runtime.h>

namespace mm {

constexpr int kMaxBatch = 16;  // fields per batched FFT launch

template <typename T> struct C2T;
template <> struct C2T<double> { using type = double2; };
template <> struct C2T<float> { using type = float2; };
template <typename T> using C2 = typename C2T<T>::type;

template <typename T> __device__ __forceinline__ C2<T> mk(T x, T y) { C2<T> r; r.x = x; r.y = y; return r; }
template <typename T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ C2<T> cconj(C2<T> a) { return mk<T>(a.x, -a.y); }
template <typename T> __device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// a * (i*s)
template <typename T> __device__ __forceinline__ C2<T> cmul_i(C2<T> a, T s) { return mk<T>(-a.y * s, a.x * s); }

// cos / sin (2 pi j / 48), exact at multiples of pi/2 and symmetric.
__device__ constexpr double kC48[48] = {
    1.0, 0.9914448613738104, 0.9659258262890683, 0.9238795325112867, 0.8660254037844386,
    0.7933533402912352, 0.7071067811865476, 0.6087614290087207, 0.5, 0.3826834323650898,
    0.25881904510252074, 0.1305261922200516, 0.0, -0.1305261922200516, -0.25881904510252074,
    -0.3826834323650898, -0.5, -0.6087614290087207, -0.7071067811865476, -0.7933533402912352,
    -0.8660254037844386, -0.9238795325112867, -0.9659258262890683, -0.9914448613738104, -1.0,
    -0.9914448613738104, -0.9659258262890683, -0.9238795325112867, -0.8660254037844386,
    -0.7933533402912352, -0.7071067811865476, -0.6087614290087207, -0.5, -0.3826834323650898,
    -0.25881904510252074, -0.1305261922200516, 0.0, 0.1305261922200516, 0.25881904510252074,
    0.3826834323650898, 0.5, 0.6087614290087207, 0.7071067811865476, 0.7933533402912352,
    0.8660254037844386, 0.9238795325112867, 0.9659258262890683, 0.9914448613738104};
__device__ constexpr double kS48[48] = {
    0.0, 0.1305261922200516, 0.25881904510252074, 0.3826834323650898, 0.5, 0.6087614290087207,
    0.7071067811865476, 0.7933533402912352, 0.8660254037844386, 0.9238795325112867,
    0.9659258262890683, 0.9914448613738104, 1.0, 0.9914448613738104, 0.9659258262890683,
    0.9238795325112867, 0.8660254037844386, 0.7933533402912352, 0.7071067811865476,
    0.6087614290087207, 0.5, 0.3826834323650898, 0.25881904510252074, 0.1305261922200516, 0.0,
    -0.1305261922200516, -0.25881904510252074, -0.3826834323650898, -0.5, -0.6087614290087207,
    -0.7071067811865476, -0.7933533402912352, -0.8660254037844386, -0.9238795325112867,
    -0.9659258262890683, -0.9914448613738104, -1.0, -0.9914448613738104, -0.9659258262890683,
    -0.9238795325112867, -0.8660254037844386, -0.7933533402912352, -0.7071067811865476,
    -0.6087614290087207, -0.5, -0.3826834323650898, -0.25881904510252074, -0.1305261922200516};

// multiply by w_R^k = exp(DIR * 2 pi i k / R), compile-time R, k (R | 48)
template <typename T, int DIR, int R, int K>
__device__ __forceinline__ C2<T> twc(C2<T> a) {
  constexpr int j = ((K % R) * (48 / R)) % 48;
  if constexpr (j == 0) {
    return a;
  } else if constexpr (j == 12) {  // exp(DIR*i*pi/2) = DIR*i
    return DIR < 0 ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x);
  } else if constexpr (j == 24) {
    return mk<T>(-a.x, -a.y);
  } else if constexpr (j == 36) {
    return DIR < 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
  } else {
    const T c = (T)kC48[j];
    const T s = (T)(DIR < 0 ? -kS48[j] : kS48[j]);
    return mk<T>(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// ------------------------------------------------------------- codelets
template <typename T, int DIR, int R> struct Dft;

template <typename T, int DIR> struct Dft<T, DIR, 1> {
  __device__ __forceinline__ static void run(C2<T> (&)[1]) {}
};

template <typename T, int DIR> struct Dft<T, DIR, 2> {
  __device__ __forceinline__ static void run(C2<T> (&a)[2]) {
    C2<T> t = a[0];
    a[0] = cadd<T>(t, a[1]);
    a[1] = csub<T>(t, a[1]);
  }
};

template <typename T, int DIR> struct Dft<T, DIR, 3> {
  __device__ __forceinline__ static void run(C2<T> (&a)[3]) {
    const T h = (T)0.8660254037844386;  // sqrt(3)/2
    C2<T> t = cadd<T>(a[1], a[2]);
    C2<T> d = csub<T>(a[1], a[2]);
    C2<T> m = mk<T>(a[0].x - (T)0.5 * t.x, a[0].y - (T)0.5 * t.y);
    // DIR*i*h*d
    C2<T> r = DIR < 0 ? mk<T>(h * d.y, -h * d.x) : mk<T>(-h * d.y, h * d.x);
    a[0] = cadd<T>(a[0], t);
    a[1] = cadd<T>(m, r);
    a[2] = csub<T>(m, r);
  }
};

// even/odd split: X_k = E_k + w_R^k O_k, X_{k+R/2} = E_k - w_R^k O_k
template <typename T, int DIR, int R, int K>
__device__ __forceinline__ void combine(C2<T> (&a)[R], const C2<T> (&e)[R / 2], const C2<T> (&o)[R / 2]) {
  if constexpr (K < R / 2) {
    C2<T> w = twc<T, DIR, R, K>(o[K]);
    a[K] = cadd<T>(e[K], w);
    a[K + R / 2] = csub<T>(e[K], w);
    combine<T, DIR, R, K + 1>(a, e, o);
  }
}

template <typename T, int DIR, int R> struct DftSplit {
  __device__ __forceinline__ static void run(C2<T> (&a)[R]) {
    C2<T> e[R / 2], o[R / 2];
#pragma unroll
    for (int k = 0; k < R / 2; ++k) { e[k] = a[2 * k]; o[k] = a[2 * k + 1]; }
    Dft<T, DIR, R / 2>::run(e);
    Dft<T, DIR, R / 2>::run(o);
    combine<T, DIR, R, 0>(a, e, o);
  }
};

template <typename T, int DIR> struct Dft<T, DIR, 4> : DftSplit<T, DIR, 4> {};
template <typename T, int DIR> struct Dft<T, DIR, 6> : DftSplit<T, DIR, 6> {};
template <typename T, int DIR> struct Dft<T, DIR, 8> : DftSplit<T, DIR, 8> {};
template <typename T, int DIR> struct Dft<T, DIR, 12> : DftSplit<T, DIR, 12> {};
template <typename T, int DIR> struct Dft<T, DIR, 16> : DftSplit<T, DIR, 16> {};

// ------------------------------------------------------------ line plans
// L = R0*R1*R2*R3, E points per thread (a multiple of every radix).
template <int L_, int E_, int R0_, int R1_, int R2_, int R3_, int R4_ = 1> struct PlanT {
  static constexpr int L = L_, E = E_, TPL = L_ / E_;
  static constexpr int R0 = R0_, R1 = R1_, R2 = R2_, R3 = R3_, R4 = R4_;
  static constexpr int NS = 1 + (R1_ > 1) + (R2_ > 1) + (R3_ > 1) + (R4_ > 1);
  static_assert(R0_ * R1_ * R2_ * R3_ * R4_ == L_, "plan product");
  static_assert(E_ % R0_ == 0 && E_ % R1_ == 0 && E_ % R2_ == 0 && E_ % R3_ == 0 && E_ % R4_ == 0,
                "plan E");
};

// default plans: few stages (one LDS exchange where the registers allow)
template <int L> struct PlanSel;
#define MM_PLAN(L_, E_, R0_, R1_, R2_, R3_) \
  template <> struct PlanSel<L_> { using type = PlanT<L_, E_, R0_, R1_, R2_, R3_>; };
MM_PLAN(4, 4, 4, 1, 1, 1)
MM_PLAN(6, 6, 6, 1, 1, 1)
MM_PLAN(8, 8, 8, 1, 1, 1)
MM_PLAN(12, 12, 4, 3, 1, 1)
MM_PLAN(16, 4, 4, 4, 1, 1)
MM_PLAN(24, 12, 4, 6, 1, 1)
MM_PLAN(32, 8, 8, 4, 1, 1)
MM_PLAN(48, 12, 4, 4, 3, 1)
MM_PLAN(64, 8, 8, 8, 1, 1)
MM_PLAN(96, 12, 4, 4, 6, 1)
MM_PLAN(128, 16, 16, 8, 1, 1)
MM_PLAN(192, 12, 4, 4, 12, 1)
MM_PLAN(256, 16, 16, 16, 1, 1)
MM_PLAN(384, 12, 4, 4, 4, 6)
MM_PLAN(512, 16, 16, 16, 2, 1)
#undef MM_PLAN
template <int L> using Plan = typename PlanSel<L>::type;

// lean plans for the fused z kernels, which keep several spectra of the
// same line in registers: at most 12 points per thread (PlanLean) or, for
// the six-spectrum elastic operation, 4 points per thread where the length
// allows (PlanLean4)
template <int L> struct PlanLeanSel { using type = Plan<L>; };
template <> struct PlanLeanSel<128> { using type = PlanT<128, 8, 8, 4, 4, 1>; };
template <> struct PlanLeanSel<256> { using type = PlanT<256, 8, 8, 8, 4, 1>; };
template <> struct PlanLeanSel<512> { using type = PlanT<512, 8, 8, 8, 8, 1>; };
template <int L> using PlanLean = typename PlanLeanSel<L>::type;

template <int L> struct PlanLean4Sel { using type = PlanLean<L>; };
template <> struct PlanLean4Sel<8> { using type = PlanT<8, 4, 4, 2, 1, 1>; };
template <> struct PlanLean4Sel<32> { using type = PlanT<32, 4, 4, 4, 2, 1>; };
template <> struct PlanLean4Sel<64> { using type = PlanT<64, 4, 4, 4, 4, 1>; };
template <> struct PlanLean4Sel<128> { using type = PlanT<128, 4, 4, 4, 4, 2>; };
template <> struct PlanLean4Sel<256> { using type = PlanT<256, 4, 4, 4, 4, 4>; };
template <> struct PlanLean4Sel<512> { using type = PlanT<512, 4, 4, 4, 4, 4, 2>; };
// radix-3 lengths: 6 points per thread (radix 6, then radix 2) instead of 12
template <> struct PlanLean4Sel<12> { using type = PlanT<12, 6, 6, 2, 1, 1>; };
template <> struct PlanLean4Sel<24> { using type = PlanT<24, 6, 6, 2, 2, 1>; };
template <> struct PlanLean4Sel<48> { using type = PlanT<48, 6, 6, 2, 2, 2>; };
template <> struct PlanLean4Sel<96> { using type = PlanT<96, 6, 6, 2, 2, 2, 2>; };
template <int L> using PlanLean4 = typename PlanLean4Sel<L>::type;

template <class P, int S> struct StageR {
  static constexpr int R = S == 0 ? P::R0 : S == 1 ? P::R1 : S == 2 ? P::R2 : S == 3 ? P::R3 : P::R4;
  static constexpr int Ls = S == 0 ? 1
                            : S == 1 ? P::R0
                            : S == 2 ? P::R0 * P::R1
                            : S == 3 ? P::R0 * P::R1 * P::R2
                                     : P::R0 * P::R1 * P::R2 * P::R3;
};

// Lines held by whole wavefronts: when a line's TPL threads are consecutive
// (thread t works on line t / TPL) and TPL divides 64, no line straddles two
// wavefronts, and the LDS exchanges between the stages of a line need only
// wavefront-level ordering (a wavefront's LDS instructions execute in order),
// not a workgroup barrier.
template <class P> constexpr bool wave_lines() { return P::TPL <= 64 && 64 % P::TPL == 0; }
template <bool WAVE> __device__ __forceinline__ void line_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Twiddle sources of the Stockham stages.  TwTable reads the forward table
// exp(-2 pi i e / L), e in [0, L) (global or LDS) at every use; TwRegs holds the
// factors this thread's butterflies use in every stage after the first (NB (R-1)
// per stage), loaded once, for kernels that run many transforms of one length.
template <typename T> struct TwTable {
  const C2<T>* __restrict__ p;
  template <int S> __device__ __forceinline__ C2<T> get(int, int, int e) const { return p[e]; }
};
template <class P> __host__ __device__ constexpr int tw_count_upto(int S) {
  int n = 0;
  for (int s = 1; s < S && s < P::NS; ++s) {
    const int R = s == 1 ? P::R1 : s == 2 ? P::R2 : s == 3 ? P::R3 : P::R4;
    n += (P::E / R) * (R - 1);
  }
  return n;
}
template <typename T, class P> struct TwRegs {
  static constexpr int N = tw_count_upto<P>(P::NS) > 0 ? tw_count_upto<P>(P::NS) : 1;
  C2<T> w[N];
  template <int S> __device__ __forceinline__ C2<T> get(int i, int m, int) const {
    constexpr int R = S == 1 ? P::R1 : S == 2 ? P::R2 : S == 3 ? P::R3 : P::R4;
    return w[tw_count_upto<P>(S) + i * (R - 1) + (m - 1)];
  }
  template <int S> __device__ __forceinline__ void load_stage(const C2<T>* __restrict__ tw, int tl) {
    if constexpr (S < P::NS) {
      constexpr int R = S == 1 ? P::R1 : S == 2 ? P::R2 : S == 3 ? P::R3 : P::R4;
      constexpr int Ls = S == 1 ? P::R0 : S == 2 ? P::R0 * P::R1 : S == 3 ? P::R0 * P::R1 * P::R2
                                                                        : P::R0 * P::R1 * P::R2 * P::R3;
#pragma unroll
      for (int i = 0; i < P::E / R; ++i) {
        const int k = (tl + P::TPL * i) % Ls;
#pragma unroll
        for (int m = 1; m < R; ++m) w[tw_count_upto<P>(S) + i * (R - 1) + (m - 1)] = tw[m * k * (P::L / (Ls * R))];
      }
      load_stage<S + 1>(tw, tl);
    }
  }
  __device__ __forceinline__ void load(const C2<T>* __restrict__ tw, int tl) { load_stage<1>(tw, tl); }
};

// One Stockham stage. `idx(e)` maps a line position to an LDS slot of this
// thread's line; tw is a twiddle source (TwTable of the forward table
// exp(-2 pi i e / L), e in [0, L), or TwRegs).
// WAVE: the lines are held by whole wavefronts (wave_lines<P>()).
template <typename T, int DIR, class P, int S, class Idx, bool WAVE, class TW>
__device__ __forceinline__ void stage(C2<T> (&v)[P::E], int tl, C2<T>* lds, const Idx& idx, const TW& tw) {
  constexpr int R = StageR<P, S>::R;
  constexpr int Ls = StageR<P, S>::Ls;
  constexpr int NB = P::E / R;  // butterflies per thread
  constexpr int L = P::L, TPL = P::TPL;
  constexpr bool last = (S == P::NS - 1);
  if constexpr (S > 0) {
    // gather this stage's inputs j + m*L/R = tl + TPL*(i + m*NB) from LDS
#pragma unroll
    for (int q = 0; q < P::E; ++q) v[q] = lds[idx(tl + TPL * q)];
    if constexpr (!last) line_sync<WAVE>();
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int j = tl + TPL * i;
    const int k = j % Ls;
    C2<T> a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = v[i + m * NB];
    if constexpr (S > 0) {
#pragma unroll
      for (int m = 1; m < R; ++m) {
        C2<T> w = tw.template get<S>(i, m, m * k * (L / (Ls * R)));
        if constexpr (DIR > 0) w.y = -w.y;
        a[m] = cmul<T>(a[m], w);
      }
    }
    Dft<T, DIR, R>::run(a);
    if constexpr (last) {
#pragma unroll
      for (int m = 0; m < R; ++m) v[i + m * NB] = a[m];
    } else {
      const int ob = (j / Ls) * (Ls * R) + k;
#pragma unroll
      for (int m = 0; m < R; ++m) lds[idx(ob + m * Ls)] = a[m];
    }
  }
  if constexpr (!last) line_sync<WAVE>();
}

template <typename T, int DIR, class P, int S, class Idx, bool WAVE, class TW>
__device__ __forceinline__ void stages(C2<T> (&v)[P::E], int tl, C2<T>* lds, const Idx& idx, const TW& tw) {
  if constexpr (S < P::NS) {
    stage<T, DIR, P, S, Idx, WAVE, TW>(v, tl, lds, idx, tw);
    stages<T, DIR, P, S + 1, Idx, WAVE, TW>(v, tl, lds, idx, tw);
  }
}

// Transform one line held as v[q] = x[tl + TPL*q]; on return v[q] = X[tl + TPL*q].
// DIR = -1: X_k = sum x_n exp(-2 pi i nk/L); DIR = +1 the unnormalised inverse.
// When the plan has more than one stage the LDS slots of this line are
// overwritten and the caller must barrier before reusing them.  WAVE = true
// only for lines held by whole wavefronts (wave_lines<P>()).
template <typename T, int DIR, class P, class Idx, bool WAVE = false>
__device__ __forceinline__ void line_fft(C2<T> (&v)[P::E], int tl, C2<T>* lds, const Idx& idx,
                                         const C2<T>* __restrict__ tw) {
  stages<T, DIR, P, 0, Idx, WAVE, TwTable<T>>(v, tl, lds, idx, TwTable<T>{tw});
}
// the same with the stage twiddles held in registers (rt.load(table, tl) first)
template <typename T, int DIR, class P, class Idx, bool WAVE = false>
__device__ __forceinline__ void line_fft(C2<T> (&v)[P::E], int tl, C2<T>* lds, const Idx& idx,
                                         const TwRegs<T, P>& rt) {
  stages<T, DIR, P, 0, Idx, WAVE, TwRegs<T, P>>(v, tl, lds, idx, rt);
}

// LDS slot maps
struct IdxLineMajor {  // slot(b, e) = b*LP + e + e/16 (breaks power-of-two strides)
  int base;
  __device__ __forceinline__ int operator()(int e) const { return base + e + (e >> 4); }
};
struct IdxElemMajor {  // slot(b, e) = e*B + b (lines of one tile are adjacent)
  int b, B;
  __device__ __forceinline__ int operator()(int e) const { return e * B + b; }
};

template <int L> __host__ __device__ constexpr int line_pitch() { return L + (L >> 4) + 1; }

}  // namespace mm

// fft_pass.h -- the three axis passes of the 3D real FFT.
//
// Real fields are (Nx, Ny, Nz) with x fastest.  Spectra use the layout
//   S[kz][kx][ky]   (ky fastest, kx in [0, Nx/2])
// so that the y pass works on contiguous lines and the z pass on columns
// that are contiguous across (kx, ky).  The x pass does the transpose
// between the two layouts on chip: it reads B whole real rows and writes, for
// every kx, the B consecutive y positions of the transposed plane.
//
//   forward:  x_r2c (real rows -> S, scaled) -> y (c2c) -> z (c2c)
//   inverse:  z (c2c) -> y (c2c) -> x_c2r (S -> real rows)
//
// x_r2c packs a real row of Nx = 2L points as L complex points
// z_n = x_{2n} + i x_{2n+1}, runs an L-point FFT and untangles
//   X_k = (Z_k + conj Z_{L-k})/2 + w_N^k (Z_k - conj Z_{L-k})/(2i),  k = 0..L.
// x_c2r inverts that with the imaginary parts of X_0 and X_L ignored (the
// c2r semantics of FFTW that P3DFFT's btran_c2r inherits).
#pragma once
#include "fft_core.h"

namespace mm {

// Field pointers of one batched launch.  gate (may be null) lets a kernel
// that was recorded into a HIP graph skip its work when the device-side
// control of the step says so (the whole workgroup returns before any barrier).
template <typename T> struct Batch {
  const void* in[kMaxBatch];
  void* out[kMaxBatch];
  const int* gate;
};
__device__ __forceinline__ bool gated_off(const int* g) { return g != nullptr && *g == 0; }

__host__ __device__ constexpr int pow2floor(int x) {
  int p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

constexpr int kLdsBudget = 65536;
constexpr int kMaxThreads = 256;

// lines per workgroup, line-major LDS (x and y passes)
template <typename T, class P> struct TileLM {
  static constexpr int L = P::L, TPL = P::TPL;
  static constexpr int B = pow2floor(cmin(kMaxThreads / TPL, kLdsBudget / (line_pitch<L>() * (int)sizeof(C2<T>))));
  static constexpr int THREADS = B * TPL;
  static constexpr int LDS = B * line_pitch<L>();
};
// columns per workgroup, element-major LDS (z pass)
template <typename T, class P> struct TileEM {
  static constexpr int L = P::L, TPL = P::TPL;
  static constexpr int B = pow2floor(cmin(kMaxThreads / TPL, kLdsBudget / (L * (int)sizeof(C2<T>))));
  static constexpr int THREADS = B * TPL;
  static constexpr int LDS = B * L;
};

// ------------------------------------------------------------------ x r2c
// grid (ceil(ny/B), nz, nfields); L = Nx/2
template <typename T, class P>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS))
k_x_r2c(Batch<T> bt, const C2<T>* __restrict__ twL, const C2<T>* __restrict__ twN, int ny, T scale) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, B = TileLM<T, P>::B, LP = line_pitch<L>();
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  if (gated_off(bt.gate)) return;
  const int f = blockIdx.z, z = blockIdx.y, y0 = blockIdx.x * B, t = threadIdx.x;
  const T* __restrict__ in = (const T*)bt.in[f];
  C2<T>* __restrict__ out = (C2<T>*)bt.out[f];
  {
    const int b = t / TPL, tl = t % TPL, y = y0 + b;
    C2<T> v[E];
    if (y < ny) {
      const C2<T>* row = (const C2<T>*)(in + ((size_t)z * ny + y) * (size_t)(2 * L));
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = row[tl + TPL * q];
    } else {
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = mk<T>(0, 0);
    }
    IdxLineMajor idx{b * LP};
    line_fft<T, -1, P, IdxLineMajor, wave_lines<P>()>(v, tl, lds, idx, twL);
    if constexpr (P::NS > 1) line_sync<wave_lines<P>()>();
#pragma unroll
    for (int q = 0; q < E; ++q) lds[idx(tl + TPL * q)] = v[q];
  }
  __syncthreads();
  {
    const int b = t % B, r = t / B, y = y0 + b;
    if (y < ny) {
      const T hs = (T)0.5 * scale;
#pragma unroll
      for (int i = 0; i <= E; ++i) {
        const int kx = r + TPL * i;
        if (kx > L) break;
        const int k1 = kx == L ? 0 : kx, k2 = kx == 0 ? 0 : L - kx;
        C2<T> zk = lds[b * LP + k1 + (k1 >> 4)];
        C2<T> zc = lds[b * LP + k2 + (k2 >> 4)];
        // E = (zk + conj zc)/2 ; O = (zk - conj zc)/(2i)
        C2<T> ev = mk<T>(zk.x + zc.x, zk.y - zc.y);
        C2<T> od = mk<T>(zk.y + zc.y, -(zk.x - zc.x));
        C2<T> w = twN[kx];
        C2<T> x = cadd<T>(ev, cmul<T>(w, od));
        out[((size_t)z * (L + 1) + kx) * ny + y] = cscale<T>(x, hs);
      }
    }
  }
}

// ------------------------------------------------------------------ x c2r
template <typename T, class P>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS))
k_x_c2r(Batch<T> bt, const C2<T>* __restrict__ twL, const C2<T>* __restrict__ twN, int ny) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, B = TileLM<T, P>::B, LP = line_pitch<L>();
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  if (gated_off(bt.gate)) return;
  const int f = blockIdx.z, z = blockIdx.y, y0 = blockIdx.x * B, t = threadIdx.x;
  const C2<T>* __restrict__ in = (const C2<T>*)bt.in[f];
  T* __restrict__ out = (T*)bt.out[f];
  {
    // kx = r + TPL*i covers 0..L in E+1 rounds; unrolled so the loads overlap
    const int b = t % B, r = t / B, y = y0 + b;
    C2<T> x[E + 1];
#pragma unroll
    for (int i = 0; i <= E; ++i) {
      const int kx = r + TPL * i;
      x[i] = (y < ny && kx <= L) ? in[((size_t)z * (L + 1) + kx) * ny + y] : mk<T>(0, 0);
    }
#pragma unroll
    for (int i = 0; i <= E; ++i) {
      const int kx = r + TPL * i;
      if (kx <= L) lds[b * LP + kx + (kx >> 4)] = x[i];
    }
  }
  __syncthreads();
  const int b = t / TPL, tl = t % TPL, y = y0 + b;
  IdxLineMajor idx{b * LP};
  C2<T> v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int k = tl + TPL * q;
    C2<T> xk = lds[idx(k)], xc = lds[idx(L - k)];
    if (k == 0) { xk.y = 0; xc.y = 0; }
    C2<T> ev = mk<T>(xk.x + xc.x, xk.y - xc.y);  // xk + conj xc
    C2<T> dd = mk<T>(xk.x - xc.x, xk.y + xc.y);  // xk - conj xc
    C2<T> w = twN[k];
    w.y = -w.y;
    C2<T> od = cmul<T>(dd, w);
    v[q] = mk<T>(ev.x - od.y, ev.y + od.x);  // ev + i*od
  }
  line_sync<wave_lines<P>()>();  // the reads above touch this line only
  line_fft<T, +1, P, IdxLineMajor, wave_lines<P>()>(v, tl, lds, idx, twL);
  if (y < ny) {
    C2<T>* row = (C2<T>*)(out + ((size_t)z * ny + y) * (size_t)(2 * L));
#pragma unroll
    for (int q = 0; q < E; ++q) row[tl + TPL * q] = v[q];
  }
}

// -------------------------------------------------- y pass (contiguous lines)
// grid (ceil(nlines/B), 1, nfields); line l = kz*(Nx/2+1) + kx
template <typename T, class P, int DIR>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS))
k_y_pass(Batch<T> bt, const C2<T>* __restrict__ twL, int nlines) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, B = TileLM<T, P>::B, LP = line_pitch<L>();
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  if (gated_off(bt.gate)) return;
  const int f = blockIdx.z, t = threadIdx.x;
  const int b = t / TPL, tl = t % TPL, line = blockIdx.x * B + b;
  // in and out may be the same buffer (in-place pass): no __restrict__
  const C2<T>* in = (const C2<T>*)bt.in[f] + (size_t)line * L;
  C2<T>* out = (C2<T>*)bt.out[f] + (size_t)line * L;
  C2<T> v[E];
  if (line < nlines) {
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = in[tl + TPL * q];
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = mk<T>(0, 0);
  }
  IdxLineMajor idx{b * LP};
  line_fft<T, DIR, P, IdxLineMajor, wave_lines<P>()>(v, tl, lds, idx, twL);
  if (line < nlines) {
#pragma unroll
    for (int q = 0; q < E; ++q) out[tl + TPL * q] = v[q];
  }
}

// --------------------------------------------- column-tiled spectrum layout
// The fused z passes read the spectra in a tiled order instead of S: the
// columns (kx, ky) are grouped in tiles of kTW consecutive ky, and a tile is
// stored contiguously over kz,
//   Z[(kx*Ny/kTW + ky/kTW)][kz][ky % kTW],
// so a workgroup that owns whole tiles reads and writes contiguous blocks
// (Nz*kTW elements) however few columns it holds.  The y pass converts:
// forward S -> Z, inverse Z -> S (out of place).  Ny must divide by kTW.
constexpr int kTW = 4;
__host__ __device__ __forceinline__ size_t zidx(int kz, int kx, int ky, int ny, int nz) {
  return (((size_t)kx * (ny / kTW) + ky / kTW) * nz + kz) * kTW + (ky % kTW);
}
// the same for a column id col = kx*Ny + ky
__host__ __device__ __forceinline__ size_t zcol(int col, int kz, int nz) {
  return ((size_t)(col / kTW) * nz + kz) * kTW + (col % kTW);
}

// y pass between the layouts: MODE 1 reads S lines and writes Z, MODE 2 reads
// Z and writes S lines.  Lines run kz fastest (line = kx*Nz + kz) so that the
// lines of a workgroup fill whole kTW-runs of the Z side.
template <typename T, class P, int DIR, int MODE>
__global__ void __launch_bounds__((TileLM<T, P>::THREADS))
k_y_pass_t(Batch<T> bt, const C2<T>* __restrict__ twL, int nz, int H) {
  constexpr int L = P::L, TPL = P::TPL, E = P::E, B = TileLM<T, P>::B, LP = line_pitch<L>();
  __shared__ C2<T> lds[TileLM<T, P>::LDS];
  __shared__ C2<T> tw[L];
  if (gated_off(bt.gate)) return;
  const int f = blockIdx.z, t = threadIdx.x;
  for (int e = t; e < L; e += TileLM<T, P>::THREADS) tw[e] = twL[e];
  __syncthreads();
  const int b = t / TPL, tl = t % TPL, line = blockIdx.x * B + b;
  const bool ok = line < nz * H;
  const int kx = ok ? line / nz : 0, kz = ok ? line % nz : 0;
  const C2<T>* __restrict__ in = (const C2<T>*)bt.in[f];
  C2<T>* __restrict__ out = (C2<T>*)bt.out[f];
  const size_t srow = ((size_t)kz * H + kx) * L;
  C2<T> v[E];
  if (ok) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int ky = tl + TPL * q;
      v[q] = MODE == 1 ? in[srow + ky] : in[zidx(kz, kx, ky, L, nz)];
    }
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = mk<T>(0, 0);
  }
  IdxLineMajor idx{b * LP};
  line_fft<T, DIR, P, IdxLineMajor, wave_lines<P>()>(v, tl, lds, idx, tw);
  if (ok) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int ky = tl + TPL * q;
      if (MODE == 1) out[zidx(kz, kx, ky, L, nz)] = v[q];
      else out[srow + ky] = v[q];
    }
  }
}

// re-order nf spectra S -> Z (MODE 1) or Z -> S (MODE 2), out of place
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_perm_sz(const C2<T>* __restrict__ in, C2<T>* __restrict__ out, int nf,
                                                 int ny, int nz, int H) {
  const size_t nc = (size_t)H * ny * nz;
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < nc * nf; i += (size_t)gridDim.x * 256) {
    const size_t f = i / nc, r = i % nc;
    const int ky = (int)(r % ny);
    const size_t l = r / ny;
    const int kx = (int)(l % H), kz = (int)(l / H);
    const size_t z = f * nc + zidx(kz, kx, ky, ny, nz);
    if (MODE == 1) out[z] = in[i];
    else out[i] = in[z];
  }
}

// ----------------------------------------------------- z pass (strided lines)
// grid (ceil(ncols/B), 1, nfields); column c = kx*Ny + ky, stride ncols
template <typename T, class P, int DIR>
__global__ void __launch_bounds__((TileEM<T, P>::THREADS))
k_z_pass(Batch<T> bt, const C2<T>* __restrict__ twL, int ncols) {
  constexpr int TPL = P::TPL, E = P::E, B = TileEM<T, P>::B;
  __shared__ C2<T> lds[TileEM<T, P>::LDS];
  if (gated_off(bt.gate)) return;
  const int f = blockIdx.z, t = threadIdx.x;
  const int b = t % B, tl = t / B, col = blockIdx.x * B + b;
  const C2<T>* in = (const C2<T>*)bt.in[f] + col;
  C2<T>* out = (C2<T>*)bt.out[f] + col;
  C2<T> v[E];
  if (col < ncols) {
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = in[(size_t)(tl + TPL * q) * ncols];
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = mk<T>(0, 0);
  }
  IdxElemMajor idx{b, B};
  line_fft<T, DIR, P>(v, tl, lds, idx, twL);
  if (col < ncols) {
#pragma unroll
    for (int q = 0; q < E; ++q) out[(size_t)(tl + TPL * q) * ncols] = v[q];
  }
}

}  // namespace mm

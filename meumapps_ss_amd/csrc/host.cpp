// host.cpp -- error strings, HIP error capture, and the host-side initial
// condition helpers (ran_2.f restated; initial fields stay a host concern).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <thread>
#include <vector>
#include <cstdio>
#include <cstring>

#include "mm_common.h"

namespace mm {

static thread_local char g_last_hip[256] = "";

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  std::snprintf(g_last_hip, sizeof(g_last_hip), "%s: %s (%s:%d)", hipGetErrorString(e), what,
                file, line);
  return MM_ERR_HIP;
}

}  // namespace mm

extern "C" const char* mm_errstr(int code) {
  switch (code) {
    case MM_OK: return "ok";
    case MM_ERR_ZSOLVE: return "Poor cond. no; halting program (flag 2: KKS interface-composition matrix)";
    case MM_ERR_QINV: return "Poor cond. no; halting program (flag 3: acoustic tensor)";
    case MM_ERR_CINV_ITER: return "Poor cond. no; halting program (flag 4: compliance mix, elastic iteration)";
    case MM_ERR_CINV_ENER: return "Poor cond. no; halting program (flag 6: compliance mix, elastic energy)";
    case MM_ERR_NEWFAC: return "newfac is too small ... trouble converging";
    case MM_ERR_ARG: return "invalid argument or unsupported configuration";
    case MM_ERR_NOMEM: return "device allocation failed";
    case MM_ERR_HIP: return mm::g_last_hip[0] ? mm::g_last_hip : "HIP runtime error";
    case MM_ERR_NODEV: return "no GPU visible";
  }
  return "unknown error";
}

// ran_2.f:1-42.  state[0] = idum2, state[1] = iy, state[2..33] = iv(1..32),
// state[34] = 1 once the DATA initialisation (idum2 = 123456789, iv = 0,
// iy = 0) has been applied.  A zero-filled state is the program-start state.
namespace {

struct Ran2 {
  int32_t idum, idum2, iy, iv[32];
  Ran2(const int32_t* idum_p, const int32_t* state) {
    idum = *idum_p;
    if (state[34] == 0) {
      idum2 = 123456789;
      iy = 0;
      for (int j = 0; j < 32; ++j) iv[j] = 0;
    } else {
      idum2 = state[0];
      iy = state[1];
      for (int j = 0; j < 32; ++j) iv[j] = state[2 + j];
    }
  }
  void store(int32_t* idum_p, int32_t* state) const {
    *idum_p = idum;
    state[0] = idum2;
    state[1] = iy;
    for (int j = 0; j < 32; ++j) state[2 + j] = iv[j];
    state[34] = 1;
  }
  double next() {
    const int32_t im1 = 2147483563, im2 = 2147483399, imm1 = im1 - 1, ia1 = 40014, ia2 = 40692,
                  iq1 = 53668, iq2 = 52774, ir1 = 12211, ir2 = 3791, ntab = 32,
                  ndiv = 1 + imm1 / ntab;
    // am = 1./im1 and rnmx = 1.-eps are evaluated with the default-real literal
    // 1. (binary32): im1 rounds to 2147483648 in binary32, so am = 2^-31.
    const double am = (double)(1.0f / (float)im1);
    const double eps = 1.2e-7, rnmx = 1.0 - eps;
    int32_t k;
    if (idum <= 0) {
      idum = std::max(-idum, 1);
      idum2 = idum;
      for (int j = ntab + 8; j >= 1; --j) {
        k = idum / iq1;
        idum = ia1 * (idum - k * iq1) - k * ir1;
        if (idum < 0) idum += im1;
        if (j <= ntab) iv[j - 1] = idum;
      }
      iy = iv[0];
    }
    k = idum / iq1;
    idum = ia1 * (idum - k * iq1) - k * ir1;
    if (idum < 0) idum += im1;
    k = idum2 / iq2;
    idum2 = ia2 * (idum2 - k * iq2) - k * ir2;
    if (idum2 < 0) idum2 += im2;
    const int32_t j = 1 + iy / ndiv;
    iy = iv[j - 1] - idum2;
    iv[j - 1] = idum;
    if (iy < 1) iy += imm1;
    return std::min(am * (double)iy, rnmx);
  }
  // ModFunc.f:77-100: v1, v2 uniform on (-1, 1); redrawn until 0 < rsq < 1;
  // the second number of the polar pair, v2 * sqrt(-2 ln rsq / rsq), times sd
  double langevin(double sd) {
    for (;;) {
      double v1 = next();
      double v2 = next();
      v1 = 2.0 * v1 - 1.0;
      v2 = 2.0 * v2 - 1.0;
      const double v1s = v1 * v1, v2s = v2 * v2;
      double rsq = v1s + v2s;
      if (rsq > 0.0 && rsq < 1.0) {
        const double lg = std::log(rsq);
        rsq = std::sqrt(-2.0 * lg / rsq);
        const double h = v2 * rsq;
        return h * sd;
      }
    }
  }
};

}  // namespace

extern "C" double mm_ran2(int32_t* idum, int32_t* state) {
  Ran2 g(idum, state);
  const double r = g.next();
  g.store(idum, state);
  return r;
}

extern "C" int mm_ran2_fill(int32_t* idum, int32_t* state, int64_t n, double* out) {
  if (!idum || !state || n < 0 || (n > 0 && !out)) return MM_ERR_ARG;
  Ran2 g(idum, state);
  for (int64_t i = 0; i < n; ++i) out[i] = g.next();
  g.store(idum, state);
  return 0;
}

extern "C" double mm_langevin(int32_t* idum, int32_t* state, double sd) {
  Ran2 g(idum, state);
  const double r = g.langevin(sd);
  g.store(idum, state);
  return r;
}

namespace {

// x * a^k mod m (exact in 64-bit: m < 2^31)
int64_t mod_pow_mul(int64_t x, int64_t a, uint64_t k, int64_t m) {
  int64_t r = 1 % m, b = a % m;
  while (k) {
    if (k & 1) r = (r * b) % m;
    b = (b * b) % m;
    k >>= 1;
  }
  return (x * r) % m;
}

int noise_threads() {
  unsigned hw = std::thread::hardware_concurrency();
  int t = hw ? (int)hw : 1;
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) t = std::min(t, v);
  }
  return std::max(1, std::min(t, 16));
}

template <class F> void parallel_for(int nt, size_t n, F&& f) {
  if (nt <= 1 || n < 4096) { f(0, n); return; }
  std::vector<std::thread> th;
  const size_t step = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const size_t lo = std::min(n, (size_t)t * step), hi = std::min(n, lo + step);
    if (lo < hi) th.emplace_back([&f, lo, hi] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

namespace mm {
// Multi_MEUMAPPS.f:554-562 for one process: points in linear (z, y, x) order,
// variants innermost; acc is SoA (V fields of n3).
//
// The draws are those of mm_langevin called V*n3 times in a row, computed in
// blocks: the two multiplicative streams of ran_2 (idum, idum2) are
// independent of the shuffle, so threads produce them from jump-ahead
// starting points (x a^k mod m, exact in 64-bit, the same integers as the
// Schrage steps); one pass in order runs the shuffle table, forms the
// uniforms and the polar pairs and keeps the accepted ones, stopping at the
// last draw needed; threads then evaluate v2 sqrt(-2 ln rsq / rsq) sd with the
// scalar libm calls.  The generator state left behind is the sequential one.
void noise_add(int32_t* idum, int32_t* state, double sd, int V, size_t n3, double* acc) {
  Ran2 g(idum, state);
  const size_t N = (size_t)V * n3;
  const int nt = noise_threads();
  if (N < ((size_t)1 << 16) || nt == 1) {
    for (size_t i = 0; i < n3; ++i)
      for (int v = 0; v < V; ++v) acc[(size_t)v * n3 + i] += g.langevin(sd);
    g.store(idum, state);
    return;
  }
  const int64_t m1 = 2147483563, m2 = 2147483399, a1 = 40014, a2 = 40692;
  const int32_t imm1 = (int32_t)(m1 - 1);
  constexpr int32_t kNdiv = 1 + (2147483563 - 1) / 32;
  const double am = (double)(1.0f / (float)m1), rnmx = 1.0 - 1.2e-7;
  size_t done = 0;  // draws produced so far, in (point, variant) order
  if (g.idum <= 0) {
    // the first draw initialises the table: take it on the scalar path
    acc[0] += g.langevin(sd);
    done = 1;
  }
  const size_t U = (size_t)1 << 22;  // uniforms per block (even)
  std::vector<int32_t> s1(U), s2(U);
  std::vector<double> pv2(U / 2 + 1), prsq(U / 2 + 1);
  while (done < N) {
    // 1. the two streams for the next U uniforms
    const int64_t x1 = g.idum, x2 = g.idum2;
    parallel_for(nt, U, [&](size_t lo, size_t hi) {
      int64_t y1 = mod_pow_mul(x1, a1, lo + 1, m1), y2 = mod_pow_mul(x2, a2, lo + 1, m2);
      for (size_t n = lo; n < hi; ++n) {
        s1[n] = (int32_t)y1;
        s2[n] = (int32_t)y2;
        y1 = (y1 * a1) % m1;
        y2 = (y2 * a2) % m2;
      }
    });
    // 2. shuffle, uniforms, polar pairs, in order; stop after the last needed draw
    // (branch-free: the sign test of iy and the acceptance test are data
    // dependent and would mispredict about half and a fifth of the time)
    // The table lives in a local array and the streams behind restrict
    // pointers, so that its stores do not order the stream loads; iy >= 0
    // here, so the table index is an unsigned division by the constant kNdiv.
    const size_t need = N - done;
    size_t used = 0, cnt = 0;
    int32_t iv[32];
    for (int j = 0; j < 32; ++j) iv[j] = g.iv[j];
    const int32_t* __restrict S1 = s1.data();
    const int32_t* __restrict S2 = s2.data();
    double* __restrict PV = pv2.data();
    double* __restrict PR = prsq.data();
    uint32_t iy = (uint32_t)g.iy;
    while (used + 2 <= U && cnt < need) {
      double u[2];
      for (int h = 0; h < 2; ++h) {
        const size_t n = used + h;
        const uint32_t j = iy / (uint32_t)kNdiv;
        int32_t t = iv[j] - S2[n];
        iv[j] = S1[n];
        t += t < 1 ? imm1 : 0;
        iy = (uint32_t)t;
        const double x = am * (double)t;
        u[h] = x < rnmx ? x : rnmx;
      }
      used += 2;
      const double v1 = 2.0 * u[0] - 1.0, v2 = 2.0 * u[1] - 1.0;
      const double rsq = v1 * v1 + v2 * v2;
      PV[cnt] = v2;
      PR[cnt] = rsq;
      cnt += (rsq > 0.0) & (rsq < 1.0);
    }
    for (int j = 0; j < 32; ++j) g.iv[j] = iv[j];
    g.iy = (int32_t)iy;
    g.idum = s1[used - 1];
    g.idum2 = s2[used - 1];
    // 3. the Gaussian numbers, added in (point, variant) order
    const size_t base = done;
    parallel_for(nt, cnt, [&](size_t lo, size_t hi) {
      for (size_t q = lo; q < hi; ++q) {
        const double lg = std::log(prsq[q]);
        const double r = std::sqrt(-2.0 * lg / prsq[q]);
        const double h = pv2[q] * r;
        const size_t k = base + q, i = k / (size_t)V;
        const int v = (int)(k % (size_t)V);
        acc[(size_t)v * n3 + i] += h * sd;
      }
    });
    done += cnt;
  }
  g.store(idum, state);
}
}  // namespace mm

extern "C" int mm_noise_accumulate(int32_t* idum, int32_t* state, double sd, int32_t nvar, int64_t n3,
                                   double* phi_noise) {
  if (!idum || !state || !phi_noise || nvar < 1 || n3 < 1) return MM_ERR_ARG;
  mm::noise_add(idum, state, sd, nvar, (size_t)n3, phi_noise);
  return 0;
}

extern "C" int mm_ic_uniform(int32_t nx, int32_t ny, int32_t nz, int32_t nvar, int32_t nel,
                             int32_t iseed, double amp, const double* cav, double* phi,
                             double* cons) {
  if (nx < 1 || ny < 1 || nz < 1 || nvar < 0 || nel < 0 || (nel > 0 && (!cav || !cons)) ||
      (nvar > 0 && !phi))
    return MM_ERR_ARG;
  int32_t state[35];
  std::memset(state, 0, sizeof(state));
  int32_t idum = iseed;
  const size_t n3 = (size_t)nx * ny * nz;
  for (size_t i = 0; i < n3; ++i) {
    for (int e = 0; e < nel; ++e) {
      double r = mm_ran2(&idum, state);
      cons[(size_t)e * n3 + i] = cav[e] + amp * (2.0 * r - 1.0);
    }
    for (int v = 0; v < nvar; ++v) {
      double r = mm_ran2(&idum, state);
      phi[(size_t)v * n3 + i] = amp * (2.0 * r - 1.0);
    }
  }
  return 0;
}

// host.cpp -- error strings, HIP error capture, and the host-side initial
// condition helpers (ran_2.f restated; initial fields stay a host concern).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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

extern "C" double mm_langevin(int32_t* idum, int32_t* state, double sd) {
  Ran2 g(idum, state);
  const double r = g.langevin(sd);
  g.store(idum, state);
  return r;
}

namespace mm {
// Multi_MEUMAPPS.f:554-562 for one process: points in linear (z, y, x) order,
// variants innermost; acc is SoA (V fields of n3)
void noise_add(int32_t* idum, int32_t* state, double sd, int V, size_t n3, double* acc) {
  Ran2 g(idum, state);
  for (size_t i = 0; i < n3; ++i)
    for (int v = 0; v < V; ++v) acc[(size_t)v * n3 + i] += g.langevin(sd);
  g.store(idum, state);
}
}  // namespace mm

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

// mm_common.h -- shared host-side helpers of the C-ABI library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/meumapps_ss.h"

namespace mm {

// Record the first HIP error seen (for mm_errstr) and map it to MM_ERR_HIP.
int hip_fail(hipError_t e, const char* what, const char* file, int line);
// Multi_MEUMAPPS.f:554-562 for one process: one langevin draw (mm_langevin) per
// point in linear (z, y, x) order and variant (innermost) added to acc[v*n3 + i]
void noise_add(int32_t* idum, int32_t* state, double sd, int V, size_t n3, double* acc);

#define MM_HIP(call)                                                  \
  do {                                                                \
    hipError_t _e = (call);                                           \
    if (_e != hipSuccess) return ::mm::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define MM_TRY(call)          \
  do {                        \
    int _rc = (call);         \
    if (_rc != 0) return _rc; \
  } while (0)

}  // namespace mm

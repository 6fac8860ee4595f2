// step.hip -- placeholder until the step pipeline lands.
#include "mm_common.h"

extern "C" int mm_create(const mm_params*, int, mm_ctx** out) { if (out) *out = nullptr; return MM_ERR_ARG; }
extern "C" void mm_destroy(mm_ctx*) {}
extern "C" int mm_upload(mm_ctx*, const double*, const double*, const double*, const double*, const double*) { return MM_ERR_ARG; }
extern "C" int mm_download(mm_ctx*, double*, double*, double*, double*, double*) { return MM_ERR_ARG; }
extern "C" int mm_step(mm_ctx*, int32_t, mm_diag*) { return MM_ERR_ARG; }
extern "C" int mm_get_step(const mm_ctx*) { return 0; }
extern "C" int mm_set_step(mm_ctx*, int32_t) { return MM_ERR_ARG; }
extern "C" int mm_field(mm_ctx*, int32_t, void**, int32_t*, int32_t*) { return MM_ERR_ARG; }
extern "C" int mm_field_to_host(mm_ctx*, int32_t, double*) { return MM_ERR_ARG; }

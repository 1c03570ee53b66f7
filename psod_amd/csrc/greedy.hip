#include "psod_internal.h"
extern "C" int psod_host_spherover(psod_ctx *, int64_t, const float *, const int64_t *, const int32_t *, int64_t, const psod_greedy_params *, int64_t, int64_t *, int64_t *, double *, double *, int32_t *, uint8_t *, int64_t *) { return psod_set_error(PSOD_E_ARG, "not built yet"); }

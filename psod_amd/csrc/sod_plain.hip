#include "psod_internal.h"
extern "C" int psod_sod_fof_plain(psod_ctx *, const psod_sod_params *, const psod_catalogue *, int64_t, const int32_t *, const int32_t *, const psod_halo_out *) { return psod_set_error(PSOD_E_ARG, "not built yet"); }

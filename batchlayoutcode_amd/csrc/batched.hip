// Batched small graphs: placeholder until the persistent batched kernel lands.
#include "bl_internal.h"

namespace bl {
int run_batched(int32_t, const int32_t *, const int64_t *, const int32_t *, const int32_t *, double *,
                const bl_params *, double *, bl_stats *) {
    return BL_E_ARG;
}
}  // namespace bl

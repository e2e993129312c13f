// Reference block Davidson on the GPU operator (placeholder until the solver lands).
#include <hip/hip_runtime.h>

#include "common.hpp"

extern "C" bd_status bd_davidson_solve_dev(bd_gamma_op*, const double*, int64_t, const double*, const double*,
                                           const bd_davidson_options*, double*, double*, double*,
                                           int64_t, int32_t*, bd_davidson_record*) {
  BD_GUARD(bd::fail(BD_ERR_APPLICATION, "Davidson solver not built"));
}

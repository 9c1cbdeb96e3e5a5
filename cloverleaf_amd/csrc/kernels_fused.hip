// Fused sweeps for CLV_PLAN_FUSED.
#include "clv_internal.h"

namespace clv {

clv_status fused_timestep_kernels(clv_chunk *, const clv_config *, double *) {
  return set_error(CLV_ERR_ARG, "CLV_PLAN_FUSED is not available in this build");
}
clv_status fused_lagrange(clv_chunk *, double) { return set_error(CLV_ERR_ARG, "CLV_PLAN_FUSED is not available in this build"); }
clv_status fused_advection(clv_chunk *, int) { return set_error(CLV_ERR_ARG, "CLV_PLAN_FUSED is not available in this build"); }

}  // namespace clv

// Learned gauge-covariant preconditioner (flexible outer solver).
#include "ii_ops.h"

namespace ii {
void learned_release(ii_context* ctx) { (void)ctx; }

ii_status learned_propagator(ii_context* ctx, double, const ii_prop_params*, const ii_field*,
                             ii_field*, ii_solver_output*) {
  return set_error(ctx, II_ERR_STATE, "no trained preconditioner on this context");
}
}  // namespace ii

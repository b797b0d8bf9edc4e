// Path-replay adjoint of the interior estimator and edge-sampling boundary
// terms (src/path_integrator.cc:80-353).
#include <hip/hip_runtime.h>

#include "render.h"

namespace psdr {

int render_backward(psdr_scene* s, const psdr_config& cfg, const psdr_camera& cam, const float* dL_dI,
                    float* grad_v, float* grad_t, float* grad_e, hipStream_t st) {
  (void)s; (void)cfg; (void)cam; (void)dL_dI; (void)grad_v; (void)grad_t; (void)grad_e; (void)st;
  set_error(PSDR_ERR_STATE, "render_backward: not built yet");
  return PSDR_ERR_STATE;
}

}  // namespace psdr

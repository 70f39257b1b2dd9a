// Path-tracing entry points (filled in by the wavefront renderer).
#include <hip/hip_runtime.h>

#include "scene_internal.h"

using namespace xyzhip;

extern "C" {

int xyz_render(xyz_scene*, const xyz_camera*, const xyz_render_params*, float*, xyz_render_stats*,
               void*) {
  set_error("xyz_render: not built");
  return XYZ_E_UNSUPPORTED;
}

int xyz_render_backward(xyz_scene*, const xyz_camera*, const xyz_render_params*, const float*,
                        float*, float*, void*) {
  set_error("xyz_render_backward: not built");
  return XYZ_E_UNSUPPORTED;
}

int xyz_render_batch(uint32_t, xyz_scene* const*, const xyz_camera*, const xyz_render_params*,
                     float* const*, const xyz_batch_tuning*, xyz_render_stats*, void*) {
  set_error("xyz_render_batch: not built");
  return XYZ_E_UNSUPPORTED;
}

int xyz_batch_autotune(uint32_t, xyz_scene* const*, const xyz_camera*, const xyz_render_params*,
                       float* const*, xyz_batch_tuning*, void*) {
  set_error("xyz_batch_autotune: not built");
  return XYZ_E_UNSUPPORTED;
}

}  // extern "C"

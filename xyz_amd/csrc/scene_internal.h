// Host-side scene object shared by the C ABI translation units.
#ifndef XYZ_SCENE_INTERNAL_H_
#define XYZ_SCENE_INTERNAL_H_

#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.h"
#include "pt_device.h"
#include "xyz_internal.h"

struct xyz_scene {
  xyzhip::DevScene dev;
  xyz_scene_info info;
  xyzhip::FlatTree tree;  // host copy (released after upload)
  std::vector<void*> allocs;
  uint64_t n_tris = 0;
  uint64_t n_texels = 0;
};

// Importon map of importance-driven path tracing: device arrays sorted by grid
// cell, plus the host copy in map order (path order, then vertex order).
struct xyz_importon_map {
  xyzhip::DevImportons dev;
  const xyz_scene* scene = nullptr;
  int device = -1;
  std::vector<float> host;  // 6 floats per importon: position, incoming direction
  std::vector<void*> allocs;
  uint32_t n_emit = 0;
  double seconds = 0;
};

namespace xyzhip {

int hip_fail(hipError_t e, const char* what);

// Stream-ordered allocation from the library's memory pool on the current
// device.  The pool keeps freed memory (release threshold: never), so the
// workspace of one call is reused by the next one on any stream instead of
// being returned to the driver at every synchronization.  Free with
// hipFreeAsync.
hipError_t pool_alloc(void** p, size_t bytes, hipStream_t stream);

// Large workspace: requests of 32 MiB and more outside stream capture come
// from blocks the library keeps per device for reuse by later calls (grown on
// demand, held until process exit); smaller ones, and any under capture, from
// the pool.  Release with big_free on the stream that last used the memory
// (also correct for pool memory).
hipError_t big_alloc(void** p, size_t bytes, hipStream_t stream);
void big_free(void* p, hipStream_t stream);

int launch_closest_hit(const DevScene& s, uint32_t n, const float* org, const float* dir,
                       int32_t* prim, float* t, float* f, float* g, uint8_t* back,
                       uint32_t* counters, unsigned long long* totals, hipStream_t stream);

}  // namespace xyzhip

#endif  // XYZ_SCENE_INTERNAL_H_

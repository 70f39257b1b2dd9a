// Internal declarations shared by the host builder and the HIP library.
#ifndef XYZ_INTERNAL_H_
#define XYZ_INTERNAL_H_

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/xyz_hip.h"

namespace xyzhip {

// Flattened kd-tree in the device layout (see xyz_kdtree_build in xyz_hip.h).
struct FlatTree {
  std::vector<uint32_t> nodes;       // 2 words per node, pre-order
  std::vector<uint32_t> leaf_prims;  // primitive id per leaf reference
  uint32_t depth = 0;                // deepest interior level + 1
  uint32_t max_leaf = 0;
  uint64_t n_leaves = 0;
  float bmin[3] = {0, 0, 0};
  float bmax[3] = {0, 0, 0};
  double build_seconds = 0;
  uint32_t max_depth_param = 16;
};

// SAH build with the reference's split rules (xyz/geom/kdtree.cpp:207-357).
// v0/v1/v2: 3*n floats.  Returns an empty string on success, else the error.
std::string build_flat_tree(uint32_t n, const float* v0, const float* v1, const float* v2,
                            uint32_t max_bucket, uint32_t max_depth, FlatTree* out);

// The same build on the current GPU (kdtree_gpu.hip); the same trees.
std::string build_flat_tree_device(uint32_t n, const float* v0, const float* v1, const float* v2,
                                   uint32_t max_bucket, uint32_t max_depth, FlatTree* out);

void set_error(const std::string& msg);

}  // namespace xyzhip

struct xyz_kdtree {
  xyzhip::FlatTree t;
  uint64_t n_tris = 0;
};

#endif  // XYZ_INTERNAL_H_

// Host computation of a triangle cluster's culling record (box, N, E), shared
// by the scene builder and the exactness test of cluster_may_hit.
#ifndef XYZ_CLUSTER_BOUNDS_H_
#define XYZ_CLUSTER_BOUNDS_H_

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace xyzhip {

// Smallest float >= x / largest float <= x.
inline float float_up(double x) {
  float f = float(x);
  if (double(f) < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}
inline float float_down(double x) {
  float f = float(x);
  if (double(f) > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}

// Records A = (v0, n.x), B = (e1, n.y), C = (e2, n.z) of n triangles ->
// lo = (box lo, N), hi = (box hi, E): the box of v0, v0 + e1, v0 + e2 (exact
// sums, rounded outward), N >= max |e1||e2|, E >= max(|e1|, |e2|).  A
// non-finite input makes N NaN, which disables culling of the cluster.
inline void cluster_bounds(const float4* A, const float4* B, const float4* C, uint32_t n,
                           float4* lo_out, float4* hi_out) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double N = 0.0, E = 0.0;
  bool bad = false;
  for (uint32_t m = 0; m < n; ++m) {
    double const av[3] = {A[m].x, A[m].y, A[m].z};
    double const bv[3] = {B[m].x, B[m].y, B[m].z};
    double const cv[3] = {C[m].x, C[m].y, C[m].z};
    for (int k = 0; k < 3; ++k) {
      double const pts[3] = {av[k], av[k] + bv[k], av[k] + cv[k]};
      for (double x : pts) {
        if (!(std::fabs(x) < 1e300)) bad = true;
        lo[k] = std::min(lo[k], x);
        hi[k] = std::max(hi[k], x);
      }
    }
    double const l1 = std::sqrt(bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2]);
    double const l2 = std::sqrt(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]);
    N = std::max(N, l1 * l2);
    E = std::max(E, std::max(l1, l2));
  }
  // 1e-12 relative slack over the double arithmetic above
  float const fN = bad ? NAN : float_up(N * (1.0 + 1e-12));
  float const fE = bad ? NAN : float_up(E * (1.0 + 1e-12));
  *lo_out = make_float4(float_down(lo[0]), float_down(lo[1]), float_down(lo[2]), fN);
  *hi_out = make_float4(float_up(hi[0]), float_up(hi[1]), float_up(hi[2]), fE);
}

}  // namespace xyzhip

#endif  // XYZ_CLUSTER_BOUNDS_H_

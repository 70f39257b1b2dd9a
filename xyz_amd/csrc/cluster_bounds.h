// Host computation of a triangle cluster's culling record, shared by the scene
// builder and the exactness test of cluster_may_hit.
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

// Records A = (v0, n.x), B = (e1, n.y), C = (e2, n.z) of n triangles -> the
// cluster's culling record lo = (box lo - K, P), hi = (box hi + K, Q).
//
// The box is that of v0, v0 + e1, v0 + e2 (exact sums, rounded outward).  With
// N >= max |e1||e2| and E >= max(|e1|, |e2|) over the triangles, u = 2^-24 and
// c = u / 1e-6, the record holds the parts of cluster_may_hit's margin that do
// not depend on the ray:
//   P >= 30 c N                 (30 lam / |d| at |det| = 1e-6)
//   Q >= 13 c N E               (13 lam E / |d| at |det| = 1e-6)
//   K >= 8 u E + 4 u (largest |box coordinate|) + 1e-30,
// so that the margin of a ray with |d| = dl and distance bound R is
// R (dl P + 8 u) + dl Q on top of the box already widened by K.  A non-finite
// input makes P NaN, which disables culling of the cluster.
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
  double const u = 0x1p-24, c = u / 1e-6;
  // 1e-12 relative slack over the double arithmetic below
  double const up = 1.0 + 1e-12;
  double big = 0.0;
  for (int k = 0; k < 3; ++k) big = std::max(big, std::max(std::fabs(lo[k]), std::fabs(hi[k])));
  double const K = (8.0 * u * E + 4.0 * u * big + 1e-30) * up;
  float const P = bad ? NAN : float_up(30.0 * c * N * up);
  float const Q = bad ? NAN : float_up(13.0 * c * N * E * up);
  *lo_out = make_float4(float_down(lo[0] - K), float_down(lo[1] - K), float_down(lo[2] - K), P);
  *hi_out = make_float4(float_up(hi[0] + K), float_up(hi[1] + K), float_up(hi[2] + K), Q);
}

}  // namespace xyzhip

#endif  // XYZ_CLUSTER_BOUNDS_H_

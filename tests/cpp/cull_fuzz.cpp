// Exactness test of the cluster culling used by the closest-hit traversal,
// run on the host with the production functions: tri_test (the reference's
// ray/triangle test, xyz/geom/triangle.cpp:127-165) and cluster_may_hit, with
// clusters built by cluster_bounds.  Whenever tri_test accepts a triangle with
// t in [t_min, t_max], cluster_may_hit must not rule out its cluster, and
// root_clip, given a root record holding the cluster's (its own record, or a
// union with a larger box and larger margins), must keep t in the query's
// narrowed interval and must not declare a miss.
//
// The rays aim at the hardest cases for the error bound: grazing rays whose
// |det| is just above the 1e-6 threshold, on triangles up to the size where
// culling switches itself off, intervals shrunk to the computed t itself,
// axis-parallel directions, origins on or inside the cluster box, smooth
// patches of a curved surface met at every tilt.  Also
// reports the largest observed (distance of o + t d from the box) / margin.
//
// usage: cull_fuzz <cases> <seed>; exit status 1 on a violation.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../xyz_amd/csrc/cluster_bounds.h"
#include "../../xyz_amd/csrc/device_common.h"

using namespace xyzhip;

namespace {

uint64_t g_state;
uint64_t next_u64() {  // splitmix64
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uni() { return double(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
double loguni(double a, double b) { return a * std::pow(b / a, uni()); }
double sym() { return 2.0 * uni() - 1.0; }

struct V { double x, y, z; };
V add(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V mul(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V unit(V a) { return mul(a, 1.0 / std::sqrt(dot(a, a))); }
V rand_dir() {
  for (;;) {
    V v{sym(), sym(), sym()};
    double const l = dot(v, v);
    if (l > 1e-4 && l <= 1.0) return mul(v, 1.0 / std::sqrt(l));
  }
}
float3 f3(V v) { return make_float3(float(v.x), float(v.y), float(v.z)); }

// Stored record of triangle (v0, v1, v2), as the scene builder makes it.
void make_record(float3 v0, float3 v1, float3 v2, float4* A, float4* B, float4* C) {
  float const b[3] = {v1.x - v0.x, v1.y - v0.y, v1.z - v0.z};
  float const c[3] = {v2.x - v0.x, v2.y - v0.y, v2.z - v0.z};
  float const nx = c[1] * b[2] - c[2] * b[1];
  float const ny = c[2] * b[0] - c[0] * b[2];
  float const nz = c[0] * b[1] - c[1] * b[0];
  *A = make_float4(v0.x, v0.y, v0.z, nx);
  *B = make_float4(b[0], b[1], b[2], ny);
  *C = make_float4(c[0], c[1], c[2], nz);
}

// What cluster_may_hit allows a computed hit point to lie outside the stored
// box (for the report only): the box's widening plus the t widening's length.
double margin(float4 L, float4 H, float3 o, float3 d) {
  double const u = 5.9604644775390625e-08;
  double const dl = std::sqrt(double(d.x) * d.x + double(d.y) * d.y + double(d.z) * d.z);
  double R2 = 0.0;
  double const oo[3] = {o.x, o.y, o.z}, lo[3] = {L.x, L.y, L.z}, hi[3] = {H.x, H.y, H.z};
  for (int k = 0; k < 3; ++k) {
    double const r = std::max(std::fabs(oo[k] - lo[k]), std::fabs(hi[k] - oo[k]));
    R2 += r * r;
  }
  double const tr = std::sqrt(R2) * (dl * L.w + 8 * u);
  return 2.0 * tr + dl * H.w;
}

double box_distance(float4 L, float4 H, V p) {
  double dx = std::max({double(L.x) - p.x, 0.0, p.x - double(H.x)});
  double dy = std::max({double(L.y) - p.y, 0.0, p.y - double(H.y)});
  double dz = std::max({double(L.z) - p.z, 0.0, p.z - double(H.z)});
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

struct Stats {
  uint64_t checks = 0, accepted = 0, violations = 0, culled = 0, cull_tries = 0, root_checks = 0;
  double worst = 0.0;  // max distance / margin over accepted hits
};

bool may_hit(float4 L, float4 H, float3 o, float3 d, float tmin, float tmax) {
  float dl;
  float3 inv;
  cull_ray(d, &dl, &inv);
  return cluster_may_hit(L, H, o, dl, inv, tmin, tmax);
}

// Tests every triangle of the cluster against the ray over [tmin, tmax].
void check(const std::vector<float4>& A, const std::vector<float4>& B, const std::vector<float4>& C,
           float3 o, float3 d, float tmin, float tmax, Stats* st) {
  float4 L, H;
  cluster_bounds(A.data(), B.data(), C.data(), uint32_t(A.size()), &L, &H);
  bool const keep = may_hit(L, H, o, d, tmin, tmax);
  st->cull_tries++;
  if (!keep) st->culled++;
  for (size_t m = 0; m < A.size(); ++m) {
    float t, f, g, det;
    st->checks++;
    if (!tri_test(o, d, A[m], B[m], C[m], tmin, tmax, &t, &f, &g, &det)) continue;
    st->accepted++;
    // the tightest interval around the accepted t must keep the cluster too
    bool const keep_t = may_hit(L, H, o, d, t, t);
    // the root interval of a query along this ray must keep t: with the
    // cluster's own record as the root, and with a union of it and a larger
    // box with larger P and Q
    bool keep_root = true;
    if (t >= kTMin) {
      for (int variant = 0; variant < 2; ++variant) {
        DevScene sc{};
        sc.root_lo = L;
        sc.root_hi = H;
        if (variant == 1) {
          float const e = float(loguni(1e-6, 10.0));
          sc.root_lo = make_float4(L.x - e * float(uni()), L.y - e * float(uni()), L.z - e * float(uni()),
                                   L.w * float(1.0 + uni()));
          sc.root_hi = make_float4(H.x + e * float(uni()), H.y + e * float(uni()), H.z + e * float(uni()),
                                   H.w * float(1.0 + uni()));
        }
        TraceState ts;
        trace_init(ts, o, d);
        root_clip(sc, ts);
        st->root_checks++;
        if (ts.in_leaf || t < ts.t_min || t > ts.t_max) keep_root = false;
      }
    }
    if (!keep || !keep_t || !keep_root) {
      st->violations++;
      if (st->violations <= 5)
        std::printf("VIOLATION: o=(%a,%a,%a) d=(%a,%a,%a) tmin=%a tmax=%a t=%a det=%a\n", o.x, o.y,
                    o.z, d.x, d.y, d.z, tmin, tmax, t, det);
    }
    double const mg = margin(L, H, o, d);
    V const p = add(V{o.x, o.y, o.z}, mul(V{d.x, d.y, d.z}, double(t)));
    double const dist = box_distance(L, H, p);
    // where culling is on: lam = |d| P / 30 <= 0.01
    if (mg > 0 && std::isfinite(mg) && L.w * std::sqrt(double(d.x) * d.x + double(d.y) * d.y +
                                                       double(d.z) * d.z) / 30.0 <= 0.01)
      st->worst = std::max(st->worst, dist / mg);
  }
}

// A grazing ray at a point of one triangle: |det| a small multiple of 1e-6.
void grazing_case(Stats* st) {
  double const S = loguni(0.01, 100.0);
  V const ctr{S * sym(), S * sym(), S * sym()};
  double const l1 = loguni(1e-4, 0.5), l2 = loguni(1e-4, 0.5);
  V const a = rand_dir();
  V b = rand_dir();
  if (uni() < 0.3) b = unit(add(a, mul(rand_dir(), loguni(1e-4, 1e-1))));  // a sliver
  float3 const v0 = f3(ctr), v1 = f3(add(ctr, mul(a, l1))), v2 = f3(add(ctr, mul(b, l2)));
  std::vector<float4> A(1), B(1), C(1);
  make_record(v0, v1, v2, &A[0], &B[0], &C[0]);
  V const n{A[0].w, B[0].w, C[0].w};
  double const nl = std::sqrt(dot(n, n));
  if (!(nl > 0)) return;
  V const nu = mul(n, 1.0 / nl);
  // target point in the triangle, direction along the plane tilted so that
  // d.n is a few times 1e-6
  double f = uni(), g = uni();
  if (f + g > 1) { f = 1 - f; g = 1 - g; }
  V const e1{B[0].x, B[0].y, B[0].z}, e2{C[0].x, C[0].y, C[0].z};
  V const P = add(V{v0.x, v0.y, v0.z}, add(mul(e1, f), mul(e2, g)));
  V inplane = rand_dir();
  inplane = unit(sub(inplane, mul(nu, dot(inplane, nu))));
  double const dscale = loguni(1e-3, 1e3);
  double const want = loguni(1e-6, 1e-4) * (uni() < 0.5 ? -1 : 1);  // target d.n
  double const tilt = want / (dscale * nl);
  V const dir = mul(unit(add(inplane, mul(nu, tilt))), dscale);
  double const dist = loguni(1e-4, 300.0);
  V const o = sub(P, mul(dir, dist / dscale));
  float3 const of = f3(o), df = f3(dir);
  check(A, B, C, of, df, 1e-5f, INFINITY, st);
  check(A, B, C, of, df, 0.0f, float(dist / dscale * (1 + 1e-4 * sym())), st);
}

// A cluster of up to 16 triangles around one point and rays from anywhere.
void cluster_case(Stats* st) {
  double const S = loguni(0.01, 50.0);
  V const ctr{S * sym(), S * sym(), S * sym()};
  double const rad = loguni(1e-3, 0.3);
  int const n = 1 + int(next_u64() % 16);
  bool const axis = uni() < 0.3;  // axis-aligned tessellated face
  std::vector<float4> A(n), B(n), C(n);
  for (int m = 0; m < n; ++m) {
    V const p = add(ctr, mul(rand_dir(), rad * uni()));
    double const l = loguni(1e-4, rad);
    V a = rand_dir(), b = rand_dir();
    if (axis) { a = {1, 0, 0}; b = {0, 0, 1}; }
    make_record(f3(p), f3(add(p, mul(a, l))), f3(add(p, mul(b, l * loguni(0.3, 3)))), &A[m], &B[m], &C[m]);
  }
  for (int r = 0; r < 8; ++r) {
    V o;
    double const kind = uni();
    if (kind < 0.3) o = add(ctr, mul(rand_dir(), rad * uni()));            // inside the box
    else if (kind < 0.5) o = V{A[0].x, A[0].y, A[0].z};                    // on a vertex
    else o = add(ctr, mul(rand_dir(), loguni(rad, 200.0)));                // outside
    V dir;
    double const dk = uni();
    if (dk < 0.5) dir = unit(sub(add(ctr, mul(rand_dir(), rad * uni())), o));  // at the cluster
    else dir = rand_dir();
    if (uni() < 0.2) {  // zero components
      int const z = int(next_u64() % 3);
      if (z == 0) dir.x = 0; else if (z == 1) dir.y = 0; else dir.z = 0;
      if (uni() < 0.3) { if (z == 0) dir.y = 0; else dir.z = 0; }
      if (dot(dir, dir) == 0) dir = {0, 1, 0};
    }
    if (axis && uni() < 0.3) dir.y = loguni(1e-7, 1e-3) * (uni() < 0.5 ? -1 : 1);  // grazing the face
    dir = mul(dir, loguni(1e-2, 1e2));
    float3 const of = f3(o), df = f3(dir);
    float const a = float(loguni(1e-5, 300.0)), b = float(loguni(1e-5, 300.0));
    check(A, B, C, of, df, 1e-5f, INFINITY, st);
    check(A, B, C, of, df, std::min(a, b), std::max(a, b), st);
  }
}

// A smooth patch: up to 16 triangles of a gently curved height field (normals
// within a few degrees of each other, the case where the normal cone bounds
// |det| far above 1e-6), rays at tilts from grazing to head-on, aimed at
// points of the patch, with intervals around the hit.
void patch_case(Stats* st) {
  double const S = loguni(0.01, 50.0);
  V const ctr{S * sym(), S * sym(), S * sym()};
  V const nz = rand_dir();
  V tx = unit(cross(nz, rand_dir()));
  V const ty = cross(nz, tx);
  double const h = loguni(1e-3, 0.5);           // cell size
  double const curv = loguni(1e-4, 1.0) / h;    // curvature of the field
  auto at = [&](double x, double y) {
    double const z = 0.5 * curv * (x * x + 0.7 * y * y);
    return add(ctr, add(add(mul(tx, x), mul(ty, y)), mul(nz, z)));
  };
  int const cells = 1 + int(next_u64() % 8);     // 2 triangles each
  std::vector<float4> A, B, C;
  for (int k = 0; k < cells; ++k) {
    double const x0 = h * (k % 4 - 2), y0 = h * (k / 4 - 1);
    V const p00 = at(x0, y0), p10 = at(x0 + h, y0), p01 = at(x0, y0 + h), p11 = at(x0 + h, y0 + h);
    A.resize(A.size() + 2); B.resize(B.size() + 2); C.resize(C.size() + 2);
    size_t const m = A.size() - 2;
    make_record(f3(p00), f3(p10), f3(p11), &A[m], &B[m], &C[m]);
    make_record(f3(p00), f3(p11), f3(p01), &A[m + 1], &B[m + 1], &C[m + 1]);
  }
  for (int r = 0; r < 6; ++r) {
    size_t const m = next_u64() % A.size();
    double f = uni(), g = uni();
    if (f + g > 1) { f = 1 - f; g = 1 - g; }
    V const P = add(V{A[m].x, A[m].y, A[m].z}, add(mul(V{B[m].x, B[m].y, B[m].z}, f),
                                                  mul(V{C[m].x, C[m].y, C[m].z}, g)));
    // tilt from the surface: log-uniform from grazing (1e-6) to head-on
    double const tilt = loguni(1e-6, 1.0);
    V const inplane = unit(add(mul(tx, sym()), mul(ty, sym())));
    V dir = unit(add(mul(inplane, std::sqrt(std::max(0.0, 1.0 - tilt * tilt))), mul(nz, tilt * (uni() < 0.5 ? -1 : 1))));
    dir = mul(dir, loguni(1e-2, 1e2));
    double const dist = loguni(1e-4, 300.0);
    V const o = sub(P, mul(dir, dist / std::sqrt(dot(dir, dir))));
    float3 const of = f3(o), df = f3(dir);
    check(A, B, C, of, df, 1e-5f, INFINITY, st);
    float const a = float(loguni(1e-5, 300.0)), b = float(loguni(1e-5, 300.0));
    check(A, B, C, of, df, std::min(a, b), std::max(a, b), st);
  }
}

}  // namespace

int main(int argc, char** argv) {
  uint64_t const cases = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 100000;
  g_state = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
  Stats st;
  for (uint64_t i = 0; i < cases; ++i) {
    grazing_case(&st);
    cluster_case(&st);
    patch_case(&st);
  }
  std::printf("checks %llu accepted %llu root checks %llu violations %llu culled %.4f worst distance/margin %.3e\n",
              (unsigned long long)st.checks, (unsigned long long)st.accepted,
              (unsigned long long)st.root_checks, (unsigned long long)st.violations,
              double(st.culled) / double(st.cull_tries), st.worst);
  return st.violations ? 1 : 0;
}

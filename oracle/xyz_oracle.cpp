// xyz_oracle.cpp -- CPU restatement of the reference renderer's hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product is
// compared against; only the test suite, the smoke check and the CPU-baseline
// leg of the benchmark load it.  Nothing in the product links or calls it.
//
// It restates, operation for operation, what the reference computes:
//   * Triangle precompute and double-faced ray/triangle test
//       xyz/geom/triangle.cpp:10-60, 75-85, 127-165; xyz/geom/triangle.hpp:122-156
//   * SAH kd-tree build (recursive, per-node event sort)
//       xyz/geom/kdtree.cpp:207-357
//   * recursive kd traversal, first-hit-leaf early exit, last-accept-wins leaves
//       xyz/geom/kdtree.cpp:82-122, 140-177, 366-372; xyz/geom/intersection.hpp:11-14
//   * brute-force closest hit                          xyz/core/scene.hpp:283-293
//   * shading basis / area sampling / light CDF        xyz/geom/triangle.cpp:194-263,
//                                                      xyz/core/scene.cpp:42-78
//   * pinhole camera                                   xyz/core/thinlenscamera.cpp:71-124
//   * path tracing with NEE + power-heuristic MIS + RR xyz/core/scene_pathtracing.cpp:11-144
//   * BSDFs: Lambert  xyz/bsdf/lambert.cpp:125-187, Mirror xyz/bsdf/mirror.cpp:50-107,
//            Glass    xyz/bsdf/glass.cpp:96-165,    Metal  xyz/bsdf/metal.cpp:68-153,
//            Light    xyz/bsdf/light.cpp:18-33, helpers xyz/stdafx.hpp:44-168
//   * progressive driver + film                        xyz/integrator/integrator.cpp:88-162
//   * bidirectional path tracing (light tracing, connections, the reference's
//     MIS weights)    xyz/core/scene_bidirectionalpathtracing.cpp:8-522,
//                     xyz/integrator/integrator.cpp:1025-1081 (see bdpt_sample())
//
// Arithmetic contract: IEEE binary32, expression order exactly as the reference
// sources (dot = (a0*b0 + a1*b1) + a2*b2, libhi/hi/math/vector3.hpp:121-122),
// no FMA contraction (built with -ffp-contract=off), correctly rounded / and sqrt.
// Where the reference mixes in double (M_1_PI, 0.5 literals) the same promotions
// are made here.
//
// Documented departures from the reference (each reduces to the reference in
// the special case named):
//   * RNG: the reference's mt19937 "next<T>" header is absent from its tree, so
//     samples come from Philox4x32-10 keyed by (pixel, seed) with counter
//     (sample, bounce, seed_hi, 0); slot use per call is fixed (see pt_path()).
//   * cos/sin of the sampled azimuth use a double-precision Taylor evaluation
//     (xyz_cos/xyz_sin below) instead of libm so that CPU and GPU agree bit for
//     bit; it is within 1 ulp of libm's float result (tested).
//   * per-light emission L_e (Phi = L_e * pi * A_total; equal L_e == reference),
//     pinhole with explicit vertical FOV, runtime depth cap kMax = D + 1,
//     RR enabled at vertices j > rr_start, optional per-texel Lambert albedo.
//   * a light-CDF upper_bound that lands on end() (undefined behaviour in the
//     reference) is clamped to the last light.
//
// Parity status of each part and the checks that pin it are listed in the
// tests (tests/test_oracle_*.py).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef float real;
const real EPSILON = real(1e-6);  // xyz/stdafx.hpp:35

// ---------------------------------------------------------------- vectors
struct V3 {
  real x[3];
  V3() { x[0] = x[1] = x[2] = 0; }
  V3(real a, real b, real c) { x[0] = a; x[1] = b; x[2] = c; }
  real& operator[](int i) { return x[i]; }
  real operator[](int i) const { return x[i]; }
};
inline V3 operator+(V3 const& a, V3 const& b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 operator-(V3 const& a, V3 const& b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 operator-(V3 const& a) { return V3(-a[0], -a[1], -a[2]); }
inline V3 operator*(V3 const& a, V3 const& b) { return V3(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
inline V3 operator*(V3 const& a, real s) { return V3(a[0] * s, a[1] * s, a[2] * s); }
inline V3 operator*(real s, V3 const& a) { return V3(s * a[0], s * a[1], s * a[2]); }
inline V3 operator/(V3 const& a, real s) { return V3(a[0] / s, a[1] / s, a[2] / s); }
// libhi/hi/math/vector3.hpp:119-131
inline real dot(V3 const& a, V3 const& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(V3 const& a, V3 const& b) {
  return V3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
// libhi/hi/math/math.hpp:146-157,235-257
inline real rcp(real x) { return real(1) / x; }
inline real rsqrt(real x) { return real(1) / std::sqrt(x); }
inline real length_squared(V3 const& a) { return dot(a, a); }
inline real length(V3 const& a) { return std::sqrt(length_squared(a)); }
inline V3 normalize(V3 const& a) { return a * rsqrt(length_squared(a)); }
// libhi/hi/math/math.hpp min/max: (x<y)?x:y and (x<y)?y:x
inline real hmin(real x, real y) { return (x < y) ? x : y; }
inline real hmax(real x, real y) { return (x < y) ? y : x; }

// --------------------------------------------------------- deterministic cos/sin
// Double-precision evaluation of cos/sin of a float angle: quadrant reduction by
// a two-part pi/2 (P1 has 33 significant bits, so k*P1 is exact for the |k| <= 8
// that float angles in [0, 4pi) give), then Horner evaluation of the Taylor
// polynomials (sin through r^17, cos through r^18; truncation < 1e-19 on
// |r| <= pi/4) with the coefficients (-1)^k/(2k+1)! and (-1)^k/(2k)! rounded
// to double.
double taylor_sin(double r) {
  double const z = r * r;
  double p = 2.8114572543455206e-15;
  p = -7.647163731819816e-13 + z * p;
  p = 1.6059043836821613e-10 + z * p;
  p = -2.505210838544172e-08 + z * p;
  p = 2.7557319223985893e-06 + z * p;
  p = -0.0001984126984126984 + z * p;
  p = 0.008333333333333333 + z * p;
  p = -0.16666666666666666 + z * p;
  return r + (r * z) * p;
}
double taylor_cos(double r) {
  double const z = r * r;
  double q = -1.5619206968586225e-16;
  q = 4.779477332387385e-14 + z * q;
  q = -1.1470745597729725e-11 + z * q;
  q = 2.08767569878681e-09 + z * q;
  q = -2.755731922398589e-07 + z * q;
  q = 2.48015873015873e-05 + z * q;
  q = -0.001388888888888889 + z * q;
  q = 0.041666666666666664 + z * q;
  q = -0.5 + z * q;
  return 1.0 + z * q;
}
void reduce_quadrant(float xf, double* r, int* q) {
  double x = double(xf);
  double k = std::floor(x * 0.63661977236758134308 + 0.5);  // 2/pi
  // pi/2 = P1 + P1t to about 4e-27
  const double P1 = 1.5707963267341256141662597656;    // 0x1.921fb544p+0
  const double P1t = 6.07710050650619224932e-11;
  double rr = (x - k * P1) - k * P1t;
  *r = rr;
  long long ki = (long long)k;
  *q = int(((ki % 4) + 4) % 4);
}
// cos and sin of a double angle in [0, 2pi] (the thin lens's disk angle,
// xyz/core/thinlenscamera.cpp:4-14, evaluates std::cos/std::sin in double):
// the same reduction and polynomials, kept in double.
void cos_sin_d(double x, double* c, double* s) {
  double const k = std::floor(x * 0.63661977236758134308 + 0.5);
  const double P1 = 1.5707963267341256141662597656;
  const double P1t = 6.07710050650619224932e-11;
  double const r = (x - k * P1) - k * P1t;
  long long const ki = (long long)k;
  int const q = int(((ki % 4) + 4) % 4);
  double const ts = taylor_sin(r), tc = taylor_cos(r);
  switch (q) {
    case 0: *c = tc; *s = ts; break;
    case 1: *c = -ts; *s = tc; break;
    case 2: *c = -tc; *s = -ts; break;
    default: *c = ts; *s = -tc; break;
  }
}
float xyz_cos(float a) {
  double r; int q;
  reduce_quadrant(a, &r, &q);
  double v;
  switch (q) {
    case 0: v = taylor_cos(r); break;
    case 1: v = -taylor_sin(r); break;
    case 2: v = -taylor_cos(r); break;
    default: v = taylor_sin(r); break;
  }
  return float(v);
}
float xyz_sin(float a) {
  double r; int q;
  reduce_quadrant(a, &r, &q);
  double v;
  switch (q) {
    case 0: v = taylor_sin(r); break;
    case 1: v = taylor_cos(r); break;
    case 2: v = -taylor_sin(r); break;
    default: v = -taylor_cos(r); break;
  }
  return float(v);
}

// Deterministic float pow for the Metal lobe (xyz/stdafx.hpp:160,
// xyz/bsdf/metal.cpp:101,146), evaluated in double so that CPU and GPU agree
// bit for bit:
//   ln x  = E ln2 + 2 atanh(z), z = (m - 1)/(m + 1), x = m 2^E, m in [sqrt(1/2), sqrt(2)],
//           atanh by its series z (1 + z^2/3 + ... + z^24/25) (|z| <= 0.1716, truncation
//           below 1e-19);
//   exp t = 2^k exp(r), r = t - k ln2 with |r| <= ln2/2, exp(r) by the Taylor
//           series through r^15 (truncation below 1e-19);
// ln2 split in two parts (ln2_hi has 32 significant bits, so E ln2_hi and
// k ln2_hi are exact).  The special cases follow C's powf.
const double kAtanhInv[13] = {1.0, 0.3333333333333333, 0.2, 0.14285714285714285, 0.1111111111111111,
                              0.09090909090909091, 0.07692307692307693, 0.06666666666666667,
                              0.058823529411764705, 0.05263157894736842, 0.047619047619047616,
                              0.043478260869565216, 0.04};
const double kExpInvFact[16] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664,
                                0.008333333333333333, 0.001388888888888889, 0.0001984126984126984,
                                2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07,
                                2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10,
                                1.1470745597729725e-11, 7.647163731819816e-13};
double dbl_from_bits(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
// x > 0 finite (normal or subnormal), y finite.
float pow_positive(float x, float y) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  uint32_t m = b & 0x7FFFFFu;
  int E = int((b >> 23) & 0xFFu) - 127;
  if ((b >> 23) == 0u) {  // subnormal: shift the leading one into the hidden bit
    E = -126;
    while (!(m & 0x800000u)) { m <<= 1; --E; }
    m &= 0x7FFFFFu;
  }
  double md = dbl_from_bits((uint64_t(0x3FF) << 52) | (uint64_t(m) << 29));
  if (md > 1.4142135623730951) { md = md * 0.5; E += 1; }
  double const ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  double const z = (md - 1.0) / (md + 1.0);
  double const z2 = z * z;
  double p = kAtanhInv[12];
  for (int k = 11; k >= 0; --k) p = kAtanhInv[k] + z2 * p;
  double const lnm = (2.0 * z) * p;
  double const L = (double(E) * ln2_hi) + ((double(E) * ln2_lo) + lnm);
  double const t = double(y) * L;
  if (t > 100.0) return std::numeric_limits<float>::infinity();
  if (t < -110.0) return 0.0f;
  double const k = std::floor(t * 1.44269504088896338700 + 0.5);
  double const r = (t - k * ln2_hi) - k * ln2_lo;
  double q = kExpInvFact[15];
  for (int n = 14; n >= 0; --n) q = kExpInvFact[n] + r * q;
  return float(q * dbl_from_bits(uint64_t(int64_t(k) + 1023) << 52));
}
float xyz_pow(float x, float y) {
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (std::isnan(x) || std::isnan(y)) return x + y;
  bool neg = false;
  if (std::signbit(x)) {
    if (std::isinf(y)) {
      if (x == -1.0f) return 1.0f;
      return ((-x < 1.0f) == (y > 0.0f)) ? 0.0f : std::numeric_limits<float>::infinity();
    }
    bool const integer = std::floor(y) == y;
    if (!integer && x != 0.0f && !std::isinf(x)) return std::numeric_limits<float>::quiet_NaN();
    neg = integer && std::fabs(y) < 16777216.0f && (int64_t(y) & 1) != 0;
    x = -x;
  }
  float r;
  if (std::isinf(y)) r = ((x < 1.0f) == (y > 0.0f)) ? 0.0f : std::numeric_limits<float>::infinity();
  else if (x == 0.0f) r = (y > 0.0f) ? 0.0f : std::numeric_limits<float>::infinity();
  else if (std::isinf(x)) r = (y > 0.0f) ? std::numeric_limits<float>::infinity() : 0.0f;
  else r = pow_positive(x, y);
  return neg ? -r : r;
}

// --------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11.
struct U4 { uint32_t v[4]; };
U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int round = 0; round < 10; ++round) {
    if (round > 0) { k0 += W0; k1 += W1; }
    uint64_t p0 = uint64_t(M0) * c.v[0];
    uint64_t p1 = uint64_t(M1) * c.v[2];
    uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    U4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
  }
  return c;
}
inline real u01(uint32_t x) { return real(x >> 8) * real(1.0f / 16777216.0f); }
struct Draw { real u[4]; };
// stream 0: pixel jitter, light and per-bounce numbers; stream 1: lens position
Draw draw(uint32_t pixel, uint64_t seed, uint32_t sample, uint32_t bounce, uint32_t stream = 0) {
  U4 c; c.v[0] = sample; c.v[1] = bounce; c.v[2] = uint32_t(seed >> 32); c.v[3] = stream;
  U4 r = philox4x32_10(c, pixel, uint32_t(seed));
  Draw d;
  for (int i = 0; i < 4; ++i) d.u[i] = u01(r.v[i]);
  return d;
}

// --------------------------------------------------------------- triangle
// xyz/geom/triangle.hpp:93-119, xyz/geom/triangle.cpp:10-60
struct Tri {
  uint32_t bsdf;
  V3 v0, e1, e2, ng, n0, s1, s2;
  real uv0[2], uv1[2], uv2[2];
  real lo(int k) const { return hmin(hmin(v0[k], v0[k] + e1[k]), v0[k] + e2[k]); }
  real hi(int k) const { return hmax(hmax(v0[k], v0[k] + e1[k]), v0[k] + e2[k]); }
  // xyz/geom/triangle.hpp:122-140 (accumulating form used for the root box)
  void acc_min(V3& v) const {
    for (int k = 0; k < 3; ++k)
      v[k] = hmin(hmin(v[k], v0[k]), hmin(v0[k] + e1[k], v0[k] + e2[k]));
  }
  void acc_max(V3& v) const {
    for (int k = 0; k < 3; ++k)
      v[k] = hmax(hmax(v[k], v0[k]), hmax(v0[k] + e1[k], v0[k] + e2[k]));
  }
  // xyz/geom/triangle.cpp:75-85
  int overlap(int axis, real value) const {
    if (hi(axis) < value) return 1;
    if (value <= lo(axis)) return 2;
    return 3;
  }
  real area() const { return length(cross(e2, e1)) * real(0.5); }
};

Tri make_tri(uint32_t bsdf, V3 p0, V3 p1, V3 p2, V3 const* ns, real const* uv) {
  Tri t;
  t.bsdf = bsdf;
  t.v0 = p0;
  t.e1 = p1 - p0;
  t.e2 = p2 - p0;
  t.ng = normalize(cross(t.e2, t.e1));
  if (ns) {
    t.n0 = ns[0];
    t.s1 = ns[1] - ns[0];
    t.s2 = ns[2] - ns[0];
  } else {
    t.n0 = t.ng;
    t.s1 = V3(0, 0, 0);
    t.s2 = V3(0, 0, 0);
  }
  for (int i = 0; i < 2; ++i) {
    t.uv0[i] = uv ? uv[i] : 0;
    t.uv1[i] = uv ? uv[2 + i] : 0;
    t.uv2[i] = uv ? uv[4 + i] : 0;
  }
  return t;
}

// xyz/stdafx.hpp:44-47 with libhi/hi/math/math.hpp:22-44 power helpers
inline real fake_fresnel(real hk) {
  real const fresnel0 = real(0.9);
  real const x = 1 - hk;
  real const x2 = x * x;
  return fresnel0 + (1 - fresnel0) * (x * (x2 * x2));
}

struct Hit {
  real t_min, t_max, f, g;
  bool back;
  void init() { t_min = real(1e-5); t_max = std::numeric_limits<real>::infinity(); }
};

// xyz/geom/triangle.cpp:127-165
bool intersect_df(Tri const& tr, V3 const& o, V3 const& d, Hit* p) {
  V3 const n = cross(tr.e2, tr.e1);
  real const det = dot(d, n);
  if (std::fabs(det) < EPSILON) return false;
  real const inv = rcp(det);
  V3 const vmo = tr.v0 - o;
  real const t = inv * dot(vmo, n);
  if ((t < p->t_min) || (p->t_max < t)) return false;
  V3 const q = cross(d, vmo);
  real const f = inv * dot(tr.e2, q);
  if ((f < 0) || (f > 1)) return false;
  real const g = inv * -dot(tr.e1, q);
  if ((g < 0) || ((f + g) > 1)) return false;
  p->t_max = t;
  p->f = f;
  p->g = g;
  p->back = det > 0;
  return true;
}

// --------------------------------------------------------------- kd-tree
// Pointer-free node pool; children[0] = [min,value), children[1] = [value,max).
struct Node {
  int kind;          // 0,1,2 = interior on that axis; 3 = leaf; 4 = null leaf
  real value;
  int child[2];
  std::vector<uint32_t> set;  // leaf triangle indices in stored order
};

struct Tree {
  std::vector<Node> nodes;
  int root = -1;
  V3 bmin, bmax;
  size_t max_bucket = 16, max_depth = 16;
  std::vector<Tri> const* tris = nullptr;
};

struct Event {
  bool type;  // false: open, true: close
  real value;
  bool operator<(Event const& e) const {
    return (value < e.value) ? true : (value > e.value) ? false : type < e.type;
  }
};

// xyz/geom/kdtree.cpp:226-357
int build_rec(Tree& tr, std::vector<uint32_t> const& set, V3 const& mn, V3 const& mx, size_t depth) {
  std::vector<Tri> const& T = *tr.tris;
  if ((set.size() <= tr.max_bucket) || (depth > tr.max_depth)) {
    Node nd;
    nd.kind = set.empty() ? 4 : 3;
    nd.value = 0;
    nd.child[0] = nd.child[1] = -1;
    nd.set = set;
    tr.nodes.push_back(std::move(nd));
    return int(tr.nodes.size()) - 1;
  }
  depth++;
  V3 const range = mx - mn;
  size_t axis = ~size_t(0);
  real value = std::numeric_limits<real>::signaling_NaN();
  real min_cost = std::numeric_limits<real>::max();
  std::vector<Event> ev(set.size() * 2);
  for (size_t k = 0; k < 3; ++k) {
    size_t const size = set.size();
    for (size_t i = 0; i < size; ++i) {
      ev[i].type = false;
      ev[i].value = T[set[i]].lo(int(k));
      ev[i + size].type = true;
      ev[i + size].value = T[set[i]].hi(int(k)) + EPSILON;
    }
    std::sort(ev.begin(), ev.end());
    real const a = range[int((k + 1) % 3)] * range[int((k + 2) % 3)];
    real const b = range[int((k + 1) % 3)] + range[int((k + 2) % 3)];
    size_t n_left = 0, n_right = set.size();
    auto update = [&](real v) {
      real const r = v - mn[int(k)];
      real const sa_left = a + b * r;
      real const sa_right = a + b * (range[int(k)] - r);
      real const cost = real(n_left) * sa_left + real(n_right) * sa_right;
      if (min_cost > cost) {
        min_cost = cost;
        axis = k;
        value = v;
      }
    };
    update(ev[0].value - EPSILON);
    for (size_t i = 0, n = ev.size(); i < n; ++i) {
      if (ev[i].type) n_right--; else n_left++;
      real v = ev[i].value;
      if ((i + 1) < n) {
        if (v == ev[i + 1].value) continue;
      } else {
        v += EPSILON;
      }
      update(v);
    }
  }
  std::vector<uint32_t> sl, sr;
  for (uint32_t id : set) {
    switch (T[id].overlap(int(axis), value)) {
      case 1: sl.push_back(id); break;
      case 2: sr.push_back(id); break;
      default: sl.push_back(id); sr.push_back(id); break;
    }
  }
  int const self = int(tr.nodes.size());
  {
    Node nd;
    nd.kind = int(axis);
    nd.value = value;
    nd.child[0] = nd.child[1] = -1;
    tr.nodes.push_back(std::move(nd));
  }
  V3 max_left(mx); max_left[int(axis)] = value;
  int const l = build_rec(tr, sl, mn, max_left, depth);
  tr.nodes[self].child[0] = l;
  V3 min_right(mn); min_right[int(axis)] = value;
  int const r = build_rec(tr, sr, min_right, mx, depth);
  tr.nodes[self].child[1] = r;
  return self;
}

// xyz/geom/kdtree.cpp:207-224
void build_tree(Tree& tr) {
  std::vector<Tri> const& T = *tr.tris;
  V3 mn(std::numeric_limits<real>::infinity(), std::numeric_limits<real>::infinity(),
        std::numeric_limits<real>::infinity());
  V3 mx(-std::numeric_limits<real>::infinity(), -std::numeric_limits<real>::infinity(),
        -std::numeric_limits<real>::infinity());
  for (Tri const& t : T) { t.acc_min(mn); t.acc_max(mx); }
  mx = mx + (mx - mn) * real(1e-5);
  tr.bmin = mn; tr.bmax = mx;
  tr.nodes.clear();
  std::vector<uint32_t> all(T.size());
  for (size_t i = 0; i < T.size(); ++i) all[i] = uint32_t(i);
  tr.root = build_rec(tr, all, mn, mx, 0);
}

struct Counters { uint32_t nodes = 0, tris = 0; };

// xyz/geom/kdtree.cpp:82-122 and 140-177 (recursive, as the reference)
int find_rec(Tree const& tr, int ni, V3 const& o, V3 const& d, Hit* p, Counters* c) {
  Node const& nd = tr.nodes[size_t(ni)];
  c->nodes++;
  if (nd.kind == 4) return -1;
  if (nd.kind == 3) {
    int s = -1;
    for (uint32_t id : nd.set) {
      c->tris++;
      if (intersect_df((*tr.tris)[id], o, d, p)) s = int(id);
    }
    return s;
  }
  int const axis = nd.kind;
  real const e = (nd.value - o[axis]) / d[axis];
  int const i = o[axis] < nd.value;
  int const nearc = nd.child[1 - i], farc = nd.child[i];
  if ((e > p->t_max) || (e < 0)) {
    return find_rec(tr, nearc, o, d, p, c);
  } else if (e < p->t_min) {
    return find_rec(tr, farc, o, d, p, c);
  } else {
    real const tmax = p->t_max;
    p->t_max = e;
    int const s = find_rec(tr, nearc, o, d, p, c);
    if (s >= 0) return s;
    p->t_min = e;
    p->t_max = tmax;
    return find_rec(tr, farc, o, d, p, c);
  }
}

// xyz/geom/kdtree.cpp:366-372
int find_intersection(Tree const& tr, V3 const& o, V3 const& d, Hit* p, Counters* c) {
  p->init();
  if (tr.root < 0) return -1;
  return find_rec(tr, tr.root, o, d, p, c);
}

// xyz/core/scene.hpp:283-293
int brute_force(std::vector<Tri> const& T, V3 const& o, V3 const& d, Hit* p) {
  p->init();
  int s = -1;
  for (size_t i = 0; i < T.size(); ++i)
    if (intersect_df(T[i], o, d, p)) s = int(i);
  return s;
}

// --------------------------------------------------------------- scene
enum { LAMBERT = 0, MIRROR = 1, GLASS = 2, METAL = 3, LIGHT = 4 };  // xyz/bsdf/bsdf.hpp:21-28

struct Bsdf {
  uint32_t type;
  V3 rgb;         // albedo (Lambert/Metal) or L_e (Light)
  real shininess;
  real ior;
  int32_t tex;
  V3 phi;         // light power Phi = L_e * pi * A_total (Light only)
};

struct Texture { uint32_t w, h; std::vector<real> texels; };

struct Scene {
  std::vector<Tri> tris;
  std::vector<Bsdf> bsdfs;
  std::vector<Texture> textures;
  std::vector<uint32_t> lights;   // triangle ids in caller order
  std::vector<real> cdf;          // cdf[0] = 0; cdf[i+1] = cdf[i] + area
  Tree tree;
  bool use_tree = true;
  real light_area() const { return cdf.back(); }
};

struct Vertex {
  int tri = -1;
  bool back = false, specular = false;
  V3 pos, dir, ng, ns, t, b;
  V3 power;
  real f_next = 0;
};

int scene_find(Scene const& sc, V3 const& o, V3 const& d, Hit* p, Counters* c) {
  if (sc.use_tree) return find_intersection(sc.tree, o, d, p, c);
  return brute_force(sc.tris, o, d, p);
}

// xyz/geom/triangle.cpp:194-219
void shading_basis(Tri const& tr, Hit const& h, Vertex* v) {
  V3 ns = normalize((tr.n0 + h.f * tr.s1) + h.g * tr.s2);
  V3 ng = tr.ng;
  if (h.back) { ng = -ng; ns = -ns; }
  V3 t(-ns[2], 0, ns[0]);
  real const l2 = length_squared(t);
  V3 b;
  if (l2 > 0) {
    t = t * rsqrt(l2);
    b = cross(t, ns);
  } else {
    b = V3(0, 0, (ns[1] > 0) ? 1 : -1);
    t = cross(ns, b);
  }
  v->ns = ns; v->ng = ng; v->t = t; v->b = b;
}

// xyz/geom/triangle.cpp:232-263 + xyz/core/scene.cpp:42-78
void sample_light(Scene const& sc, real s, real t, Vertex* lv, int* light_index) {
  if (sc.lights.empty()) { lv->tri = -1; *light_index = -1; return; }
  real const u = s * sc.cdf.back();
  size_t it = size_t(std::upper_bound(sc.cdf.begin(), sc.cdf.end(), u) - sc.cdf.begin());
  size_t n = it - 1;
  if (n >= sc.lights.size()) n = sc.lights.size() - 1;  // clamp (reference: UB)
  real const v = (u - sc.cdf[n]) / (sc.cdf[n + 1] - sc.cdf[n]);
  Tri const& tr = sc.tris[sc.lights[n]];
  real const ss = std::sqrt(v);
  real const tt = t;
  real const uu = ss * (1 - tt);
  real const vv = ss * tt;
  lv->tri = int(sc.lights[n]);
  lv->pos = (tr.v0 + uu * tr.e1) + vv * tr.e2;
  lv->ns = normalize((tr.n0 + uu * tr.s1) + vv * tr.s2);
  lv->ng = tr.ng;
  lv->f_next = rcp(sc.light_area());
  lv->specular = false;
  *light_index = int(n);
}

struct Camera {
  V3 eye, back, right, up;
  real film, aspect, v;  // film size F, aspect A, lens-film distance V
  bool thin = false;     // thin-lens model (the reference's default camera)
  real focus, lens_r, focal_plane;  // L, R = f/(2 F#), F L / V
};

// xyz/core/thinlenscamera.cpp:71-90.  cam = eye[3], at[3], up[3], fov_y_deg,
// focal_length, f_number.  focal_length 0: pinhole with an explicit vertical
// field of view (V = 0.0125 / tan(fov/2)); focal_length > 0: the reference's
// thin lens, V = L f / (L - f), focus distance L = |eye - at|, fov unused.
Camera make_camera(float const* cam, uint32_t W, uint32_t H) {
  Camera c;
  V3 eye(cam[0], cam[1], cam[2]), at(cam[3], cam[4], cam[5]), up(cam[6], cam[7], cam[8]);
  real const fov = cam[9];
  c.eye = eye;
  c.film = real(0.025);
  c.aspect = static_cast<real>(int(W)) / int(H);
  V3 const view = eye - at;
  if (cam[10] > 0.0f) {
    real const f = cam[10], fnum = cam[11];
    c.thin = true;
    c.focus = std::sqrt(dot(view, view));     // hi::length
    c.v = c.focus * f / (c.focus - f);         // thinlenscamera.cpp:74-75
    c.lens_r = f / (2 * fnum);                 // :76
    c.focal_plane = c.film * c.focus / c.v;    // :192-193
  } else {
    c.v = real(double(c.film) * 0.5 / std::tan(double(fov) * 3.14159265358979323846 / 360.0));
    c.focus = c.lens_r = c.focal_plane = 0;
  }
  c.back = normalize(view);
  c.right = normalize(cross(up, c.back));
  c.up = cross(c.back, c.right);
  return c;
}

// xyz/core/thinlenscamera.cpp:111-124
V3 primary_dir(Camera const& c, real s, real t) {
  real const lx = real((double(s) - 0.5) * double(c.film) * double(c.aspect));
  real const ly = real((double(t) - 0.5) * double(c.film));
  real const lz = -c.v;
  return normalize((lx * c.right + ly * c.up) + lz * c.back);
}

// Primary ray of either model.  Thin lens: SampleDisk (thinlenscamera.cpp:4-14)
// with (ls, lt), origin on the lens (:159-181), direction through the focus
// plane (:186-205).
void primary_ray(Camera const& c, real s, real t, real ls, real lt, V3* org, V3* dir) {
  if (!c.thin) {
    *org = c.eye;
    *dir = primary_dir(c, s, t);
    return;
  }
  real const radius = c.lens_r * std::sqrt(ls);
  double cs, sn;
  cos_sin_d((2 * 3.14159265358979323846) * double(lt), &cs, &sn);
  real const lx = real(double(radius) * cs);
  real const ly = real(double(radius) * sn);
  *org = (c.eye + lx * c.right) + ly * c.up;
  real const dx = real((double(s) - 0.5) * double(c.focal_plane) * double(c.aspect) - double(lx));
  real const dy = real((double(t) - 0.5) * double(c.focal_plane) - double(ly));
  real const dz = -c.focus;
  *dir = normalize((dx * c.right + dy * c.up) + dz * c.back);
}

V3 albedo_at(Scene const& sc, Bsdf const& b, Tri const& tr, Hit const& h, int* texel) {
  *texel = -1;
  if (b.tex < 0) return b.rgb;
  Texture const& tx = sc.textures[size_t(b.tex)];
  // uv = uv0 + f (uv1-uv0) + g (uv2-uv0), same barycentrics as triangle.hpp:194-199
  real u = (tr.uv0[0] + h.f * (tr.uv1[0] - tr.uv0[0])) + h.g * (tr.uv2[0] - tr.uv0[0]);
  real v = (tr.uv0[1] + h.f * (tr.uv1[1] - tr.uv0[1])) + h.g * (tr.uv2[1] - tr.uv0[1]);
  int ix = int(u * real(tx.w));
  int iy = int(v * real(tx.h));
  ix = ix < 0 ? 0 : (ix >= int(tx.w) ? int(tx.w) - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= int(tx.h) ? int(tx.h) - 1 : iy);
  int const idx = iy * int(tx.w) + ix;
  *texel = idx;
  return V3(tx.texels[size_t(idx) * 3 + 0], tx.texels[size_t(idx) * 3 + 1], tx.texels[size_t(idx) * 3 + 2]);
}

// xyz/stdafx.hpp:50-63 (cos/sin as the path tracer's)
real diffused(real mu_t, real mu_p, V3 const& t, V3 const& n, V3 const& b, V3* out) {
  real const phi = real(2 * M_PI) * mu_p;
  real const ct = std::sqrt(mu_t);
  real const st = std::sqrt(1 - mu_t);
  *out = (t * (st * xyz_cos(phi)) + b * (st * xyz_sin(phi))) + n * ct;
  return ct;
}

// xyz/stdafx.hpp:154-168
real glossy_half(real mu_t, real mu_p, real shin, V3 const& t, V3 const& n, V3 const& b, V3* h) {
  real const phi = real(2 * M_PI) * mu_p;
  real const ct = xyz_pow(mu_t, rcp(shin));
  real const st = std::sqrt(real(1) - ct * ct);
  *h = (t * (st * xyz_cos(phi)) + b * (st * xyz_sin(phi))) + n * ct;
  return ct;
}

struct PTParams {
  uint32_t kmax;        // reference XYZ_CONFIG_kMaxRandomWalkDepth = max depth D + 1
  uint32_t rr_start;    // RR applies at vertices j > rr_start (0 = reference)
  real rr_threshold;    // CONFIG_TERMINATE_THRESHOLD = 0.2
  uint64_t seed;
};

// What the reference gradient needs of one sample's path.  A sample's colour
// is a sum of terms, each the path throughput times a local factor; vertex i
// (a hit that reaches the NEE step, scene_pathtracing.cpp:111-142) multiplies
// the throughput by its albedo a_i (Lambert, Metal; 1 for specular BSDFs),
// adds its NEE term beta_i a_i Y_i, then scales the throughput by sigma_i
// (roulette's 1/q, Metal's Fresnel weight).  A path that ends on a light adds
// beta_n E.  With R_n = E and R_i = a_i (Y_i + sigma_i R_{i+1}), the colour's
// derivative with respect to a_i is beta_i (Y_i + sigma_i R_{i+1}) -- no
// division by a_i.  Every emission term is linear in its light's L_e through
// Phi = L_e pi A_lights, so dcolour/dL_e is the term with pi A_lights in place
// of Phi.
struct VertexRecord {
  int texture;  // -1 unless a textured Lambert
  int texel;
  V3 albedo;    // a_i
  V3 beta;      // throughput before a_i
  V3 nee;       // Y_i (0 when the NEE sample is not visible)
  real sigma;   // throughput scale after the NEE step
};
struct PathRecord {
  std::vector<VertexRecord> v;
  V3 emission;                    // E of the light the path ends on, per unit throughput
  std::vector<double> d_emission;  // 3 per BSDF: dcolour / dL_e
};

struct PathStats { uint32_t queries = 0; uint32_t vertices = 0; Counters c; };

// ------------------------------------------------------------------ IDPT
// Importance-driven path tracing (Scene::EvaluteImportanceDrivenPathTracing,
// xyz/core/scene_importancedrivenpathtracing.cpp:11-167): a map of importons
// (position, incoming direction) left by light paths on Lambert surfaces right
// after a specular bounce (Scene::CreateLightImportanceMap,
// xyz/core/scene_photonmap.cpp:272-334, with Bsdf::BoundsImportance) guides
// the direction sampling and NEE weights of Lambert eye vertices reached
// through a specular bounce, through a 2 x 8 histogram of the importons within
// 0.1 around the vertex (ImportanceQuery, xyz/core/photonmap.cpp:321-512).
struct Importon { V3 pos, dir; };
constexpr uint32_t kIdptStreamEmit = 4;
constexpr int kU = 2, kV = 8, kN = kU * kV;  // CONFIG_U_SIZE, CONFIG_V_SIZE
const real kSearchR2 = real(0.1) * real(0.1);  // square_of(CONFIG_SEARCH_RADIUS)

// Bsdf::BoundsImportance; v.dir holds the incoming direction and becomes the
// outgoing one.  u: x choice / Russian roulette, y factor test (glass), y z
// direction (Lambert, Metal).
bool bounds_importance(Bsdf const& b, bool back, V3 const& ns, V3 const& ng, V3 const& t,
                       V3 const& bn, Draw const& u, V3* dir, bool* specular) {
  switch (b.type) {
    case LAMBERT: {  // lambert.cpp:16-40
      real const cs = -dot(*dir, ns);
      if (cs <= 0) return false;
      real const cg = -dot(*dir, ns);  // the reference takes the shading normal twice
      real const f = cs / cg;
      if (u.u[0] >= b.rgb[1] * f) return false;
      *specular = false;
      real const phi = real(2 * M_PI) * u.u[2];
      real const ct = std::sqrt(u.u[1]);
      real const st = std::sqrt(1 - u.u[1]);
      *dir = (t * (st * xyz_cos(phi)) + bn * (st * xyz_sin(phi))) + ng * ct;
      return true;
    }
    case MIRROR: {  // mirror.cpp:18-48
      real const cs = -dot(*dir, ns);
      if (cs <= 0) return false;
      real const cg = -dot(*dir, ng);
      *dir = *dir + (real(2) * cs) * ns;
      real const co = dot(*dir, ng);
      if (co <= 0) return false;
      if (u.u[0] >= co / cg) return false;
      *specular = true;
      return true;
    }
    case GLASS: {  // glass.cpp:16-91
      real const cs = -dot(*dir, ns);
      if (cs <= 0) return false;
      real const cg = -dot(*dir, ng);
      real const r = back ? rcp(b.ior) : b.ior;
      real const ior_sq = r * r;
      real const nk = cs;
      real const bb = ior_sq + nk * nk - 1;
      *specular = true;
      if (bb > 0) {
        real const g = std::sqrt(bb);
        real const gpc = g + nk, gmc = g - nk;
        real const alpha = gmc / gpc;
        real const beta = (nk * gpc - 1) / (nk * gmc + 1);
        real const fr = real(0.5) * alpha * alpha * (beta * beta + 1);
        if (u.u[0] < fr) {
          *dir = *dir + (real(2) * nk) * ns;
          real const co = dot(*dir, ng);
          if (co <= 0) return false;
          if (u.u[1] >= co / cg) return false;
        } else {
          *dir = (*dir - gmc * ns) / std::sqrt(ior_sq);
          real const co = -dot(*dir, ng);
          if (co <= 0) return false;
          if (u.u[1] >= co / cg) return false;
        }
      } else {
        *dir = *dir + (real(2) * nk) * ns;
        real const co = dot(*dir, ng);
        if (co <= 0) return false;
        if (u.u[1] >= co / cg) return false;
      }
      return true;
    }
    case METAL: {  // metal.cpp:22-66
      real const nk1 = -dot(*dir, ns);
      if (nk1 <= 0) return false;
      real const cg = -dot(*dir, ng);
      V3 hv;
      glossy_half(u.u[1], u.u[2], b.shininess, t, ns, bn, &hv);
      real const hk = -dot(hv, *dir);
      *dir = *dir + (2 * hk) * hv;
      real const nk2 = dot(*dir, ns);
      if (nk2 <= 0) return false;
      real const co = dot(ng, *dir);
      if (co <= 0) return false;
      real const f = fake_fresnel(hk) / hmax(nk1, nk2) * nk2 * (co / cg);
      if (u.u[0] >= b.rgb[1] * f) return false;
      *specular = false;
      return true;
    }
    default:
      return false;  // light.cpp:11-16
  }
}

// Scene::CreateLightImportanceMap with n_emit light paths of at most kmax
// segments; random numbers: Philox stream 4 keyed (path index, seed),
// vertex 0 the light sample (x, y) and emission direction (z, w).
void emit_importons(Scene const& sc, uint32_t n_emit, uint32_t kmax, uint64_t seed,
                    std::vector<Importon>* out) {
  out->clear();
  if (sc.lights.empty()) return;
  for (uint32_t i = 0; i < n_emit; ++i) {
    Draw const d0 = draw(i, seed, 0, 0, kIdptStreamEmit);
    Vertex lv;
    int n = -1;
    sample_light(sc, d0.u[0], d0.u[1], &lv, &n);
    V3 t(-lv.ng[2], 0, lv.ng[0]), bn;
    real const l2 = length_squared(t);
    if (l2 > 0) {
      t = t * rsqrt(l2);
      bn = cross(t, lv.ng);
    } else {
      bn = V3(0, 0, (lv.ng[1] > 0) ? 1 : -1);
      t = cross(lv.ng, bn);
    }
    real const phi = real(2 * M_PI) * d0.u[3];
    real const ct = std::sqrt(d0.u[2]);
    real const st = std::sqrt(1 - d0.u[2]);
    V3 dir = (t * (st * xyz_cos(phi)) + bn * (st * xyz_sin(phi))) + lv.ng * ct;
    V3 pos = lv.pos;
    bool specular = false;
    for (uint32_t depth = 0; depth < kmax; ++depth) {
      Hit h;
      Counters c;
      int const tri = scene_find(sc, pos, dir, &h, &c);
      if (tri < 0) break;
      Tri const& tr = sc.tris[size_t(tri)];
      Bsdf const& b = sc.bsdfs[tr.bsdf];
      Vertex v;
      if (b.type == LAMBERT) {
        // Triangle::GeometricBasis
        V3 ns = normalize((tr.n0 + h.f * tr.s1) + h.g * tr.s2);
        V3 ng = tr.ng;
        if (h.back) { ng = -ng; ns = -ns; }
        V3 tt(-ng[2], 0, ng[0]);
        real const q2 = length_squared(tt);
        V3 bb;
        if (q2 > 0) {
          tt = tt * rsqrt(q2);
          bb = cross(tt, ng);
        } else {
          bb = V3(0, 0, (ng[1] > 0) ? 1 : -1);
          tt = cross(ng, bb);
        }
        v.ns = ns; v.ng = ng; v.t = tt; v.b = bb;
        pos = pos + dir * h.t_max;
        if (depth > 0 && specular) out->push_back(Importon{pos, dir});
      } else {
        shading_basis(tr, h, &v);
        pos = pos + dir * h.t_max;
      }
      if (!bounds_importance(b, h.back, v.ns, v.ng, v.t, v.b,
                             draw(i, seed, 0, depth + 1, kIdptStreamEmit), &dir, &specular))
        break;
    }
  }
}

// Octant of the azimuth theta = atan2(y, x) in [0, 2 pi), an edge belonging to
// the octant above it: the reference's v = floor(phi * 8 / pi) with
// phi = atan2(y, x) / 2 (+ pi when negative) (photonmap.cpp:351-360) in exact
// arithmetic.  The reference rounds atan2 to float first, which can move a
// direction within about 1e-7 rad of an octant edge to the neighbouring bin.
int azimuth_octant(real x, real y) {
  if (x == 0 && y == 0) return 0;
  if (y > 0 || (y == 0 && x > 0)) {  // theta in [0, pi)
    if (x > y) return 0;
    if (x > 0) return 1;
    return (y > -x) ? 2 : 3;
  }
  if (x < 0) return (y > x) ? 4 : 5;  // theta in [pi, 2 pi)
  return (y < -x) ? 6 : 7;
}

// Bin of a direction with cosine ct to the shading normal and components x, y
// along -tangent, -binormal: u = int(ct^2 * 2) (photonmap.cpp:345-349), v the
// azimuth octant.
int importance_bin(real ct, real x, real y) {
  int u = int((ct * ct) * kU);
  if (u >= kU) u = kU - 1;
  return u * kV + azimuth_octant(x, y);
}

// ImportanceQuery: cdf[0..16] of the importons within reach of vertex v.
void importance_cdf(std::vector<Importon> const& imp, Vertex const& v, real cdf[kN + 1]) {
  cdf[0] = 0;
  for (int i = 1; i <= kN; ++i) cdf[i] = real(kN);
  for (Importon const& m : imp) {
    V3 const d = m.pos - v.pos;
    if (!(length_squared(d) < kSearchR2)) continue;
    real const ct = -dot(m.dir, v.ns);
    if (ct <= 0) continue;
    V3 const pt = d + m.dir * dot(d, v.ns);
    if (length_squared(pt) >= kSearchR2) continue;
    cdf[1 + importance_bin(ct, -dot(m.dir, v.t), -dot(m.dir, v.b))] += 1;
  }
  for (int i = 2; i <= kN; ++i) cdf[i] += cdf[i - 1];
}

// ImportanceQuery::GetDensity (photonmap.cpp:430-457)
real importance_density(real const cdf[kN + 1], Vertex const& v, V3 const& dir) {
  real const ct = -dot(dir, v.ns);
  if (ct <= 0) return 0;
  int const n = importance_bin(ct, -dot(dir, v.t), -dot(dir, v.b));
  return kN * (cdf[n + 1] - cdf[n]) / cdf[kN];
}

// Lambert::NextScatteringDirection with importons (lambert.cpp:243-259) and
// ImportanceQuery::SetDiffusedVector (photonmap.cpp:459-494): a bin from the
// cdf with u_bin, then cos^2 theta uniform in the bin's half (u_t) and the
// azimuth uniform in its octant (u_p).  Returns the bin's density; *f_next =
// the direction's density per solid angle, density cos theta / pi.
real guided_sample(real const cdf[kN + 1], Vertex const& v, real u_bin, real u_t, real u_p,
                   V3* dir, real* f_next) {
  real const value = cdf[kN] * u_bin;
  int n = int(std::upper_bound(cdf, cdf + kN + 1, value) - cdf) - 1;
  if (n < 0) n = 0;
  if (n >= kN) n = kN - 1;
  int const bu = n / kV, bv = n % kV;
  real const mu_u = (real(bu) + u_t) / real(kU);
  real const mu_v = (real(bv) + u_p) / real(kV);
  real const cos_t = std::sqrt(mu_u);
  real const sin_t = std::sqrt(1 - mu_u);
  real const phi = real(2 * M_PI) * mu_v;
  *dir = (v.t * (sin_t * xyz_cos(phi)) + v.b * (sin_t * xyz_sin(phi))) + v.ns * cos_t;
  real const density = kN * (cdf[n + 1] - cdf[n]) / cdf[kN];
  *f_next = real(double(density * cos_t) * M_1_PI);
  return density;
}

// xyz/core/scene_pathtracing.cpp:11-144.  Returns the sample's colour.
// Optional: records what the reference gradient needs (PathRecord).
V3 pt_path(Scene const& sc, Camera const& cam, PTParams const& pp, uint32_t W, uint32_t H,
           uint32_t px, uint32_t py, uint32_t sample, PathStats* st,
           PathRecord* rec,
           std::vector<Importon> const* guide = nullptr) {
  uint32_t const pixel = py * W + px;
  Draw d0 = draw(pixel, pp.seed, sample, 0);
  // xyz/integrator/integrator.cpp:124-125
  real const sx = (real(int(px)) + d0.u[0]) / real(int(W));
  real const sy = (real(int(py)) + d0.u[1]) / real(int(H));

  Vertex lv;
  int light_n = -1;
  sample_light(sc, d0.u[2], d0.u[3], &lv, &light_n);
  if (light_n >= 0) lv.power = sc.bsdfs[sc.tris[size_t(lv.tri)].bsdf].phi;

  Vertex ev;
  ev.power = V3(1, 1, 1);
  ev.f_next = 0;
  if (cam.thin) {
    // two more numbers after the light's (scene_pathtracing.cpp:24-26)
    Draw const dl = draw(pixel, pp.seed, sample, 0, 1);
    primary_ray(cam, sx, sy, dl.u[0], dl.u[1], &ev.pos, &ev.dir);
  } else {
    primary_ray(cam, sx, sy, 0, 0, &ev.pos, &ev.dir);
  }
  ev.specular = true;

  V3 color(0, 0, 0);
  real const A = sc.cdf.back();
  for (uint32_t k = 1;;) {
    Hit h;
    if (st) st->queries++;
    Counters c;
    ev.tri = scene_find(sc, ev.pos, ev.dir, &h, &c);
    if (st) { st->c.nodes += c.nodes; st->c.tris += c.tris; }
    if (ev.tri < 0) break;
    Tri const& tr = sc.tris[size_t(ev.tri)];
    shading_basis(tr, h, &ev);
    ev.back = h.back;
    if (-dot(ev.dir, ev.ns) <= 0.0) break;
    ev.pos = ev.pos + ev.dir * h.t_max;

    Bsdf const& b = sc.bsdfs[tr.bsdf];
    if ((b.type == LIGHT) && !ev.back) {
      V3 const lp = b.phi;
      V3 add;
      double wfac;  // add = (Phi * power) * wfac
      if (ev.specular) {
        double const w = M_1_PI / double(A);
        V3 const pe = lp * ev.power;
        add = V3(real(double(pe[0]) * w), real(double(pe[1]) * w), real(double(pe[2]) * w));
        wfac = w;
      } else {
        real const cos_t = std::fabs(dot(ev.dir, ev.ns));
        real const dist2 = h.t_max * h.t_max;
        real const p_imp = ev.f_next * cos_t / dist2;
        real const p_exp = rcp(A);
        real const dens = p_exp / p_imp;
        double const w1 = M_1_PI / double(A);
        double const w2 = double(rcp(1 + dens * dens));
        V3 const pe = lp * ev.power;
        add = V3(real(double(pe[0]) * w1 * w2), real(double(pe[1]) * w1 * w2),
                 real(double(pe[2]) * w1 * w2));
        wfac = w1 * w2;
      }
      if (rec) {
        // the light that was hit: dPhi/dL_e = pi A_lights
        double const unit = M_PI * double(A);
        for (int c = 0; c < 3; ++c) {
          rec->emission[c] = real(double(lp[c]) * wfac);
          rec->d_emission[3 * size_t(tr.bsdf) + size_t(c)] += double(ev.power[c]) * unit * wfac;
        }
      }
      color = color + add;
      break;
    }
    if (++k >= pp.kmax) break;
    if (st) st->vertices++;

    uint32_t const j = k - 1;  // vertex index, 1-based
    Draw dj = draw(pixel, pp.seed, sample, j);

    // Importance-driven branch (scene_importancedrivenpathtracing.cpp:122-152):
    // a Lambert vertex reached through a specular bounce, not the first hit.
    bool const guided = guide && (b.type == LAMBERT) && (k > 2) && ev.specular;
    real cdf[kN + 1];
    if (guided) importance_cdf(*guide, ev, cdf);

    VertexRecord* vr = nullptr;
    if (rec) {
      rec->v.push_back(VertexRecord{-1, -1, V3(1, 1, 1), ev.power, V3(0, 0, 0), real(1)});
      vr = &rec->v.back();
    }

    // ---- NEE: Bsdf::CalculateWeightedDirectLighting
    if (b.type == LAMBERT || b.type == METAL) {
      int texel = -1;
      V3 const alb = albedo_at(sc, b, tr, h, &texel);
      if (vr) {
        vr->texture = (b.type == LAMBERT) ? b.tex : -1;
        vr->texel = texel;
        vr->albedo = alb;
      }
      ev.power = ev.power * alb;
      if (lv.tri >= 0) {
        V3 const w = normalize(ev.pos - lv.pos);
        real const nk3 = dot(w, lv.ng);
        bool ok = !((nk3 <= 0) || (dot(w, lv.ns) <= 0));
        real nk2 = 0;
        if (ok) {
          nk2 = -dot(w, ev.ns);
          ok = !((nk2 <= 0) || (-dot(w, ev.ng) <= 0));
        }
        if (ok) {
          Hit sh;
          if (st) st->queries++;
          Counters c2;
          int const hit = scene_find(sc, lv.pos, w, &sh, &c2);
          if (st) { st->c.nodes += c2.nodes; st->c.tris += c2.tris; }
          if (hit == ev.tri) {
            V3 add;
            double nee_w = 0;
            if (b.type == LAMBERT) {
              // xyz/bsdf/lambert.cpp:157-172
              real const G = nk2 * nk3 / (sh.t_max * sh.t_max);
              // lambert.cpp:228-229 scales the implicit density by the histogram's
              real const p_imp = guided ? real((M_1_PI * double(G)) *
                                               double(importance_density(cdf, ev, w)))
                                        : real(M_1_PI * double(G));
              real const p_exp = rcp(A);
              real const dens = p_imp / p_exp;
              double const wgt = (double(G) * (M_1_PI * M_1_PI)) * double(rcp(1 + dens * dens));
              V3 const pe = lv.power * ev.power;
              add = V3(real(double(pe[0]) * wgt), real(double(pe[1]) * wgt), real(double(pe[2]) * wgt));
              nee_w = wgt;
            } else {
              // xyz/bsdf/metal.cpp:100-116
              real const wgt = nk3 / (sh.t_max * sh.t_max);
              real const G = nk2 * wgt;
              V3 const hv = -normalize(w + ev.dir);
              real const hk = -dot(hv, w);
              real const nh = dot(ev.ns, hv);
              real const nk1 = -dot(ev.ns, ev.dir);
              real const fsn = real(double(b.shininess + 1) * (M_1_PI / 8) *
                                    double(xyz_pow(nh, b.shininess)) / double(hk));
              real const dens = fsn * wgt * A;
              real const fres = fake_fresnel(hk);
              V3 const pe = lv.power * ev.power;
              double const s1 = double(G) * M_1_PI;
              double const mx = double(hmax(nk1, nk2));
              double const rw = double(rcp(1 + dens * dens));
              for (int ch = 0; ch < 3; ++ch)
                add[ch] = real(((((double(pe[ch]) * s1) * double(fsn)) / mx) * double(fres)) * rw);
              nee_w = (((s1 * double(fsn)) / mx) * double(fres)) * rw;
            }
            if (vr) {
              // add = (Phi * power) * nee_w: per unit throughput Phi nee_w; the
              // sampled light's dPhi/dL_e = pi A_lights
              uint32_t const lb = sc.tris[size_t(lv.tri)].bsdf;
              double const unit = M_PI * double(A);
              for (int c = 0; c < 3; ++c) {
                vr->nee[c] = real(double(lv.power[c]) * nee_w);
                rec->d_emission[3 * size_t(lb) + size_t(c)] += double(ev.power[c]) * unit * nee_w;
              }
            }
            color = color + add;
          }
        }
      }
    }

    // ---- Russian roulette: xyz/core/scene_pathtracing.cpp:128-137
    if (j > pp.rr_start) {
      real const t = ev.power[1];
      if (t < pp.rr_threshold) {
        real const q = t / pp.rr_threshold;
        if (dj.u[0] >= q) break;
        real const inv = rcp(q);
        ev.power = ev.power * inv;
        if (vr) vr->sigma = vr->sigma * inv;
      }
    }

    // ---- Bsdf::NextScatteringDirection
    if (guided) {
      real const density = guided_sample(cdf, ev, dj.u[3], dj.u[1], dj.u[2], &ev.dir, &ev.f_next);
      ev.power = ev.power * rcp(density);
      if (vr) vr->sigma = vr->sigma * rcp(density);
      ev.specular = false;
      if (!(dot(ev.dir, ev.ng) > 0)) break;
    } else if (b.type == LAMBERT) {
      // xyz/bsdf/lambert.cpp:176-187, xyz/stdafx.hpp:50-63
      real const phi = real(2 * M_PI) * dj.u[2];
      real const cos_t = std::sqrt(dj.u[1]);
      real const sin_t = std::sqrt(1 - dj.u[1]);
      ev.dir = (ev.t * (sin_t * xyz_cos(phi)) + ev.b * (sin_t * xyz_sin(phi))) + ev.ns * cos_t;
      ev.specular = false;
      ev.f_next = real(M_1_PI * double(cos_t));
      if (!(dot(ev.dir, ev.ng) > 0)) break;
    } else if (b.type == MIRROR) {
      // xyz/bsdf/mirror.cpp:56-105
      real const c = -dot(ev.dir, ev.ns);
      ev.dir = ev.dir + (real(2) * c) * ev.ns;
      if (dot(ev.dir, ev.ng) <= 0) break;
      ev.specular = true;
      ev.f_next = 1;
    } else if (b.type == GLASS) {
      // xyz/bsdf/glass.cpp:104-165
      real const ior = b.ior;
      real const r = ev.back ? rcp(ior) : ior;
      real const ior_sq = r * r;
      real const nk = -dot(ev.dir, ev.ns);
      real const bb = ior_sq + nk * nk - 1;
      if (bb > 0) {
        real const g = std::sqrt(bb);
        real const gpc = g + nk;
        real const gmc = g - nk;
        real const alpha = gmc / gpc;
        real const beta = (nk * gpc - 1) / (nk * gmc + 1);
        real const fr = real(0.5) * alpha * alpha * (beta * beta + 1);
        if (dj.u[1] < fr) {
          ev.dir = ev.dir + (real(2) * nk) * ev.ns;
          if (dot(ev.dir, ev.ng) <= 0) break;
        } else {
          ev.dir = (ev.dir - gmc * ev.ns) / std::sqrt(ior_sq);
          if (dot(ev.dir, ev.ng) >= 0) break;
        }
      } else {
        ev.dir = ev.dir + (real(2) * nk) * ev.ns;
        if (dot(ev.dir, ev.ng) <= 0) break;
      }
      ev.specular = true;
      ev.f_next = 1;
    } else if (b.type == METAL) {
      // xyz/bsdf/metal.cpp:122-153, xyz/stdafx.hpp:154-168
      V3 hv;
      real const cos_h = glossy_half(dj.u[1], dj.u[2], b.shininess, ev.t, ev.ns, ev.b, &hv);
      real const nk1 = -dot(ev.ns, ev.dir);
      real const hk = -dot(hv, ev.dir);
      ev.dir = ev.dir + (2 * hk) * hv;
      real const nk2 = dot(ev.ns, ev.dir);
      if ((nk2 <= 0) || (dot(ev.ng, ev.dir) <= 0)) break;
      real const fres = fake_fresnel(hk);
      real const sc1 = fres / hmax(nk1, nk2) * nk2;
      ev.power = ev.power * sc1;
      if (vr) vr->sigma = vr->sigma * sc1;
      ev.specular = false;
      ev.f_next = real(double(b.shininess + 1) * (M_1_PI / 8) * double(xyz_pow(cos_h, b.shininess)) / double(hk));
    } else {
      break;  // Light seen from behind: xyz/bsdf/light.cpp:27-31
    }
  }
  return color;
}

// ------------------------------------------------------------------ BDPT
// Scene::EvaluateBidirectionalPathTracing (xyz/core/scene_bidirectionalpathtracing.cpp:8-522)
// with the bidirectional BSDF methods (xyz/bsdf/{lambert,mirror,glass,metal,light}.cpp
// LightsToCamera/CameraToLightsScatteringDirection) and the pinhole camera's
// film mapping (xyz/core/thinlenscamera.hpp:38-90, thinlenscamera.cpp:125-157).
// Reference behaviour kept as it is:
//   * hi::single_stack (libhi/hi/util.hpp:24-36) assigns its value for good (its
//     destructor writes the new value back), so each connection's temporary
//     density stays in the vertex and later strategies read it;
//   * Mirror::CameraToLightsScatteringDirection ends the eye path when the
//     reflected direction is ABOVE the geometric surface (mirror.cpp:235-239);
//   * connections weigh every non-specular vertex as Lambertian (M_1_PI).
// Vertex fields the reference would read before any write in the same sample
// (it reuses the vertex arrays across samples) start from zero / false here.
struct BVert {
  int tri = -1;
  bool back = false, specular = false;
  V3 pos, win, wout, ng, ns, t, b, power;
  real bsdf_x_ipdf = 0, s_prev = 0, s_next = 0, cos_in_sh = 0, cos_out_geo = 0, g = 0;
};

constexpr uint32_t kBdptStreamLight = 2, kBdptStreamEye = 3;

// Triangle::GeometricBasis (xyz/geom/triangle.cpp:167-192)
void geometric_basis(Tri const& tr, Hit const& h, BVert* v) {
  V3 ns = normalize((tr.n0 + h.f * tr.s1) + h.g * tr.s2);
  V3 ng = tr.ng;
  if (h.back) { ng = -ng; ns = -ns; }
  V3 t(-ng[2], 0, ng[0]);
  real const l2 = length_squared(t);
  V3 b;
  if (l2 > 0) {
    t = t * rsqrt(l2);
    b = cross(t, ng);
  } else {
    b = V3(0, 0, (ng[1] > 0) ? 1 : -1);
    t = cross(ng, b);
  }
  v->ns = ns; v->ng = ng; v->t = t; v->b = b;
}

real glass_ior_sq(Bsdf const& b, bool back) {
  real const r = back ? rcp(b.ior) : b.ior;
  return r * r;
}

real min1(real x) { return (x < real(1)) ? x : real(1); }  // std::min(float_t(1), x)

// Bsdf::LightsToCameraScatteringDirection; u = (dir u1, dir u2, choice)
bool lights_to_camera(Scene const& sc, Bsdf const& bs, Tri const& tr, Hit const& h, BVert* v,
                      Draw const& u) {
  switch (bs.type) {
    case LAMBERT: {  // lambert.cpp:314-365
      int texel;
      V3 const alb = albedo_at(sc, bs, tr, h, &texel);
      real const cig = dot(v->win, v->ng);
      real const alpha = v->cos_in_sh / cig;
      real const pp = min1(alb[1]), pn = min1(alb[1] * alpha);
      v->specular = false;
      v->s_prev = real(M_1_PI * double(pp));
      v->s_next = real(M_1_PI * double(pn));
      v->power = v->power * alb;
      v->power = v->power * alpha;
      v->bsdf_x_ipdf = rcp(pn);
      if (u.u[2] >= pn) return false;
      v->cos_out_geo = diffused(u.u[0], u.u[1], v->t, v->ng, v->b, &v->wout);
      if (dot(v->wout, v->ns) <= 0) return false;
      return true;
    }
    case MIRROR: {  // mirror.cpp:172-203
      v->specular = true;
      v->wout = -v->win;
      v->wout = v->wout + (real(2) * v->cos_in_sh) * v->ns;
      v->cos_out_geo = dot(v->wout, v->ng);
      if (v->cos_out_geo <= 0) return false;
      v->s_prev = rcp(v->cos_in_sh);
      v->s_next = rcp(v->cos_out_geo);
      real const cig = dot(v->win, v->ng);
      v->bsdf_x_ipdf = v->cos_out_geo / cig;
      return true;
    }
    case GLASS: {  // glass.cpp:239-323
      real const ior_sq = glass_ior_sq(bs, v->back);
      real const nk = dot(v->win, v->ns);
      real const bb = ior_sq + nk * nk - 1;
      v->wout = -v->win;
      v->specular = true;
      if (bb > 0) {
        real const g = std::sqrt(bb);
        real const gpc = g + nk, gmc = g - nk;
        real const alpha = gmc / gpc;
        real const beta = (nk * gpc - 1) / (nk * gmc + 1);
        real const fr = real(0.5) * alpha * alpha * (beta * beta + 1);
        if (u.u[2] >= fr) {
          v->s_prev = v->s_next = 1 - fr;
          v->wout = (v->wout - gmc * v->ns) / std::sqrt(ior_sq);
          v->cos_out_geo = -dot(v->wout, v->ng);
          if (v->cos_out_geo <= 0) return false;
          real const cig = dot(v->win, v->ng);
          v->bsdf_x_ipdf = v->cos_out_geo / cig;
          return true;
        }
        v->s_prev = v->s_next = fr;
      } else {
        v->s_prev = v->s_next = 1;
      }
      v->wout = v->wout + (real(2) * nk) * v->ns;
      v->cos_out_geo = dot(v->wout, v->ng);
      if (v->cos_out_geo <= 0) return false;
      real const cig = dot(v->win, v->ng);
      v->bsdf_x_ipdf = v->cos_out_geo / cig;
      return true;
    }
    case METAL: {  // metal.cpp:227-278
      v->specular = false;
      V3 hv;
      real const nh = glossy_half(u.u[0], u.u[1], bs.shininess, v->t, v->ns, v->b, &hv);
      real const hk = dot(hv, v->win);
      v->wout = (2 * hk) * hv - v->win;
      v->cos_out_geo = dot(v->ng, v->wout);
      if (v->cos_out_geo <= 0) return false;
      real const cos_sh = dot(v->ns, v->wout);
      if (cos_sh <= 0) return false;
      real const cig = dot(v->ng, v->win);
      v->bsdf_x_ipdf = fake_fresnel(hk) * v->cos_in_sh / hmax(cos_sh, v->cos_in_sh) *
                       v->cos_out_geo / cig;
      v->s_prev = v->s_next = real(double(bs.shininess + 1) * (M_1_PI / 8) *
                                   double(xyz_pow(nh, bs.shininess)) / double(hk));
      v->s_prev *= rcp(v->cos_in_sh);
      v->s_next *= rcp(v->cos_out_geo);
      return true;
    }
    default:
      return false;  // light.cpp:44-51
  }
}

// Bsdf::CameraToLightsScatteringDirection; u = (dir u1, dir u2, choice)
bool camera_to_lights(Scene const& sc, Bsdf const& bs, Tri const& tr, Hit const& h, BVert* v,
                      Draw const& u) {
  switch (bs.type) {
    case LAMBERT: {  // lambert.cpp:380-424
      v->specular = false;
      v->cos_in_sh = diffused(u.u[0], u.u[1], v->t, v->ns, v->b, &v->win);
      if (dot(v->win, v->ng) <= 0) return false;
      real const cog = dot(v->wout, v->ng);
      int texel;
      V3 const alb = albedo_at(sc, bs, tr, h, &texel);
      real const alpha = v->cos_in_sh / cog;
      real const pp = min1(alb[1]), pn = min1(alb[1] * alpha);
      v->s_prev = real(M_1_PI * double(pp));
      v->s_next = real(M_1_PI * double(pn));
      v->power = v->power * alb;
      v->bsdf_x_ipdf = rcp(pp);
      if (u.u[2] >= pp) return false;
      return true;
    }
    case MIRROR: {  // mirror.cpp:217-248
      v->specular = true;
      v->cos_in_sh = dot(v->wout, v->ns);
      v->win = -v->wout;
      v->win = v->win + (real(2) * v->cos_in_sh) * v->ns;
      if (dot(v->win, v->ng) > 0) return false;
      v->s_prev = rcp(v->cos_in_sh);
      v->s_next = rcp(v->cos_out_geo);
      v->bsdf_x_ipdf = 1;
      return true;
    }
    case GLASS: {  // glass.cpp:339-405
      real const ior_sq = glass_ior_sq(bs, v->back);
      real const nk = dot(v->wout, v->ns);
      real const bb = ior_sq + nk * nk - 1;
      v->win = -v->wout;
      v->specular = true;
      if (bb > 0) {
        real const g = std::sqrt(bb);
        real const gpc = g + nk, gmc = g - nk;
        real const alpha = gmc / gpc;
        real const beta = (nk * gpc - 1) / (nk * gmc + 1);
        real const fr = real(0.5) * alpha * alpha * (beta * beta + 1);
        if (u.u[2] >= fr) {
          v->s_prev = v->s_next = 1 - fr;
          v->win = (v->win - gmc * v->ns) / std::sqrt(ior_sq);
          v->cos_in_sh = -dot(v->win, v->ng);
          if (v->cos_in_sh <= 0) return false;
          v->bsdf_x_ipdf = 1;
          return true;
        }
        v->s_prev = v->s_next = fr;
      } else {
        v->s_prev = v->s_next = 1;
      }
      v->win = v->win + (real(2) * nk) * v->ns;
      v->cos_in_sh = dot(v->win, v->ng);
      if (v->cos_in_sh <= 0) return false;
      v->bsdf_x_ipdf = 1;
      return true;
    }
    case METAL: {  // metal.cpp:292-340
      v->specular = false;
      V3 hv;
      real const nh = glossy_half(u.u[0], u.u[1], bs.shininess, v->t, v->ns, v->b, &hv);
      real const hk = dot(hv, v->wout);
      v->win = (2 * hk) * hv - v->wout;
      if (dot(v->ng, v->win) <= 0) return false;
      v->cos_in_sh = dot(v->ns, v->win);
      if (v->cos_in_sh <= 0) return false;
      real const cos_sh = dot(v->ns, v->wout);
      v->bsdf_x_ipdf = fake_fresnel(hk) * v->cos_in_sh / hmax(cos_sh, v->cos_in_sh);
      v->s_prev = v->s_next = real(double(bs.shininess + 1) * (M_1_PI / 8) *
                                   double(xyz_pow(nh, bs.shininess)) / double(hk));
      v->s_prev *= rcp(v->cos_in_sh);
      v->s_next *= rcp(v->cos_out_geo);
      return true;
    }
    default:
      return false;
  }
}

// GetConstFactor() = (f / F)^2 / A (thinlenscamera.hpp:83-85); a pinhole given by
// its field of view uses its lens-film distance V for f.
real bdpt_const_factor(Camera const& c, float focal_length) {
  real const f = (focal_length > 0) ? real(focal_length) : c.v;
  real const r = f / c.film;
  return (r * r) / c.aspect;
}

// Pinhole GetFilmPosition (thinlenscamera.cpp:125-157)
bool film_position(Camera const& c, V3 const& d, real* fx, real* fy) {
  real const lx = dot(d, c.right), ly = dot(d, c.up), lz = dot(d, c.back);
  real const k = (c.v / -lz) / c.film;
  *fx = real(0.5) + lx * k / c.aspect;
  *fy = real(0.5) + ly * k;
  return (0 <= *fx) && (*fx < 1) && (0 <= *fy) && (*fy < 1);
}

struct Splat { uint32_t pixel; V3 value; };

inline real fourth(real x) { real const s = x * x; return s * s; }

// One sample of pixel (px, py): returns the pixel's own colour; light-tracing
// (t = 1) contributions to other pixels are appended to *splats in s order.
V3 bdpt_sample(Scene const& sc, Camera const& cam, real C, uint32_t kmax, uint64_t seed,
               uint32_t W, uint32_t H, uint32_t px, uint32_t py, uint32_t smp,
               std::vector<Splat>* splats, PathStats* st) {
  uint32_t const pixel = py * W + px;
  BVert L[20], E[20];
  real const A = sc.light_area();
  auto find = [&](V3 const& o, V3 const& d, Hit* h) {
    if (st) st->queries++;
    Counters c;
    int const r = scene_find(sc, o, d, h, &c);
    if (st) { st->c.nodes += c.nodes; st->c.tris += c.tris; }
    return r;
  };
  Draw const de = draw(pixel, seed, smp, 0, kBdptStreamEye);
  real const x = (real(int(px)) + de.u[0]) / real(int(W));
  real const y = (real(int(py)) + de.u[1]) / real(int(H));
  Draw const dl = draw(pixel, seed, smp, 0, kBdptStreamLight);
  // SampleLightPosition (scene.cpp:42-78) + Triangle::Sample's basis (triangle.cpp:232-263)
  {
    Vertex lv;
    int n = -1;
    sample_light(sc, dl.u[0], dl.u[1], &lv, &n);
    L[1].tri = lv.tri;
    if (lv.tri >= 0) {
      L[1].pos = lv.pos;
      L[1].ns = lv.ns;
      L[1].ng = lv.ng;
      V3 t(-lv.ng[2], 0, lv.ng[0]);
      real const l2 = length_squared(t);
      if (l2 > 0) {
        t = t * rsqrt(l2);
        L[1].b = cross(t, lv.ng);
      } else {
        L[1].b = V3(0, 0, (lv.ng[1] > 0) ? 1 : -1);
        t = cross(lv.ng, L[1].b);
      }
      L[1].t = t;
      L[1].s_next = rcp(A);
      L[1].specular = false;
      L[1].power = sc.bsdfs[sc.tris[size_t(lv.tri)].bsdf].phi;
    }
  }
  L[2].tri = -1;
  // SampleLensPosition (scene.cpp:80-98): eye, camera basis, normal = -back
  E[1].tri = -1;
  E[1].pos = cam.eye;
  E[1].t = cam.right;
  E[1].b = cam.up;
  E[1].ng = E[1].ns = -cam.back;
  E[1].s_next = 0;
  E[1].power = V3(1, 1, 1);

  // BuildLightsPath
  if (L[1].tri >= 0) {
    L[0].s_next = rcp(A);
    L[1].cos_in_sh = 1;
    L[1].cos_out_geo = diffused(dl.u[2], dl.u[3], L[1].t, L[1].ng, L[1].b, &L[1].wout);
    L[1].s_prev = 1;
    L[1].s_next = real(M_1_PI);
    L[1].g = 1;
    L[1].bsdf_x_ipdf = 1;
    for (uint32_t s = 1;;) {
      BVert const& o = L[s];
      BVert& nw = L[s + 1];
      Hit h;
      nw.tri = find(o.pos, o.wout, &h);
      if (nw.tri < 0) break;
      Tri const& tr = sc.tris[size_t(nw.tri)];
      Bsdf const& bs = sc.bsdfs[tr.bsdf];
      if (bs.type == LIGHT) { nw.tri = -1; break; }
      geometric_basis(tr, h, &nw);
      nw.back = h.back;
      nw.win = -o.wout;
      nw.cos_in_sh = dot(nw.win, nw.ns);
      if (nw.cos_in_sh <= 0) { nw.tri = -1; break; }
      nw.power = o.power * o.bsdf_x_ipdf;
      nw.pos = o.wout * h.t_max + o.pos;
      nw.g = o.cos_out_geo * nw.cos_in_sh / (h.t_max * h.t_max);
      if (++s >= kmax) { L[s + 1].tri = -1; break; }
      if (!lights_to_camera(sc, bs, tr, h, &L[s], draw(pixel, seed, smp, s, kBdptStreamLight))) {
        L[s + 1].tri = -1;
        break;
      }
    }
  }

  // EvaluatePathS1: t = 1, light vertices seen by the camera
  for (uint32_t s = 2; s < kmax; ++s) {
    BVert const& lv = L[s];
    if (lv.tri < 0) break;
    uint32_t const ty = sc.bsdfs[sc.tris[size_t(lv.tri)].bsdf].type;
    if (ty == MIRROR || ty == GLASS) continue;
    V3 const d = normalize(lv.pos - E[1].pos);
    real const cog = -dot(d, lv.ng);
    if (cog <= 0) continue;
    if (-dot(d, lv.ns) <= 0) continue;
    real const cis = dot(d, E[1].ns);
    if (cis <= 0) continue;
    real fx, fy;
    if (!film_position(cam, lv.pos - E[1].pos, &fx, &fy)) continue;
    Hit h;
    if (find(E[1].pos, d, &h) != lv.tri) continue;
    // fx < 1 can still round fx * W up to W (one past the row in the
    // reference): keep the last column / row
    uint32_t const jx = std::min(uint32_t(fx * real(int(W))), W - 1u);
    uint32_t const jy = std::min(uint32_t(fy * real(int(H))), H - 1u);
    real const flux = C / fourth(cis);
    real const G = cis * cog / (h.t_max * h.t_max);
    real wst = 1;
    L[s + 1].s_prev = flux;
    real pst = 1;
    for (uint32_t i = s; i > 0; --i) {
      pst *= L[i + 1].s_prev / L[i - 1].s_next;
      if (!L[i].specular && !L[i - 1].specular) {
        real const q = pst * G / L[i].g;
        wst += q * q;
      }
    }
    wst = rcp(wst);
    double const sc1 = double(flux * G) * M_1_PI * double(wst);
    V3 const pw = lv.power * E[1].power;
    if (splats) splats->push_back(Splat{jy * W + jx, pw * real(sc1)});
  }

  // BuildTheEyePath (pinhole)
  E[1].win = primary_dir(cam, x, y);
  E[1].power = V3(1, 1, 1);
  E[1].cos_in_sh = dot(E[1].win, E[1].ns);
  E[1].cos_out_geo = 1;
  E[1].s_prev = C / fourth(E[1].cos_in_sh);
  E[1].s_next = 1;
  E[1].g = 1;
  E[1].bsdf_x_ipdf = 1;
  for (uint32_t t = 1;;) {
    BVert const& o = E[t];
    BVert& nw = E[t + 1];
    Hit h;
    nw.tri = find(o.pos, o.win, &h);
    if (nw.tri < 0) break;
    Tri const& tr = sc.tris[size_t(nw.tri)];
    Vertex sv;
    shading_basis(tr, h, &sv);
    nw.ns = sv.ns; nw.ng = sv.ng; nw.t = sv.t; nw.b = sv.b;
    nw.back = h.back;
    nw.wout = -o.win;
    if (dot(nw.wout, nw.ns) <= 0) { nw.tri = -1; break; }
    nw.power = o.power;
    nw.power = nw.power * o.bsdf_x_ipdf;
    nw.pos = o.win * h.t_max + o.pos;
    nw.cos_out_geo = dot(nw.wout, nw.ng);
    nw.g = nw.cos_out_geo * o.cos_in_sh / (h.t_max * h.t_max);
    Bsdf const& bs = sc.bsdfs[tr.bsdf];
    if (++t >= kmax) { E[t + 1].tri = -1; break; }
    if (bs.type == LIGHT) { E[t + 1].tri = -1; break; }
    if (!camera_to_lights(sc, bs, tr, h, &E[t], draw(pixel, seed, smp, t, kBdptStreamEye))) {
      E[t + 1].tri = -1;
      break;
    }
  }

  // EvaluatePathST
  V3 color(0, 0, 0);
  for (uint32_t t = 2; t < kmax; ++t) {
    BVert& ev = E[t];
    if (ev.tri < 0) break;
    Bsdf const& eb = sc.bsdfs[sc.tris[size_t(ev.tri)].bsdf];
    if (eb.type == LIGHT) {
      real wst = 1;
      if (t > 2) {
        E[t + 1].s_next = L[0].s_next;
        ev.s_next = real(M_1_PI);
        real pst = 1;
        for (uint32_t i = t; i > 1; --i) {
          pst *= E[i + 1].s_next / E[i - 1].s_prev;
          if (!E[i].specular && !E[i - 1].specular) {
            real const q = pst / E[i].g;
            wst += q * q;
          }
        }
        wst = rcp(wst);
      }
      double const sc1 = double(rcp(A)) * M_1_PI * double(wst);
      V3 const pw = eb.phi * ev.power;  // the power of the light that was hit
      color = color + pw * real(sc1);
      break;
    }
    if (eb.type == MIRROR || eb.type == GLASS) continue;
    for (uint32_t s = 1; (s + t - 1) < kmax; ++s) {
      BVert const& lv = L[s];
      if (lv.tri < 0) break;
      uint32_t const lt = sc.bsdfs[sc.tris[size_t(lv.tri)].bsdf].type;
      if (lt == MIRROR || lt == GLASS) continue;
      V3 const d = normalize(ev.pos - lv.pos);
      real const cog = dot(d, lv.ng);
      if (cog <= 0) continue;
      if (dot(d, lv.ns) <= 0) continue;
      if (-dot(d, ev.ng) <= 0) continue;
      real const cis = -dot(d, ev.ns);
      if (cis <= 0) continue;
      Hit h;
      if (find(lv.pos, d, &h) != ev.tri) continue;
      real const G = cis * cog / (h.t_max * h.t_max);
      real wst = 1;
      L[s + 1].s_prev = real(M_1_PI);
      real pst = 1;
      for (uint32_t i = s; i > 0; --i) {
        pst *= L[i + 1].s_prev / L[i - 1].s_next;
        if (!L[i].specular && !L[i - 1].specular) {
          real const q = pst * G / L[i].g;
          wst += q * q;
        }
      }
      E[t + 1].s_next = real(M_1_PI);
      pst = 1;
      for (uint32_t i = t; i > 1; --i) {
        pst *= E[i + 1].s_next / E[i - 1].s_prev;
        if (!E[i].specular && !E[i - 1].specular) {
          real const q = pst * G / E[i].g;
          wst += q * q;
        }
      }
      wst = rcp(wst);
      double const sc1 = double(G) * (M_1_PI * M_1_PI) * double(wst);
      V3 const pw = lv.power * ev.power;
      color = color + pw * real(sc1);
    }
  }
  return color;
}

int n_threads(int req) {
#ifdef _OPENMP
  return req > 0 ? req : omp_get_max_threads();
#else
  (void)req;
  return 1;
#endif
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

// Scene with triangles in caller order (caller order defines primitive ids).
void* orc_scene_create(uint32_t n_tris, const float* v0, const float* v1, const float* v2,
                       const float* n0, const float* n1, const float* n2, const float* uv,
                       const uint32_t* tri_bsdf, uint32_t n_bsdfs, const uint32_t* bsdf_type,
                       const float* bsdf_rgb, const float* bsdf_shininess, const float* bsdf_ior,
                       const int32_t* bsdf_tex, uint32_t n_tex, const uint32_t* tex_wh,
                       const float* texels, uint32_t max_bucket, uint32_t max_depth, int use_tree) {
  Scene* sc = new Scene();
  sc->tris.reserve(n_tris);
  for (uint32_t i = 0; i < n_tris; ++i) {
    V3 p0(v0[3 * i], v0[3 * i + 1], v0[3 * i + 2]);
    V3 p1(v1[3 * i], v1[3 * i + 1], v1[3 * i + 2]);
    V3 p2(v2[3 * i], v2[3 * i + 1], v2[3 * i + 2]);
    V3 ns[3];
    if (n0) {
      ns[0] = V3(n0[3 * i], n0[3 * i + 1], n0[3 * i + 2]);
      ns[1] = V3(n1[3 * i], n1[3 * i + 1], n1[3 * i + 2]);
      ns[2] = V3(n2[3 * i], n2[3 * i + 1], n2[3 * i + 2]);
    }
    sc->tris.push_back(make_tri(tri_bsdf ? tri_bsdf[i] : 0, p0, p1, p2, n0 ? ns : nullptr,
                                uv ? uv + 6 * size_t(i) : nullptr));
  }
  sc->bsdfs.resize(n_bsdfs);
  for (uint32_t b = 0; b < n_bsdfs; ++b) {
    Bsdf& bs = sc->bsdfs[b];
    bs.type = bsdf_type[b];
    bs.rgb = V3(bsdf_rgb[3 * b], bsdf_rgb[3 * b + 1], bsdf_rgb[3 * b + 2]);
    bs.shininess = bsdf_shininess ? bsdf_shininess[b] : real(0);
    bs.ior = bsdf_ior ? bsdf_ior[b] : real(1.5);
    bs.tex = bsdf_tex ? bsdf_tex[b] : -1;
  }
  size_t off = 0;
  for (uint32_t t = 0; t < n_tex; ++t) {
    Texture tx;
    tx.w = tex_wh[2 * t];
    tx.h = tex_wh[2 * t + 1];
    tx.texels.assign(texels + off, texels + off + size_t(tx.w) * tx.h * 3);
    off += size_t(tx.w) * tx.h * 3;
    sc->textures.push_back(std::move(tx));
  }
  // light list and area CDF in caller order: xyz/core/scene_mqo.cpp:509-518
  sc->cdf.push_back(0);
  for (uint32_t i = 0; i < n_tris; ++i) {
    if (sc->bsdfs[sc->tris[i].bsdf].type == LIGHT) {
      sc->lights.push_back(i);
      sc->cdf.push_back(sc->cdf.back() + sc->tris[i].area());
    }
  }
  real const A = sc->cdf.back();
  for (Bsdf& bs : sc->bsdfs) {
    if (bs.type == LIGHT) {
      double const s = 3.14159265358979323846 * double(A);
      bs.phi = V3(real(double(bs.rgb[0]) * s), real(double(bs.rgb[1]) * s), real(double(bs.rgb[2]) * s));
    }
  }
  sc->use_tree = use_tree != 0;
  sc->tree.tris = &sc->tris;
  sc->tree.max_bucket = max_bucket;
  sc->tree.max_depth = max_depth;
  if (sc->use_tree) build_tree(sc->tree);
  return sc;
}

void orc_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

float orc_light_area(void* s) { return static_cast<Scene*>(s)->light_area(); }

// Tree structure: counts[0] = nodes, counts[1] = leaves (incl. null), counts[2] = refs,
// counts[3] = max depth reached.  Bounds: 6 floats (min, upper bound).
void orc_tree_counts(void* s, uint64_t* counts, float* bounds) {
  Tree const& tr = static_cast<Scene*>(s)->tree;
  uint64_t leaves = 0, refs = 0;
  for (Node const& n : tr.nodes) {
    if (n.kind >= 3) { leaves++; refs += n.set.size(); }
  }
  counts[0] = tr.nodes.size();
  counts[1] = leaves;
  counts[2] = refs;
  if (bounds) {
    for (int k = 0; k < 3; ++k) { bounds[k] = tr.bmin[k]; bounds[3 + k] = tr.bmax[k]; }
  }
}

// Pre-order dump: kind[i] (0..2 axis, 3 leaf, 4 null leaf), split[i], and for
// leaves the range [leaf_off[i], leaf_off[i]+leaf_cnt[i]) into refs.
static void dump_rec(Tree const& tr, int ni, uint32_t* kind, float* split, uint32_t* loff,
                     uint32_t* lcnt, uint32_t* refs, uint64_t* pos, uint64_t* rpos) {
  Node const& n = tr.nodes[size_t(ni)];
  uint64_t const i = (*pos)++;
  kind[i] = uint32_t(n.kind);
  split[i] = n.value;
  loff[i] = 0;
  lcnt[i] = 0;
  if (n.kind >= 3) {
    loff[i] = uint32_t(*rpos);
    lcnt[i] = uint32_t(n.set.size());
    for (uint32_t id : n.set) refs[(*rpos)++] = id;
    return;
  }
  dump_rec(tr, n.child[0], kind, split, loff, lcnt, refs, pos, rpos);
  dump_rec(tr, n.child[1], kind, split, loff, lcnt, refs, pos, rpos);
}

void orc_tree_dump(void* s, uint32_t* kind, float* split, uint32_t* leaf_off, uint32_t* leaf_cnt,
                   uint32_t* refs) {
  Tree const& tr = static_cast<Scene*>(s)->tree;
  uint64_t pos = 0, rpos = 0;
  if (tr.root >= 0) dump_rec(tr, tr.root, kind, split, leaf_off, leaf_cnt, refs, &pos, &rpos);
}

// Closest hit for n rays (AoS float3 origin/direction).  prim = -1 on a miss,
// t = +inf on a miss.  Optional outputs may be NULL.  Returns wall seconds.
double orc_closest_hit(void* s, uint64_t n, const float* org, const float* dir, int32_t* prim,
                       float* t, float* f, float* g, uint8_t* back, uint32_t* cnt_nodes,
                       uint32_t* cnt_tris, int threads) {
  Scene const& sc = *static_cast<Scene*>(s);
  auto t0 = std::chrono::steady_clock::now();
  int const nt = n_threads(threads);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    V3 o(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
    V3 d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    Hit h;
    Counters c;
    int const r = scene_find(sc, o, d, &h, &c);
    if (prim) prim[i] = r;
    if (t) t[i] = r >= 0 ? h.t_max : std::numeric_limits<float>::infinity();
    if (f) f[i] = r >= 0 ? h.f : 0.0f;
    if (g) g[i] = r >= 0 ? h.g : 0.0f;
    if (back) back[i] = r >= 0 ? uint8_t(h.back) : 0;
    if (cnt_nodes) cnt_nodes[i] = c.nodes;
    if (cnt_tris) cnt_tris[i] = c.tris;
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Brute-force closest hit (xyz/core/scene.hpp:283-293), regardless of use_tree.
void orc_brute_force(void* s, uint64_t n, const float* org, const float* dir, int32_t* prim,
                     float* t, int threads) {
  Scene const& sc = *static_cast<Scene*>(s);
  int const nt = n_threads(threads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    V3 o(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
    V3 d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    Hit h;
    int const r = brute_force(sc.tris, o, d, &h);
    prim[i] = r;
    t[i] = r >= 0 ? h.t_max : std::numeric_limits<float>::infinity();
  }
}

// Single-triangle test (xyz/geom/triangle.cpp:127-165) with a caller interval.
int orc_intersect_triangle(const float* v0, const float* v1, const float* v2, const float* o,
                           const float* d, float t_min, float t_max, float* out4) {
  Tri tr = make_tri(0, V3(v0[0], v0[1], v0[2]), V3(v1[0], v1[1], v1[2]), V3(v2[0], v2[1], v2[2]),
                    nullptr, nullptr);
  Hit h;
  h.t_min = t_min;
  h.t_max = t_max;
  h.f = h.g = 0;
  h.back = false;
  bool const ok = intersect_df(tr, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), &h);
  out4[0] = h.t_max; out4[1] = h.f; out4[2] = h.g; out4[3] = h.back ? 1.0f : 0.0f;
  return ok ? 1 : 0;
}

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  U4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  U4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void orc_cos_sin(uint64_t n, const float* x, float* c, float* s) {
  for (uint64_t i = 0; i < n; ++i) { c[i] = xyz_cos(x[i]); s[i] = xyz_sin(x[i]); }
}
// The Metal lobe's half-vector sampler (xyz/stdafx.hpp:154-168) in the basis
// t = +x, n = +y, b = +z: out = n * 3 floats.
void orc_glossy_half(uint64_t n, const float* mu_t, const float* mu_p, float shininess, float* out) {
  V3 const t(1, 0, 0), nn(0, 1, 0), b(0, 0, 1);
  for (uint64_t i = 0; i < n; ++i) {
    V3 h;
    glossy_half(mu_t[i], mu_p[i], shininess, t, nn, b, &h);
    for (int c = 0; c < 3; ++c) out[3 * i + c] = h[c];
  }
}
void orc_pow(uint64_t n, const float* x, const float* y, float* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = xyz_pow(x[i], y[i]);
}

// Shading basis (xyz/geom/triangle.cpp:194-219) of triangle `tri` at (f, g, back).
void orc_shading_basis(void* s, uint32_t tri, float f, float g, int back, float* out12) {
  Scene const& sc = *static_cast<Scene*>(s);
  Hit h; h.f = f; h.g = g; h.back = back != 0;
  Vertex v;
  shading_basis(sc.tris[tri], h, &v);
  for (int k = 0; k < 3; ++k) {
    out12[k] = v.ns[k]; out12[3 + k] = v.ng[k]; out12[6 + k] = v.t[k]; out12[9 + k] = v.b[k];
  }
}

// Light sample (xyz/core/scene.cpp:42-78).  out: pos3, ns3, ng3, light tri.
int orc_sample_light(void* s, float u, float v, float* out9) {
  Scene const& sc = *static_cast<Scene*>(s);
  Vertex lv; int n = -1;
  sample_light(sc, u, v, &lv, &n);
  if (lv.tri < 0) return -1;
  for (int k = 0; k < 3; ++k) { out9[k] = lv.pos[k]; out9[3 + k] = lv.ns[k]; out9[6 + k] = lv.ng[k]; }
  return lv.tri;
}

// cam: 12 floats (see make_camera); (ls, lt) = lens numbers, unused by the pinhole.
void orc_primary_ray(const float* cam, uint32_t W, uint32_t H, float s, float t, float ls,
                     float lt, float* org3, float* dir3) {
  Camera const c = make_camera(cam, W, H);
  V3 o, d;
  primary_ray(c, s, t, ls, lt, &o, &d);
  for (int k = 0; k < 3; ++k) { org3[k] = o[k]; dir3[k] = d[k]; }
}

void orc_primary_dir(const float* cam, uint32_t W, uint32_t H, float s, float t, float* out3) {
  Camera c = make_camera(cam, W, H);
  V3 d = primary_dir(c, s, t);
  for (int k = 0; k < 3; ++k) out3[k] = d[k];
}

// Progressive path tracing (xyz/integrator/integrator.cpp:119-128) over rows
// [y0, y1) for samples [sample0, sample0 + spp).  film (W*H*3) receives the
// per-pixel SUM of sample colours, added sample pass by sample pass.
// stats[0] = closest-hit queries, stats[1] = nodes visited, stats[2] = tris tested,
// stats[3] = shaded vertices (hits that reach the NEE step).  Returns wall seconds.
double orc_render(void* s, const float* cam, uint32_t W, uint32_t H, uint32_t spp,
                  uint32_t sample0, uint32_t max_depth, uint32_t rr_start, float rr_threshold,
                  uint64_t seed, uint32_t y0, uint32_t y1, float* film, uint64_t* stats,
                  int threads) {
  Scene const& sc = *static_cast<Scene*>(s);
  Camera const c = make_camera(cam, W, H);
  PTParams pp;
  pp.kmax = max_depth + 1;
  pp.rr_start = rr_start;
  pp.rr_threshold = rr_threshold;
  pp.seed = seed;
  int const nt = n_threads(threads);
  std::atomic<uint64_t> q(0), nn(0), ntr(0), nv(0);
  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t y = int64_t(y0); y < int64_t(y1); ++y) {
    PathStats ps;
    for (uint32_t x = 0; x < W; ++x) {
      float* px = film + (size_t(y) * W + x) * 3;
      for (uint32_t smp = sample0; smp < sample0 + spp; ++smp) {
        V3 col = pt_path(sc, c, pp, W, H, x, uint32_t(y), smp, stats ? &ps : nullptr, nullptr);
        px[0] += col[0]; px[1] += col[1]; px[2] += col[2];
      }
    }
    if (stats) { q += ps.queries; nn += ps.c.nodes; ntr += ps.c.tris; nv += ps.vertices; }
  }
  if (stats) { stats[0] = q; stats[1] = nn; stats[2] = ntr; stats[3] = nv; }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Bidirectional path tracing (BidirectionalPathTracer::run,
// xyz/integrator/integrator.cpp:1025-1081): for each pass (sample index), every
// pixel's sample adds its own colour to own_film and its light-tracing
// contributions to splat_film, in the reference's pass / row / column / s
// order.  Films are W*H*3 sums (row 0 = bottom).  stats: queries, nodes, tris.
double orc_render_bdpt(void* s, const float* cam, uint32_t W, uint32_t H, uint32_t spp,
                       uint32_t sample0, uint32_t max_depth, uint64_t seed, float* own_film,
                       float* splat_film, uint64_t* stats, int threads) {
  Scene const& sc = *static_cast<Scene*>(s);
  Camera const c = make_camera(cam, W, H);
  real const C = bdpt_const_factor(c, cam[10]);
  uint32_t const kmax = max_depth + 1;
  if (kmax > 17) return -1.0;  // max_depth <= 16, as the device
  int const nt = n_threads(threads);
  std::atomic<uint64_t> q(0), nn(0), ntr(0);
  std::vector<std::vector<Splat>> rows(H);
  auto t0 = std::chrono::steady_clock::now();
  for (uint32_t smp = sample0; smp < sample0 + spp; ++smp) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t y = 0; y < int64_t(H); ++y) {
      PathStats ps;
      rows[size_t(y)].clear();
      for (uint32_t x = 0; x < W; ++x) {
        V3 const col = bdpt_sample(sc, c, C, kmax, seed, W, H, x, uint32_t(y), smp,
                                   &rows[size_t(y)], stats ? &ps : nullptr);
        float* px = own_film + (size_t(y) * W + x) * 3;
        px[0] += col[0]; px[1] += col[1]; px[2] += col[2];
      }
      if (stats) { q += ps.queries; nn += ps.c.nodes; ntr += ps.c.tris; }
    }
    for (uint32_t y = 0; y < H; ++y)
      for (Splat const& sp : rows[y]) {
        float* px = splat_film + size_t(sp.pixel) * 3;
        px[0] += sp.value[0]; px[1] += sp.value[1]; px[2] += sp.value[2];
      }
  }
  if (stats) { stats[0] = q; stats[1] = nn; stats[2] = ntr; }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Reference gradient of the image loss sum_p w_p . S_p, where S_p is the SUM of
// the `spp` sample colours of pixel p (the film) and w = d_film (W*H*3) =
// dLoss/dS, with respect to every texel of every texture (d_tex, concatenated
// like the scene's texels) and to each BSDF's emission L_e (d_emit, 3 per
// BSDF; non-zero for lights only).  Each sample's path is recorded
// (PathRecord) and swept from its end: R = E, then for vertex i from the last
// to the first g_i = Y_i + sigma_i R, dS/da_i += beta_i g_i, R = a_i g_i.
// Russian-roulette probabilities and discrete choices are held fixed.
double orc_render_backward(void* s, const float* cam, uint32_t W, uint32_t H, uint32_t spp,
                           uint32_t sample0, uint32_t max_depth, uint32_t rr_start,
                           float rr_threshold, uint64_t seed, const float* d_film, double* d_tex,
                           double* d_emit, int threads) {
  Scene const& sc = *static_cast<Scene*>(s);
  Camera const c = make_camera(cam, W, H);
  PTParams pp;
  pp.kmax = max_depth + 1;
  pp.rr_start = rr_start;
  pp.rr_threshold = rr_threshold;
  pp.seed = seed;
  std::vector<size_t> tex_off(sc.textures.size() + 1, 0);
  for (size_t i = 0; i < sc.textures.size(); ++i)
    tex_off[i + 1] = tex_off[i] + size_t(sc.textures[i].w) * sc.textures[i].h * 3;
  int const nt = n_threads(threads);
  auto t0 = std::chrono::steady_clock::now();
  size_t const nb = sc.bsdfs.size();
  std::vector<std::vector<double>> gt(size_t(nt), std::vector<double>(tex_off.back(), 0.0));
  std::vector<std::vector<double>> ge(size_t(nt), std::vector<double>(nb * 3, 0.0));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t y = 0; y < int64_t(H); ++y) {
#ifdef _OPENMP
    int const tid = omp_get_thread_num();
#else
    int const tid = 0;
#endif
    PathRecord rec;
    for (uint32_t x = 0; x < W; ++x) {
      size_t const p = size_t(y) * W + x;
      double const w[3] = {d_film[3 * p], d_film[3 * p + 1], d_film[3 * p + 2]};
      for (uint32_t smp = sample0; smp < sample0 + spp; ++smp) {
        rec.v.clear();
        rec.emission = V3(0, 0, 0);
        rec.d_emission.assign(nb * 3, 0.0);
        pt_path(sc, c, pp, W, H, x, uint32_t(y), smp, nullptr, &rec);
        double R[3] = {double(rec.emission[0]), double(rec.emission[1]), double(rec.emission[2])};
        for (size_t i = rec.v.size(); i-- > 0;) {
          VertexRecord const& r = rec.v[i];
          double g[3];
          for (int ch = 0; ch < 3; ++ch) g[ch] = double(r.nee[ch]) + double(r.sigma) * R[ch];
          if (r.texture >= 0) {
            size_t const base = tex_off[size_t(r.texture)] + size_t(r.texel) * 3;
            for (int ch = 0; ch < 3; ++ch) gt[size_t(tid)][base + ch] += w[ch] * double(r.beta[ch]) * g[ch];
          }
          for (int ch = 0; ch < 3; ++ch) R[ch] = double(r.albedo[ch]) * g[ch];
        }
        for (size_t b = 0; b < nb; ++b) {
          if (sc.bsdfs[b].type != LIGHT) continue;
          for (int ch = 0; ch < 3; ++ch) ge[size_t(tid)][3 * b + ch] += w[ch] * rec.d_emission[3 * b + ch];
        }
      }
    }
  }
  for (int t = 0; t < nt; ++t) {
    for (size_t i = 0; i < tex_off.back(); ++i) d_tex[i] += gt[size_t(t)][i];
    for (size_t i = 0; i < nb * 3; ++i) d_emit[i] += ge[size_t(t)][i];
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Importon map (Scene::CreateLightImportanceMap, scene_photonmap.cpp:272-334):
// n_emit light paths of at most max_depth + 1 segments.  The map keeps the
// importons in path order, then vertex order.
void* orc_importon_map_create(void* s, uint32_t n_emit, uint32_t max_depth, uint64_t seed) {
  auto* m = new std::vector<Importon>();
  emit_importons(*static_cast<Scene*>(s), n_emit, max_depth + 1, seed, m);
  return m;
}
uint64_t orc_importon_count(void* m) { return static_cast<std::vector<Importon>*>(m)->size(); }
// out: count * 6 floats (position, incoming direction)
void orc_importon_export(void* m, float* out) {
  auto const& v = *static_cast<std::vector<Importon>*>(m);
  for (size_t i = 0; i < v.size(); ++i)
    for (int c = 0; c < 3; ++c) {
      out[6 * i + c] = v[i].pos[c];
      out[6 * i + 3 + c] = v[i].dir[c];
    }
}
void orc_azimuth_octant(uint64_t n, const float* x, const float* y, int32_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = azimuth_octant(x[i], y[i]);
}
// Importon map from caller importons (n * 6 floats: position, direction).
void* orc_importon_map_from(const float* data, uint64_t n) {
  auto* m = new std::vector<Importon>(n);
  for (uint64_t i = 0; i < n; ++i) {
    (*m)[i].pos = V3(data[6 * i], data[6 * i + 1], data[6 * i + 2]);
    (*m)[i].dir = V3(data[6 * i + 3], data[6 * i + 4], data[6 * i + 5]);
  }
  return m;
}
// guided_sample for n draws (u: 3 per draw: bin, cos^2 theta, azimuth) at the
// vertex basis12 = pos, ns, t, b with cdf (17 floats).  out: 5 per draw:
// direction, bin density, density per solid angle.
void orc_guided_sample(const float* cdf, const float* basis12, uint64_t n, const float* u,
                       float* out) {
  Vertex q;
  q.pos = V3(basis12[0], basis12[1], basis12[2]);
  q.ns = V3(basis12[3], basis12[4], basis12[5]);
  q.t = V3(basis12[6], basis12[7], basis12[8]);
  q.b = V3(basis12[9], basis12[10], basis12[11]);
  real c[kN + 1];
  for (int i = 0; i <= kN; ++i) c[i] = cdf[i];
  for (uint64_t i = 0; i < n; ++i) {
    V3 d;
    real fn;
    real const dens = guided_sample(c, q, u[3 * i], u[3 * i + 1], u[3 * i + 2], &d, &fn);
    out[5 * i] = d[0];
    out[5 * i + 1] = d[1];
    out[5 * i + 2] = d[2];
    out[5 * i + 3] = dens;
    out[5 * i + 4] = fn;
  }
}
// importance_density of n directions for a given cdf (the query without a map)
void orc_importance_density(const float* cdf, const float* basis12, uint64_t n, const float* dirs,
                            float* out) {
  Vertex q;
  q.ns = V3(basis12[3], basis12[4], basis12[5]);
  q.t = V3(basis12[6], basis12[7], basis12[8]);
  q.b = V3(basis12[9], basis12[10], basis12[11]);
  real c[kN + 1];
  for (int i = 0; i <= kN; ++i) c[i] = cdf[i];
  for (uint64_t i = 0; i < n; ++i)
    out[i] = importance_density(c, q, V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}
void orc_importon_map_destroy(void* m) { delete static_cast<std::vector<Importon>*>(m); }

// ImportanceQuery at a vertex: cdf (17 floats) for position / shading normal /
// tangent / binormal in basis12 = pos, ns, t, b; density of each dir in dirs.
void orc_importance_query(void* m, const float* basis12, uint32_t n_dirs, const float* dirs,
                          float* cdf, float* density) {
  auto const& v = *static_cast<std::vector<Importon>*>(m);
  Vertex q;
  q.pos = V3(basis12[0], basis12[1], basis12[2]);
  q.ns = V3(basis12[3], basis12[4], basis12[5]);
  q.t = V3(basis12[6], basis12[7], basis12[8]);
  q.b = V3(basis12[9], basis12[10], basis12[11]);
  importance_cdf(v, q, cdf);
  for (uint32_t i = 0; i < n_dirs; ++i)
    density[i] = importance_density(cdf, q, V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

// Importance-driven path tracing (Scene::EvaluteImportanceDrivenPathTracing,
// scene_importancedrivenpathtracing.cpp:11-179): orc_render with the guided
// Lambert vertices.  film: per-pixel sums in sample order.
double orc_render_idpt(void* s, void* m, const float* cam, uint32_t W, uint32_t H, uint32_t spp,
                       uint32_t sample0, uint32_t max_depth, uint32_t rr_start, float rr_threshold,
                       uint64_t seed, uint32_t y0, uint32_t y1, float* film, uint64_t* stats,
                       int threads) {
  Scene const& sc = *static_cast<Scene*>(s);
  auto const* guide = static_cast<std::vector<Importon>*>(m);
  Camera const c = make_camera(cam, W, H);
  PTParams pp;
  pp.kmax = max_depth + 1;
  pp.rr_start = rr_start;
  pp.rr_threshold = rr_threshold;
  pp.seed = seed;
  int const nt = n_threads(threads);
  std::atomic<uint64_t> q(0), nn(0), ntr(0);
  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t y = int64_t(y0); y < int64_t(y1); ++y) {
    PathStats ps;
    for (uint32_t x = 0; x < W; ++x) {
      float* px = film + (size_t(y) * W + x) * 3;
      for (uint32_t smp = sample0; smp < sample0 + spp; ++smp) {
        V3 col = pt_path(sc, c, pp, W, H, x, uint32_t(y), smp, stats ? &ps : nullptr, nullptr,
                         guide);
        px[0] += col[0]; px[1] += col[1]; px[2] += col[2];
      }
    }
    if (stats) { q += ps.queries; nn += ps.c.nodes; ntr += ps.c.tris; }
  }
  if (stats) { stats[0] = q; stats[1] = nn; stats[2] = ntr; }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int orc_max_threads(void) { return n_threads(0); }

}  // extern "C"

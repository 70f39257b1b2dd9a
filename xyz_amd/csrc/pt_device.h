// Device functions of the unidirectional path tracer, shared by the wavefront
// renderer, the path-replay adjoint and the persistent batched kernel.
//
// Every function restates one piece of the reference's integrator in the same
// binary32 operation order (double where the reference promotes via M_1_PI or
// 0.5 literals), so that a path traced here is bit-identical to the same path
// traced on the CPU in that order.
#ifndef XYZ_PT_DEVICE_H_
#define XYZ_PT_DEVICE_H_

#include "device_common.h"

namespace xyzhip {

constexpr double kInvPi = 0.31830988618379067154;  // M_1_PI
constexpr double kPi = 3.14159265358979323846;

// Camera basis and lens-film distance, computed on the host
// (xyz/core/thinlenscamera.cpp:71-90 with an explicit vertical FOV).
struct DevCamera {
  float eye[3], right[3], up[3], back[3];
  float film, aspect, v;
  uint32_t width, height;
  uint32_t thin;                     // thin-lens model
  float focus, lens_r, focal_plane;  // L, f / (2 F#), F L / V
};

struct PTConst {
  uint32_t kmax;        // XYZ_CONFIG_kMaxRandomWalkDepth = max_depth + 1
  uint32_t rr_start;    // Russian roulette at vertices j > rr_start
  float rr_threshold;   // CONFIG_TERMINATE_THRESHOLD
  uint32_t seed_lo, seed_hi;
};

// ---------------------------------------------------------------- vectors
__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float3 scl3(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 scl3l(float s, float3 a) { return f3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float3 neg3(float3 a) { return f3(-a.x, -a.y, -a.z); }
// libhi/hi/math/vector3.hpp:119-131
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// libhi/hi/math/math.hpp:146-157,254-257: normalize(a) = a * (1 / sqrt(dot(a, a)))
__device__ __forceinline__ float rcpf(float x) { return recip_exact(x); }
__device__ __forceinline__ float3 normalize3(float3 a) { return scl3(a, recip_exact(sqrtf(dot3(a, a)))); }
__device__ __forceinline__ float3 ld3(const float4& v) { return f3(v.x, v.y, v.z); }

// ------------------------------------------------------- deterministic cos/sin
// cos/sin of a float angle in double precision: two-part pi/2 reduction and
// Horner evaluation of the Taylor polynomials (sin through r^17, cos through
// r^18, coefficients rounded to double).  The same operations run on the CPU
// oracle.
__device__ __forceinline__ double taylor_sin(double r) {
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
__device__ __forceinline__ double taylor_cos(double r) {
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
// cos and sin of a double angle (the thin lens's disk angle, evaluated in double
// by xyz/core/thinlenscamera.cpp:4-14), results kept in double.
__device__ __forceinline__ void cos_sin_d(double x, double* c, double* s) {
  double const k = floor(x * 0.63661977236758134308 + 0.5);
  double const P1 = 1.5707963267341256141662597656;
  double const P1t = 6.07710050650619224932e-11;
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
__device__ __forceinline__ void cos_sin(float a, float* c, float* s) {
  double const x = double(a);
  double const k = floor(x * 0.63661977236758134308 + 0.5);
  double const P1 = 1.5707963267341256141662597656;
  double const P1t = 6.07710050650619224932e-11;
  double const r = (x - k * P1) - k * P1t;
  long long const ki = (long long)k;
  int const q = int(((ki % 4) + 4) % 4);
  double const ts = taylor_sin(r), tc = taylor_cos(r);
  double cv, sv;
  switch (q) {
    case 0: cv = tc; sv = ts; break;
    case 1: cv = -ts; sv = tc; break;
    case 2: cv = -tc; sv = -ts; break;
    default: cv = ts; sv = -tc; break;
  }
  *c = float(cv);
  *s = float(sv);
}

// ------------------------------------------------------- deterministic pow
// powf of the Metal lobe (xyz/stdafx.hpp:160, xyz/bsdf/metal.cpp:101,146) in
// double, so that every build computes the same bits:
//   ln x  = E ln2 + 2 atanh(z), z = (m - 1)/(m + 1), x = m 2^E with m in
//           [sqrt(1/2), sqrt(2)]; atanh by its series through z^25
//           (|z| <= 0.1716, truncation below 1e-19);
//   exp t = 2^k exp(r), r = t - k ln2, |r| <= ln2/2; exp(r) by its Taylor
//           series through r^15 (truncation below 1e-19).
// ln2 is split in two (ln2_hi has 32 significant bits, so E ln2_hi and
// k ln2_hi are exact).  Special cases follow C's powf.
__device__ __forceinline__ float pow_positive(float x, float y) {
  uint32_t const b = __float_as_uint(x);
  uint32_t m = b & 0x7FFFFFu;
  int E = int((b >> 23) & 0xFFu) - 127;
  if ((b >> 23) == 0u) {  // subnormal: the leading one moved into the hidden bit
    int const sh = __clz(m) - 8;
    m = (m << sh) & 0x7FFFFFu;
    E = -126 - sh;
  }
  double md = __longlong_as_double((0x3FFll << 52) | ((long long)m << 29));
  if (md > 1.4142135623730951) {
    md = md * 0.5;
    E += 1;
  }
  double const ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  double const z = (md - 1.0) / (md + 1.0);
  double const z2 = z * z;
  double p = 0.04;  // 1/25, then 1/23, ..., 1/3, 1
  p = 0.043478260869565216 + z2 * p;
  p = 0.047619047619047616 + z2 * p;
  p = 0.05263157894736842 + z2 * p;
  p = 0.058823529411764705 + z2 * p;
  p = 0.06666666666666667 + z2 * p;
  p = 0.07692307692307693 + z2 * p;
  p = 0.09090909090909091 + z2 * p;
  p = 0.1111111111111111 + z2 * p;
  p = 0.14285714285714285 + z2 * p;
  p = 0.2 + z2 * p;
  p = 0.3333333333333333 + z2 * p;
  p = 1.0 + z2 * p;
  double const lnm = (2.0 * z) * p;
  double const L = (double(E) * ln2_hi) + ((double(E) * ln2_lo) + lnm);
  double const t = double(y) * L;
  if (t > 100.0) return __int_as_float(0x7F800000);
  if (t < -110.0) return 0.0f;
  double const k = floor(t * 1.44269504088896338700 + 0.5);
  double const r = (t - k * ln2_hi) - k * ln2_lo;
  double q = 7.647163731819816e-13;  // 1/15!, then 1/14!, ..., 1
  q = 1.1470745597729725e-11 + r * q;
  q = 1.6059043836821613e-10 + r * q;
  q = 2.08767569878681e-09 + r * q;
  q = 2.505210838544172e-08 + r * q;
  q = 2.755731922398589e-07 + r * q;
  q = 2.7557319223985893e-06 + r * q;
  q = 2.48015873015873e-05 + r * q;
  q = 0.0001984126984126984 + r * q;
  q = 0.001388888888888889 + r * q;
  q = 0.008333333333333333 + r * q;
  q = 0.041666666666666664 + r * q;
  q = 0.16666666666666666 + r * q;
  q = 0.5 + r * q;
  q = 1.0 + r * q;
  q = 1.0 + r * q;
  return float(q * __longlong_as_double(((long long)(k) + 1023ll) << 52));
}
__device__ __forceinline__ float pow_det(float x, float y) {
  float const inf = __int_as_float(0x7F800000);
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (isnan(x) || isnan(y)) return x + y;
  bool neg = false;
  if (signbit(x)) {
    if (isinf(y)) {
      if (x == -1.0f) return 1.0f;
      return ((-x < 1.0f) == (y > 0.0f)) ? 0.0f : inf;
    }
    bool const integer = floorf(y) == y;
    if (!integer && x != 0.0f && !isinf(x)) return __int_as_float(0x7FC00000);
    neg = integer && fabsf(y) < 16777216.0f && ((long long)(y) & 1ll) != 0;
    x = -x;
  }
  float r;
  if (isinf(y)) r = ((x < 1.0f) == (y > 0.0f)) ? 0.0f : inf;
  else if (x == 0.0f) r = (y > 0.0f) ? 0.0f : inf;
  else if (isinf(x)) r = (y > 0.0f) ? inf : 0.0f;
  else r = pow_positive(x, y);
  return neg ? -r : r;
}

// ----------------------------------------------------------- Philox4x32-10
// Salmon et al., SC'11.  key = (pixel, seed_lo), counter = (sample, bounce, seed_hi, 0).
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t const hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    uint32_t const hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}
__device__ __forceinline__ float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }
// stream 0: pixel jitter, light and per-bounce numbers; stream 1: lens position
__device__ __forceinline__ float4 draw4(const PTConst& c, uint32_t pixel, uint32_t sample,
                                        uint32_t bounce, uint32_t stream = 0u) {
  uint4 const r = philox(make_uint4(sample, bounce, c.seed_hi, stream), pixel, c.seed_lo);
  return make_float4(u01(r.x), u01(r.y), u01(r.z), u01(r.w));
}

// ----------------------------------------------------------------- camera
// xyz/core/thinlenscamera.cpp:111-124 (pinhole) and integrator.cpp:124-125
__device__ __forceinline__ float3 primary_dir(const DevCamera& cam, uint32_t px, uint32_t py,
                                              float jx, float jy) {
  float const s = (float(int(px)) + jx) / float(int(cam.width));
  float const t = (float(int(py)) + jy) / float(int(cam.height));
  float const lx = float((double(s) - 0.5) * double(cam.film) * double(cam.aspect));
  float const ly = float((double(t) - 0.5) * double(cam.film));
  float const lz = -cam.v;
  float3 const r = f3(cam.right[0], cam.right[1], cam.right[2]);
  float3 const u = f3(cam.up[0], cam.up[1], cam.up[2]);
  float3 const b = f3(cam.back[0], cam.back[1], cam.back[2]);
  return normalize3(add3(add3(scl3l(lx, r), scl3l(ly, u)), scl3l(lz, b)));
}

// Primary ray of either camera model for pixel (px, py) of sample `sample`:
// jitter from draw4(.., 0).xy; the thin lens samples the disk
// (thinlenscamera.cpp:4-14) with stream 1, starts on the lens (:159-181) and
// aims through the focus plane (:186-205).
__device__ __forceinline__ void primary_ray(const DevCamera& cam, const PTConst& c, uint32_t pixel,
                                            uint32_t sample, float jx, float jy, float3* org,
                                            float3* dir) {
  uint32_t const px = pixel % cam.width, py = pixel / cam.width;
  float3 const eye = f3(cam.eye[0], cam.eye[1], cam.eye[2]);
  if (!cam.thin) {
    *org = eye;
    *dir = primary_dir(cam, px, py, jx, jy);
    return;
  }
  float4 const ul = draw4(c, pixel, sample, 0u, 1u);
  float const s = (float(int(px)) + jx) / float(int(cam.width));
  float const t = (float(int(py)) + jy) / float(int(cam.height));
  float const radius = cam.lens_r * sqrtf(ul.x);
  double cs, sn;
  cos_sin_d((2 * 3.14159265358979323846) * double(ul.y), &cs, &sn);
  float const lx = float(double(radius) * cs);
  float const ly = float(double(radius) * sn);
  float3 const r = f3(cam.right[0], cam.right[1], cam.right[2]);
  float3 const u = f3(cam.up[0], cam.up[1], cam.up[2]);
  float3 const b = f3(cam.back[0], cam.back[1], cam.back[2]);
  *org = add3(add3(eye, scl3l(lx, r)), scl3l(ly, u));
  float const dx = float((double(s) - 0.5) * double(cam.focal_plane) * double(cam.aspect) - double(lx));
  float const dy = float((double(t) - 0.5) * double(cam.focal_plane) - double(ly));
  float const dz = -cam.focus;
  *dir = normalize3(add3(add3(scl3l(dx, r), scl3l(dy, u)), scl3l(dz, b)));
}

// ------------------------------------------------------------ light sample
struct LightVertex {
  int32_t prim;  // -1: no light in the scene
  float3 pos, ns, ng;
  float3 power;  // Phi of the sampled light
  uint32_t bsdf;
};

// xyz/core/scene.cpp:42-78 and xyz/geom/triangle.cpp:232-263
__device__ __forceinline__ LightVertex sample_light(const DevScene& s, float u0, float u1) {
  LightVertex lv;
  if (s.n_lights == 0) {
    lv.prim = -1;
    lv.pos = lv.ns = lv.ng = lv.power = f3(0.f, 0.f, 0.f);
    lv.bsdf = 0;
    return lv;
  }
  float const total = gp(s.light_cdf)[s.n_lights];
  float const u = u0 * total;
  // std::upper_bound over cdf[0..n]: first index with cdf[i] > u
  uint32_t lo = 0, hi = s.n_lights + 1;
  while (lo < hi) {
    uint32_t const mid = (lo + hi) >> 1;
    if (u < gp(s.light_cdf)[mid]) hi = mid; else lo = mid + 1;
  }
  uint32_t n = lo - 1;
  if (n >= s.n_lights) n = s.n_lights - 1;
  float const v = (u - gp(s.light_cdf)[n]) / (gp(s.light_cdf)[n + 1] - gp(s.light_cdf)[n]);
  uint32_t const p = gp(s.light_prim)[n];
  float3 const v0 = ld3(gp(s.prim_v0)[p]), e1 = ld3(gp(s.prim_e1)[p]), e2 = ld3(gp(s.prim_e2)[p]);
  float const ss = sqrtf(v);
  float const tt = u1;
  float const uu = ss * (1.0f - tt);
  float const vv = ss * tt;
  lv.prim = int32_t(p);
  lv.pos = add3(add3(v0, scl3l(uu, e1)), scl3l(vv, e2));
  float4 const ngb = gp(s.prim_ng)[p];
  lv.ng = ld3(ngb);
  if (s.prim_n0) {
    float3 const n0 = ld3(gp(s.prim_n0)[p]), s1 = ld3(gp(s.prim_s1)[p]), s2 = ld3(gp(s.prim_s2)[p]);
    lv.ns = normalize3(add3(add3(n0, scl3l(uu, s1)), scl3l(vv, s2)));
  } else {
    float3 const z = f3(0.f, 0.f, 0.f);
    lv.ns = normalize3(add3(add3(lv.ng, scl3l(uu, z)), scl3l(vv, z)));
  }
  lv.bsdf = __float_as_uint(ngb.w);
  auto const& b = gp(s.bsdfs)[lv.bsdf];
  lv.power = f3(b.phi[0], b.phi[1], b.phi[2]);
  return lv;
}

// ------------------------------------------------------------ shading basis
struct Basis {
  float3 ns, ng, t, b;
};

// xyz/geom/triangle.cpp:194-219
__device__ __forceinline__ Basis shading_basis(const DevScene& s, int32_t prim, float f, float g,
                                               bool back, uint32_t* bsdf) {
  float4 const ngb = gp(s.prim_ng)[prim];
  *bsdf = __float_as_uint(ngb.w);
  float3 n0, s1, s2;
  if (s.prim_n0) {
    n0 = ld3(gp(s.prim_n0)[prim]);
    s1 = ld3(gp(s.prim_s1)[prim]);
    s2 = ld3(gp(s.prim_s2)[prim]);
  } else {
    n0 = ld3(ngb);
    s1 = s2 = f3(0.f, 0.f, 0.f);
  }
  Basis B;
  B.ns = normalize3(add3(add3(n0, scl3l(f, s1)), scl3l(g, s2)));
  B.ng = ld3(ngb);
  if (back) {
    B.ng = neg3(B.ng);
    B.ns = neg3(B.ns);
  }
  float3 t = f3(-B.ns.z, 0.0f, B.ns.x);
  float const l2 = dot3(t, t);
  if (l2 > 0.0f) {
    t = scl3(t, recip_exact(sqrtf(l2)));
    B.b = cross3(t, B.ns);
  } else {
    B.b = f3(0.0f, 0.0f, (B.ns.y > 0.0f) ? 1.0f : -1.0f);
    t = cross3(B.ns, B.b);
  }
  B.t = t;
  return B;
}

// Lambert albedo, optionally from a texture at the hit's uv (nearest texel).
__device__ __forceinline__ float3 albedo_at(const DevScene& s, const DevBsdf& b, int32_t prim,
                                            float f, float g, int32_t* texel) {
  *texel = -1;
  if (b.texture < 0) return f3(b.rgb[0], b.rgb[1], b.rgb[2]);
  DevTex const tx = gp(s.tex)[b.texture];
  float2 const a = gp(s.prim_uv)[3 * prim], c1 = gp(s.prim_uv)[3 * prim + 1], c2 = gp(s.prim_uv)[3 * prim + 2];
  float const u = (a.x + f * (c1.x - a.x)) + g * (c2.x - a.x);
  float const v = (a.y + f * (c1.y - a.y)) + g * (c2.y - a.y);
  int ix = int(u * float(tx.w));
  int iy = int(v * float(tx.h));
  ix = ix < 0 ? 0 : (ix >= int(tx.w) ? int(tx.w) - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= int(tx.h) ? int(tx.h) - 1 : iy);
  int const idx = iy * int(tx.w) + ix;
  *texel = idx;
  auto const p = gp(s.texels) + tx.offset + 3 * size_t(idx);
  return f3(p[0], p[1], p[2]);
}

// xyz/stdafx.hpp:44-47, libhi/hi/math/math.hpp:41-44
__device__ __forceinline__ float fake_fresnel(float hk) {
  float const fresnel0 = 0.9f;
  float const x = 1.0f - hk;
  float const x2 = x * x;
  return fresnel0 + (1.0f - fresnel0) * (x * (x2 * x2));
}

// --------------------------------------------------------------- path state
struct PathState {
  float3 pos, dir, power, color;
  float f_next;
  uint32_t k;     // the reference's vertex counter (starts at 1)
  bool specular;
};

// An NEE visibility query and what is needed to finish its contribution.
struct ShadowReq {
  float3 o, w;       // from the light sample toward the shading point
  int32_t expected;  // the shading point's triangle
  uint32_t type;     // LAMBERT or METAL
  uint32_t light_bsdf;
  float nk2, nk3;
  float3 pe;         // light power * path power
  float fsn, fres, mx;  // METAL only
  float gd;          // LAMBERT: the importance histogram's density of w (1 when unguided)
  uint32_t vj;       // index of the shading vertex on its path (0 = first hit)
  float t_target;    // distance of the shading point's triangle along w (+inf if NaN)
};

// Contribution of a visible NEE sample (xyz/bsdf/lambert.cpp:157-172,
// xyz/bsdf/metal.cpp:100-116), given the shadow ray's hit distance t.  *wfac:
// the contribution per unit of r.pe (for the adjoint).
__device__ __forceinline__ float3 nee_contribution(const ShadowReq& r, float t, float A,
                                                   double* wfac = nullptr) {
  if (r.type == XYZ_LAMBERT) {
    float const G = r.nk2 * r.nk3 / (t * t);
    // lambert.cpp:161 and, guided, :228-229 (M_1_PI * G * density)
    float const p_imp = float((kInvPi * double(G)) * double(r.gd));
    float const p_exp = rcpf(A);
    float const dens = p_imp / p_exp;
    double const wgt = (double(G) * (kInvPi * kInvPi)) * double(rcpf(1.0f + dens * dens));
    if (wfac) *wfac = wgt;
    return f3(float(double(r.pe.x) * wgt), float(double(r.pe.y) * wgt),
              float(double(r.pe.z) * wgt));
  }
  float const wg = r.nk3 / (t * t);
  float const G = r.nk2 * wg;
  float const dens = r.fsn * wg * A;
  double const s1 = double(G) * kInvPi;
  double const mx = double(r.mx);
  double const rw = double(rcpf(1.0f + dens * dens));
  double const fr = double(r.fres);
  double const fs = double(r.fsn);
  if (wfac) *wfac = (((s1 * fs) / mx) * fr) * rw;
  return f3(float(((((double(r.pe.x) * s1) * fs) / mx) * fr) * rw),
            float(((((double(r.pe.y) * s1) * fs) / mx) * fr) * rw),
            float(((((double(r.pe.z) * s1) * fs) / mx) * fr) * rw));
}

// ------------------------------------------ importance-driven path tracing
// Importon map on the device (Scene::CreateLightImportanceMap,
// xyz/core/scene_photonmap.cpp:272-334): positions and incoming directions
// sorted by the cell of a uniform grid whose cells are at least the search
// radius wide; cell_start[c] .. cell_start[c + 1] are cell c's importons, with
// cells numbered x fastest.
struct DevImportons {
  const float4* pos;
  const float4* dir;
  const uint32_t* cell_start;
  float lo[3];
  float inv_cell;
  float reach;  // half-width of a query's cell range (radius plus a margin)
  float r2;     // square of the search radius, float(0.1f * 0.1f)
  uint32_t dim[3];
  uint32_t n;
};

constexpr uint32_t kImpU = 2, kImpV = 8, kImpN = 16;  // CONFIG_U_SIZE, CONFIG_V_SIZE

// Octant of the azimuth atan2(y, x) in [0, 2 pi), an edge belonging to the
// octant above it: ImportanceQuery's v = floor(phi * 8 / pi) with
// phi = atan2(y, x) / 2 (+ pi when negative) (photonmap.cpp:351-360) in exact
// arithmetic, found with comparisons instead of atan2.
__host__ __device__ __forceinline__ uint32_t azimuth_octant(float x, float y) {
  if (x == 0.0f && y == 0.0f) return 0u;
  if (y > 0.0f || (y == 0.0f && x > 0.0f)) {
    if (x > y) return 0u;
    if (x > 0.0f) return 1u;
    return (y > -x) ? 2u : 3u;
  }
  if (x < 0.0f) return (y > x) ? 4u : 5u;
  return (y < -x) ? 6u : 7u;
}

// Histogram bin of a direction with cosine ct to the shading normal and
// components x = -dot(dir, t), y = -dot(dir, b) (photonmap.cpp:345-360).
__device__ __forceinline__ uint32_t importance_bin(float ct, float x, float y) {
  int u = int((ct * ct) * float(kImpU));
  if (u >= int(kImpU)) u = int(kImpU) - 1;
  return uint32_t(u) * kImpV + azimuth_octant(x, y);
}

// Grid cell of coordinate v (host and device: the map is sorted on the host).
__host__ __device__ __forceinline__ int imp_cell(float v, float lo, float inv, uint32_t dim) {
  float const f = fminf(fmaxf(floorf((v - lo) * inv), 0.0f), float(dim - 1u));
  return int(f);
}

// ImportanceQuery over the importons within the search radius of p
// (hi::array_kdtree::kNN_query, libhi/hi/sds/array_kdtree.inl:211-243, and
// ImportanceQuery::push_back, photonmap.cpp:331-363): per-bin counts c.
// The bins' cdf is cdf[i] = 16 i + c[0] + ... + c[i - 1] (begin_query's 16 per
// bin, end_query's prefix sums).
__device__ __forceinline__ void importance_counts(const DevImportons& m, float3 p, const Basis& B,
                                                  uint32_t c[kImpN]) {
#pragma unroll
  for (uint32_t q = 0; q < kImpN; ++q) c[q] = 0u;
  if (m.n == 0u) return;
  // the margin covers the rounding of p +- reach for coordinates of any size
  float const rx = m.reach + fabsf(p.x) * (1.0f / 4194304.0f);
  float const ry = m.reach + fabsf(p.y) * (1.0f / 4194304.0f);
  float const rz = m.reach + fabsf(p.z) * (1.0f / 4194304.0f);
  int const x0 = imp_cell(p.x - rx, m.lo[0], m.inv_cell, m.dim[0]);
  int const x1 = imp_cell(p.x + rx, m.lo[0], m.inv_cell, m.dim[0]);
  int const y0 = imp_cell(p.y - ry, m.lo[1], m.inv_cell, m.dim[1]);
  int const y1 = imp_cell(p.y + ry, m.lo[1], m.inv_cell, m.dim[1]);
  int const z0 = imp_cell(p.z - rz, m.lo[2], m.inv_cell, m.dim[2]);
  int const z1 = imp_cell(p.z + rz, m.lo[2], m.inv_cell, m.dim[2]);
  for (int z = z0; z <= z1; ++z) {
    for (int y = y0; y <= y1; ++y) {
      uint32_t const row = (uint32_t(z) * m.dim[1] + uint32_t(y)) * m.dim[0];
      uint32_t const b0 = gp(m.cell_start)[row + uint32_t(x0)];
      uint32_t const b1 = gp(m.cell_start)[row + uint32_t(x1) + 1u];
      for (uint32_t i = b0; i < b1; ++i) {
        float4 const P = gp(m.pos)[i];
        float3 const d = f3(P.x - p.x, P.y - p.y, P.z - p.z);
        if (!(dot3(d, d) < m.r2)) continue;
        float3 const w = ld3(gp(m.dir)[i]);
        float const ct = -dot3(w, B.ns);
        if (ct <= 0.0f) continue;
        float3 const vp = add3(d, scl3(w, dot3(d, B.ns)));
        if (dot3(vp, vp) >= m.r2) continue;
        uint32_t const bin = importance_bin(ct, -dot3(w, B.t), -dot3(w, B.b));
#pragma unroll
        for (uint32_t q = 0; q < kImpN; ++q) c[q] += (bin == q) ? 1u : 0u;
      }
    }
  }
}

// importance_counts for the lanes of a wave that want it, all lanes working on
// one group of queries at a time: the first pending query and every pending
// query whose grid cells are the same.  The group's candidate importons are
// loaded 64 at a time, once, and tested against each member query (its point
// and frame broadcast); a bin's count grows by the population of a ballot.  A
// lane near a caustic may own thousands of candidates while its neighbours
// own none; shared, such a query takes candidates / 64 steps, and neighbouring
// queries share their loads.  The counts are those of importance_counts.
// Every lane of the wave must call it (with want false for no query).
// Bin of candidate importon (P, W) for the query at qp with frame (ns, tt, bb),
// or kImpN when it is not counted (radius, front and tangent-disc tests).
__device__ __forceinline__ uint32_t importance_candidate_bin(float4 P, float4 W, float3 qp,
                                                            float3 ns, float3 tt, float3 bb,
                                                            float r2) {
  float3 const d = f3(P.x - qp.x, P.y - qp.y, P.z - qp.z);
  if (!(dot3(d, d) < r2)) return kImpN;
  float3 const w = ld3(W);
  float const ct = -dot3(w, ns);
  if (ct <= 0.0f) return kImpN;
  float3 const vp = add3(d, scl3(w, dot3(d, ns)));
  if (dot3(vp, vp) >= r2) return kImpN;
  return importance_bin(ct, -dot3(w, tt), -dot3(w, bb));
}

constexpr uint32_t kImpRows = 16;  // grid rows of a query read at once (4 x 4)

__device__ __forceinline__ float3 shfl3(float3 v, uint32_t lane) {
  return f3(__shfl(v.x, int(lane)), __shfl(v.y, int(lane)), __shfl(v.z, int(lane)));
}

__device__ __forceinline__ void imp_range(const DevImportons& m, float3 p, int r[6]) {
  // the margin covers the rounding of p +- reach for coordinates of any size
  float const rx = m.reach + fabsf(p.x) * (1.0f / 4194304.0f);
  float const ry = m.reach + fabsf(p.y) * (1.0f / 4194304.0f);
  float const rz = m.reach + fabsf(p.z) * (1.0f / 4194304.0f);
  r[0] = imp_cell(p.x - rx, m.lo[0], m.inv_cell, m.dim[0]);
  r[1] = imp_cell(p.x + rx, m.lo[0], m.inv_cell, m.dim[0]);
  r[2] = imp_cell(p.y - ry, m.lo[1], m.inv_cell, m.dim[1]);
  r[3] = imp_cell(p.y + ry, m.lo[1], m.inv_cell, m.dim[1]);
  r[4] = imp_cell(p.z - rz, m.lo[2], m.inv_cell, m.dim[2]);
  r[5] = imp_cell(p.z + rz, m.lo[2], m.inv_cell, m.dim[2]);
}

// Adds, for every query lane of `members`, the bins of the candidates this lane
// holds (bin kImpN: none) to that lane's counts.
__device__ __forceinline__ void add_member_bins(unsigned long long members, float4 P, float4 W,
                                                bool have, float3 p, const Basis& B, float r2,
                                                uint32_t lane, uint32_t c[kImpN]) {
  for (unsigned long long mm = members; mm != 0ull; mm &= mm - 1ull) {
    uint32_t const q = uint32_t(__builtin_ctzll(mm));
    float3 const qp = shfl3(p, q);
    float3 const ns = shfl3(B.ns, q), tt = shfl3(B.t, q), bb = shfl3(B.b, q);
    uint32_t const bin = have ? importance_candidate_bin(P, W, qp, ns, tt, bb, r2) : kImpN;
#pragma unroll
    for (uint32_t k = 0; k < kImpN; ++k) {
      uint32_t const n = uint32_t(__popcll(__ballot(bin == k)));
      c[k] += (lane == q) ? n : 0u;
    }
  }
}

__device__ __forceinline__ void wave_importance_counts(const DevImportons& m, bool want, float3 p,
                                                       const Basis& B, uint32_t c[kImpN]) {
#pragma unroll
  for (uint32_t q = 0; q < kImpN; ++q) c[q] = 0u;
  unsigned long long pending = __ballot(want);
  if (m.n == 0u) return;
  uint32_t const lane = __lane_id();
  int rng[6];
  imp_range(m, p, rng);
  while (pending != 0ull) {
    // the first pending query and every pending query with the same cells
    uint32_t const owner = uint32_t(__builtin_ctzll(pending));
    int o[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) o[a] = __builtin_amdgcn_readlane(rng[a], int(owner));
    bool same = ((pending >> lane) & 1ull) != 0ull;
#pragma unroll
    for (int a = 0; a < 6; ++a) same = same && (rng[a] == o[a]);
    unsigned long long const members = __ballot(same) | (1ull << owner);
    pending &= ~members;
    int const x0 = o[0], x1 = o[1], y0 = o[2], y1 = o[3], z0 = o[4], z1 = o[5];
    uint32_t const ny = uint32_t(y1 - y0 + 1), nrows = ny * uint32_t(z1 - z0 + 1);
    if (nrows <= kImpRows) {
      // lane r < nrows reads row r's importon range; the rows' candidates are
      // then taken as one list, 64 at a time, position and direction together,
      // and tested against every member query
      uint32_t rb0 = 0u, rlen = 0u;
      if (lane < nrows) {
        uint32_t const row = ((uint32_t(z0) + lane / ny) * m.dim[1] + uint32_t(y0) + lane % ny) *
                             m.dim[0];
        rb0 = gp(m.cell_start)[row + uint32_t(x0)];
        rlen = gp(m.cell_start)[row + uint32_t(x1) + 1u] - rb0;
      }
      uint32_t b0s[kImpRows], pre[kImpRows + 1];
      pre[0] = 0u;
#pragma unroll
      for (uint32_t r = 0; r < kImpRows; ++r) {
        b0s[r] = uint32_t(__builtin_amdgcn_readlane(int(rb0), int(r)));
        pre[r + 1] = pre[r] + uint32_t(__builtin_amdgcn_readlane(int(rlen), int(r)));
      }
      uint32_t const total = pre[kImpRows];
      for (uint32_t base = 0; base < total; base += 64u) {
        uint32_t const k = base + lane;
        bool const have = k < total;
        float4 P = make_float4(0.f, 0.f, 0.f, 0.f), W = P;
        if (have) {
          uint32_t i = 0u;
#pragma unroll
          for (uint32_t r = 0; r < kImpRows; ++r)
            i = (k >= pre[r] && k < pre[r + 1]) ? b0s[r] + (k - pre[r]) : i;
          P = gp(m.pos)[i];
          W = gp(m.dir)[i];
        }
        add_member_bins(members, P, W, have, p, B, m.r2, lane, c);
      }
    } else {  // a range of many cells (coordinates far beyond the grid's scale)
      for (int z = z0; z <= z1; ++z) {
        for (int y = y0; y <= y1; ++y) {
          uint32_t const row = (uint32_t(z) * m.dim[1] + uint32_t(y)) * m.dim[0];
          uint32_t const b0 = gp(m.cell_start)[row + uint32_t(x0)];
          uint32_t const b1 = gp(m.cell_start)[row + uint32_t(x1) + 1u];
          for (uint32_t base = b0; base < b1; base += 64u) {
            uint32_t const i = base + lane;
            bool const have = i < b1;
            float4 P = make_float4(0.f, 0.f, 0.f, 0.f), W = P;
            if (have) {
              P = gp(m.pos)[i];
              W = gp(m.dir)[i];
            }
            add_member_bins(members, P, W, have, p, B, m.r2, lane, c);
          }
        }
      }
    }
  }
}

// ImportanceQuery::GetDensity (photonmap.cpp:430-457): 16 (cdf[n+1] - cdf[n]) / cdf[16].
__device__ __forceinline__ float importance_density(const uint32_t c[kImpN], float3 dir,
                                                    const Basis& B) {
  float const ct = -dot3(dir, B.ns);
  if (ct <= 0.0f) return 0.0f;
  uint32_t const n = importance_bin(ct, -dot3(dir, B.t), -dot3(dir, B.b));
  uint32_t cn = 0u, tot = 0u;
#pragma unroll
  for (uint32_t q = 0; q < kImpN; ++q) {
    cn = (q == n) ? c[q] : cn;
    tot += c[q];
  }
  return float(kImpN) * float(kImpN + cn) / float(kImpN * kImpN + tot);
}

// Lambert::NextScatteringDirection with importons (lambert.cpp:243-259) and
// ImportanceQuery::SetDiffusedVector (photonmap.cpp:459-494): a bin from the
// cdf with u.w, a direction uniform in the bin's cos^2 theta (u.y) and azimuth
// (u.z) ranges; density cos / pi times the bin's density; the path power
// divided by the bin's density.
__device__ __forceinline__ bool guided_scatter(const uint32_t c[kImpN], float4 u, const Basis& B,
                                               PathState& ps) {
  uint32_t tot = 0u;
#pragma unroll
  for (uint32_t q = 0; q < kImpN; ++q) tot += c[q];
  float const value = float(kImpN * kImpN + tot) * u.w;
  // std::upper_bound over cdf[0..16]: n = #{i in 1..16 : cdf[i] <= value}
  uint32_t n = 0u, pre = 0u, cn = 0u;
#pragma unroll
  for (uint32_t i = 1; i <= kImpN; ++i) {
    pre += c[i - 1];
    n += (float(kImpN * i + pre) <= value) ? 1u : 0u;
  }
  if (n >= kImpN) n = kImpN - 1u;
#pragma unroll
  for (uint32_t q = 0; q < kImpN; ++q) cn = (q == n) ? c[q] : cn;
  uint32_t const bu = n / kImpV, bv = n % kImpV;
  float const mu_u = (float(bu) + u.y) / float(kImpU);
  float const mu_v = (float(bv) + u.z) / float(kImpV);
  float const cos_t = sqrtf(mu_u);
  float const sin_t = sqrtf(1.0f - mu_u);
  float const phi = 6.28318530717958647692f * mu_v;
  float cp, sp;
  cos_sin(phi, &cp, &sp);
  ps.dir = add3(add3(scl3(B.t, sin_t * cp), scl3(B.b, sin_t * sp)), scl3(B.ns, cos_t));
  float const density = float(kImpN) * float(kImpN + cn) / float(kImpN * kImpN + tot);
  ps.f_next = float(double(density * cos_t) * kInvPi);
  ps.power = scl3(ps.power, rcpf(density));
  ps.specular = false;
  return dot3(ps.dir, B.ng) > 0.0f;
}

// Gradient hooks of the adjoint's replay; the primal uses NoGrad.
//   emission(bsdf, power, wfac, b, k): the path ends on light `bsdf` and adds
//       (b.phi * power) * wfac; k is the vertex counter;
//   vertex(vj, power, alb, texture, texel): shading vertex vj multiplies the
//       throughput `power` by its albedo (1 for specular BSDFs; texture >= 0:
//       a textured Lambert, albedo from `texel`);
//   scale(s): the throughput is scaled by s after that (roulette, Metal).
struct NoGrad {
  __device__ __forceinline__ void emission(uint32_t, float3, double, const DevBsdf&, uint32_t) {}
  __device__ __forceinline__ void vertex(uint32_t, float3, float3, int32_t, int32_t) {}
  __device__ __forceinline__ void scale(float) {}
};

// One vertex of Scene::EvalutePathTracing (xyz/core/scene_pathtracing.cpp:52-142)
// after the extension ray found `h`.  Updates the path; emits at most one
// shadow request.  Returns false when the path ends.
// kGuide: importance-driven path tracing (Scene::EvaluteImportanceDrivenPathTracing,
// xyz/core/scene_importancedrivenpathtracing.cpp:11-179) with the importon map
// imp: a Lambert vertex reached through a specular bounce, other than the first
// hit, takes its NEE weight and its next direction from the importance histogram
// around it (:122-152).
// The vertex is shaded in two steps so that a kernel can run the importance
// query of its guided lanes between them: shade_begin (hit, basis, emission,
// depth cap) and shade_end (NEE, Russian roulette, scattering).
struct VertexCtx {
  Basis B;
  uint32_t bsdf_id;
  float4 u;      // the vertex's numbers (stream 0)
  bool guided;   // importance-driven vertex (kGuide only)
};

// Returns false when the path ends at this hit.
template <class Grad, bool kGuide = false>
__device__ __forceinline__ bool shade_begin(const DevScene& s, const PTConst& c, PathState& ps,
                                            const HitRec& h, uint32_t pixel, uint32_t sample,
                                            VertexCtx& x, Grad& grad) {
  x.guided = false;
  if (h.prim < 0) return false;  // through the air
  x.B = shading_basis(s, h.prim, h.f, h.g, h.back, &x.bsdf_id);
  Basis const& B = x.B;
  uint32_t const bsdf_id = x.bsdf_id;
  if (-dot3(ps.dir, B.ns) <= 0.0f) return false;
  ps.pos = add3(ps.pos, scl3(ps.dir, h.t));
  auto const& b = gp(s.bsdfs)[bsdf_id];
  float const A = s.light_area;
  if (b.type == XYZ_LIGHT && !h.back) {
    float3 const pe = mul3(f3(b.phi[0], b.phi[1], b.phi[2]), ps.power);
    float3 add;
    double wfac;
    if (ps.specular) {
      double const w = kInvPi / double(A);
      add = f3(float(double(pe.x) * w), float(double(pe.y) * w), float(double(pe.z) * w));
      wfac = w;
    } else {
      float const cos_t = fabsf(dot3(ps.dir, B.ns));
      float const dist2 = h.t * h.t;
      float const p_imp = ps.f_next * cos_t / dist2;
      float const p_exp = rcpf(A);
      float const dens = p_exp / p_imp;
      double const w1 = kInvPi / double(A);
      double const w2 = double(rcpf(1.0f + dens * dens));
      add = f3(float(double(pe.x) * w1 * w2), float(double(pe.y) * w1 * w2),
               float(double(pe.z) * w1 * w2));
      wfac = w1 * w2;
    }
    grad.emission(bsdf_id, ps.power, wfac, b, ps.k);
    ps.color = add3(ps.color, add);
    return false;
  }
  if (++ps.k >= c.kmax) return false;
  x.u = draw4(c, pixel, sample, ps.k - 1);
  if constexpr (kGuide) x.guided = (b.type == XYZ_LAMBERT) && (ps.k > 2u) && ps.specular;
  return true;
}

// hc: the importance histogram's counts when x.guided.  Returns false when the
// path ends.
template <class Grad, bool kGuide = false>
__device__ __forceinline__ bool shade_end(const DevScene& s, const PTConst& c, PathState& ps,
                                          const HitRec& h, const VertexCtx& x,
                                          const LightVertex& lv, ShadowReq* sr, bool* has_sr,
                                          Grad& grad, const uint32_t* hc) {
  *has_sr = false;
  Basis const& B = x.B;
  auto const& b = gp(s.bsdfs)[x.bsdf_id];
  uint32_t const j = ps.k - 1;
  float4 const u = x.u;
  bool const guided = kGuide && x.guided;

  // NEE: Bsdf::CalculateWeightedDirectLighting
  uint32_t const vj = ps.k - 2u;  // this vertex on the path (first hit: 0)
  if (b.type == XYZ_LAMBERT || b.type == XYZ_METAL) {
    int32_t texel;
    float3 const alb = albedo_at(s, b, h.prim, h.f, h.g, &texel);
    grad.vertex(vj, ps.power, alb, b.type == XYZ_LAMBERT ? b.texture : -1, texel);
    ps.power = mul3(ps.power, alb);
    if (lv.prim >= 0) {
      float3 const w = normalize3(sub3(ps.pos, lv.pos));
      float const nk3 = dot3(w, lv.ng);
      bool ok = !((nk3 <= 0.0f) || (dot3(w, lv.ns) <= 0.0f));
      float nk2 = 0.0f;
      if (ok) {
        nk2 = -dot3(w, B.ns);
        ok = !((nk2 <= 0.0f) || (-dot3(w, B.ng) <= 0.0f));
      }
      if (ok) {
        // The shadow query (a closest hit from lv.pos along w) can find the
        // shading point's triangle only at that triangle's own tri_test
        // distance T: where tri_test rejects it over [kTMin, +inf] no query
        // finds it and the sample adds nothing, so no query is made.
        float tf, tg, tdet;
        ok = tri_test(lv.pos, w, gp(s.prim_v0)[h.prim], gp(s.prim_e1)[h.prim], gp(s.prim_e2)[h.prim], kTMin,
                      __builtin_inff(), &sr->t_target, &tf, &tg, &tdet);
        if (!(sr->t_target == sr->t_target)) sr->t_target = __builtin_inff();
      }
      if (ok) {
        sr->o = lv.pos;
        sr->w = w;
        sr->expected = h.prim;
        sr->type = b.type;
        sr->light_bsdf = lv.bsdf;
        sr->nk2 = nk2;
        sr->nk3 = nk3;
        sr->pe = mul3(lv.power, ps.power);
        sr->fsn = sr->fres = sr->mx = 0.0f;
        sr->gd = 1.0f;
        sr->vj = vj;
        if constexpr (kGuide) {
          if (guided) sr->gd = importance_density(hc, w, B);
        }
        if (b.type == XYZ_METAL) {
          float3 const hv = neg3(normalize3(add3(w, ps.dir)));
          float const hk = -dot3(hv, w);
          float const nh = dot3(B.ns, hv);
          float const nk1 = -dot3(B.ns, ps.dir);
          sr->fsn = float(double(b.shininess + 1.0f) * (kInvPi / 8) *
                          double(pow_det(nh, b.shininess)) / double(hk));
          sr->fres = fake_fresnel(hk);
          sr->mx = (nk1 < nk2) ? nk2 : nk1;
        }
        *has_sr = true;
      }
    }
  } else {
    grad.vertex(vj, ps.power, f3(1.0f, 1.0f, 1.0f), -1, -1);  // specular: no albedo factor
  }

  // Russian roulette: xyz/core/scene_pathtracing.cpp:128-137
  if (j > c.rr_start) {
    float const t = ps.power.y;
    if (t < c.rr_threshold) {
      float const q = t / c.rr_threshold;
      if (u.x >= q) return false;
      float const inv = rcpf(q);
      ps.power = scl3(ps.power, inv);
      grad.scale(inv);
    }
  }

  // Bsdf::NextScatteringDirection
  if constexpr (kGuide) {
    if (guided) return guided_scatter(hc, u, B, ps);
  }
  if (b.type == XYZ_LAMBERT) {
    // xyz/bsdf/lambert.cpp:176-187, xyz/stdafx.hpp:50-63
    float const phi = 6.28318530717958647692f * u.z;
    float const cos_t = sqrtf(u.y);
    float const sin_t = sqrtf(1.0f - u.y);
    float cp, sp;
    cos_sin(phi, &cp, &sp);
    ps.dir = add3(add3(scl3(B.t, sin_t * cp), scl3(B.b, sin_t * sp)), scl3(B.ns, cos_t));
    ps.specular = false;
    ps.f_next = float(kInvPi * double(cos_t));
    return dot3(ps.dir, B.ng) > 0.0f;
  }
  if (b.type == XYZ_MIRROR) {
    // xyz/bsdf/mirror.cpp:56-105
    float const cth = -dot3(ps.dir, B.ns);
    ps.dir = add3(ps.dir, scl3l(2.0f * cth, B.ns));
    if (dot3(ps.dir, B.ng) <= 0.0f) return false;
    ps.specular = true;
    ps.f_next = 1.0f;
    return true;
  }
  if (b.type == XYZ_GLASS) {
    // xyz/bsdf/glass.cpp:104-165
    float const r = h.back ? rcpf(b.ior) : b.ior;
    float const ior_sq = r * r;
    float const nk = -dot3(ps.dir, B.ns);
    float const bb = ior_sq + nk * nk - 1.0f;
    if (bb > 0.0f) {
      float const g = sqrtf(bb);
      float const gpc = g + nk;
      float const gmc = g - nk;
      float const alpha = gmc / gpc;
      float const beta = (nk * gpc - 1.0f) / (nk * gmc + 1.0f);
      float const fr = 0.5f * alpha * alpha * (beta * beta + 1.0f);
      if (u.y < fr) {
        ps.dir = add3(ps.dir, scl3l(2.0f * nk, B.ns));
        if (dot3(ps.dir, B.ng) <= 0.0f) return false;
      } else {
        float const sq = sqrtf(ior_sq);
        float3 const v = sub3(ps.dir, scl3l(gmc, B.ns));
        ps.dir = f3(v.x / sq, v.y / sq, v.z / sq);
        if (dot3(ps.dir, B.ng) >= 0.0f) return false;
      }
    } else {
      ps.dir = add3(ps.dir, scl3l(2.0f * nk, B.ns));
      if (dot3(ps.dir, B.ng) <= 0.0f) return false;
    }
    ps.specular = true;
    ps.f_next = 1.0f;
    return true;
  }
  if (b.type == XYZ_METAL) {
    // xyz/bsdf/metal.cpp:122-153, xyz/stdafx.hpp:154-168
    float const phi = 6.28318530717958647692f * u.z;
    float const cos_h = pow_det(u.y, rcpf(b.shininess));
    float const sin_h = sqrtf(1.0f - cos_h * cos_h);
    float cp, sp;
    cos_sin(phi, &cp, &sp);
    float3 const hv = add3(add3(scl3(B.t, sin_h * cp), scl3(B.b, sin_h * sp)), scl3(B.ns, cos_h));
    float const nk1 = -dot3(B.ns, ps.dir);
    float const hk = -dot3(hv, ps.dir);
    ps.dir = add3(ps.dir, scl3l(2.0f * hk, hv));
    float const nk2 = dot3(B.ns, ps.dir);
    if ((nk2 <= 0.0f) || (dot3(B.ng, ps.dir) <= 0.0f)) return false;
    float const fres = fake_fresnel(hk);
    float const mx = (nk1 < nk2) ? nk2 : nk1;
    float const sc = fres / mx * nk2;
    ps.power = scl3(ps.power, sc);
    grad.scale(sc);
    ps.specular = false;
    ps.f_next = float(double(b.shininess + 1.0f) * (kInvPi / 8) * double(pow_det(cos_h, b.shininess)) /
                      double(hk));
    return true;
  }
  return false;  // LIGHT seen from behind: xyz/bsdf/light.cpp:27-31
}

template <class Grad, bool kGuide = false>
__device__ __forceinline__ bool shade_vertex(const DevScene& s, const PTConst& c, PathState& ps,
                                             const HitRec& h, uint32_t pixel, uint32_t sample,
                                             const LightVertex& lv, ShadowReq* sr, bool* has_sr,
                                             Grad& grad, const DevImportons* imp = nullptr) {
  *has_sr = false;
  VertexCtx x;
  if (!shade_begin<Grad, kGuide>(s, c, ps, h, pixel, sample, x, grad)) return false;
  uint32_t hc[kImpN];
  if constexpr (kGuide) {
    if (x.guided) importance_counts(*imp, ps.pos, x.B, hc);
  }
  return shade_end<Grad, kGuide>(s, c, ps, h, x, lv, sr, has_sr, grad, hc);
}

}  // namespace xyzhip

#endif  // XYZ_PT_DEVICE_H_

// Device code of the bidirectional path tracer: one sample of one pixel per
// call, Scene::EvaluateBidirectionalPathTracing
// (xyz/core/scene_bidirectionalpathtracing.cpp:8-522) with the bidirectional
// BSDF methods (LightsToCamera / CameraToLightsScatteringDirection of
// xyz/bsdf/{lambert,mirror,glass,metal,light}.cpp) and the pinhole camera's
// film mapping and constant factor (xyz/core/thinlenscamera.hpp:38-90,
// thinlenscamera.cpp:125-157), in the reference's binary32 operation order.
//
// The reference's own behaviour is kept:
//   * hi::single_stack (libhi/hi/util.hpp:24-36) keeps the value it assigns
//     (its destructor writes the new value back), so each connection's
//     temporary density stays in the vertex for the strategies after it;
//   * Mirror::CameraToLightsScatteringDirection ends the eye path when the
//     mirrored direction lies above the geometric surface (mirror.cpp:235-239);
//   * connections and light tracing weigh every non-specular vertex as
//     Lambertian (the M_1_PI factors);
//   * the eye path starts at the pinhole (the lens-film distance V of the
//     camera, also for a thin-lens camera).
// Vertex fields the reference reads before writing them within a sample (it
// reuses its vertex arrays across samples) start from zero / false here.
// Random numbers: Philox streams 2 (light path) and 3 (eye path); draw (pixel,
// sample, vertex index, stream) gives (direction u1, u2, choice / Russian
// roulette) at a vertex, vertex 0 of stream 2 the light sample and its
// emission direction, vertex 0 of stream 3 the pixel jitter.
#ifndef XYZ_BDPT_DEVICE_H_
#define XYZ_BDPT_DEVICE_H_

#include "pt_device.h"

namespace xyzhip {

constexpr uint32_t kBdptMaxDepth = 16;  // vertex arrays hold kmax + 2 <= 19 entries
constexpr uint32_t kBdptStreamLight = 2, kBdptStreamEye = 3;

struct BdptConst {
  PTConst c;   // kmax, seeds (Russian roulette settings unused)
  float C;     // GetConstFactor() = (f / F)^2 / A
  float inv_area;
};

// A path vertex as the strategies read it once both paths are built (kept in
// the per-lane vertex arrays, 68 bytes).
struct BVert {
  int32_t tri;
  bool specular;
  float3 pos, ng, ns, power;
  float s_prev, s_next, g;
};

// The rest of the reference's PathVertex, needed only while the path is
// extended from this vertex: one instance per path, for its last vertex.
struct BWork {
  bool back;
  float3 win, wout, t, b;
  float bsdf_x_ipdf, cos_in_sh, cos_out_geo;
};

// The fields a sample can read before writing them (the flags of the dummy
// vertices and of a path's last vertex, and densities of vertices whose BSDF
// returned early); every other field is written before it is read.
__device__ __forceinline__ void bvert_clear(BVert& v) {
  v.tri = -1;
  v.specular = false;
  v.s_prev = v.s_next = 0.0f;
}

// xyz/stdafx.hpp:50-63
__device__ __forceinline__ float bd_diffused(float mu_t, float mu_p, float3 t, float3 n, float3 b,
                                             float3* out) {
  float const phi = 6.28318530717958647692f * mu_p;
  float const ct = sqrtf(mu_t);
  float const st = sqrtf(1.0f - mu_t);
  float cp, sp;
  cos_sin(phi, &cp, &sp);
  *out = add3(add3(scl3(t, st * cp), scl3(b, st * sp)), scl3(n, ct));
  return ct;
}

// xyz/stdafx.hpp:154-168
__device__ __forceinline__ float bd_glossy_half(float mu_t, float mu_p, float shin, float3 t,
                                                float3 n, float3 b, float3* h) {
  float const phi = 6.28318530717958647692f * mu_p;
  float const ct = pow_det(mu_t, rcpf(shin));
  float const st = sqrtf(1.0f - ct * ct);
  float cp, sp;
  cos_sin(phi, &cp, &sp);
  *h = add3(add3(scl3(t, st * cp), scl3(b, st * sp)), scl3(n, ct));
  return ct;
}

// tangent and binormal about a normal (xyz/geom/triangle.cpp:181-191, 232-263)
__device__ __forceinline__ void basis_about(float3 n, float3* t, float3* b) {
  float3 tt = f3(-n.z, 0.0f, n.x);
  float const l2 = dot3(tt, tt);
  if (l2 > 0.0f) {
    tt = scl3(tt, recip_exact(sqrtf(l2)));
    *b = cross3(tt, n);
  } else {
    *b = f3(0.0f, 0.0f, (n.y > 0.0f) ? 1.0f : -1.0f);
    tt = cross3(n, *b);
  }
  *t = tt;
}

// Triangle::GeometricBasis (xyz/geom/triangle.cpp:167-192)
__device__ __forceinline__ void geometric_basis(const DevScene& s, int32_t prim, float f, float g,
                                                bool back, BVert& v, BWork& w) {
  uint32_t bsdf;
  Basis const B = shading_basis(s, prim, f, g, back, &bsdf);  // ns, ng as ShadingBasis
  v.ns = B.ns;
  v.ng = B.ng;
  basis_about(B.ng, &w.t, &w.b);
}

__device__ __forceinline__ float min1(float x) { return (x < 1.0f) ? x : 1.0f; }
__device__ __forceinline__ float hmaxf(float x, float y) { return (x < y) ? y : x; }
__device__ __forceinline__ float fourth(float x) {
  float const s = x * x;
  return s * s;
}

__device__ __forceinline__ float glass_ior_sq(const DevBsdf& b, bool back) {
  float const r = back ? rcpf(b.ior) : b.ior;
  return r * r;
}

// Bsdf::LightsToCameraScatteringDirection
__device__ __forceinline__ bool lights_to_camera(const DevScene& s, const DevBsdf& bs, int32_t prim,
                                                 float hf, float hg, BVert& v, BWork& w,
                                                 float4 u) {
  switch (bs.type) {
    case XYZ_LAMBERT: {  // lambert.cpp:314-365
      int32_t texel;
      float3 const alb = albedo_at(s, bs, prim, hf, hg, &texel);
      float const cig = dot3(w.win, v.ng);
      float const alpha = w.cos_in_sh / cig;
      float const pp = min1(alb.y), pn = min1(alb.y * alpha);
      v.specular = false;
      v.s_prev = float(kInvPi * double(pp));
      v.s_next = float(kInvPi * double(pn));
      v.power = mul3(v.power, alb);
      v.power = scl3(v.power, alpha);
      w.bsdf_x_ipdf = rcpf(pn);
      if (u.z >= pn) return false;
      w.cos_out_geo = bd_diffused(u.x, u.y, w.t, v.ng, w.b, &w.wout);
      return dot3(w.wout, v.ns) > 0.0f;
    }
    case XYZ_MIRROR: {  // mirror.cpp:172-203
      v.specular = true;
      w.wout = neg3(w.win);
      w.wout = add3(w.wout, scl3l(2.0f * w.cos_in_sh, v.ns));
      w.cos_out_geo = dot3(w.wout, v.ng);
      if (w.cos_out_geo <= 0.0f) return false;
      v.s_prev = rcpf(w.cos_in_sh);
      v.s_next = rcpf(w.cos_out_geo);
      float const cig = dot3(w.win, v.ng);
      w.bsdf_x_ipdf = w.cos_out_geo / cig;
      return true;
    }
    case XYZ_GLASS: {  // glass.cpp:239-323
      float const ior_sq = glass_ior_sq(bs, w.back);
      float const nk = dot3(w.win, v.ns);
      float const bb = ior_sq + nk * nk - 1.0f;
      w.wout = neg3(w.win);
      v.specular = true;
      if (bb > 0.0f) {
        float const g = sqrtf(bb);
        float const gpc = g + nk, gmc = g - nk;
        float const alpha = gmc / gpc;
        float const beta = (nk * gpc - 1.0f) / (nk * gmc + 1.0f);
        float const fr = 0.5f * alpha * alpha * (beta * beta + 1.0f);
        if (u.z >= fr) {
          v.s_prev = v.s_next = 1.0f - fr;
          float const q = sqrtf(ior_sq);
          float3 const r = sub3(w.wout, scl3l(gmc, v.ns));
          w.wout = f3(r.x / q, r.y / q, r.z / q);
          w.cos_out_geo = -dot3(w.wout, v.ng);
          if (w.cos_out_geo <= 0.0f) return false;
          float const cig = dot3(w.win, v.ng);
          w.bsdf_x_ipdf = w.cos_out_geo / cig;
          return true;
        }
        v.s_prev = v.s_next = fr;
      } else {
        v.s_prev = v.s_next = 1.0f;
      }
      w.wout = add3(w.wout, scl3l(2.0f * nk, v.ns));
      w.cos_out_geo = dot3(w.wout, v.ng);
      if (w.cos_out_geo <= 0.0f) return false;
      float const cig = dot3(w.win, v.ng);
      w.bsdf_x_ipdf = w.cos_out_geo / cig;
      return true;
    }
    case XYZ_METAL: {  // metal.cpp:227-278
      v.specular = false;
      float3 hv;
      float const nh = bd_glossy_half(u.x, u.y, bs.shininess, w.t, v.ns, w.b, &hv);
      float const hk = dot3(hv, w.win);
      w.wout = sub3(scl3l(2.0f * hk, hv), w.win);
      w.cos_out_geo = dot3(v.ng, w.wout);
      if (w.cos_out_geo <= 0.0f) return false;
      float const cos_sh = dot3(v.ns, w.wout);
      if (cos_sh <= 0.0f) return false;
      float const cig = dot3(v.ng, w.win);
      w.bsdf_x_ipdf = fake_fresnel(hk) * w.cos_in_sh / hmaxf(cos_sh, w.cos_in_sh) *
                      w.cos_out_geo / cig;
      v.s_prev = v.s_next = float(double(bs.shininess + 1.0f) * (kInvPi / 8) *
                                  double(pow_det(nh, bs.shininess)) / double(hk));
      v.s_prev *= rcpf(w.cos_in_sh);
      v.s_next *= rcpf(w.cos_out_geo);
      return true;
    }
    default:
      return false;  // light.cpp:44-51
  }
}

// Bsdf::CameraToLightsScatteringDirection
__device__ __forceinline__ bool camera_to_lights(const DevScene& s, const DevBsdf& bs, int32_t prim,
                                                 float hf, float hg, BVert& v, BWork& w,
                                                 float4 u) {
  switch (bs.type) {
    case XYZ_LAMBERT: {  // lambert.cpp:380-424
      v.specular = false;
      w.cos_in_sh = bd_diffused(u.x, u.y, w.t, v.ns, w.b, &w.win);
      if (dot3(w.win, v.ng) <= 0.0f) return false;
      float const cog = dot3(w.wout, v.ng);
      int32_t texel;
      float3 const alb = albedo_at(s, bs, prim, hf, hg, &texel);
      float const alpha = w.cos_in_sh / cog;
      float const pp = min1(alb.y), pn = min1(alb.y * alpha);
      v.s_prev = float(kInvPi * double(pp));
      v.s_next = float(kInvPi * double(pn));
      v.power = mul3(v.power, alb);
      w.bsdf_x_ipdf = rcpf(pp);
      return !(u.z >= pp);
    }
    case XYZ_MIRROR: {  // mirror.cpp:217-248
      v.specular = true;
      w.cos_in_sh = dot3(w.wout, v.ns);
      w.win = neg3(w.wout);
      w.win = add3(w.win, scl3l(2.0f * w.cos_in_sh, v.ns));
      if (dot3(w.win, v.ng) > 0.0f) return false;
      v.s_prev = rcpf(w.cos_in_sh);
      v.s_next = rcpf(w.cos_out_geo);
      w.bsdf_x_ipdf = 1.0f;
      return true;
    }
    case XYZ_GLASS: {  // glass.cpp:339-405
      float const ior_sq = glass_ior_sq(bs, w.back);
      float const nk = dot3(w.wout, v.ns);
      float const bb = ior_sq + nk * nk - 1.0f;
      w.win = neg3(w.wout);
      v.specular = true;
      if (bb > 0.0f) {
        float const g = sqrtf(bb);
        float const gpc = g + nk, gmc = g - nk;
        float const alpha = gmc / gpc;
        float const beta = (nk * gpc - 1.0f) / (nk * gmc + 1.0f);
        float const fr = 0.5f * alpha * alpha * (beta * beta + 1.0f);
        if (u.z >= fr) {
          v.s_prev = v.s_next = 1.0f - fr;
          float const q = sqrtf(ior_sq);
          float3 const r = sub3(w.win, scl3l(gmc, v.ns));
          w.win = f3(r.x / q, r.y / q, r.z / q);
          w.cos_in_sh = -dot3(w.win, v.ng);
          if (w.cos_in_sh <= 0.0f) return false;
          w.bsdf_x_ipdf = 1.0f;
          return true;
        }
        v.s_prev = v.s_next = fr;
      } else {
        v.s_prev = v.s_next = 1.0f;
      }
      w.win = add3(w.win, scl3l(2.0f * nk, v.ns));
      w.cos_in_sh = dot3(w.win, v.ng);
      if (w.cos_in_sh <= 0.0f) return false;
      w.bsdf_x_ipdf = 1.0f;
      return true;
    }
    case XYZ_METAL: {  // metal.cpp:292-340
      v.specular = false;
      float3 hv;
      float const nh = bd_glossy_half(u.x, u.y, bs.shininess, w.t, v.ns, w.b, &hv);
      float const hk = dot3(hv, w.wout);
      w.win = sub3(scl3l(2.0f * hk, hv), w.wout);
      if (dot3(v.ng, w.win) <= 0.0f) return false;
      w.cos_in_sh = dot3(v.ns, w.win);
      if (w.cos_in_sh <= 0.0f) return false;
      float const cos_sh = dot3(v.ns, w.wout);
      w.bsdf_x_ipdf = fake_fresnel(hk) * w.cos_in_sh / hmaxf(cos_sh, w.cos_in_sh);
      v.s_prev = v.s_next = float(double(bs.shininess + 1.0f) * (kInvPi / 8) *
                                  double(pow_det(nh, bs.shininess)) / double(hk));
      v.s_prev *= rcpf(w.cos_in_sh);
      v.s_next *= rcpf(w.cos_out_geo);
      return true;
    }
    default:
      return false;
  }
}

// Pinhole GetFilmPosition (thinlenscamera.cpp:125-157)
__device__ __forceinline__ bool film_position(const DevCamera& cam, float3 d, float* fx, float* fy) {
  float const lx = dot3(d, f3(cam.right[0], cam.right[1], cam.right[2]));
  float const ly = dot3(d, f3(cam.up[0], cam.up[1], cam.up[2]));
  float const lz = dot3(d, f3(cam.back[0], cam.back[1], cam.back[2]));
  float const k = (cam.v / -lz) / cam.film;
  *fx = 0.5f + lx * k / cam.aspect;
  *fy = 0.5f + ly * k;
  return (0.0f <= *fx) && (*fx < 1.0f) && (0.0f <= *fy) && (*fy < 1.0f);
}

// The strategies' weights: one vertex array side of the MIS sum
// (scene_bidirectionalpathtracing.cpp:174-193, 460-507).  `G` is the
// connection's geometric factor (1 for s = 0, where it does not appear).
__device__ __forceinline__ float mis_light_side(const BVert* L, uint32_t s, float G, float wst) {
  float pst = 1.0f;
  for (uint32_t i = s; i > 0; --i) {
    pst *= L[i + 1].s_prev / L[i - 1].s_next;
    if (!L[i].specular && !L[i - 1].specular) {
      float const q = pst * G / L[i].g;
      wst += q * q;
    }
  }
  return wst;
}
__device__ __forceinline__ float mis_eye_side(const BVert* E, uint32_t t, float G, float wst) {
  float pst = 1.0f;
  for (uint32_t i = t; i > 1; --i) {
    pst *= E[i + 1].s_next / E[i - 1].s_prev;
    if (!E[i].specular && !E[i - 1].specular) {
      float const q = pst * G / E[i].g;
      wst += q * q;
    }
  }
  return wst;
}

// The answer of a query that finds nothing.
__device__ __forceinline__ HitRec bdpt_no_hit() {
  HitRec h;
  h.prim = -1;
  h.t = __builtin_inff();
  h.f = h.g = 0.0f;
  h.back = false;
  return h;
}

// The set-up of one BDPT sample of pixel (px, py): Scene::EvaluateBidirectionalPathTracing,
// and the starts of BuildLightsPath and BuildTheEyePath
// (scene_bidirectionalpathtracing.cpp:8-126, 251-347) into L and E (kmax + 2
// entries each) and the working vertices wl, we.  Returns whether the light
// path has a first ray (from L[1].pos along wl.wout); the eye path always has
// one (from E[1].pos along we.win).
__device__ __forceinline__ bool bdpt_begin(const DevScene& s, const DevCamera& cam, const BdptConst& bc,
                                           uint32_t px, uint32_t py, uint32_t smp, BVert* L, BVert* E,
                                           BWork& wl, BWork& we) {
  uint32_t const W = cam.width;
  uint32_t const pixel = py * W + px;
  uint32_t const kmax = bc.c.kmax;
  for (uint32_t i = 0; i < kmax + 2; ++i) {
    bvert_clear(L[i]);
    bvert_clear(E[i]);
  }
  float4 const de = draw4(bc.c, pixel, smp, 0u, kBdptStreamEye);
  float4 const dl = draw4(bc.c, pixel, smp, 0u, kBdptStreamLight);
  // SampleLightPosition and Triangle::Sample's basis.  wl: the working part of
  // the light path's last vertex.
  {
    LightVertex const lv = sample_light(s, dl.x, dl.y);
    L[1].tri = lv.prim;
    if (lv.prim >= 0) {
      L[1].pos = lv.pos;
      L[1].ns = lv.ns;
      L[1].ng = lv.ng;
      basis_about(lv.ng, &wl.t, &wl.b);
      L[1].s_next = bc.inv_area;
      L[1].specular = false;
      L[1].power = lv.power;
    }
  }
  L[2].tri = -1;
  // SampleLensPosition: the pinhole, normal = -back
  float3 const eye = f3(cam.eye[0], cam.eye[1], cam.eye[2]);
  float3 const fwd = neg3(f3(cam.back[0], cam.back[1], cam.back[2]));
  E[1].pos = eye;
  E[1].ng = E[1].ns = fwd;
  E[1].power = f3(1.f, 1.f, 1.f);

  // BuildLightsPath's set-up
  bool const light = L[1].tri >= 0;
  if (light) {
    L[0].s_next = bc.inv_area;
    wl.cos_in_sh = 1.0f;
    wl.cos_out_geo = bd_diffused(dl.z, dl.w, wl.t, L[1].ng, wl.b, &wl.wout);
    L[1].s_prev = 1.0f;
    L[1].s_next = float(kInvPi);
    L[1].g = 1.0f;
    wl.bsdf_x_ipdf = 1.0f;
  }

  // BuildTheEyePath's set-up.  we: the working part of the eye path's last vertex.
  we.win = primary_dir(cam, px, py, de.x, de.y);
  we.cos_in_sh = dot3(we.win, E[1].ns);
  we.cos_out_geo = 1.0f;
  E[1].s_prev = bc.C / fourth(we.cos_in_sh);
  E[1].s_next = 1.0f;
  E[1].g = 1.0f;
  we.bsdf_x_ipdf = 1.0f;
  return light;
}

// One step of BuildLightsPath: the vertex after L[k] from h, the hit of the
// ray from L[k].pos along wl.wout.  Returns whether the path goes on (then k
// has advanced and the next ray leaves L[k].pos along wl.wout).
__device__ __forceinline__ bool bdpt_light_step(const DevScene& s, const BdptConst& bc, uint32_t pixel,
                                                uint32_t smp, BVert* L, BWork& wl, uint32_t& k,
                                                const HitRec& h) {
  uint32_t const kmax = bc.c.kmax;
  BVert& nw = L[k + 1];
  BWork wn;
  nw.tri = h.prim;
  if (h.prim < 0) return false;
  uint32_t bsdf;
  {
    float4 const ngb = gp(s.prim_ng)[h.prim];
    bsdf = __float_as_uint(ngb.w);
  }
  DevBsdf const bs = gp(s.bsdfs)[bsdf];
  if (bs.type == XYZ_LIGHT) {
    nw.tri = -1;
    return false;
  }
  geometric_basis(s, h.prim, h.f, h.g, h.back, nw, wn);
  wn.back = h.back;
  wn.win = neg3(wl.wout);
  wn.cos_in_sh = dot3(wn.win, nw.ns);
  if (wn.cos_in_sh <= 0.0f) {
    nw.tri = -1;
    return false;
  }
  nw.power = scl3(L[k].power, wl.bsdf_x_ipdf);
  nw.pos = add3(scl3(wl.wout, h.t), L[k].pos);
  nw.g = wl.cos_out_geo * wn.cos_in_sh / (h.t * h.t);
  wl = wn;
  if (++k >= kmax) {
    L[k + 1].tri = -1;
    return false;
  }
  if (!lights_to_camera(s, bs, h.prim, h.f, h.g, L[k], wl, draw4(bc.c, pixel, smp, k, kBdptStreamLight))) {
    L[k + 1].tri = -1;
    return false;
  }
  return true;
}

// One step of BuildTheEyePath: the vertex after E[k] from h, the hit of the
// ray from E[k].pos along we.win.  Returns whether the path goes on (then k
// has advanced and the next ray leaves E[k].pos along we.win).
__device__ __forceinline__ bool bdpt_eye_step(const DevScene& s, const BdptConst& bc, uint32_t pixel,
                                              uint32_t smp, BVert* E, BWork& we, uint32_t& k,
                                              const HitRec& h) {
  uint32_t const kmax = bc.c.kmax;
  BVert& nw = E[k + 1];
  BWork wn;
  nw.tri = h.prim;
  if (h.prim < 0) return false;
  uint32_t bsdf;
  Basis const B = shading_basis(s, h.prim, h.f, h.g, h.back, &bsdf);
  nw.ns = B.ns;
  nw.ng = B.ng;
  wn.t = B.t;
  wn.b = B.b;
  wn.back = h.back;
  wn.wout = neg3(we.win);
  if (dot3(wn.wout, nw.ns) <= 0.0f) {
    nw.tri = -1;
    return false;
  }
  nw.power = scl3(E[k].power, we.bsdf_x_ipdf);
  nw.pos = add3(scl3(we.win, h.t), E[k].pos);
  wn.cos_out_geo = dot3(wn.wout, nw.ng);
  nw.g = wn.cos_out_geo * we.cos_in_sh / (h.t * h.t);
  we = wn;
  DevBsdf const bs = gp(s.bsdfs)[bsdf];
  if (++k >= kmax) {
    E[k + 1].tri = -1;
    return false;
  }
  if (bs.type == XYZ_LIGHT) {
    E[k + 1].tri = -1;
    return false;
  }
  if (!camera_to_lights(s, bs, h.prim, h.f, h.g, E[k], we, draw4(bc.c, pixel, smp, k, kBdptStreamEye))) {
    E[k + 1].tri = -1;
    return false;
  }
  return true;
}

// The two subpaths of one BDPT sample of pixel (px, py) with per-thread
// closest-hit queries: bdpt_begin, then the light path's and the eye path's
// steps.
template <bool kCount>
__device__ __forceinline__ void bdpt_build(const DevScene& s, const DevCamera& cam,
                                           const BdptConst& bc, uint32_t px, uint32_t py,
                                           uint32_t smp, uint32_t* stk, uint32_t stride,
                                           BVert* L, BVert* E, uint32_t* queries) {
  uint32_t const pixel = py * cam.width + px;
  BWork wl, we;
  bool const light = bdpt_begin(s, cam, bc, px, py, smp, L, E, wl, we);
  if (light) {
    for (uint32_t k = 1;;) {
      if (kCount) ++*queries;
      HitRec const h = trace_lane(s, L[k].pos, wl.wout, stk, stride);
      if (!bdpt_light_step(s, bc, pixel, smp, L, wl, k, h)) break;
    }
  }
  for (uint32_t k = 1;;) {
    if (kCount) ++*queries;
    HitRec const h = trace_lane(s, E[k].pos, we.win, stk, stride);
    if (!bdpt_eye_step(s, bc, pixel, smp, E, we, k, h)) break;
  }
}

// The two subpaths of the items of one wave, built with the cooperative
// traversal; all 64 lanes call it together (act: the lane has an item).  Each
// round, every lane with a pending ray (its light path's, then, once that has
// ended, its eye path's) traces it -- coop_step's wave-shared triangle and box
// batches, which lanes without a ray help with -- and takes that subpath's
// next step.  The steps, their order within each subpath and the hits are the
// per-thread build's, so L and E are the same.  keys and own: this wave's LDS
// tables of coop_step.
template <bool kCount>
__device__ __forceinline__ void bdpt_build_wave(const DevScene& s, const DevCamera& cam,
                                                const BdptConst& bc, bool act, uint32_t px,
                                                uint32_t py, uint32_t smp, uint32_t* stk,
                                                uint32_t stride, unsigned long long* keys, uint8_t* own,
                                                BVert* L, BVert* E, uint32_t* queries) {
  uint32_t const pixel = py * cam.width + px;
  BWork wl, we;
  bool light = false;
  if (act) light = bdpt_begin(s, cam, bc, px, py, smp, L, E, wl, we);
  uint32_t phase = act ? (light ? 0u : 1u) : 2u;  // 0 light path, 1 eye path, 2 done
  uint32_t kl = 1, ke = 1;
  while (__ballot(phase < 2u) != 0) {
    bool const pend = phase < 2u;
    float3 o = f3(0.0f, 0.0f, 0.0f), d = f3(0.0f, 0.0f, 1.0f);
    if (phase == 0u) {
      o = L[kl].pos;
      d = wl.wout;
    } else if (phase == 1u) {
      o = E[ke].pos;
      d = we.win;
    }
    if (kCount && pend) ++*queries;
    TraceState st;
    trace_init(st, o, d);
    HitRec h = bdpt_no_hit();
    bool q = pend;
    while (__ballot(q) != 0) {
      bool const fin = coop_step<false>(s, st, q, stk, stride, keys, own);
      if (q && fin) {
        h = trace_result(s, st);
        q = false;
      }
    }
    if (phase == 0u) {
      if (!bdpt_light_step(s, bc, pixel, smp, L, wl, kl, h)) phase = 1u;
    } else if (phase == 1u) {
      if (!bdpt_eye_step(s, bc, pixel, smp, E, we, ke, h)) phase = 2u;
    }
  }
}

// The connections of one sample over its built subpaths: EvaluatePathS1 (light
// tracing, t = 1) and EvaluatePathST (scene_bidirectionalpathtracing.cpp:127-250,
// 348-522)
// in the reference's order.  Every visibility query goes through q.find(o, d),
// in that order; which queries are made depends on the vertices alone (the
// strategies' densities the connections overwrite are not among their
// conditions), so a pass that only records them makes the same ones.  The
// light-tracing contributions go to `splat` (atomically), the own colour is
// returned.
template <class Q>
__device__ __forceinline__ float3 bdpt_connect(const DevScene& s, const DevCamera& cam,
                                               const BdptConst& bc, BVert* L, BVert* E,
                                               float* splat, Q& q) {
  uint32_t const W = cam.width, H = cam.height;
  uint32_t const kmax = bc.c.kmax;
  // EvaluatePathS1 (t = 1)
  for (uint32_t k = 2; k < kmax; ++k) {
    BVert const& lv = L[k];
    if (lv.tri < 0) break;
    uint32_t const ty = gp(s.bsdfs)[__float_as_uint(gp(s.prim_ng)[lv.tri].w)].type;
    if (ty == XYZ_MIRROR || ty == XYZ_GLASS) continue;
    float3 const d = normalize3(sub3(lv.pos, E[1].pos));
    float const cog = -dot3(d, lv.ng);
    if (cog <= 0.0f) continue;
    if (-dot3(d, lv.ns) <= 0.0f) continue;
    float const cis = dot3(d, E[1].ns);
    if (cis <= 0.0f) continue;
    float fx, fy;
    if (!film_position(cam, sub3(lv.pos, E[1].pos), &fx, &fy)) continue;
    HitRec const h = q.find(E[1].pos, d);
    if (h.prim != lv.tri) continue;
    // fx < 1 can still round fx * W up to W: keep the last column / row
    uint32_t const jx = min(uint32_t(fx * float(int(W))), W - 1u);
    uint32_t const jy = min(uint32_t(fy * float(int(H))), H - 1u);
    float const flux = bc.C / fourth(cis);
    float const G = cis * cog / (h.t * h.t);
    L[k + 1].s_prev = flux;
    float const wst = rcpf(mis_light_side(L, k, G, 1.0f));
    double const sc1 = double(flux * G) * kInvPi * double(wst);
    float3 const pw = mul3(lv.power, E[1].power);
    float const w = float(sc1);
    float* dst = splat + 3 * size_t(jy * W + jx);
    atomicAdd(dst + 0, pw.x * w);
    atomicAdd(dst + 1, pw.y * w);
    atomicAdd(dst + 2, pw.z * w);
  }

  // EvaluatePathST
  float3 color = f3(0.f, 0.f, 0.f);
  for (uint32_t t = 2; t < kmax; ++t) {
    BVert& ev = E[t];
    if (ev.tri < 0) break;
    DevBsdf const eb = gp(s.bsdfs)[__float_as_uint(gp(s.prim_ng)[ev.tri].w)];
    if (eb.type == XYZ_LIGHT) {
      float wst = 1.0f;
      if (t > 2) {
        E[t + 1].s_next = L[0].s_next;
        ev.s_next = float(kInvPi);
        wst = rcpf(mis_eye_side(E, t, 1.0f, 1.0f));
      }
      double const sc1 = double(bc.inv_area) * kInvPi * double(wst);
      float3 const pw = mul3(f3(eb.phi[0], eb.phi[1], eb.phi[2]), ev.power);
      color = add3(color, scl3(pw, float(sc1)));
      break;
    }
    if (eb.type == XYZ_MIRROR || eb.type == XYZ_GLASS) continue;
    for (uint32_t k = 1; (k + t - 1) < kmax; ++k) {
      BVert const& lv = L[k];
      if (lv.tri < 0) break;
      uint32_t const lt = gp(s.bsdfs)[__float_as_uint(gp(s.prim_ng)[lv.tri].w)].type;
      if (lt == XYZ_MIRROR || lt == XYZ_GLASS) continue;
      float3 const d = normalize3(sub3(ev.pos, lv.pos));
      float const cog = dot3(d, lv.ng);
      if (cog <= 0.0f) continue;
      if (dot3(d, lv.ns) <= 0.0f) continue;
      if (-dot3(d, ev.ng) <= 0.0f) continue;
      float const cis = -dot3(d, ev.ns);
      if (cis <= 0.0f) continue;
      HitRec const h = q.find(lv.pos, d);
      if (h.prim != ev.tri) continue;
      float const G = cis * cog / (h.t * h.t);
      L[k + 1].s_prev = float(kInvPi);
      float wst = mis_light_side(L, k, G, 1.0f);
      E[t + 1].s_next = float(kInvPi);
      wst = rcpf(mis_eye_side(E, t, G, wst));
      double const sc1 = double(G) * (kInvPi * kInvPi) * double(wst);
      float3 const pw = mul3(lv.power, ev.power);
      color = add3(color, scl3(pw, float(sc1)));
    }
  }
  return color;
}

// Visibility queries traced at once by the calling thread.
template <bool kCount>
struct BdptTrace {
  const DevScene* s;
  uint32_t* stk;
  uint32_t stride;
  uint32_t* queries;
  __device__ __forceinline__ HitRec find(float3 o, float3 d) {
    if (kCount) ++*queries;
    return trace_lane(*s, o, d, stk, stride);
  }
};


// Counts a sample's visibility queries (every answer a miss).
struct BdptCountQueries {
  uint32_t n;
  __device__ __forceinline__ HitRec find(float3, float3) {
    ++n;
    return bdpt_no_hit();
  }
};

// Writes a sample's visibility rays to consecutive queue entries from `pos`
// (origins and directions as 3 floats each; every answer a miss); entries at
// or past `cap` are not written (the caller sizes the queue for the most
// queries a sample can make and checks the total).
struct BdptRecordQueries {
  float* org;
  float* dir;
  uint32_t pos;
  uint32_t cap;
  __device__ __forceinline__ HitRec find(float3 o, float3 d) {
    if (pos >= cap) return bdpt_no_hit();
    size_t const k = 3 * size_t(pos++);
    org[k] = o.x;
    org[k + 1] = o.y;
    org[k + 2] = o.z;
    dir[k] = d.x;
    dir[k + 1] = d.y;
    dir[k + 2] = d.z;
    return bdpt_no_hit();
  }
};

// Answers a sample's visibility queries, in order, from the closest hits of
// its recorded rays.
struct BdptReplayQueries {
  const int32_t* prim;
  const float* t;
  uint32_t pos;
  __device__ __forceinline__ HitRec find(float3, float3) {
    HitRec h = bdpt_no_hit();
    h.prim = prim[pos];
    h.t = t[pos];
    ++pos;
    return h;
  }
};

// One BDPT sample of pixel (px, py) traced by one thread: returns the pixel's
// own colour and adds the light-tracing (t = 1) contributions to `splat`.
template <bool kCount>
__device__ __forceinline__ float3 bdpt_sample(const DevScene& s, const DevCamera& cam,
                                              const BdptConst& bc, uint32_t px, uint32_t py,
                                              uint32_t smp, uint32_t* stk, uint32_t stride,
                                              float* splat, uint32_t* queries) {
  BVert L[kBdptMaxDepth + 3], E[kBdptMaxDepth + 3];
  bdpt_build<kCount>(s, cam, bc, px, py, smp, stk, stride, L, E, queries);
  BdptTrace<kCount> q{&s, stk, stride, queries};
  return bdpt_connect(s, cam, bc, L, E, splat, q);
}

// Stored subpath vertices: entry k of path p (0 light, 1 eye) of item i at
// v[((p * entries + k) * 4 + f) * items + i] for its four float4 fields f, so
// that neighbouring items' fields are neighbours.  (pos, tri + 1 | specular
// << 31), (ng, g), (ns, s_prev), (power, s_next).
__device__ __forceinline__ void bvert_store(float4* v, size_t fstride, const BVert& b) {
  uint32_t const w = uint32_t(b.tri + 1) | (b.specular ? 0x80000000u : 0u);
  v[0] = make_float4(b.pos.x, b.pos.y, b.pos.z, __uint_as_float(w));
  v[fstride] = make_float4(b.ng.x, b.ng.y, b.ng.z, b.g);
  v[2 * fstride] = make_float4(b.ns.x, b.ns.y, b.ns.z, b.s_prev);
  v[3 * fstride] = make_float4(b.power.x, b.power.y, b.power.z, b.s_next);
}
__device__ __forceinline__ void bvert_load(const float4* v, size_t fstride, BVert& b) {
  float4 const a = v[0], g = v[fstride], n = v[2 * fstride], p = v[3 * fstride];
  uint32_t const w = __float_as_uint(a.w);
  b.tri = int32_t(w & 0x7FFFFFFFu) - 1;
  b.specular = (w >> 31) != 0;
  b.pos = f3(a.x, a.y, a.z);
  b.ng = f3(g.x, g.y, g.z);
  b.g = g.w;
  b.ns = f3(n.x, n.y, n.z);
  b.s_prev = n.w;
  b.power = f3(p.x, p.y, p.z);
  b.s_next = p.w;
}

// The entries of a built subpath that a connection pass can read: up to and
// including the first entry from `first` on past the path's end (tri < 0), at
// most kmax + 1.  The light path's vertices start at entry 1, the eye path's at
// entry 2 (entry 1 is the camera, which has no triangle).
__device__ __forceinline__ uint32_t bdpt_path_len(const BVert* P, uint32_t kmax, uint32_t first) {
  uint32_t k = first;
  while (k < kmax + 1 && P[k].tri >= 0) ++k;
  return k;
}

}  // namespace xyzhip

#endif  // XYZ_BDPT_DEVICE_H_

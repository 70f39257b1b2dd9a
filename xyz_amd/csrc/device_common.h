// Device-side data layout and the closest-hit traversal shared by every kernel.
//
// IEEE binary32 contract (built with -ffp-contract=off, correctly rounded
// division and square root, denormals kept): every expression below is written
// in the reference's operation order so that results are bit-identical.
#ifndef XYZ_DEVICE_COMMON_H_
#define XYZ_DEVICE_COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xyzhip {

// Per-material record on the device.
struct DevBsdf {
  uint32_t type;
  int32_t texture;
  float rgb[3];     // albedo or L_e
  float phi[3];     // LIGHT: power Phi = L_e * pi * A_lights
  float shininess;
  float ior;
};

// Texture directory entry.
struct DevTex {
  uint32_t w, h, offset, pad;
};

// Everything a kernel needs to know about one scene; passed by value.
//
// HBM layout (all arrays 16-byte aligned):
//   nodes     uint2[n_nodes]   pre-order kd nodes, 8 bytes each
//   triA/B/C  float4[n_refs]   leaf-ordered triangle records:
//             A = (v0.xyz, n.x), B = (e1.xyz, n.y), C = (e2.xyz, n.z),
//             n = cross(e2, e1) exactly as the reference recomputes it per test
//   ref_prim  int32[n_refs]    primitive id of each record
//   prim_*    per primitive    shading data, read once per hit
struct DevScene {
  const uint2* nodes;
  const float4* triA;
  const float4* triB;
  const float4* triC;
  const int32_t* ref_prim;
  uint32_t stack_cap;  // kd_max_depth + 1 entries of 8 bytes per lane
  // shading
  const float4* prim_ng;    // (ng.xyz, bsdf index as bits)
  const float4* prim_n0;    // NULL for flat shading, else n0.xyz
  const float4* prim_s1;
  const float4* prim_s2;
  const float2* prim_uv;    // NULL or 3 per primitive
  const float4* prim_v0;    // v0, e1, e2 per primitive (for light sampling)
  const float4* prim_e1;
  const float4* prim_e2;
  // lights (caller order of LIGHT triangles)
  uint32_t n_lights;
  const float* light_cdf;   // n_lights + 1
  const uint32_t* light_prim;
  float light_area;
  // materials and textures
  const DevBsdf* bsdfs;
  uint32_t n_bsdfs;
  const float* texels;
  const DevTex* tex;
  uint32_t n_tex;
};

struct HitRec {
  int32_t prim;  // -1 = miss
  float t, f, g;
  bool back;
};

__device__ __forceinline__ float sel3(float3 v, uint32_t a) {
  return a == 0 ? v.x : (a == 1 ? v.y : v.z);
}

// Closest-hit query, the reference's recursive traversal (xyz/geom/kdtree.cpp:82-150,
// 366-372) run iteratively.  The stack holds (far node, e) pairs; t_max after a
// pop is the e of the new top (or +inf) because only pushes change t_max before
// a hit ends the query.  `stk` points at this lane's first stack slot; slots are
// `stride` uint2 apart (LDS, interleaved by lane).
template <bool kCount>
__device__ __forceinline__ HitRec trace_closest(const DevScene& s, float3 o, float3 d,
                                                uint2* stk, uint32_t stride,
                                                uint32_t* cnt_nodes, uint32_t* cnt_tris) {
  const float kInf = __builtin_inff();
  float t_min = 1e-5f;  // Intersection::Init, xyz/geom/intersection.hpp:11-14
  float t_max = kInf;
  uint32_t sp = 0;
  uint32_t node = 0;
  uint32_t nn = 0, nt = 0;
  HitRec h;
  h.prim = -1;
  h.t = kInf;
  h.f = 0.0f;
  h.g = 0.0f;
  h.back = false;
  for (;;) {
    uint2 w = s.nodes[node];
    if (kCount) nn++;
    // descend to a leaf: xyz/geom/kdtree.cpp:82-122
    while ((w.x & 3u) != 3u) {
      uint32_t const axis = w.x & 3u;
      float const split = __uint_as_float(w.y);
      float const oa = sel3(o, axis);
      float const da = sel3(d, axis);
      float const e = (split - oa) / da;
      bool const i = oa < split;
      uint32_t const left = node + 1u, right = w.x >> 2;
      uint32_t const nearc = i ? left : right;
      uint32_t const farc = i ? right : left;
      if ((e > t_max) || (e < 0.0f)) {
        node = nearc;
      } else if (e < t_min) {
        node = farc;
      } else {
        stk[sp * stride] = make_uint2(farc, __float_as_uint(e));
        sp++;
        t_max = e;
        node = nearc;
      }
      w = s.nodes[node];
      if (kCount) nn++;
    }
    // leaf: every triangle in stored order, last accept wins (kdtree.cpp:140-150)
    uint32_t const cnt = w.x >> 2;
    uint32_t const off = w.y;
    int32_t hit = -1;
    for (uint32_t j = 0; j < cnt; ++j) {
      if (kCount) nt++;
      float4 const A = s.triA[off + j];
      // xyz/geom/triangle.cpp:127-165 (FindIntersectionDF)
      float4 const B = s.triB[off + j];
      float4 const C = s.triC[off + j];
      float const nx = A.w, ny = B.w, nz = C.w;
      float const det = d.x * nx + d.y * ny + d.z * nz;
      if (fabsf(det) < 1e-6f) continue;
      float const inv = 1.0f / det;
      float const mx = A.x - o.x, my = A.y - o.y, mz = A.z - o.z;
      float const t = inv * (mx * nx + my * ny + mz * nz);
      if ((t < t_min) || (t_max < t)) continue;
      float const qx = d.y * mz - d.z * my;
      float const qy = d.z * mx - d.x * mz;
      float const qz = d.x * my - d.y * mx;
      float const f = inv * (C.x * qx + C.y * qy + C.z * qz);
      if ((f < 0.0f) || (f > 1.0f)) continue;
      float const g = inv * -(B.x * qx + B.y * qy + B.z * qz);
      if ((g < 0.0f) || ((f + g) > 1.0f)) continue;
      t_max = t;
      h.f = f;
      h.g = g;
      h.back = det > 0.0f;
      hit = int32_t(off + j);
    }
    if (hit >= 0) {
      h.prim = s.ref_prim[hit];
      h.t = t_max;
      break;
    }
    if (sp == 0) break;
    sp--;
    uint2 const top = stk[sp * stride];
    node = top.x;
    t_min = __uint_as_float(top.y);
    t_max = (sp > 0) ? __uint_as_float(stk[(sp - 1) * stride].y) : kInf;
  }
  if (kCount) {
    *cnt_nodes = nn;
    *cnt_tris = nt;
  }
  return h;
}

}  // namespace xyzhip

#endif  // XYZ_DEVICE_COMMON_H_

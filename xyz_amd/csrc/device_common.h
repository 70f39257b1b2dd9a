// Device-side data layout and the closest-hit traversal shared by every kernel.
//
// IEEE binary32 contract (built with -ffp-contract=off, correctly rounded
// division and square root, denormals kept): every expression below is written
// in the reference's operation order so that results are bit-identical.
#ifndef XYZ_DEVICE_COMMON_H_
#define XYZ_DEVICE_COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/xyz_hip.h"

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

// Recomputes what the traversal decided at interior node `w`: the distance e to
// its split plane and its far child (xyz/geom/kdtree.cpp:110-122).
__device__ __forceinline__ float split_distance(uint2 w, float3 o, float3 d, uint32_t node,
                                                uint32_t* farc, uint32_t* nearc) {
  uint32_t const axis = w.x & 3u;
  float const split = __uint_as_float(w.y);
  float const oa = sel3(o, axis);
  float const da = sel3(d, axis);
  bool const i = oa < split;
  uint32_t const left = node + 1u, right = w.x >> 2;
  *nearc = i ? left : right;
  *farc = i ? right : left;
  return (split - oa) / da;
}

// Closest-hit query, the reference's recursive traversal (xyz/geom/kdtree.cpp:82-150,
// 366-372) run iteratively.
//
// Stack: one 4-byte entry per pending far child, holding the interior node whose
// far child is pending.  Only pushes change t_max before a hit ends the query,
// so t_max always equals the split distance e of the top entry (+inf when the
// stack is empty).  A pop therefore sets t_min = t_max (the popped entry's e)
// and recomputes t_max from the new top; both recomputations repeat the same
// IEEE division and give the same bits.  `stk` points at this lane's first
// stack slot; slots are `stride` words apart (LDS, interleaved by lane).
template <bool kCount>
__device__ __forceinline__ HitRec trace_closest(const DevScene& s, float3 o, float3 d,
                                                uint32_t* stk, uint32_t stride,
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
    // descend to a leaf: xyz/geom/kdtree.cpp:82-108
    while ((w.x & 3u) != 3u) {
      uint32_t farc, nearc;
      float const e = split_distance(w, o, d, node, &farc, &nearc);
      if ((e > t_max) || (e < 0.0f)) {
        node = nearc;
      } else if (e < t_min) {
        node = farc;
      } else {
        stk[sp * stride] = node;
        sp++;
        t_max = e;
        node = nearc;
      }
      w = s.nodes[node];
      if (kCount) nn++;
    }
    // leaf: every triangle in stored order, last accept wins (kdtree.cpp:140-150).
    // The next record's loads are issued before the current one is tested.
    uint32_t const cnt = w.x >> 2;
    uint32_t const off = w.y;
    int32_t hit = -1;
    if (cnt > 0) {
      float4 A = s.triA[off], B = s.triB[off], C = s.triC[off];
      for (uint32_t j = 0; j < cnt; ++j) {
        uint32_t const jn = (j + 1 < cnt) ? j + 1 : j;
        float4 const nA = s.triA[off + jn], nB = s.triB[off + jn], nC = s.triC[off + jn];
        if (kCount) nt++;
        // xyz/geom/triangle.cpp:127-165 (FindIntersectionDF)
        float const nx = A.w, ny = B.w, nz = C.w;
        float const det = d.x * nx + d.y * ny + d.z * nz;
        if (!(fabsf(det) < 1e-6f)) {
          float const inv = 1.0f / det;
          float const mx = A.x - o.x, my = A.y - o.y, mz = A.z - o.z;
          float const t = inv * (mx * nx + my * ny + mz * nz);
          if (!((t < t_min) || (t_max < t))) {
            float const qx = d.y * mz - d.z * my;
            float const qy = d.z * mx - d.x * mz;
            float const qz = d.x * my - d.y * mx;
            float const f = inv * (C.x * qx + C.y * qy + C.z * qz);
            if (!((f < 0.0f) || (f > 1.0f))) {
              float const g = inv * -(B.x * qx + B.y * qy + B.z * qz);
              if (!((g < 0.0f) || ((f + g) > 1.0f))) {
                t_max = t;
                h.f = f;
                h.g = g;
                h.back = det > 0.0f;
                hit = int32_t(off + j);
              }
            }
          }
        }
        A = nA;
        B = nB;
        C = nC;
      }
    }
    if (hit >= 0) {
      h.prim = s.ref_prim[hit];
      h.t = t_max;
      break;
    }
    if (sp == 0) break;
    // pop: continue at the far child of the top entry with [e, e_below]
    sp--;
    uint32_t const p = stk[sp * stride];
    uint32_t farc, nearc;
    float const e = split_distance(s.nodes[p], o, d, p, &farc, &nearc);
    node = farc;
    t_min = e;
    if (sp > 0) {
      uint32_t const pb = stk[(sp - 1) * stride];
      uint32_t f2, n2;
      t_max = split_distance(s.nodes[pb], o, d, pb, &f2, &n2);
    } else {
      t_max = kInf;
    }
  }
  if (kCount) {
    *cnt_nodes = nn;
    *cnt_tris = nt;
  }
  return h;
}

}  // namespace xyzhip

#endif  // XYZ_DEVICE_COMMON_H_

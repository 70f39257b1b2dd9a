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
// 366-372) run iteratively, as a step machine so that a lane can be refilled with
// a new ray as soon as its current one is finished.
//
// Stack: one 4-byte entry per pending far child, holding the interior node whose
// far child is pending.  Only pushes change t_max before a hit ends the query,
// so t_max always equals the split distance e of the top entry (+inf when the
// stack is empty).  A pop therefore sets t_min = t_max (the popped entry's e)
// and recomputes t_max from the new top; both recomputations repeat the same
// IEEE division and give the same bits.  The stack lives in LDS, interleaved by
// lane (`stride` words apart).
//
// A step either descends to the next leaf or tests up to kLeafChunk of the
// current leaf's triangles in stored order with t_max carried from one to the
// next; after the leaf's last triangle it ends the query (hit, or empty stack)
// or pops.  Each ray therefore runs exactly the reference's operations.
#ifndef XYZ_LEAF_CHUNK
#define XYZ_LEAF_CHUNK 32
#endif
constexpr uint32_t kLeafChunk = XYZ_LEAF_CHUNK;

struct TraceState {
  float3 o, d;
  float t_min, t_max;
  uint32_t sp, node, lj, lend;
  int32_t hit;
  float f, g;
  bool back, in_leaf;
  uint32_t nn, nt;
};

__device__ __forceinline__ void trace_init(TraceState& st, float3 o, float3 d) {
  st.o = o;
  st.d = d;
  st.t_min = 1e-5f;  // Intersection::Init, xyz/geom/intersection.hpp:11-14
  st.t_max = __builtin_inff();
  st.sp = 0;
  st.node = 0;
  st.lj = st.lend = 0;
  st.hit = -1;
  st.f = st.g = 0.0f;
  st.back = false;
  st.in_leaf = false;
  st.nn = st.nt = 0;
}

// One double-faced ray/triangle test, xyz/geom/triangle.cpp:127-165, written
// without branches: every quantity is computed in the reference's order and
// the accept flag combines the reference's reject tests.
__device__ __forceinline__ bool tri_test(float3 o, float3 d, float4 A, float4 B, float4 C,
                                         float t_min, float t_max, float* t_out, float* f_out,
                                         float* g_out, float* det_out) {
  float const nx = A.w, ny = B.w, nz = C.w;
  float const det = d.x * nx + d.y * ny + d.z * nz;
  float const inv = 1.0f / det;
  float const mx = A.x - o.x, my = A.y - o.y, mz = A.z - o.z;
  float const t = inv * (mx * nx + my * ny + mz * nz);
  float const qx = d.y * mz - d.z * my;
  float const qy = d.z * mx - d.x * mz;
  float const qz = d.x * my - d.y * mx;
  float const f = inv * (C.x * qx + C.y * qy + C.z * qz);
  float const g = inv * -(B.x * qx + B.y * qy + B.z * qz);
  bool const ok = !(fabsf(det) < 1e-6f) && !((t < t_min) || (t_max < t)) &&
                  !((f < 0.0f) || (f > 1.0f)) && !((g < 0.0f) || ((f + g) > 1.0f));
  *t_out = t;
  *f_out = f;
  *g_out = g;
  *det_out = det;
  return ok;
}

// Advances the query by one step; returns true when it has finished.
template <bool kCount>
__device__ __forceinline__ bool trace_step(const DevScene& s, TraceState& st, uint32_t* stk,
                                           uint32_t stride) {
  if (!st.in_leaf) {
    // descend to a leaf: xyz/geom/kdtree.cpp:82-108
    uint2 w = s.nodes[st.node];
    if (kCount) st.nn++;
    while ((w.x & 3u) != 3u) {
      uint32_t farc, nearc;
      float const e = split_distance(w, st.o, st.d, st.node, &farc, &nearc);
      if ((e > st.t_max) || (e < 0.0f)) {
        st.node = nearc;
      } else if (e < st.t_min) {
        st.node = farc;
      } else {
        stk[st.sp * stride] = st.node;
        st.sp++;
        st.t_max = e;
        st.node = nearc;
      }
      w = s.nodes[st.node];
      if (kCount) st.nn++;
    }
    st.lj = w.y;
    st.lend = w.y + (w.x >> 2);
    st.in_leaf = true;
  }
  // up to kLeafChunk triangles of the leaf, in stored order (kdtree.cpp:140-150)
  uint32_t const stop = (st.lend - st.lj > kLeafChunk) ? st.lj + kLeafChunk : st.lend;
  if (st.lj < stop) {
    float4 A = s.triA[st.lj], B = s.triB[st.lj], C = s.triC[st.lj];
    for (uint32_t j = st.lj; j < stop; ++j) {
      uint32_t const jn = (j + 1 < stop) ? j + 1 : j;
      float4 const nA = s.triA[jn], nB = s.triB[jn], nC = s.triC[jn];
      if (kCount) st.nt++;
      float t, f, g, det;
      if (tri_test(st.o, st.d, A, B, C, st.t_min, st.t_max, &t, &f, &g, &det)) {
        st.t_max = t;
        st.f = f;
        st.g = g;
        st.back = det > 0.0f;
        st.hit = int32_t(j);
      }
      A = nA;
      B = nB;
      C = nC;
    }
    st.lj = stop;
  }
  if (st.lj < st.lend) return false;  // more of this leaf next step
  st.in_leaf = false;
  if (st.hit >= 0 || st.sp == 0) return true;
  // pop: continue at the far child of the top entry with [e, e_below]
  st.sp--;
  uint32_t const p = stk[st.sp * stride];
  uint32_t farc, nearc;
  float const e = split_distance(s.nodes[p], st.o, st.d, p, &farc, &nearc);
  st.node = farc;
  st.t_min = e;
  if (st.sp > 0) {
    uint32_t const pb = stk[(st.sp - 1) * stride];
    uint32_t f2, n2;
    st.t_max = split_distance(s.nodes[pb], st.o, st.d, pb, &f2, &n2);
  } else {
    st.t_max = __builtin_inff();
  }
  return false;
}

// ---------------------------------------------------------------------------
// Wave-cooperative variant of trace_step, for persistent kernels in which all
// 64 lanes of a wave call it together (`active` marks lanes with a query).
//
// Within one kd leaf the reference's sequential loop ("accept when t_min <= t
// <= t_max, then t_max = t"; kdtree.cpp:140-150, triangle.cpp:143) returns the
// valid triangle with the smallest t, ties going to the later triangle, and
// that triangle's t -- for any interval bounds that are not NaN and any t that
// is not NaN.  So the (lane, triangle) pairs of all lanes' current leaf
// chunks can be tested in any order by any lane: the wave spreads them over
// its 64 lanes and each ray keeps the minimum of the 64-bit key
// (bits(t) << 32) | ~index in LDS (t >= t_min > 0, so its bits order like t).
// A lane whose bounds are NaN runs its chunk sequentially; a NaN t stores key 0
// and makes its lane redo the chunk sequentially.  f, g and the side are
// recomputed from the winning triangle (the same operations give the same
// bits).
#ifndef XYZ_COOP_CHUNK
#define XYZ_COOP_CHUNK 128
#endif
constexpr uint32_t kCoopChunk = XYZ_COOP_CHUNK;

template <bool kCount>
__device__ __forceinline__ void seq_chunk(const DevScene& s, TraceState& st, uint32_t stop) {
  for (uint32_t j = st.lj; j < stop; ++j) {
    if (kCount) st.nt++;
    float t, f, g, det;
    if (tri_test(st.o, st.d, s.triA[j], s.triB[j], s.triC[j], st.t_min, st.t_max, &t, &f, &g,
                 &det)) {
      st.t_max = t;
      st.f = f;
      st.g = g;
      st.back = det > 0.0f;
      st.hit = int32_t(j);
    }
  }
  st.lj = stop;
}

// Returns true when this lane's (active) query finished in this step.
// `keys` is this wave's 64-entry LDS array (index = lane).
template <bool kCount>
__device__ __forceinline__ bool coop_step(const DevScene& s, TraceState& st, bool active,
                                          uint32_t* stk, uint32_t stride,
                                          unsigned long long* keys) {
  uint32_t const lane = __lane_id();
  if (active && !st.in_leaf) {
    uint2 w = s.nodes[st.node];
    if (kCount) st.nn++;
    while ((w.x & 3u) != 3u) {
      uint32_t farc, nearc;
      float const e = split_distance(w, st.o, st.d, st.node, &farc, &nearc);
      if ((e > st.t_max) || (e < 0.0f)) {
        st.node = nearc;
      } else if (e < st.t_min) {
        st.node = farc;
      } else {
        stk[st.sp * stride] = st.node;
        st.sp++;
        st.t_max = e;
        st.node = nearc;
      }
      w = s.nodes[st.node];
      if (kCount) st.nn++;
    }
    st.lj = w.y;
    st.lend = w.y + (w.x >> 2);
    st.in_leaf = true;
    st.hit = -1;
  }
  bool const nan_bounds = (st.t_min != st.t_min) || (st.t_max != st.t_max);
  bool const coop = active && !nan_bounds;
  uint32_t const rem = st.lend - st.lj;
  uint32_t const c = coop ? (rem < kCoopChunk ? rem : kCoopChunk) : 0u;
  // exclusive prefix of c over the wave
  uint32_t incl = c;
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    uint32_t const v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  uint32_t const excl = incl - c;
  uint32_t const total = __shfl(incl, 63);
  keys[lane] = ~0ull;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t base = 0; base < total; base += 64) {
    uint32_t const k = base + lane;
    // owner: the last lane whose prefix start is <= k
    uint32_t lo = 0, hi = 64;
#pragma unroll
    for (int it = 0; it < 6; ++it) {
      uint32_t const mid = (lo + hi) >> 1;
      uint32_t const pm = __shfl(excl, mid);
      if (pm <= k) lo = mid; else hi = mid;
    }
    uint32_t const owner = lo;
    float3 const o = make_float3(__shfl(st.o.x, owner), __shfl(st.o.y, owner), __shfl(st.o.z, owner));
    float3 const d = make_float3(__shfl(st.d.x, owner), __shfl(st.d.y, owner), __shfl(st.d.z, owner));
    float const tmin = __shfl(st.t_min, owner);
    float const tmax = __shfl(st.t_max, owner);
    uint32_t const j = __shfl(st.lj, owner) + (k - __shfl(excl, owner));
    if (k < total) {
      float t, f, g, det;
      bool const ok = tri_test(o, d, s.triA[j], s.triB[j], s.triC[j], tmin, tmax, &t, &f, &g, &det);
      if (t != t && !(fabsf(det) < 1e-6f)) {
        atomicMin(&keys[owner], 0ull);
      } else if (ok) {
        unsigned long long const key =
            (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | (0xFFFFFFFFu - j);
        atomicMin(&keys[owner], key);
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (!active) return false;
  if (coop) {
    unsigned long long const key = keys[lane];
    if (key == 0ull) {
      seq_chunk<kCount>(s, st, st.lj + c);  // a NaN t: the sequential semantics
    } else {
      if (key != ~0ull) {
        st.t_max = __uint_as_float(uint32_t(key >> 32));
        st.hit = int32_t(0xFFFFFFFFu - uint32_t(key & 0xFFFFFFFFull));
      }
      if (kCount) st.nt += c;
      st.lj += c;
    }
  } else {
    uint32_t const stop = (rem > kLeafChunk) ? st.lj + kLeafChunk : st.lend;
    seq_chunk<kCount>(s, st, stop);
  }
  if (st.lj < st.lend) return false;
  st.in_leaf = false;
  if (st.hit >= 0) {
    if (coop) {
      float t, f, g, det;
      uint32_t const j = uint32_t(st.hit);
      (void)tri_test(st.o, st.d, s.triA[j], s.triB[j], s.triC[j], st.t_min, st.t_max, &t, &f, &g, &det);
      st.f = f;
      st.g = g;
      st.back = det > 0.0f;
    }
    return true;
  }
  if (st.sp == 0) return true;
  st.sp--;
  uint32_t const p = stk[st.sp * stride];
  uint32_t farc, nearc;
  float const e = split_distance(s.nodes[p], st.o, st.d, p, &farc, &nearc);
  st.node = farc;
  st.t_min = e;
  if (st.sp > 0) {
    uint32_t const pb = stk[(st.sp - 1) * stride];
    uint32_t f2, n2;
    st.t_max = split_distance(s.nodes[pb], st.o, st.d, pb, &f2, &n2);
  } else {
    st.t_max = __builtin_inff();
  }
  return false;
}

__device__ __forceinline__ HitRec trace_result(const DevScene& s, const TraceState& st) {
  HitRec h;
  if (st.hit >= 0) {
    h.prim = s.ref_prim[st.hit];
    h.t = st.t_max;
    h.f = st.f;
    h.g = st.g;
    h.back = st.back;
  } else {
    h.prim = -1;
    h.t = __builtin_inff();
    h.f = h.g = 0.0f;
    h.back = false;
  }
  return h;
}

template <bool kCount>
__device__ __forceinline__ HitRec trace_closest(const DevScene& s, float3 o, float3 d,
                                                uint32_t* stk, uint32_t stride,
                                                uint32_t* cnt_nodes, uint32_t* cnt_tris) {
  TraceState st;
  trace_init(st, o, d);
  while (!trace_step<kCount>(s, st, stk, stride)) {
  }
  if (kCount) {
    *cnt_nodes = st.nn;
    *cnt_tris = st.nt;
  }
  return trace_result(s, st);
}

#ifndef XYZ_COOP
#define XYZ_COOP 1
#endif

// Per-wave ray source for persistent kernels.  A wave takes batches of 64
// consecutive work indices from a global counter (one atomic per batch, so
// neighbouring -- coherent -- rays stay in one wave) and hands them to lanes as
// the lanes finish.  All lanes must call refill() together; the batch fields
// are wave-uniform.
struct WaveFeeder {
  uint32_t base, count, next;
  __device__ __forceinline__ void init() { base = count = next = 0; }
  // Lanes with `need` receive an index (returned) or learn that the work is
  // exhausted (returns UINT32_MAX and sets *exhausted).  At most `room` lanes
  // of the wave receive one; indices are claimed `batch` (1..64) at a time.
  __device__ __forceinline__ uint32_t refill(bool need, uint32_t* counter, uint32_t total,
                                            bool* exhausted, uint32_t room = 64u,
                                            uint32_t batch = 64u) {
    uint32_t const lane = __lane_id();
    unsigned long long const below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t got = 0xFFFFFFFFu;
    for (;;) {
      unsigned long long const m = __ballot(need);
      if (m == 0 || room == 0) break;
      if (next == count) {
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(counter, batch);
        b = __shfl(b, 0);
        base = b;
        count = (b >= total) ? 0u : ((total - b < batch) ? total - b : batch);
        next = 0;
        if (count == 0) {
          if (need) *exhausted = true;
          break;
        }
      }
      uint32_t const avail = count - next;
      uint32_t const rank = uint32_t(__popcll(m & below));
      uint32_t const want = uint32_t(__popcll(m));
      uint32_t take = want < avail ? want : avail;
      take = take < room ? take : room;
      if (need && rank < take) {
        got = base + next + rank;
        need = false;
      }
      next += take;
      room -= take;
    }
    return got;
  }
};

// Wave-level work distribution for persistent kernels: the lanes with `need`
// take consecutive indices from `counter` (one atomic per wave); returns the
// lane's index, or UINT32_MAX when the lane does not need one.
__device__ __forceinline__ uint32_t wave_fetch(bool need, uint32_t* counter) {
  uint32_t const lane = __lane_id();
  unsigned long long const m = __ballot(need);
  if (m == 0) return 0xFFFFFFFFu;
  uint32_t const leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(m)));
  base = __shfl(base, leader);
  unsigned long long const below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  return need ? base + uint32_t(__popcll(m & below)) : 0xFFFFFFFFu;
}

}  // namespace xyzhip

#endif  // XYZ_DEVICE_COMMON_H_

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

// Triangles of a leaf are stored in clusters of kClusterSize records; a leaf's
// records occupy whole clusters (the last one may be partly used).
#ifndef XYZ_CLUSTER_SHIFT
#define XYZ_CLUSTER_SHIFT 4
#endif
#ifndef XYZ_CLUSTERS_PER_STEP
#define XYZ_CLUSTERS_PER_STEP 16
#endif
constexpr uint32_t kClusterShift = XYZ_CLUSTER_SHIFT;
constexpr uint32_t kClusterSize = 1u << kClusterShift;
// Clusters a lane examines per cooperative step (<= 16: survivor indices are
// packed 4 bits each in one 64-bit word).
constexpr uint32_t kClustersPerStep = XYZ_CLUSTERS_PER_STEP;
static_assert(kClustersPerStep >= 1 && kClustersPerStep <= 16, "survivor list holds 16 entries");
// Interior nodes a lane may visit per cooperative step (0: no limit); a lane
// that has not reached a leaf continues in the next step.
#ifndef XYZ_DESCENT_BUDGET
#define XYZ_DESCENT_BUDGET 8
#endif

// Everything a kernel needs to know about one scene; passed by value.
//
// HBM layout (all arrays 16-byte aligned):
//   nodes     uint2[n_nodes]   pre-order kd nodes, 8 bytes each; a leaf is
//                              (count << 2 | 3, first cluster)
//   clo/chi   float4[n_cl]     cluster culling records: (lo.xyz, P), (hi.xyz, Q),
//                              the triangles' box widened by the ray-independent
//                              part of the culling margin, and the margin's
//                              coefficients (cluster_bounds.h)
//   triA/B/C  float4[n_cl*16]  triangle records in cluster order:
//             A = (v0.xyz, n.x), B = (e1.xyz, n.y), C = (e2.xyz, n.z),
//             n = cross(e2, e1) exactly as the reference recomputes it per test
//   ref_prim  int32[n_cl*16]   primitive id of each record
//   rec_rank  uint32[n_cl*16]  the record's position in the reference's leaf
//                              order (the tie-break of equal distances)
//   rank_pos  uint32[n_cl*16]  per leaf, record position of each rank
//   prim_*    per primitive    shading data, read once per hit
struct DevScene {
  const uint2* nodes;
  const float4* clo;
  const float4* chi;
  const float4* triA;
  const float4* triB;
  const float4* triC;
  const int32_t* ref_prim;
  const uint32_t* rec_rank;
  const uint32_t* rank_pos;
  uint32_t stack_cap;  // kd_max_depth + 1 stack entries (one 32-bit word each) per lane
  // the root record: the union of the cluster boxes, (lo, largest P) and
  // (hi, largest Q); P NaN when no clip may be made (root_clip)
  float4 root_lo, root_hi;
  // shading
  const float4* prim_ng;    // (ng.xyz, bsdf index as bits)
  const float4* prim_n0;    // NULL for flat shading, else n0.xyz
  const float4* prim_s1;
  const float4* prim_s2;
  const float2* prim_uv;    // NULL or 3 per primitive
  const float4* prim_v0;    // per primitive: the triangle record's A, B, C
  const float4* prim_e1;    // ((v0, n.x), (e1, n.y), (e2, n.z)), for light
  const float4* prim_e2;    // sampling and shadow queries
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

// A scene array as a pointer the compiler knows addresses global memory.  A
// kernel that finds its scene in a job table in memory (not in a kernel
// argument) cannot tell which address space the table's pointers address and
// would read through flat loads, which make a wave wait for its LDS and
// memory operations together; every array of a scene is device memory.
#if defined(__HIP_DEVICE_COMPILE__)
#define XYZ_GLOBAL_AS __attribute__((address_space(1)))
#else
#define XYZ_GLOBAL_AS
#endif
template <typename T>
__device__ __forceinline__ const XYZ_GLOBAL_AS T* gp(const T* p) {
  return (const XYZ_GLOBAL_AS T*)p;
}

// Element i of one of the scene's record arrays, addressed by a 32-bit byte
// offset (scene creation keeps every record array below 4 GiB), so that the
// load takes the array's address as a uniform base plus a per-lane offset.
template <typename T>
__device__ __forceinline__ T rec_at(const T* base, uint32_t i) {
  auto const b = (const XYZ_GLOBAL_AS char*)base;
  return *(const XYZ_GLOBAL_AS T*)(b + i * uint32_t(sizeof(T)));
}

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

// The far child of interior node `node` (word w) for a ray from o: the child
// split_distance() returns as the far one.
__device__ __forceinline__ uint32_t far_child(uint2 w, float3 o, uint32_t node) {
  uint32_t const axis = w.x & 3u;
  float const split = __uint_as_float(w.y);
  return (sel3(o, axis) < split) ? (w.x >> 2) : node + 1u;
}

// Closest-hit query, the reference's recursive traversal (xyz/geom/kdtree.cpp:82-150,
// 366-372) run iteratively, as a step machine so that a lane can be refilled with
// a new ray as soon as its current one is finished.
//
// Stack: one 4-byte entry per pending far child, holding the interior node whose
// far child is pending.  Only pushes change t_max before a hit ends the query,
// so t_max always equals the split distance e of the top entry (t_cap, the
// root interval's t_max, when the stack is empty: +inf for a plain query).  A pop therefore sets t_min = t_max (the popped entry's e)
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
  float t_cap;              // t_max of the root interval (+inf but for capped shadow queries)
  uint32_t sp, node;
  uint32_t lj, lend;        // cooperative: next / end cluster; sequential: next rank / count
  uint32_t lc0, ln;         // current leaf: first cluster, triangle count
  unsigned long long key;   // cooperative: best (bits(t) << 32 | ~rank) in the leaf
  int32_t hit;              // record position of the hit
  float f, g;
  bool back, in_leaf;
  uint32_t nn, nt;  // counting mode: nodes visited, reference triangle tests
  uint32_t xt, xb;  // counting mode: triangle tests and cluster box tests executed
};

constexpr float kTMin = 1e-5f;  // Intersection::Init, xyz/geom/intersection.hpp:11-14

__host__ __device__ __forceinline__ void trace_init(TraceState& st, float3 o, float3 d) {
  st.o = o;
  st.d = d;
  st.t_min = kTMin;
  st.t_max = __builtin_inff();
  st.t_cap = __builtin_inff();
  st.sp = 0;
  st.node = 0;
  st.lj = st.lend = 0;
  st.lc0 = st.ln = 0;
  st.key = ~0ull;
  st.hit = -1;
  st.f = st.g = 0.0f;
  st.back = false;
  st.in_leaf = false;
  st.nn = st.nt = 0;
  st.xt = st.xb = 0;
}

// 1/det as the reference's IEEE division gives it, wherever tri_test uses the
// value.  For 1e-6 <= |det| < 2^126 one Newton step from the hardware
// reciprocal yields the correctly rounded quotient: xyz_check_reciprocal
// compares the two over every binary32 value of that range.  Below 1e-6
// tri_test rejects the triangle and ignores the value; 2^126 and above,
// infinities and NaN take the division.
constexpr float kRecipFastMax = 0x1p126f;
__host__ __device__ __forceinline__ float recip_fast(float det) {
#if defined(__HIP_DEVICE_COMPILE__)
  float const y = __builtin_amdgcn_rcpf(det);
  return __builtin_fmaf(__builtin_fmaf(-det, y, 1.0f), y, y);
#else
  return 1.0f / det;
#endif
}
__host__ __device__ __forceinline__ float recip_det(float det) {
  if (fabsf(det) < kRecipFastMax) return recip_fast(det);
  return 1.0f / det;
}
// 1/x with the bits of the IEEE quotient for every x: the fast form on the
// range where it is checked to equal the quotient, the division elsewhere.
__host__ __device__ __forceinline__ float recip_exact(float x) {
  float const ax = fabsf(x);
  if (ax >= 1e-6f && ax < kRecipFastMax) return recip_fast(x);
  return 1.0f / x;
}

// One double-faced ray/triangle test, xyz/geom/triangle.cpp:127-165, written
// without branches (but recip_det's division for out-of-range det): every
// quantity is computed in the reference's order and the accept flag combines
// the reference's reject tests.
__host__ __device__ __forceinline__ bool tri_test(float3 o, float3 d, float4 A, float4 B, float4 C,
                                         float t_min, float t_max, float* t_out, float* f_out,
                                         float* g_out, float* det_out) {
  float const nx = A.w, ny = B.w, nz = C.w;
  float const det = d.x * nx + d.y * ny + d.z * nz;
  float const inv = recip_det(det);
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
    uint2 w = gp(s.nodes)[st.node];
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
      w = gp(s.nodes)[st.node];
      if (kCount) st.nn++;
    }
    st.lc0 = w.y;
    st.ln = w.x >> 2;
    st.lj = 0;
    st.lend = st.ln;
    st.in_leaf = true;
  }
  // up to kLeafChunk triangles of the leaf, in the reference's order (kdtree.cpp:140-150)
  uint32_t const stop = (st.lend - st.lj > kLeafChunk) ? st.lj + kLeafChunk : st.lend;
  auto const pos = gp(s.rank_pos) + (size_t(st.lc0) << kClusterShift);
  for (uint32_t r = st.lj; r < stop; ++r) {
    uint32_t const j = pos[r];
    if (kCount) st.nt++;
    float t, f, g, det;
    if (tri_test(st.o, st.d, gp(s.triA)[j], gp(s.triB)[j], gp(s.triC)[j], st.t_min, st.t_max, &t, &f, &g,
                 &det)) {
      st.t_max = t;
      st.f = f;
      st.g = g;
      st.back = det > 0.0f;
      st.hit = int32_t(j);
    }
  }
  st.lj = stop;
  if (st.lj < st.lend) return false;  // more of this leaf next step
  st.in_leaf = false;
  if (st.hit >= 0 || st.sp == 0) return true;
  // pop: continue at the far child of the top entry with [e, e_below]
  st.sp--;
  uint32_t const p = stk[st.sp * stride];
  uint32_t farc, nearc;
  float const e = split_distance(gp(s.nodes)[p], st.o, st.d, p, &farc, &nearc);
  st.node = farc;
  st.t_min = e;
  if (st.sp > 0) {
    uint32_t const pb = stk[(st.sp - 1) * stride];
    uint32_t f2, n2;
    st.t_max = split_distance(gp(s.nodes)[pb], st.o, st.d, pb, &f2, &n2);
  } else {
    st.t_max = st.t_cap;
  }
  return false;
}

// ---------------------------------------------------------------------------
// Cluster culling.  cluster_may_hit() returns false only when no triangle of
// the cluster can be accepted by tri_test with t in [t_min, t_max].
//
// Error bound.  Let n* = e2 x e1 (exact), D* = d . n*, m* = v0 - o and (T, F, G)
// the exact solution of o + T d = v0 + F e1 + G e2 (T = m* . n* / D*,
// F = e2 . (d x m*) / D*, G = -e1 . (d x m*) / D*, the reference's formulas
// without rounding).  With u = 2^-24, every operation of tri_test rounded once,
// the stored n within 2.9 u |e1||e2| of n*, det, f and g as computed:
//   det = D* + eD,  |eD| <= 5.91 u |d| N
//   fl(e2 . q) = e2 . (d x m*) + eF,  |eF| <= 6.9 u |e2||d||m*|  (q = fl(d x m)
//   within 3.9 u |d||m*| of d x m*), and alike for e1 and for fl(m . n)
//   (|eT| <= 6.91 u |m*| N),
// and, exactly, f - F = F p + (1 + p)(eF - F eD) / det with |p| <= 2.01 u (the
// reciprocal and the product), alike for g and t.  Here N >= |e1||e2|,
// E >= max(|e1|, |e2|), lam = u |d| N / |det| (an accepted |det| is at least
// 1e-6), R >= |m*| and R >= |T d| - (the excess below).  For an accepted
// triangle (f, g >= 0, fl(f + g) <= 1) the point v0 + f e1 + g e2 lies within
// 2 u E of the triangle, so the point of the ray's line at T lies within
//   |f - F||e1| + |g - G||e2| + 2 u E <= 15 lam R + 6.5 lam E + 4.2 u E
// of the triangle, and the computed t is within
//   |t - T| <= (14.6 lam R + 2.7 u R) / |d|
// of T; both include the second-order terms while lam <= 0.01 (|F|, |G| <= 1
// + their errors, |T d| <= 1.3 R).  So the ray's line must meet the box
// widened by the first bound at a parameter within the second bound of
// [t_min, t_max].  The test widens the box by twice the lam terms plus 8 u R
// and 8 u E + 4 u (largest box coordinate) + 1e-30 (that part, which does not
// depend on the ray, is built into the stored box), i.e. by
//   D = R (|d| P + 8 u) + |d| Q,   P = 30 u N / 1e-6,  Q = 13 u N E / 1e-6
// (cluster_bounds.h), widens the slab interval by 6e-7 of its ends (at least
// 5 ulp, the box test's own rounding), and widens [t_min, t_max] by
// R (|d| P + 8 u) / |d|, with 1 / |d| bounded above by min |1 / d_i|.  Rays or
// clusters outside the range where no operation of tri_test can overflow
// (R < 1e12, 1e-12 < |d| < 1e12) or with lam > 0.01 are never culled.
//
// A zero direction component has an infinite reciprocal: the slab of that axis
// then gives (-inf, +inf) when o lies strictly inside it and an empty interval
// (culled) otherwise.  Where o lies exactly on a face of the widened box the
// product 0 * inf is NaN and fminf / fmaxf take the other end, which culls:
// the ray's points keep that coordinate, which lies D + K beyond the
// triangles' box, farther than the first bound.
// Scale of the culling margin: 1 in the library; the exactness test builds a
// copy with 0 to show that it detects a missing margin.
#ifndef XYZ_CULL_MARGIN
#define XYZ_CULL_MARGIN 1.0f
#endif

// sqrt and reciprocal within 1 ulp (the hardware instructions on the device);
// cluster_may_hit's margins cover their error.
__host__ __device__ __forceinline__ float approx_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
__host__ __device__ __forceinline__ float approx_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// The ray's part of the box test: |d| (within an ulp) and 1/d per axis (within
// an ulp; a zero component gives an infinity of its sign).  |d| is returned as
// +inf, which makes cluster_may_hit keep every cluster, for rays outside the
// range the margin covers: |d| not in (1e-12, 1e12) (NaN included), or a
// non-zero component whose reciprocal is not below 3e38.
__host__ __device__ __forceinline__ void cull_ray(float3 d, float* dlen, float3* inv) {
  float const dl = approx_sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
  float3 const iv = make_float3(approx_rcp(d.x), approx_rcp(d.y), approx_rcp(d.z));
  bool const ok = (dl > 1e-12f) && (dl < 1e12f) && (d.x == 0.0f || fabsf(iv.x) < 3e38f) &&
                  (d.y == 0.0f || fabsf(iv.y) < 3e38f) && (d.z == 0.0f || fabsf(iv.z) < 3e38f);
  *dlen = ok ? dl : __builtin_inff();
  *inv = iv;
}

// L, H: the cluster's record; o, dlen, inv: the ray (cull_ray).
__host__ __device__ __forceinline__ bool cluster_may_hit(float4 L, float4 H, float3 o, float dlen,
                                                         float3 inv, float t_min, float t_max) {
  float const u = 5.9604645e-08f;
  float const a = dlen * L.w;  // 30 lam
  // the largest distance from o to a point of the box
  float const rx = fmaxf(fabsf(o.x - L.x), fabsf(H.x - o.x));
  float const ry = fmaxf(fabsf(o.y - L.y), fabsf(H.y - o.y));
  float const rz = fmaxf(fabsf(o.z - L.z), fabsf(H.z - o.z));
  float const R = approx_sqrt(rx * rx + ry * ry + rz * rz) * 1.000002f;
  float const tr = R * (a + 8.0f * u);  // also the t widening, times |d|
  float const D = XYZ_CULL_MARGIN * (tr + dlen * H.w) * 1.00001f;
  // outside the covered range every cluster is kept (computed without a
  // branch, so that the record's loads are not split around one)
  bool const covered = (a <= 0.3f) && (R < 1e12f) && (D < 1e30f);
  float const px = ((L.x - D) - o.x) * inv.x, qx = ((H.x + D) - o.x) * inv.x;
  float const py = ((L.y - D) - o.y) * inv.y, qy = ((H.y + D) - o.y) * inv.y;
  float const pz = ((L.z - D) - o.z) * inv.z, qz = ((H.z + D) - o.z) * inv.z;
  float te = fmaxf(fmaxf(fminf(px, qx), fminf(py, qy)), fminf(pz, qz));
  float tx = fminf(fminf(fmaxf(px, qx), fmaxf(py, qy)), fmaxf(pz, qz));
  // widen by 6e-7 relative (at least 5 ulp); multiplications keep infinities
  te = te * ((te > 0.0f) ? 0.9999994f : 1.0000006f);
  tx = tx * ((tx > 0.0f) ? 1.0000006f : 0.9999994f);
  // the computed t's distance from the line's parameter, 1 / |d| <= min |1 / d_i|
  float const dt = XYZ_CULL_MARGIN * tr *
                   fminf(fminf(fabsf(inv.x), fabsf(inv.y)), fabsf(inv.z)) * 1.00001f;
  return !covered || !((te - dt > t_max) || (tx + dt < t_min) || (te > tx));
}

// The root interval of a fresh query narrowed to where a hit can lie.  The
// root record (the union of the cluster boxes with the largest P and Q) makes
// every quantity of cluster_may_hit at least as large as for any cluster, and
// each of its steps is monotone in them, so the record's widened slab
// interval [te - dt, tx + dt] holds every t at which tri_test can accept a
// triangle of the scene (cluster_may_hit over [t, t] keeps its cluster), and
// an empty one means the ray meets none.  The reference's traversal (root
// interval [1e-5, +inf]) run over [max(1e-5, te - dt), min(t_max, tx + dt)]
// instead visits the same leaves with intervals cut to that range, leaves
// out leaves wholly outside it, where nothing can be accepted, and so finds
// the same triangle at the same t (the argument of the capped shadow query,
// at both ends).  A query that certainly misses is put in an empty leaf with
// an empty stack, which its next step ends as a miss.  Rays or scenes outside
// cluster_may_hit's covered range keep their interval.
#ifndef XYZ_ROOT_CLIP
#define XYZ_ROOT_CLIP 1
#endif
__host__ __device__ __forceinline__ void root_clip(const DevScene& s, TraceState& st) {
  float dlen;
  float3 inv;
  cull_ray(st.d, &dlen, &inv);
  float4 const L = s.root_lo, H = s.root_hi;
  float3 const o = st.o;
  float const u = 5.9604645e-08f;
  float const a = dlen * L.w;
  float const rx = fmaxf(fabsf(o.x - L.x), fabsf(H.x - o.x));
  float const ry = fmaxf(fabsf(o.y - L.y), fabsf(H.y - o.y));
  float const rz = fmaxf(fabsf(o.z - L.z), fabsf(H.z - o.z));
  float const R = approx_sqrt(rx * rx + ry * ry + rz * rz) * 1.000002f;
  float const tr = R * (a + 8.0f * u);
  float const D = XYZ_CULL_MARGIN * (tr + dlen * H.w) * 1.00001f;
  bool const covered = (a <= 0.3f) && (R < 1e12f) && (D < 1e30f);
  if (!covered) return;
  float const px = ((L.x - D) - o.x) * inv.x, qx = ((H.x + D) - o.x) * inv.x;
  float const py = ((L.y - D) - o.y) * inv.y, qy = ((H.y + D) - o.y) * inv.y;
  float const pz = ((L.z - D) - o.z) * inv.z, qz = ((H.z + D) - o.z) * inv.z;
  float te = fmaxf(fmaxf(fminf(px, qx), fminf(py, qy)), fminf(pz, qz));
  float tx = fminf(fminf(fmaxf(px, qx), fmaxf(py, qy)), fmaxf(pz, qz));
  te = te * ((te > 0.0f) ? 0.9999994f : 1.0000006f);
  tx = tx * ((tx > 0.0f) ? 1.0000006f : 0.9999994f);
  float const dt = XYZ_CULL_MARGIN * tr *
                   fminf(fminf(fabsf(inv.x), fabsf(inv.y)), fabsf(inv.z)) * 1.00001f;
  float const lo = te - dt, hi = tx + dt;
  if ((te > tx) || (lo > st.t_max) || (hi < st.t_min)) {
    st.in_leaf = true;  // an empty leaf (lj == lend), nothing pending: a miss
    st.lc0 = st.ln = st.lj = st.lend = 0;
    st.sp = 0;
    return;
  }
  if (lo > st.t_min) st.t_min = lo;
  if (hi < st.t_max) st.t_max = st.t_cap = hi;
}

// ---------------------------------------------------------------------------
// Wave-cooperative variant of trace_step, for persistent kernels in which all
// 64 lanes of a wave call it together (`active` marks lanes with a query).
//
// Within one kd leaf the reference's sequential loop ("accept when t_min <= t
// <= t_max, then t_max = t"; kdtree.cpp:140-150, triangle.cpp:143) returns the
// valid triangle with the smallest t, ties going to the later triangle, and
// that triangle's t -- for interval bounds that are not NaN and any t that is
// not NaN.  So the (lane, triangle) pairs can be tested in any order by any
// lane: each ray keeps the minimum of the 64-bit key (bits(t) << 32) | ~rank
// in LDS across the steps of one leaf (t >= t_min > 0, so its bits order like
// t; rank = position in the reference's leaf order).
//
// A step takes up to kClustersPerStep clusters of each lane's leaf; the lane
// drops the clusters that cluster_may_hit rules out, and the wave spreads the
// remaining clusters' triangles over its 64 lanes.  A lane whose bounds are
// NaN, or that met a NaN t, runs the whole leaf sequentially in the
// reference's order.  f, g and the side are recomputed from the winning
// triangle (the same operations give the same bits).

#ifndef XYZ_COOP_BOX
#define XYZ_COOP_BOX 1
#endif

// Orders this wave's LDS accesses before and after the call.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive prefix sum of c over the wave and its total (all 64 lanes call
// it).  Inclusive scan by data-parallel primitives: within each row of 16
// lanes, the values 1, 2, 4 and 8 lanes below are added (a lane with no such
// lane in its row adds 0); then row_bcast:15 adds row r's last lane to row
// r + 1 (rows 1 and 3), and row_bcast:31 adds lane 31 to rows 2 and 3.
__device__ __forceinline__ void wave_prefix(uint32_t c, uint32_t* excl, uint32_t* total) {
  int v = int(c);
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  uint32_t const incl = uint32_t(v);
  *excl = incl - c;
  *total = uint32_t(__builtin_amdgcn_readlane(v, 63));
}

// The lane that owns item base + lane of a wave_prefix distribution (excl, c),
// for a batch with base < total, found without a search.  `own` is this wave's
// LDS table of 64 one-byte entries kStride bytes apart (a stride of 8 lets a
// kernel address it from its 64-bit key array's lane address with a constant
// offset, which saves a register); it holds 0xFF in every entry on entry and
// again on return.  Each lane whose items meet the batch marks the batch position of
// its first item there with its lane id; position 0 is always marked (by the
// lane holding item `base`), and the owner of position p is the lane that
// marked the last marked position <= p.  All 64 lanes must call it together.
template <uint32_t kStride>
__device__ __forceinline__ uint32_t batch_owner(uint8_t* own, uint32_t excl, uint32_t c,
                                                uint32_t base) {
  uint32_t const lane = __lane_id();
  if (c > 0u && excl < base + 64u && excl + c > base)
    own[kStride * (excl > base ? excl - base : 0u)] = uint8_t(lane);
  wave_sync();
  uint32_t const v = own[kStride * lane];
  // the marks at positions <= lane, shifted to the top: the leading zeros
  // count the positions from the last of them up to this lane
  unsigned long long const marks = __ballot(v != 0xFFu) << (63u - lane);
  uint32_t const p = lane - uint32_t(__builtin_clzll(marks));
  own[kStride * lane] = 0xFFu;  // after this lane's read of the same entry
  return uint32_t(__shfl(int(v), int(p)));
}

// The current leaf in the reference's order, with the sequential semantics.
__device__ __forceinline__ void leaf_sequential(const DevScene& s, TraceState& st) {
  auto const pos = gp(s.rank_pos) + (size_t(st.lc0) << kClusterShift);
  for (uint32_t r = 0; r < st.ln; ++r) {
    uint32_t const j = pos[r];
    float t, f, g, det;
    if (tri_test(st.o, st.d, gp(s.triA)[j], gp(s.triB)[j], gp(s.triC)[j], st.t_min, st.t_max, &t, &f, &g,
                 &det)) {
      st.t_max = t;
      st.f = f;
      st.g = g;
      st.back = det > 0.0f;
      st.hit = int32_t(j);
    }
  }
  st.lj = st.lend;
}

// t_max of the current interval: the split distance of the top stack entry.
__device__ __forceinline__ float interval_t_max(const DevScene& s, const TraceState& st,
                                                const uint32_t* stk, uint32_t stride) {
  if (st.sp == 0) return st.t_cap;
  uint32_t const pb = stk[(st.sp - 1) * stride];
  uint32_t f2, n2;
  return split_distance(rec_at(s.nodes, pb), st.o, st.d, pb, &f2, &n2);
}

// Returns true when this lane's (active) query finished in this step.
// `keys` is this wave's 64-entry LDS array (index = lane), `own` its table for
// batch_owner (64 entries kOwnStride bytes apart).
// kFG: the caller needs f, g and the side of the hit (otherwise only the
// primitive and t are filled in).
template <bool kCount, bool kFG = true, uint32_t kOwnStride = 8>
__device__ __forceinline__ bool coop_step(const DevScene& s, TraceState& st, bool active,
                                          uint32_t* stk, uint32_t stride,
                                          unsigned long long* keys, uint8_t* own) {
  uint32_t const lane = __lane_id();
  if (active && !st.in_leaf) {
    uint2 w = rec_at(s.nodes, st.node);
    if (kCount) st.nn++;
    uint32_t budget = XYZ_DESCENT_BUDGET;
    bool leaf = (w.x & 3u) == 3u;
    while (!leaf) {
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
      if (XYZ_DESCENT_BUDGET != 0 && --budget == 0) break;  // st.node is loaded next step
      w = rec_at(s.nodes, st.node);
      if (kCount) st.nn++;
      leaf = (w.x & 3u) == 3u;
    }
    if (leaf) {
      st.lc0 = w.y;
      st.ln = w.x >> 2;
      st.lj = st.lc0;
      st.lend = st.lc0 + ((st.ln + kClusterSize - 1) >> kClusterShift);
      st.in_leaf = true;
      st.hit = -1;
      st.key = ~0ull;
      if (kCount) st.nt += st.ln;  // the reference tests every triangle of the leaf
    }
  }
  bool const nan_bounds = (st.t_min != st.t_min) || (st.t_max != st.t_max);
  bool const coop = active && st.in_leaf && !nan_bounds;
  // this step's clusters of the lane's leaf, and which of them may hold a hit
  uint32_t const left = st.lend - st.lj;
  uint32_t const ncl = coop ? (left < kClustersPerStep ? left : kClustersPerStep) : 0u;
  unsigned long long list = 0;
  uint32_t nsurv = 0, c = 0;
#if XYZ_COOP_BOX
  {
    // the (lane, cluster) pairs are spread over the wave like the triangle
    // pairs below; a cluster that may hold a hit sets its bit in the owner's
    // LDS word (the first half of the owner's key slot)
    uint32_t bexcl, btotal;
    wave_prefix(ncl, &bexcl, &btotal);
    float dlen;
    float3 inv;
    cull_ray(st.d, &dlen, &inv);
    uint32_t* const masks = reinterpret_cast<uint32_t*>(keys);
    masks[2 * lane] = 0u;
    own[kOwnStride * lane] = 0xFFu;
    wave_sync();
    for (uint32_t base = 0; base < btotal; base += 64) {
      uint32_t const k = base + lane;
      uint32_t const owner = batch_owner<kOwnStride>(own, bexcl, ncl, base);
      float3 const o = make_float3(__shfl(st.o.x, owner), __shfl(st.o.y, owner), __shfl(st.o.z, owner));
      float3 const iv = make_float3(__shfl(inv.x, owner), __shfl(inv.y, owner), __shfl(inv.z, owner));
      float const dl = __shfl(dlen, owner);
      float const tmin = __shfl(st.t_min, owner);
      float const tmax = __shfl(st.t_max, owner);
      uint32_t const i = k - __shfl(bexcl, owner);
      uint32_t const ci = __shfl(st.lj, owner) + i;
      if (k < btotal) {
        if (cluster_may_hit(rec_at(s.clo, ci), rec_at(s.chi, ci), o, dl, iv, tmin, tmax))
          atomicOr(&masks[2 * owner], 1u << i);
      }
    }
    wave_sync();
    // survivors in ascending order (only bits below ncl are set); every
    // cluster holds kClusterSize records but the leaf's last, which may be
    // partly used
    uint32_t mask = masks[2 * lane];
    nsurv = uint32_t(__popc(mask));
    c = nsurv << kClusterShift;
    uint32_t const last = st.lend - 1u - st.lj;  // the leaf's last cluster, as a bit of this step
    if (ncl > 0 && last < ncl && ((mask >> last) & 1u))
      c -= kClusterSize - (st.ln - ((st.lend - 1u - st.lc0) << kClusterShift));
    for (uint32_t k = 0; mask != 0u; ++k) {
      uint32_t const i = uint32_t(__ffs(mask)) - 1u;
      mask &= mask - 1u;
      list |= static_cast<unsigned long long>(i) << (4u * k);
    }
    wave_sync();
  }
#else
  if (ncl > 0) {
    float dlen;
    float3 inv;
    cull_ray(st.d, &dlen, &inv);
    for (uint32_t i = 0; i < ncl; ++i) {
      uint32_t const ci = st.lj + i;
      if (cluster_may_hit(gp(s.clo)[ci], gp(s.chi)[ci], st.o, dlen, inv, st.t_min, st.t_max)) {
        list |= static_cast<unsigned long long>(i) << (4u * nsurv);
        ++nsurv;
        c += (ci + 1 == st.lend) ? st.ln - ((ci - st.lc0) << kClusterShift) : kClusterSize;
      }
    }
  }
#endif
  if (kCount) {
    st.xb += ncl;
    st.xt += c;
  }
  uint32_t excl, total;
  wave_prefix(c, &excl, &total);
  keys[lane] = coop ? st.key : ~0ull;
  own[kOwnStride * lane] = 0xFFu;
  wave_sync();
  for (uint32_t base = 0; base < total; base += 64) {
    uint32_t const k = base + lane;
    uint32_t const owner = batch_owner<kOwnStride>(own, excl, c, base);
    float3 const o = make_float3(__shfl(st.o.x, owner), __shfl(st.o.y, owner), __shfl(st.o.z, owner));
    float3 const d = make_float3(__shfl(st.d.x, owner), __shfl(st.d.y, owner), __shfl(st.d.z, owner));
    float const tmin = __shfl(st.t_min, owner);
    float const tmax = __shfl(st.t_max, owner);
    uint32_t const r = k - __shfl(excl, owner);
    uint32_t const lj_o = __shfl(st.lj, owner);
    unsigned long long const list_o = __shfl(list, owner);
    uint32_t const ci = lj_o + uint32_t((list_o >> (4u * (r >> kClusterShift))) & 15ull);
    uint32_t const j = (ci << kClusterShift) + (r & (kClusterSize - 1u));
    if (k < total) {
      float t, f, g, det;
      float4 const A = rec_at(s.triA, j), B = rec_at(s.triB, j), C = rec_at(s.triC, j);
      // the tie-break rank, loaded with the record
      uint32_t const rank = rec_at(s.rec_rank, j);
      bool const ok = tri_test(o, d, A, B, C, tmin, tmax, &t, &f, &g, &det);
      if (t != t && !(fabsf(det) < 1e-6f)) {
        atomicMin(&keys[owner], 0ull);
      } else if (ok) {
        unsigned long long const key =
            (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | (0xFFFFFFFFu - rank);
        atomicMin(&keys[owner], key);
      }
    }
  }
  wave_sync();
  if (!active || !st.in_leaf) return false;
  if (coop) {
    unsigned long long const key = keys[lane];
    if (key == 0ull) {
      // a NaN t: the sequential semantics over the whole leaf, from its interval
      st.t_max = interval_t_max(s, st, stk, stride);
      st.hit = -1;
      if (kCount) st.xt += st.ln;
      leaf_sequential(s, st);
    } else {
      st.key = key;
      st.lj += ncl;
      if (key != ~0ull) {
        st.t_max = __uint_as_float(uint32_t(key >> 32));
        if (st.lj == st.lend) {
          uint32_t const rank = 0xFFFFFFFFu - uint32_t(key & 0xFFFFFFFFull);
          uint32_t const j = gp(s.rank_pos)[(size_t(st.lc0) << kClusterShift) + rank];
          st.hit = int32_t(j);
          if constexpr (kFG) {
            float t, f, g, det;
            (void)tri_test(st.o, st.d, gp(s.triA)[j], gp(s.triB)[j], gp(s.triC)[j], st.t_min, st.t_max, &t,
                           &f, &g, &det);
            st.f = f;
            st.g = g;
            st.back = det > 0.0f;
          }
        }
      }
    }
  } else {
    if (kCount) st.xt += st.ln;
    leaf_sequential(s, st);
  }
  if (st.lj < st.lend) return false;
  st.in_leaf = false;
  if (st.hit >= 0) return true;
  if (st.sp == 0) return true;
  // pop: the far child of the top entry, with t_min = that entry's split
  // distance, which t_max holds (no hit changed it: a hit ends the query), and
  // t_max = the split distance of the entry below
  st.sp--;
  uint32_t const p = stk[st.sp * stride];
  st.node = far_child(rec_at(s.nodes, p), st.o, p);
  st.t_min = st.t_max;
  st.t_max = interval_t_max(s, st, stk, stride);
  return false;
}

__device__ __forceinline__ HitRec trace_result(const DevScene& s, const TraceState& st) {
  HitRec h;
  if (st.hit >= 0) {
    h.prim = gp(s.ref_prim)[st.hit];
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

// Closest hit of one lane's ray, for kernels whose lanes trace on their own
// (divergent callers such as subpath construction).  Descent and pops are
// trace_step's; a leaf is evaluated like coop_step's, by one lane: the least
// key (bits(t) << 32) | ~rank over the triangles of the clusters that
// cluster_may_hit keeps, which is the reference's sequential result, with the
// bound tightened to the best t found so far (a later triangle at the same t
// still wins on its rank).  A leaf with NaN bounds or a NaN t is run
// sequentially in the reference's order.
__device__ __forceinline__ HitRec trace_closest_culled(const DevScene& s, float3 o, float3 d,
                                                       uint32_t* stk, uint32_t stride) {
  TraceState st;
  trace_init(st, o, d);
  float dlen;
  float3 inv;
  cull_ray(d, &dlen, &inv);
  for (;;) {
    // descend to a leaf: xyz/geom/kdtree.cpp:82-108
    uint2 w = rec_at(s.nodes, st.node);
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
      w = rec_at(s.nodes, st.node);
    }
    st.lc0 = w.y;
    st.ln = w.x >> 2;
    st.lend = st.lc0 + ((st.ln + kClusterSize - 1) >> kClusterShift);
    st.hit = -1;
    bool seq = (st.t_min != st.t_min) || (st.t_max != st.t_max);
    if (!seq) {
      unsigned long long best = ~0ull;
      float tmax = st.t_max;
      for (uint32_t ci = st.lc0; ci < st.lend && !seq; ++ci) {
        if (!cluster_may_hit(rec_at(s.clo, ci), rec_at(s.chi, ci), st.o, dlen, inv, st.t_min, tmax)) continue;
        uint32_t const n = (ci + 1u == st.lend) ? st.ln - ((ci - st.lc0) << kClusterShift) : kClusterSize;
        for (uint32_t m = 0; m < n; ++m) {
          uint32_t const j = (ci << kClusterShift) + m;
          float t, f, g, det;
          bool const ok = tri_test(st.o, st.d, rec_at(s.triA, j), rec_at(s.triB, j), rec_at(s.triC, j), st.t_min,
                                   tmax, &t, &f, &g, &det);
          if (t != t && !(fabsf(det) < 1e-6f)) {
            seq = true;
            break;
          }
          if (ok) {
            unsigned long long const key = (static_cast<unsigned long long>(__float_as_uint(t)) << 32) |
                                           (0xFFFFFFFFu - rec_at(s.rec_rank, j));
            if (key < best) {
              best = key;
              tmax = t;
              st.hit = int32_t(j);
              st.f = f;
              st.g = g;
              st.back = det > 0.0f;
            }
          }
        }
      }
      if (!seq && st.hit >= 0) st.t_max = tmax;
    }
    if (seq) {
      st.hit = -1;
      leaf_sequential(s, st);
    }
    if (st.hit >= 0 || st.sp == 0) break;
    // pop: the far child of the top entry, with [e, e_below]
    st.sp--;
    uint32_t const p = stk[st.sp * stride];
    st.node = far_child(rec_at(s.nodes, p), st.o, p);
    st.t_min = st.t_max;
    st.t_max = interval_t_max(s, st, stk, stride);
  }
  return trace_result(s, st);
}

// The closest hit a lane traces on its own.
#ifndef XYZ_LANE_CULL
#define XYZ_LANE_CULL 1
#endif
__device__ __forceinline__ HitRec trace_lane(const DevScene& s, float3 o, float3 d, uint32_t* stk,
                                             uint32_t stride) {
  if (XYZ_LANE_CULL) return trace_closest_culled(s, o, d, stk, stride);
  return trace_closest<false>(s, o, d, stk, stride, nullptr, nullptr);
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

  // The same, over work split into kFeedParts contiguous parts [total k / P,
  // total (k + 1) / P) with a counter each (counters[k]).  The wave starts on
  // part `part` (a block's XCD under round-robin dispatch, so that each XCD's
  // L2 holds the scene data of its own rays) and moves to the next part when
  // one is used up; after all parts it is exhausted.  Each index is handed
  // out once; each wave visits each part at most once.
  static constexpr uint32_t kFeedParts = 8;
  uint32_t part, parts_left;
  __device__ __forceinline__ void init_parts(uint32_t first) {
    init();
    part = first % kFeedParts;
    parts_left = kFeedParts;
  }
  __device__ __forceinline__ uint32_t refill_parts(bool need, uint32_t* counters, uint32_t total,
                                                  bool* exhausted) {
    uint32_t const lane = __lane_id();
    unsigned long long const below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t got = 0xFFFFFFFFu;
    for (;;) {
      unsigned long long const m = __ballot(need);
      if (m == 0) break;
      if (next == count) {
        if (parts_left == 0) {
          if (need) *exhausted = true;
          break;
        }
        uint32_t const pt = __builtin_amdgcn_readfirstlane(part);
        uint32_t const lo = uint32_t((uint64_t(total) * pt) / kFeedParts);
        uint32_t const size = uint32_t((uint64_t(total) * (pt + 1)) / kFeedParts) - lo;
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&counters[pt], 64u);
        b = __builtin_amdgcn_readfirstlane(b);
        base = lo + b;
        count = (b >= size) ? 0u : ((size - b < 64u) ? size - b : 64u);
        next = 0;
        if (count == 0) {
          part = (pt + 1) % kFeedParts;
          parts_left = __builtin_amdgcn_readfirstlane(parts_left) - 1u;
          continue;
        }
      }
      uint32_t const avail = count - next;
      uint32_t const rank = uint32_t(__popcll(m & below));
      uint32_t const want = uint32_t(__popcll(m));
      uint32_t const take = want < avail ? want : avail;
      if (need && rank < take) {
        got = base + next + rank;
        need = false;
      }
      next += take;
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

// GPU SAH kd-tree build: the same trees as the host builder (kdtree_build.cpp),
// which follows the reference's split rules (xyz/geom/kdtree.cpp:207-357).
//
// Level-synchronous: all nodes of one depth are processed together.
//   * Events (open at min_k, close at max_k + 1e-6) are sorted once per axis
//     by radix sort on (value, close) and from then on stably partitioned into
//     the children, so each node's events stay sorted.  A node's events are
//     one contiguous segment, with the same segment layout on the three axes.
//   * Sweep: a scan gives every event the number of opens and closes up to it
//     within its node; the last event of each run of equal values (and the
//     first value - 1e-6, the last value + 1e-6) is a candidate whose cost is
//     the reference's expression in binary32.  Each (node, axis) keeps the
//     first strict minimum through a 64-bit atomicMin of (cost, position in
//     the sweep); -0 is canonicalised, so equal costs tie on position.  The
//     axes are then combined in order 0, 1, 2 with the same strict test.
//   * Triangles go left / right / both by the same bound tests; ids and events
//     are partitioned with scans.  Children that are leaves (n <= max_bucket or
//     depth > max_depth) move their ids, in caller order, to a leaf pool.
// The breadth-first node table is laid out in pre-order on the host.  The root
// box is accumulated on the host in triangle order, as the host builder does.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "scene_internal.h"
#include "xyz_internal.h"

namespace xyzhip {
namespace {

constexpr uint32_t kB = 256;
constexpr float kEpsG = 1e-6f;  // xyz/stdafx.hpp:35
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr float kFltMax = 3.40282347e+38f;

__device__ __forceinline__ float dmin(float x, float y) { return (x < y) ? x : y; }  // hi::min
__device__ __forceinline__ float dmax(float x, float y) { return (x < y) ? y : x; }  // hi::max

// Unsigned key that orders like < on non-NaN floats; -0 is mapped to +0.
__device__ __forceinline__ uint32_t order_key(float f) {
  uint32_t const u = __float_as_uint(f + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float order_inv(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

uint32_t blocks(size_t n) { return uint32_t((n + kB - 1) / kB); }

// Triangle::min/max over v0, v0 + e1, v0 + e2 (xyz/geom/triangle.hpp:122-156);
// lo/hi hold one array of n per axis.
__global__ void k_bounds(uint32_t n, const float* __restrict__ v0, const float* __restrict__ v1,
                         const float* __restrict__ v2, float* lo, float* hi) {
  uint32_t const i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; ++k) {
    float const a = v0[3 * size_t(i) + k], b = v1[3 * size_t(i) + k], c = v2[3 * size_t(i) + k];
    float const e1 = b - a, e2 = c - a;
    float const p1 = a + e1, p2 = a + e2;
    lo[size_t(k) * n + i] = dmin(dmin(a, p1), p2);
    hi[size_t(k) * n + i] = dmax(dmax(a, p1), p2);
  }
}

// Root events of one axis: sort keys (value, close) and tt = triangle << 1 | close.
__global__ void k_events(uint32_t n, const float* __restrict__ lo, const float* __restrict__ hi,
                         unsigned long long* keys, uint32_t* tt) {
  uint32_t const i = blockIdx.x * kB + threadIdx.x;
  if (i >= 2 * n) return;
  bool const close = i >= n;
  uint32_t const t = close ? i - n : i;
  float const v = close ? hi[t] + kEpsG : lo[t];
  keys[i] = (static_cast<unsigned long long>(order_key(v)) << 1) | (close ? 1ull : 0ull);
  tt[i] = (t << 1) | (close ? 1u : 0u);
}

__global__ void k_event_values(size_t m, const uint32_t* __restrict__ tt, const float* __restrict__ lo,
                               const float* __restrict__ hi, float* v) {
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  if (i >= m) return;
  uint32_t const e = tt[i];
  v[i] = (e & 1u) ? hi[e >> 1] + kEpsG : lo[e >> 1];
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
  uint32_t const i = blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = i;
}
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  if (i < n) p[i] = v;
}

// Nodes of one level.
struct Level {
  uint32_t* n;      // triangles
  uint32_t* eoff;   // first event (2 n events per axis)
  uint32_t* ioff;   // first id
  float* box;       // 6 per node: min[3], max[3]
  uint32_t* depth;
  uint32_t* bfs;    // breadth-first node index
  int32_t* axis;    // -1: no finite cost, the node is a leaf
  float* value;
};

// Per-event open / close flags (low / high 32 bits) for the counting scan.
__global__ void k_oc_flags(size_t m, const uint32_t* __restrict__ tt, unsigned long long* f) {
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  if (i < m) f[i] = (tt[i] & 1u) ? (1ull << 32) : 1ull;
}

__device__ __forceinline__ unsigned long long cand_key(float cost, uint32_t pos) {
  // the sweep starts from FLT_MAX and replaces on a strict decrease: costs that
  // are NaN or >= FLT_MAX never win
  if (!(cost < kFltMax)) return ~0ull;
  return (static_cast<unsigned long long>(order_key(cost)) << 32) | pos;
}

// Candidates of every node on axis k (xyz/geom/kdtree.cpp:261-312).  A block
// reduces the keys of equal nodes before its atomics: near the root all events
// of a block belong to one node.
__global__ __launch_bounds__(kB) void k_sweep(size_t m, int k, const float* __restrict__ ev,
                                              const uint32_t* __restrict__ enode,
                                              const unsigned long long* __restrict__ scan, Level lv,
                                              unsigned long long* best) {
  __shared__ unsigned long long red_key[kB / 64];
  __shared__ uint32_t red_node[kB / 64];
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  uint32_t a = kNone;
  unsigned long long key = ~0ull;
  if (i < m) {
    a = enode[i];
    uint32_t const s = lv.eoff[a], n = lv.n[a];
    size_t const e = size_t(s) + 2 * size_t(n);
    unsigned long long const base = (s > 0) ? scan[s - 1] : 0ull;
    unsigned long long const c = scan[i] - base;
    uint32_t const opens = uint32_t(c & 0xFFFFFFFFull), closes = uint32_t(c >> 32);
    float mn[3], range[3];
    for (int q = 0; q < 3; ++q) {
      mn[q] = lv.box[6 * size_t(a) + q];
      range[q] = lv.box[6 * size_t(a) + 3 + q] - mn[q];
    }
    float const A = range[(k + 1) % 3] * range[(k + 2) % 3];
    float const B = range[(k + 1) % 3] + range[(k + 2) % 3];
    auto cost_at = [&](float v, uint32_t n_left, uint32_t n_right) {
      float const r = v - mn[k];
      float const sa_left = A + B * r;
      float const sa_right = A + B * (range[k] - r);
      return float(n_left) * sa_left + float(n_right) * sa_right;
    };
    if (i == s) key = cand_key(cost_at(ev[i] - kEpsG, 0u, n), 0u);
    float v = ev[i];
    bool const last = (i + 1 == e);
    if (last || v != ev[i + 1]) {  // the last event of a run of equal values
      if (last) v += kEpsG;
      key = std::min(key, cand_key(cost_at(v, opens, n - closes), uint32_t(i - s) + 1u));
    }
  }
  // events are grouped by node: reduce a wave whose lanes share one node
  uint32_t const a0 = __builtin_amdgcn_readfirstlane(a);
  bool const uniform = __all(a == a0 || a == kNone);
  uint32_t const w = threadIdx.x >> 6;
  if (uniform) {
    for (int off = 32; off > 0; off >>= 1) key = std::min(key, __shfl_xor(key, off));
    if ((threadIdx.x & 63) == 0) {
      red_key[w] = key;
      red_node[w] = a0;
    }
  } else {
    if (key != ~0ull) atomicMin(best + 3 * size_t(a) + k, key);
    if ((threadIdx.x & 63) == 0) red_node[w] = kNone;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t cur = kNone;
    unsigned long long acc = ~0ull;
    for (uint32_t q = 0; q < kB / 64; ++q) {
      uint32_t const nq = red_node[q];
      if (nq == kNone) continue;
      if (nq != cur) {
        if (cur != kNone && acc != ~0ull) atomicMin(best + 3 * size_t(cur) + k, acc);
        cur = nq;
        acc = ~0ull;
      }
      acc = std::min(acc, red_key[q]);
    }
    if (cur != kNone && acc != ~0ull) atomicMin(best + 3 * size_t(cur) + k, acc);
  }
}

// Per node: the first strict minimum over axes 0, 1, 2 (kdtree.cpp:313-327)
// and the split value of the winning candidate.
__global__ void k_decide(uint32_t L, Level lv, const unsigned long long* __restrict__ best,
                         const float* ev0, const float* ev1, const float* ev2) {
  uint32_t const a = blockIdx.x * kB + threadIdx.x;
  if (a >= L) return;
  float min_cost = kFltMax;
  int axis = -1;
  uint32_t pos = 0;
  for (int k = 0; k < 3; ++k) {
    unsigned long long const b = best[3 * size_t(a) + k];
    if (b == ~0ull) continue;
    float const cost = order_inv(uint32_t(b >> 32));
    if (min_cost > cost) {
      min_cost = cost;
      axis = k;
      pos = uint32_t(b & 0xFFFFFFFFull);
    }
  }
  lv.axis[a] = axis;
  if (axis < 0) return;
  const float* ev = axis == 0 ? ev0 : (axis == 1 ? ev1 : ev2);
  uint32_t const s = lv.eoff[a];
  size_t const e = size_t(s) + 2 * size_t(lv.n[a]);
  float v;
  if (pos == 0) {
    v = ev[s] - kEpsG;
  } else {
    size_t const i = size_t(s) + pos - 1;
    v = ev[i];
    if (i + 1 == e) v += kEpsG;
  }
  lv.value[a] = v;
}

// 1 left, 2 right, 3 both (Triangle::Overlap, xyz/geom/triangle.cpp:75-85).
__device__ __forceinline__ uint32_t side_of(const float* lo, const float* hi, uint32_t n_tris,
                                            int axis, float value, uint32_t t) {
  if (hi[size_t(axis) * n_tris + t] < value) return 1u;
  if (value <= lo[size_t(axis) * n_tris + t]) return 2u;
  return 3u;
}

// Left / right flags (low / high 32 bits) of the ids or events of splitting
// nodes; `shift` turns an entry into its triangle.
__global__ void k_side_flags(size_t m, const uint32_t* __restrict__ tri, uint32_t shift,
                             const uint32_t* __restrict__ node, Level lv, const float* lo,
                             const float* hi, uint32_t n_tris, unsigned long long* f) {
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  if (i >= m) return;
  uint32_t const a = node[i];
  int const axis = lv.axis[a];
  unsigned long long v = 0;
  if (axis >= 0) {
    uint32_t const s = side_of(lo, hi, n_tris, axis, lv.value[a], tri[i] >> shift);
    v = ((s & 1u) ? 1ull : 0ull) | ((s & 2u) ? (1ull << 32) : 0ull);
  }
  f[i] = v;
}

// Children of the level: slot 2a (left) and 2a + 1 (right) of node a.
struct Kids {
  uint32_t* n;       // 2 L: child triangle counts
  uint32_t* active;  // 2 L: 1 if the child is split at the next level
  uint32_t* leafn;   // 3 L: ids sent to the leaf pool by each child, then by each node itself
  uint32_t* nsz;     // 2 L: n of active children, 0 otherwise
};

struct BfsTable {
  uint32_t* w0;  // interior: axis; leaf: kNone
  float* split;
  uint32_t* left;
  uint32_t* right;
  uint32_t* leaf_off;
  uint32_t* leaf_n;
};

__global__ void k_children(uint32_t L, Level lv, const unsigned long long* __restrict__ iscan,
                           uint32_t max_bucket, uint32_t max_depth, uint32_t bfs_next, Kids kd,
                           BfsTable bt) {
  uint32_t const a = blockIdx.x * kB + threadIdx.x;
  if (a >= L) return;
  uint32_t const node = lv.bfs[a];
  uint32_t const n = lv.n[a];
  if (lv.axis[a] < 0) {
    for (int side = 0; side < 2; ++side) {
      kd.n[2 * a + side] = 0;
      kd.active[2 * a + side] = 0;
      kd.leafn[2 * a + side] = 0;
      kd.nsz[2 * a + side] = 0;
    }
    kd.leafn[2 * L + a] = n;
    bt.w0[node] = kNone;
    bt.leaf_n[node] = n;
    return;
  }
  uint32_t const s = lv.ioff[a];
  unsigned long long const base = (s > 0) ? iscan[s - 1] : 0ull;
  unsigned long long const c = iscan[s + n - 1] - base;
  uint32_t const cn[2] = {uint32_t(c & 0xFFFFFFFFull), uint32_t(c >> 32)};
  uint32_t const depth = lv.depth[a] + 1;
  kd.leafn[2 * L + a] = 0;
  for (int side = 0; side < 2; ++side) {
    bool const leaf = cn[side] <= max_bucket || depth > max_depth;  // kdtree.cpp:230-238
    kd.n[2 * a + side] = cn[side];
    kd.active[2 * a + side] = leaf ? 0u : 1u;
    kd.leafn[2 * a + side] = leaf ? cn[side] : 0u;
    kd.nsz[2 * a + side] = leaf ? 0u : cn[side];
    if (leaf) {
      uint32_t const child = bfs_next + 2 * a + side;
      bt.w0[child] = kNone;
      bt.leaf_n[child] = cn[side];
    }
  }
  bt.w0[node] = uint32_t(lv.axis[a]);
  bt.split[node] = lv.value[a];
  bt.left[node] = bfs_next + 2 * a;
  bt.right[node] = bfs_next + 2 * a + 1;
}

// Next-level entries of the active children, leaf-pool offsets of the rest.
// Threads 0 .. 2L-1 are child slots, 2L .. 3L-1 the nodes that became leaves.
__global__ void k_next_level(uint32_t L, Level lv, Kids kd, const uint32_t* __restrict__ act_x,
                             const uint32_t* __restrict__ n_x, const uint32_t* __restrict__ leaf_x,
                             uint32_t bfs_next, uint32_t pool_base, Level nx, BfsTable bt,
                             uint32_t* child_slot, uint32_t* child_off) {
  uint32_t const j = blockIdx.x * kB + threadIdx.x;
  if (j < 2 * L) {
    uint32_t const a = j >> 1;
    uint32_t const side = j & 1u;
    if (kd.active[j]) {
      uint32_t const b = act_x[j];
      child_slot[j] = b;
      child_off[j] = n_x[j];
      nx.n[b] = kd.n[j];
      nx.ioff[b] = n_x[j];
      nx.eoff[b] = 2 * n_x[j];
      nx.depth[b] = lv.depth[a] + 1;
      nx.bfs[b] = bfs_next + j;
      float const* pb = lv.box + 6 * size_t(a);
      float box[6] = {pb[0], pb[1], pb[2], pb[3], pb[4], pb[5]};
      int const axis = lv.axis[a];
      if (side == 0) box[3 + axis] = lv.value[a]; else box[axis] = lv.value[a];  // kdtree.cpp:345-354
      for (int q = 0; q < 6; ++q) nx.box[6 * size_t(b) + q] = box[q];
    } else {
      child_slot[j] = kNone;
      child_off[j] = pool_base + leaf_x[j];
      if (lv.axis[a] >= 0) bt.leaf_off[bfs_next + j] = pool_base + leaf_x[j];
    }
  } else if (j < 3 * L) {
    uint32_t const a = j - 2 * L;
    if (lv.axis[a] < 0) bt.leaf_off[lv.bfs[a]] = pool_base + leaf_x[j];
  }
}

// Ids go to the active children or to the leaf pool, order kept.
__global__ void k_part_ids(size_t m, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ node,
                           Level lv, const unsigned long long* __restrict__ iscan,
                           const uint32_t* __restrict__ child_slot, const uint32_t* __restrict__ child_off,
                           const uint32_t* __restrict__ leaf_x, uint32_t pool_base, uint32_t L,
                           const float* lo, const float* hi, uint32_t n_tris, uint32_t* ids_next,
                           uint32_t* node_next, uint32_t* pool) {
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  if (i >= m) return;
  uint32_t const a = node[i];
  uint32_t const t = ids[i];
  uint32_t const s = lv.ioff[a];
  if (lv.axis[a] < 0) {
    pool[pool_base + leaf_x[2 * L + a] + uint32_t(i - s)] = t;
    return;
  }
  unsigned long long const base = (s > 0) ? iscan[s - 1] : 0ull;
  unsigned long long const before = iscan[i] - base;
  uint32_t const sd = side_of(lo, hi, n_tris, lv.axis[a], lv.value[a], t);
  for (uint32_t side = 0; side < 2; ++side) {
    if (!(sd & (1u << side))) continue;
    uint32_t const rank = uint32_t(side == 0 ? (before & 0xFFFFFFFFull) : (before >> 32)) - 1u;
    uint32_t const j = 2 * a + side;
    if (child_slot[j] != kNone) {
      ids_next[child_off[j] + rank] = t;
      node_next[child_off[j] + rank] = child_slot[j];
    } else {
      pool[child_off[j] + rank] = t;
    }
  }
}

// Events of one axis go to the active children, order kept.
__global__ void k_part_events(size_t m, const uint32_t* __restrict__ tt, const float* __restrict__ ev,
                              const uint32_t* __restrict__ node, Level lv,
                              const unsigned long long* __restrict__ escan,
                              const uint32_t* __restrict__ child_slot,
                              const uint32_t* __restrict__ child_off, const float* lo,
                              const float* hi, uint32_t n_tris, uint32_t* tt_next, float* ev_next,
                              uint32_t* node_next) {
  size_t const i = size_t(blockIdx.x) * kB + threadIdx.x;
  if (i >= m) return;
  uint32_t const a = node[i];
  if (lv.axis[a] < 0) return;
  uint32_t const s = lv.eoff[a];
  unsigned long long const base = (s > 0) ? escan[s - 1] : 0ull;
  unsigned long long const before = escan[i] - base;
  uint32_t const e = tt[i];
  uint32_t const sd = side_of(lo, hi, n_tris, lv.axis[a], lv.value[a], e >> 1);
  for (uint32_t side = 0; side < 2; ++side) {
    if (!(sd & (1u << side))) continue;
    uint32_t const j = 2 * a + side;
    if (child_slot[j] == kNone) continue;
    uint32_t const rank = uint32_t(side == 0 ? (before & 0xFFFFFFFFull) : (before >> 32)) - 1u;
    size_t const dst = size_t(2) * child_off[j] + rank;
    tt_next[dst] = e;
    ev_next[dst] = ev[i];
    if (node_next) node_next[dst] = child_slot[j];
  }
}

// ------------------------------------------------------------------ host
// Stream-ordered device memory of one build, from the library pool; buffers
// grow on demand (references multiply where triangles straddle splits).
// Root box of triangles [i0, i1), accumulated in triangle order as the host
// builder does: hi::min keeps the later of two equal values and hi::max the
// earlier, which matters when -0 meets +0.  bad: the first triangle with a
// non-finite coordinate, or -1.
struct BoxPart {
  float mn[3], mx[3];
  int64_t bad;
};

BoxPart box_part(const float* v0, const float* v1, const float* v2, uint32_t i0, uint32_t i1) {
  BoxPart r;
  for (int k = 0; k < 3; ++k) {
    r.mn[k] = std::numeric_limits<float>::infinity();
    r.mx[k] = -std::numeric_limits<float>::infinity();
  }
  r.bad = -1;
  auto fmn = [](float x, float y) { return (x < y) ? x : y; };
  auto fmx = [](float x, float y) { return (x < y) ? y : x; };
  for (uint32_t i = i0; i < i1; ++i) {
    for (int k = 0; k < 3; ++k) {
      float const a = v0[3 * size_t(i) + k], b = v1[3 * size_t(i) + k], c = v2[3 * size_t(i) + k];
      if (!std::isfinite(a) || !std::isfinite(b) || !std::isfinite(c)) {
        r.bad = int64_t(i);
        return r;
      }
      float const p1 = a + (b - a), p2 = a + (c - a);
      r.mn[k] = fmn(fmn(r.mn[k], a), fmn(p1, p2));
      r.mx[k] = fmx(fmx(r.mx[k], a), fmx(p1, p2));
    }
  }
  return r;
}

// Combines the parts in triangle order with the same min / max, which gives
// the sequential accumulation's bits (later equal values win the minimum,
// earlier ones the maximum, in both).
std::string combine_box_parts(std::vector<std::future<BoxPart>>& parts, float bmin[3],
                              float bmax[3]) {
  for (int k = 0; k < 3; ++k) {
    bmin[k] = std::numeric_limits<float>::infinity();
    bmax[k] = -std::numeric_limits<float>::infinity();
  }
  std::string err;
  for (auto& f : parts) {
    BoxPart const r = f.get();
    if (!err.empty()) continue;
    if (r.bad >= 0) {
      err = "non-finite vertex coordinate at triangle " + std::to_string(r.bad);
      continue;
    }
    for (int k = 0; k < 3; ++k) {
      bmin[k] = (bmin[k] < r.mn[k]) ? bmin[k] : r.mn[k];
      bmax[k] = (bmax[k] < r.mx[k]) ? r.mx[k] : bmax[k];
    }
  }
  parts.clear();
  return err;
}

struct Mem {
  hipStream_t st = nullptr;
  hipError_t err = hipSuccess;
  std::vector<void*> live;
  void* get(size_t bytes) {
    void* p = nullptr;
    if (err != hipSuccess) return nullptr;
    err = pool_alloc(&p, std::max<size_t>(bytes, 256), st);
    if (err != hipSuccess) return nullptr;
    live.push_back(p);
    return p;
  }
  void put(void* p) {
    auto it = std::find(live.begin(), live.end(), p);
    if (it == live.end()) return;
    live.erase(it);
    (void)hipFreeAsync(p, st);
  }
  ~Mem() {
    for (void* p : live) (void)hipFreeAsync(p, st);
  }
};

template <typename T>
struct DVec {
  T* p = nullptr;
  size_t cap = 0;
};

// Makes v hold at least `need` elements, keeping the first `keep` (those that
// it holds).
template <typename T>
bool reserve(Mem& mem, DVec<T>& v, size_t need, size_t keep = 0) {
  if (need <= v.cap) return true;
  size_t const cap = std::max(need, v.cap + v.cap / 2);
  T* q = static_cast<T*>(mem.get(cap * sizeof(T)));
  if (!q) return false;
  keep = std::min(keep, v.cap);
  if (keep > 0) {
    mem.err = hipMemcpyAsync(q, v.p, keep * sizeof(T), hipMemcpyDeviceToDevice, mem.st);
    if (mem.err != hipSuccess) return false;
  }
  if (v.p) mem.put(v.p);
  v.p = q;
  v.cap = cap;
  return true;
}

struct LevelBufs {
  DVec<uint32_t> n, eoff, ioff, depth, bfs;
  DVec<float> box, value;
  DVec<int32_t> axis;
  bool reserve_all(Mem& mem, size_t L) {
    return reserve(mem, n, L) && reserve(mem, eoff, L) && reserve(mem, ioff, L) &&
           reserve(mem, depth, L) && reserve(mem, bfs, L) && reserve(mem, box, 6 * L) &&
           reserve(mem, value, L) && reserve(mem, axis, L);
  }
  Level view() const { return Level{n.p, eoff.p, ioff.p, box.p, depth.p, bfs.p, axis.p, value.p}; }
};

struct EventBufs {
  DVec<uint32_t> tt[3];
  DVec<float> ev[3];
  DVec<uint32_t> enode, ids, inode;
  bool reserve_all(Mem& mem, size_t m_events, size_t m_ids) {
    for (int k = 0; k < 3; ++k)
      if (!reserve(mem, tt[k], m_events) || !reserve(mem, ev[k], m_events)) return false;
    return reserve(mem, enode, m_events) && reserve(mem, ids, m_ids) && reserve(mem, inode, m_ids);
  }
};

template <typename In, typename Out>
hipError_t inclusive_sum(Mem& mem, DVec<char>& tmp, In in, Out out, size_t m) {
  size_t bytes = 0;
  hipError_t e = hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, m, mem.st);
  if (e != hipSuccess) return e;
  if (!reserve(mem, tmp, bytes)) return mem.err;
  return hipcub::DeviceScan::InclusiveSum(tmp.p, bytes, in, out, m, mem.st);
}
template <typename In, typename Out>
hipError_t exclusive_sum(Mem& mem, DVec<char>& tmp, In in, Out out, size_t m) {
  size_t bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, m, mem.st);
  if (e != hipSuccess) return e;
  if (!reserve(mem, tmp, bytes)) return mem.err;
  return hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, m, mem.st);
}

template <typename T>
hipError_t download(std::vector<T>* h, const T* d, size_t n, hipStream_t st) {
  h->resize(n);
  if (n == 0) return hipSuccess;
  return hipMemcpyAsync(h->data(), d, n * sizeof(T), hipMemcpyDeviceToHost, st);
}

std::string fail(const char* what, hipError_t e) {
  return std::string("device kd build, ") + what + ": " + hipGetErrorString(e);
}

}  // namespace

std::string build_flat_tree_device(uint32_t n, const float* v0, const float* v1, const float* v2,
                                   uint32_t max_bucket, uint32_t max_depth, FlatTree* out) {
  auto t0 = std::chrono::steady_clock::now();
  if (max_bucket == 0) max_bucket = 16;
  if (max_depth == 0) max_depth = 16;
  if (max_depth > 60) return "kd_max_depth must be <= 60";
  if (n >= (1u << 29)) return "too many triangles (limit 2^29)";
  if (n > 0 && (!v0 || !v1 || !v2)) return "NULL vertex array";
  // validation and the root box on host threads, overlapping the uploads and
  // the root event sorts below
  uint32_t const parts = (n >= 65536) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
  std::vector<std::future<BoxPart>> box_parts;
  for (uint32_t q = 0; q < parts; ++q) {
    uint32_t const i0 = uint32_t(uint64_t(n) * q / parts), i1 = uint32_t(uint64_t(n) * (q + 1) / parts);
    box_parts.push_back(std::async(std::launch::async, box_part, v0, v1, v2, i0, i1));
  }
  float bmin[3], bmax[3];
  auto root_box = [&]() -> std::string {
    std::string const err = combine_box_parts(box_parts, bmin, bmax);
    if (!err.empty()) return err;
    for (int k = 0; k < 3; ++k) bmax[k] = bmax[k] + (bmax[k] - bmin[k]) * 1e-5f;
    for (int k = 0; k < 3; ++k) {
      out->bmin[k] = bmin[k];
      out->bmax[k] = bmax[k];
    }
    return std::string();
  };
  out->nodes.clear();
  out->leaf_prims.clear();
  out->max_depth_param = max_depth;
  out->depth = 0;
  if (n <= max_bucket) {  // a single leaf
    std::string const err = root_box();
    if (!err.empty()) return err;
    out->nodes = {(n << 2) | 3u, 0u};
    out->leaf_prims.resize(n);
    for (uint32_t i = 0; i < n; ++i) out->leaf_prims[i] = i;
    out->max_leaf = n;
    out->n_leaves = 1;
    out->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return std::string();
  }

  Mem mem;
  hipStream_t const st = mem.st;
  hipError_t e = hipSuccess;
  float* dv[3];
  for (int q = 0; q < 3; ++q) dv[q] = static_cast<float*>(mem.get(3 * size_t(n) * sizeof(float)));
  float* lo = static_cast<float*>(mem.get(3 * size_t(n) * sizeof(float)));
  float* hi = static_cast<float*>(mem.get(3 * size_t(n) * sizeof(float)));
  if (mem.err != hipSuccess) return fail("allocation", mem.err);
  const float* hv[3] = {v0, v1, v2};
  for (int q = 0; q < 3 && e == hipSuccess; ++q)
    e = hipMemcpyAsync(dv[q], hv[q], 3 * size_t(n) * sizeof(float), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return fail("upload", e);
  hipLaunchKernelGGL(k_bounds, dim3(blocks(n)), dim3(kB), 0, st, n, dv[0], dv[1], dv[2], lo, hi);
  for (int q = 0; q < 3; ++q) mem.put(dv[q]);

  // root events, sorted per axis by (value, close)
  size_t const m0 = 2 * size_t(n);
  EventBufs eb[2];
  if (!eb[0].reserve_all(mem, m0, n)) return fail("allocation", mem.err);
  DVec<char> tmp;
  {
    auto* keys = static_cast<unsigned long long*>(mem.get(m0 * 8));
    auto* keys2 = static_cast<unsigned long long*>(mem.get(m0 * 8));
    auto* tt_in = static_cast<uint32_t*>(mem.get(m0 * 4));
    if (mem.err != hipSuccess) return fail("allocation", mem.err);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
      hipLaunchKernelGGL(k_events, dim3(blocks(m0)), dim3(kB), 0, st, n, lo + size_t(k) * n,
                         hi + size_t(k) * n, keys, tt_in);
      size_t bytes = 0;
      e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, keys2, tt_in, eb[0].tt[k].p, m0, 0,
                                             33, st);
      if (e == hipSuccess && !reserve(mem, tmp, bytes)) e = mem.err;
      if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, keys, keys2, tt_in, eb[0].tt[k].p, m0, 0,
                                               33, st);
      hipLaunchKernelGGL(k_event_values, dim3(blocks(m0)), dim3(kB), 0, st, m0, eb[0].tt[k].p,
                         lo + size_t(k) * n, hi + size_t(k) * n, eb[0].ev[k].p);
    }
    mem.put(keys);
    mem.put(keys2);
    mem.put(tt_in);
    if (e != hipSuccess) return fail("event sort", e);
  }
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(m0)), dim3(kB), 0, st, eb[0].enode.p, m0, 0u);
  hipLaunchKernelGGL(k_iota, dim3(blocks(n)), dim3(kB), 0, st, eb[0].ids.p, n);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(n)), dim3(kB), 0, st, eb[0].inode.p, size_t(n), 0u);

  LevelBufs lb[2];
  if (!lb[0].reserve_all(mem, 1)) return fail("allocation", mem.err);
  {
    std::string const err = root_box();
    if (!err.empty()) {
      (void)hipStreamSynchronize(st);  // the uploads above read the caller's arrays
      return err;
    }
    uint32_t const root[5] = {n, 0u, 0u, 0u, 0u};  // n, eoff, ioff, depth, bfs
    float const box[6] = {bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2]};
    e = hipMemcpyAsync(lb[0].n.p, &root[0], 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(lb[0].eoff.p, &root[1], 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(lb[0].ioff.p, &root[2], 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(lb[0].depth.p, &root[3], 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(lb[0].bfs.p, &root[4], 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(lb[0].box.p, box, sizeof(box), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail("root", e);
  }
  DVec<uint32_t> w0, left, right, leaf_off, leaf_n, pool;
  DVec<float> split;
  DVec<unsigned long long> scan_a, scan_b, iscan, best;
  DVec<uint32_t> kn, kact, kleafn, knsz, act_x, n_x, leaf_x, child_slot, child_off;

  uint32_t L = 1;
  size_t mi = n, m = m0;  // ids and events per axis of the level
  uint32_t bfs_next = 1, pool_used = 0;
  int cur = 0;
  while (L > 0) {
    int const nxt = cur ^ 1;
    EventBufs& ec = eb[cur];
    Level const lc = lb[cur].view();
    size_t const L2 = 2 * size_t(L), L3 = 3 * size_t(L), nb = bfs_next + L2;
    if (nb >= (size_t(1) << 30)) return "device kd build: too many nodes";
    if (!reserve(mem, scan_a, m) || !reserve(mem, scan_b, m) || !reserve(mem, iscan, mi) ||
        !reserve(mem, best, L3) || !reserve(mem, kn, L2) || !reserve(mem, kact, L2) ||
        !reserve(mem, kleafn, L3) || !reserve(mem, knsz, L2) || !reserve(mem, act_x, L2) ||
        !reserve(mem, n_x, L2) || !reserve(mem, leaf_x, L3) || !reserve(mem, child_slot, L2) ||
        !reserve(mem, child_off, L2) || !reserve(mem, w0, nb, bfs_next) ||
        !reserve(mem, split, nb, bfs_next) || !reserve(mem, left, nb, bfs_next) ||
        !reserve(mem, right, nb, bfs_next) || !reserve(mem, leaf_off, nb, bfs_next) ||
        !reserve(mem, leaf_n, nb, bfs_next))
      return fail("allocation", mem.err);
    BfsTable const bt{w0.p, split.p, left.p, right.p, leaf_off.p, leaf_n.p};
    Kids const kd{kn.p, kact.p, kleafn.p, knsz.p};

    // SAH sweep on the three axes, then the decision per node
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks(L3)), dim3(kB), 0, st, best.p, L3, ~0ull);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
      hipLaunchKernelGGL(k_oc_flags, dim3(blocks(m)), dim3(kB), 0, st, m, ec.tt[k].p, scan_a.p);
      e = inclusive_sum(mem, tmp, scan_a.p, scan_b.p, m);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_sweep, dim3(blocks(m)), dim3(kB), 0, st, m, k, ec.ev[k].p, ec.enode.p,
                           scan_b.p, lc, best.p);
    }
    if (e != hipSuccess) return fail("sweep", e);
    hipLaunchKernelGGL(k_decide, dim3(blocks(L)), dim3(kB), 0, st, L, lc, best.p, ec.ev[0].p,
                       ec.ev[1].p, ec.ev[2].p);

    // child sizes and the next level's layout
    hipLaunchKernelGGL(k_side_flags, dim3(blocks(mi)), dim3(kB), 0, st, mi, ec.ids.p, 0u, ec.inode.p,
                       lc, lo, hi, n, scan_a.p);
    e = inclusive_sum(mem, tmp, scan_a.p, iscan.p, mi);
    if (e == hipSuccess)
      hipLaunchKernelGGL(k_children, dim3(blocks(L)), dim3(kB), 0, st, L, lc, iscan.p, max_bucket,
                         max_depth, bfs_next, kd, bt);
    if (e == hipSuccess) e = exclusive_sum(mem, tmp, kact.p, act_x.p, L2);
    if (e == hipSuccess) e = exclusive_sum(mem, tmp, kleafn.p, leaf_x.p, L3);
    if (e == hipSuccess) e = exclusive_sum(mem, tmp, knsz.p, n_x.p, L2);
    if (e != hipSuccess) return fail("child layout", e);
    uint32_t tail[6];
    e = hipMemcpyAsync(&tail[0], act_x.p + L2 - 1, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&tail[1], kact.p + L2 - 1, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&tail[2], n_x.p + L2 - 1, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&tail[3], knsz.p + L2 - 1, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&tail[4], leaf_x.p + L3 - 1, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&tail[5], kleafn.p + L3 - 1, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail("level totals", e);
    uint32_t const L_next = tail[0] + tail[1];
    size_t const mi_next = size_t(tail[2]) + tail[3];
    size_t const leaf_total = size_t(tail[4]) + tail[5];
    if (size_t(pool_used) + leaf_total >= (size_t(1) << 32) || 2 * mi_next >= (size_t(1) << 32))
      return "device kd build: more than 2^32 references";
    if (!lb[nxt].reserve_all(mem, std::max<size_t>(L_next, 1)) ||
        !eb[nxt].reserve_all(mem, std::max<size_t>(2 * mi_next, 1), std::max<size_t>(mi_next, 1)) ||
        !reserve(mem, pool, std::max<size_t>(pool_used + leaf_total, 1), pool_used))
      return fail("allocation", mem.err);
    Level const ln = lb[nxt].view();
    EventBufs& en = eb[nxt];

    // partition ids and events
    hipLaunchKernelGGL(k_next_level, dim3(blocks(L3)), dim3(kB), 0, st, L, lc, kd, act_x.p, n_x.p,
                       leaf_x.p, bfs_next, pool_used, ln, bt, child_slot.p, child_off.p);
    hipLaunchKernelGGL(k_part_ids, dim3(blocks(mi)), dim3(kB), 0, st, mi, ec.ids.p, ec.inode.p, lc,
                       iscan.p, child_slot.p, child_off.p, leaf_x.p, pool_used, L, lo, hi, n, en.ids.p,
                       en.inode.p, pool.p);
    for (int k = 0; k < 3 && e == hipSuccess && mi_next > 0; ++k) {
      hipLaunchKernelGGL(k_side_flags, dim3(blocks(m)), dim3(kB), 0, st, m, ec.tt[k].p, 1u, ec.enode.p,
                         lc, lo, hi, n, scan_a.p);
      e = inclusive_sum(mem, tmp, scan_a.p, scan_b.p, m);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_part_events, dim3(blocks(m)), dim3(kB), 0, st, m, ec.tt[k].p, ec.ev[k].p,
                           ec.enode.p, lc, scan_b.p, child_slot.p, child_off.p, lo, hi, n, en.tt[k].p,
                           en.ev[k].p, k == 0 ? en.enode.p : nullptr);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail("partition", e);
    pool_used += uint32_t(leaf_total);
    bfs_next += uint32_t(L2);
    L = L_next;
    mi = mi_next;
    m = 2 * mi_next;
    cur = nxt;
  }

  // pre-order layout of the breadth-first table
  std::vector<uint32_t> hw0, hleft, hright, hloff, hln, hpool;
  std::vector<float> hsplit;
  e = download(&hw0, w0.p, bfs_next, st);
  if (e == hipSuccess) e = download(&hleft, left.p, bfs_next, st);
  if (e == hipSuccess) e = download(&hright, right.p, bfs_next, st);
  if (e == hipSuccess) e = download(&hloff, leaf_off.p, bfs_next, st);
  if (e == hipSuccess) e = download(&hln, leaf_n.p, bfs_next, st);
  if (e == hipSuccess) e = download(&hsplit, split.p, bfs_next, st);
  if (e == hipSuccess) e = download(&hpool, static_cast<const uint32_t*>(pool.p), pool_used, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail("download", e);
  // nodes reachable from the root, parents before children
  std::vector<uint32_t> order;
  order.reserve(bfs_next);
  order.push_back(0);
  for (size_t q = 0; q < order.size(); ++q) {
    uint32_t const b = order[q];
    if (hw0[b] != kNone) {
      order.push_back(hleft[b]);
      order.push_back(hright[b]);
    }
  }
  std::vector<uint32_t> size(bfs_next, 0), height(bfs_next, 0), pos(bfs_next, 0);
  for (size_t q = order.size(); q-- > 0;) {
    uint32_t const b = order[q];
    if (hw0[b] != kNone) {
      size[b] = 1 + size[hleft[b]] + size[hright[b]];
      height[b] = 1 + std::max(height[hleft[b]], height[hright[b]]);
    } else {
      size[b] = 1;
    }
  }
  for (uint32_t b : order) {  // left child at +1, right after the left subtree
    if (hw0[b] == kNone) continue;
    pos[hleft[b]] = pos[b] + 1;
    pos[hright[b]] = pos[b] + 1 + size[hleft[b]];
  }
  size_t const nn = order.size();
  std::vector<uint32_t> by_pos(nn);
  for (uint32_t b : order) by_pos[pos[b]] = b;
  out->nodes.assign(2 * nn, 0u);
  out->leaf_prims.reserve(pool_used);
  uint32_t max_leaf = 0;
  uint64_t leaves = 0;
  for (size_t p = 0; p < nn; ++p) {
    uint32_t const b = by_pos[p];
    if (hw0[b] == kNone) {
      out->nodes[2 * p] = (hln[b] << 2) | 3u;
      out->nodes[2 * p + 1] = uint32_t(out->leaf_prims.size());
      out->leaf_prims.insert(out->leaf_prims.end(), hpool.begin() + hloff[b],
                             hpool.begin() + hloff[b] + hln[b]);
      max_leaf = std::max(max_leaf, hln[b]);
      ++leaves;
    } else {
      uint32_t bits;
      std::memcpy(&bits, &hsplit[b], 4);
      out->nodes[2 * p] = (pos[hright[b]] << 2) | hw0[b];
      out->nodes[2 * p + 1] = bits;
    }
  }
  out->depth = height[0];
  out->max_leaf = max_leaf;
  out->n_leaves = leaves;
  out->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return std::string();
}

}  // namespace xyzhip

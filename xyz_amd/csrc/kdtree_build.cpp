// Host SAH kd-tree builder producing the flat device layout.
//
// Split decisions are the reference's (xyz/geom/kdtree.cpp:207-357), evaluated
// in the same IEEE binary32 operation order:
//   * root box from Triangle::min/max accumulation (xyz/geom/triangle.hpp:122-140)
//     with max += (max - min) * 1e-5 (kdtree.cpp:215-216);
//   * leaf when n <= max_bucket or depth > max_depth (kdtree.cpp:230-238);
//   * per axis k = 0,1,2: open events at min_k(tri), close events at
//     max_k(tri) + 1e-6, ordered by (value, open-before-close) (kdtree.cpp:251-282);
//     candidates: first value - 1e-6, every last-of-equal value, last value + 1e-6;
//     cost = n_left*(a + b*r) + n_right*(a + b*(range_k - r)), the first strict
//     minimum wins (kdtree.cpp:261-312);
//   * triangles go left / right / both by Triangle::Overlap (triangle.cpp:75-85);
//     children get the clipped boxes (kdtree.cpp:345-354).
// What differs is only the bookkeeping: events are sorted once per axis at the
// root and stably partitioned into the children (a child's event list is the
// parent's, filtered, because triangle boxes are never clipped), so a node costs
// O(n) instead of O(n log n), and large subtrees are built on worker threads.
// Leaves keep the caller order of their triangles, as the reference's
// set_left/set_right vectors do (kdtree.cpp:329-339).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

#include "xyz_internal.h"

namespace xyzhip {
namespace {

const float kEps = 1e-6f;  // xyz/stdafx.hpp:35

inline float fmin_ref(float x, float y) { return (x < y) ? x : y; }  // hi::min
inline float fmax_ref(float x, float y) { return (x < y) ? y : x; }  // hi::max

// Event: value and (triangle << 1 | close).
struct Ev {
  float v;
  uint32_t tt;
};
inline bool ev_less(Ev const& a, Ev const& b) {
  return (a.v < b.v) ? true : (a.v > b.v) ? false : (a.tt & 1u) < (b.tt & 1u);
}

struct Ctx {
  const float* lo[3];
  const float* hi[3];
  uint32_t max_bucket;
  uint32_t max_depth;
  size_t n_tris;
  std::atomic<int> workers{0};
  int max_workers = 1;
};

struct Sub {
  std::vector<uint32_t> nodes;  // 2 words per node; child/leaf offsets local
  std::vector<uint32_t> refs;
  uint32_t depth = 0;
  uint32_t max_leaf = 0;
  uint64_t leaves = 0;
};

void append(Sub& dst, Sub&& src) {
  uint32_t const node_base = uint32_t(dst.nodes.size() / 2);
  uint32_t const ref_base = uint32_t(dst.refs.size());
  for (size_t i = 0; i < src.nodes.size(); i += 2) {
    uint32_t w0 = src.nodes[i], w1 = src.nodes[i + 1];
    if ((w0 & 3u) == 3u) {
      w1 += ref_base;
    } else {
      w0 = (((w0 >> 2) + node_base) << 2) | (w0 & 3u);
    }
    dst.nodes.push_back(w0);
    dst.nodes.push_back(w1);
  }
  dst.refs.insert(dst.refs.end(), src.refs.begin(), src.refs.end());
  dst.depth = std::max(dst.depth, src.depth);
  dst.max_leaf = std::max(dst.max_leaf, src.max_leaf);
  dst.leaves += src.leaves;
}

struct AxisBest {
  float cost;
  float value;
  bool found;
};

// One axis of the SAH sweep, kdtree.cpp:261-312 with a local first-minimum.
AxisBest sweep(std::vector<Ev> const& E, size_t n, int k, float const mn[3], float const range[3]) {
  AxisBest best{std::numeric_limits<float>::max(), 0.0f, false};
  float const a = range[(k + 1) % 3] * range[(k + 2) % 3];
  float const b = range[(k + 1) % 3] + range[(k + 2) % 3];
  size_t n_left = 0, n_right = n;
  auto update = [&](float v) {
    float const r = v - mn[k];
    float const sa_left = a + b * r;
    float const sa_right = a + b * (range[k] - r);
    float const cost = float(n_left) * sa_left + float(n_right) * sa_right;
    if (best.cost > cost) {
      best.cost = cost;
      best.value = v;
      best.found = true;
    }
  };
  update(E[0].v - kEps);
  size_t const size = E.size();
  for (size_t i = 0; i < size; ++i) {
    if (E[i].tt & 1u) n_right--; else n_left++;
    float v = E[i].v;
    if ((i + 1) < size) {
      if (v == E[i + 1].v) continue;
    } else {
      v += kEps;
    }
    update(v);
  }
  return best;
}

Sub build(Ctx& cx, std::vector<uint32_t>&& ids, std::vector<Ev> (&ev)[3], float const mn[3],
          float const mx[3], uint32_t depth) {
  Sub out;
  size_t const n = ids.size();
  if (n <= cx.max_bucket || depth > cx.max_depth) {
    out.nodes.push_back((uint32_t(n) << 2) | 3u);
    out.nodes.push_back(0);
    out.refs = std::move(ids);
    out.max_leaf = uint32_t(n);
    out.leaves = 1;
    out.depth = 0;
    for (auto& e : ev) { std::vector<Ev>().swap(e); }
    return out;
  }
  depth++;
  float range[3];
  for (int k = 0; k < 3; ++k) range[k] = mx[k] - mn[k];

  AxisBest ab[3];
  if (n > 200000) {
    std::future<AxisBest> f1 = std::async(std::launch::async, [&] { return sweep(ev[1], n, 1, mn, range); });
    std::future<AxisBest> f2 = std::async(std::launch::async, [&] { return sweep(ev[2], n, 2, mn, range); });
    ab[0] = sweep(ev[0], n, 0, mn, range);
    ab[1] = f1.get();
    ab[2] = f2.get();
  } else {
    for (int k = 0; k < 3; ++k) ab[k] = sweep(ev[k], n, k, mn, range);
  }
  // sequential combination over axes preserves the reference's first minimum
  float min_cost = std::numeric_limits<float>::max();
  int axis = -1;
  float value = 0.0f;
  for (int k = 0; k < 3; ++k) {
    if (ab[k].found && min_cost > ab[k].cost) {
      min_cost = ab[k].cost;
      axis = k;
      value = ab[k].value;
    }
  }
  if (axis < 0) {
    // No finite cost (the reference would index with axis ~0): make a leaf.
    out.nodes.push_back((uint32_t(n) << 2) | 3u);
    out.nodes.push_back(0);
    out.refs = std::move(ids);
    out.max_leaf = uint32_t(n);
    out.leaves = 1;
    for (auto& e : ev) { std::vector<Ev>().swap(e); }
    return out;
  }

  // classify: 1 left, 2 right, 3 both (Triangle::Overlap).  The side table is
  // indexed by triangle id and private to the building thread; it is consumed
  // before this thread recurses.
  thread_local std::vector<uint8_t> side;
  if (side.size() < cx.n_tris) side.resize(cx.n_tris);
  std::vector<uint32_t> left_ids, right_ids;
  left_ids.reserve(n);
  right_ids.reserve(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t const id = ids[i];
    uint8_t s;
    if (cx.hi[axis][id] < value) s = 1;
    else if (value <= cx.lo[axis][id]) s = 2;
    else s = 3;
    side[id] = s;
    if (s & 1) left_ids.push_back(id);
    if (s & 2) right_ids.push_back(id);
  }
  uint8_t const* side_tab = side.data();
  std::vector<Ev> lev[3], rev[3];
  auto split_axis = [&](int k) {
    std::vector<Ev>& E = ev[k];
    lev[k].reserve(left_ids.size() * 2);
    rev[k].reserve(right_ids.size() * 2);
    for (Ev const& e : E) {
      uint8_t const s = side_tab[e.tt >> 1];
      if (s & 1) lev[k].push_back(e);
      if (s & 2) rev[k].push_back(e);
    }
    std::vector<Ev>().swap(E);
  };
  if (n > 200000) {
    std::future<void> f1 = std::async(std::launch::async, [&] { split_axis(1); });
    std::future<void> f2 = std::async(std::launch::async, [&] { split_axis(2); });
    split_axis(0);
    f1.get();
    f2.get();
  } else {
    for (int k = 0; k < 3; ++k) split_axis(k);
  }
  std::vector<uint32_t>().swap(ids);

  float max_left[3] = {mx[0], mx[1], mx[2]};
  max_left[axis] = value;
  float min_right[3] = {mn[0], mn[1], mn[2]};
  min_right[axis] = value;

  Sub L, R;
  bool spawn = left_ids.size() > 20000 && right_ids.size() > 20000 &&
               cx.workers.fetch_add(1) < cx.max_workers;
  if (spawn) {
    std::future<Sub> fl = std::async(std::launch::async, [&] {
      return build(cx, std::move(left_ids), lev, mn, max_left, depth);
    });
    R = build(cx, std::move(right_ids), rev, min_right, mx, depth);
    L = fl.get();
    cx.workers.fetch_sub(1);
  } else {
    if (left_ids.size() > 20000 && right_ids.size() > 20000) cx.workers.fetch_sub(1);
    L = build(cx, std::move(left_ids), lev, mn, max_left, depth);
    R = build(cx, std::move(right_ids), rev, min_right, mx, depth);
  }

  uint32_t bits;
  std::memcpy(&bits, &value, 4);
  out.nodes.reserve(2 + L.nodes.size() + R.nodes.size());
  out.nodes.push_back(0);  // patched below
  out.nodes.push_back(bits);
  Sub tmp;
  tmp.nodes.swap(out.nodes);
  // children are stored after this node: left at +1, right after the left subtree
  Sub self;
  self.nodes = std::move(tmp.nodes);
  append(self, std::move(L));
  uint32_t const right_index = uint32_t(self.nodes.size() / 2);
  append(self, std::move(R));
  self.nodes[0] = (right_index << 2) | uint32_t(axis);
  self.depth = self.depth + 1;
  return self;
}

}  // namespace

std::string build_flat_tree(uint32_t n, const float* v0, const float* v1, const float* v2,
                            uint32_t max_bucket, uint32_t max_depth, FlatTree* out) {
  auto t0 = std::chrono::steady_clock::now();
  if (max_bucket == 0) max_bucket = 16;
  if (max_depth == 0) max_depth = 16;
  if (max_depth > 60) return "kd_max_depth must be <= 60";
  if (n >= (1u << 29)) return "too many triangles (limit 2^29)";
  if (n > 0 && (!v0 || !v1 || !v2)) return "NULL vertex array";
  std::vector<float> lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k].resize(n);
    hi[k].resize(n);
  }
  float bmin[3], bmax[3];
  for (int k = 0; k < 3; ++k) {
    bmin[k] = std::numeric_limits<float>::infinity();
    bmax[k] = -std::numeric_limits<float>::infinity();
  }
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      float const a = v0[3 * size_t(i) + k], b = v1[3 * size_t(i) + k], c = v2[3 * size_t(i) + k];
      if (!std::isfinite(a) || !std::isfinite(b) || !std::isfinite(c))
        return "non-finite vertex coordinate at triangle " + std::to_string(i);
      // Triangle stores e1 = v1 - v0, e2 = v2 - v0 (triangle.cpp:18-19); its
      // bounds use v0 + e1, v0 + e2 (triangle.hpp:142-156).
      float const e1 = b - a, e2 = c - a;
      float const p1 = a + e1, p2 = a + e2;
      lo[k][i] = fmin_ref(fmin_ref(a, p1), p2);
      hi[k][i] = fmax_ref(fmax_ref(a, p1), p2);
      bmin[k] = fmin_ref(fmin_ref(bmin[k], a), fmin_ref(p1, p2));
      bmax[k] = fmax_ref(fmax_ref(bmax[k], a), fmax_ref(p1, p2));
    }
  }
  for (int k = 0; k < 3; ++k) bmax[k] = bmax[k] + (bmax[k] - bmin[k]) * 1e-5f;

  Ctx cx;
  for (int k = 0; k < 3; ++k) {
    cx.lo[k] = lo[k].data();
    cx.hi[k] = hi[k].data();
  }
  cx.max_bucket = max_bucket;
  cx.max_depth = max_depth;
  cx.n_tris = n;
  cx.max_workers = std::max(1, int(std::thread::hardware_concurrency()) - 1);

  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; ++i) ids[i] = i;
  std::vector<Ev> ev[3];
  if (n > max_bucket) {
    auto make = [&](int k) {
      ev[k].resize(2 * size_t(n));
      for (uint32_t i = 0; i < n; ++i) {
        ev[k][i] = Ev{lo[k][i], i << 1};
        ev[k][size_t(n) + i] = Ev{hi[k][i] + kEps, (i << 1) | 1u};
      }
      std::sort(ev[k].begin(), ev[k].end(), ev_less);
    };
    std::future<void> f1 = std::async(std::launch::async, [&] { make(1); });
    std::future<void> f2 = std::async(std::launch::async, [&] { make(2); });
    make(0);
    f1.get();
    f2.get();
  }
  Sub root = build(cx, std::move(ids), ev, bmin, bmax, 0);
  out->nodes = std::move(root.nodes);
  out->leaf_prims = std::move(root.refs);
  out->depth = root.depth;
  out->max_leaf = root.max_leaf;
  out->n_leaves = root.leaves;
  for (int k = 0; k < 3; ++k) {
    out->bmin[k] = bmin[k];
    out->bmax[k] = bmax[k];
  }
  out->max_depth_param = max_depth;
  out->build_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return std::string();
}

}  // namespace xyzhip

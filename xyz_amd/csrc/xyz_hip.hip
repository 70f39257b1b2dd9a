// C ABI of the gfx950 library: trees, scenes and the closest-hit entry point.
// The path-tracing entry points live in render.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cluster_bounds.h"
#include "device_common.h"
#include "scene_internal.h"
#include "xyz_internal.h"

namespace xyzhip {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return XYZ_E_DEVICE;
}

namespace {
std::mutex g_pool_mu;
std::map<int, hipMemPool_t> g_pools;
}  // namespace

hipError_t pool_alloc(void** p, size_t bytes, hipStream_t stream) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipMemPool_t pool = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    auto it = g_pools.find(dev);
    if (it != g_pools.end()) {
      pool = it->second;
    } else {
      hipMemPoolProps props = {};
      props.allocType = hipMemAllocationTypePinned;
      props.handleTypes = hipMemHandleTypeNone;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = dev;
      e = hipMemPoolCreate(&pool, &props);
      if (e != hipSuccess) return e;
      uint64_t keep = UINT64_MAX;
      e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      if (e != hipSuccess) return e;
      g_pools[dev] = pool;
    }
  }
  return hipMallocFromPoolAsync(p, bytes, pool, stream);
}

namespace {
// Blocks of large workspace (wave buffers, BDPT chunks) cached per device.
// The stream-ordered pool hands a freed block out whole to any later request
// it can hold, so renders whose workspace differs in size from call to call
// (tens of GB at the largest) kept growing it: a second pass over the 64-job
// render sequence left the GPU idle for two thirds of its time.  A block here
// is held by one call at a time; its release is recorded as an event on the
// caller's stream.  A request takes the smallest idle block that holds it and
// was last used on the same stream (stream order makes it free); else a new
// block while the cache stays within its budget, so calls on different
// streams can run at once; else another stream's idle block, after waiting
// for that block's event.
constexpr size_t kBigMin = size_t(32) << 20;
struct BigBlock {
  void* p = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;  // recorded at the last release
  hipStream_t last = nullptr;
  bool held = false;
};
struct BigCache {
  std::vector<BigBlock> blocks;
  size_t bytes = 0;     // held and idle
  size_t quarter = 0;   // a quarter of the device's memory
};
std::mutex g_big_mu;
std::map<int, BigCache> g_big;

// bytes rounded up to a multiple of an eighth of the next power of two
size_t big_round(size_t bytes) {
  size_t pw = size_t(1) << 21;
  while (pw < bytes) pw <<= 1;
  size_t const step = pw >> 3;
  return (bytes + step - 1) / step * step;
}

// The cache's budget: XYZ_WORKSPACE_CACHE_BYTES when set, else a quarter of
// the device's memory.
size_t big_budget(BigCache& c) {
  if (char const* v = std::getenv("XYZ_WORKSPACE_CACHE_BYTES")) return size_t(std::strtoull(v, nullptr, 10));
  if (!c.quarter) {
    size_t fr = 0, tot = 0;
    c.quarter = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 4 : size_t(16) << 30;
  }
  return c.quarter;
}

// Frees the idle blocks of `c` smaller than `below` (waits for their last use).
void big_trim(BigCache& c, size_t below) {
  for (auto it = c.blocks.begin(); it != c.blocks.end();) {
    if (it->held || it->bytes >= below) {
      ++it;
      continue;
    }
    (void)hipEventSynchronize(it->done);
    (void)hipFree(it->p);
    (void)hipEventDestroy(it->done);
    c.bytes -= it->bytes;
    it = c.blocks.erase(it);
  }
}
}  // namespace

hipError_t big_alloc(void** p, size_t bytes, hipStream_t stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(stream, &cs);
  if (e != hipSuccess) return e;
  if (bytes < kBigMin || cs != hipStreamCaptureStatusNone) return pool_alloc(p, bytes, stream);
  int dev = 0;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(g_big_mu);
  BigCache& c = g_big[dev];
  BigBlock *same = nullptr, *other = nullptr;
  for (BigBlock& b : c.blocks) {
    if (b.held || b.bytes < bytes) continue;
    BigBlock*& slot = b.last == stream ? same : other;
    if (!slot || b.bytes < slot->bytes) slot = &b;
  }
  size_t const want = big_round(bytes);
  bool const within = c.bytes + want <= big_budget(c);
  BigBlock* const take = same ? same : (!within ? other : nullptr);
  if (take) {
    if (take != same) {
      e = hipStreamWaitEvent(stream, take->done, 0);
      if (e != hipSuccess) return e;
    }
    take->held = true;
    *p = take->p;
    return hipSuccess;
  }
  if (!within) big_trim(c, bytes);
  BigBlock b;
  b.bytes = want;
  e = hipMalloc(&b.p, b.bytes);
  if (e == hipErrorOutOfMemory) {
    // every idle block and the pool's idle memory go back to the driver, then
    // one more try
    (void)hipGetLastError();
    big_trim(c, SIZE_MAX);
    (void)hipDeviceSynchronize();
    {
      std::lock_guard<std::mutex> pl(g_pool_mu);
      auto it = g_pools.find(dev);
      if (it != g_pools.end()) (void)hipMemPoolTrimTo(it->second, 0);
    }
    e = hipMalloc(&b.p, b.bytes);
  }
  if (e != hipSuccess) return e;
  e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipFree(b.p);
    return e;
  }
  b.held = true;
  b.last = stream;
  c.blocks.push_back(b);
  c.bytes += b.bytes;
  *p = b.p;
  return hipSuccess;
}

void big_free(void* p, hipStream_t stream) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lock(g_big_mu);
    for (auto& kv : g_big)
      for (BigBlock& b : kv.second.blocks)
        if (b.p == p && b.held) {
          (void)hipEventRecord(b.done, stream);
          b.last = stream;
          b.held = false;
          return;
        }
  }
  (void)hipFreeAsync(p, stream);
}

// ------------------------------------------------------------ closest hit
constexpr uint32_t kHitBlock = 256;
// Persistent workgroups per CU: 7 is what fits at once (67 VGPRs give 7 waves
// per SIMD; the LDS stacks of 7 workgroups take 150 of the CU's 160 KiB).
#ifndef XYZ_HIT_BLOCKS_PER_CU
#define XYZ_HIT_BLOCKS_PER_CU 7
#endif
constexpr uint32_t kHitBlocksPerCu = XYZ_HIT_BLOCKS_PER_CU;

// Persistent closest-hit kernel: every lane keeps a ray in flight; when its
// query finishes the lane takes the next ray of its wave's batch (batches of 64
// consecutive rays), so a wave's SIMD lanes stay busy whatever the spread of
// per-ray cost.  The rays are split into 8 contiguous parts with a counter
// each; a block starts on part blockIdx.x % 8, the XCD that round-robin
// dispatch gives it, so each XCD's L2 serves neighbouring rays, and moves on to
// the other parts when its own is used up.
// Waves per SIMD the kernel is compiled for: 7, what its LDS allows (7
// workgroups of 4 waves per CU), which keeps it within 72 VGPRs (0: the
// compiler's choice, which took 73 with root_clip and left 6 waves).
#ifndef XYZ_HIT_WAVES_PER_EU
#define XYZ_HIT_WAVES_PER_EU 7
#endif
template <bool kCount, bool kFG>
__global__ void __launch_bounds__(kHitBlock)
#if XYZ_HIT_WAVES_PER_EU
__attribute__((amdgpu_waves_per_eu(XYZ_HIT_WAVES_PER_EU, XYZ_HIT_WAVES_PER_EU)))
#endif
closest_hit_kernel(DevScene s, uint32_t n, const float* __restrict__ org,
                   const float* __restrict__ dir, int32_t* __restrict__ prim,
                   float* __restrict__ tout, float* __restrict__ fout, float* __restrict__ gout,
                   uint8_t* __restrict__ bout, uint32_t* __restrict__ counters,
                   unsigned long long* __restrict__ totals, uint32_t* __restrict__ work) {
  extern __shared__ uint32_t lds_stack[];
  // per wave: 64 keys, and batch_owner's table in the low bytes of the
  // block's second half
  __shared__ unsigned long long keys[2 * kHitBlock];
  uint32_t* const stk = lds_stack + threadIdx.x;
  WaveFeeder feed;
  feed.init_parts(blockIdx.x);
  TraceState st;
  uint32_t r = 0xFFFFFFFFu;
  bool exhausted = false;
  unsigned long long sum[4] = {0, 0, 0, 0};  // counting mode: this lane's work
  for (;;) {
    bool const need = (r == 0xFFFFFFFFu) && !exhausted;
    uint32_t const got = feed.refill_parts(need, work, n, &exhausted);
    if (got != 0xFFFFFFFFu) {
      r = got;
      trace_init(st, make_float3(org[3 * size_t(r)], org[3 * size_t(r) + 1], org[3 * size_t(r) + 2]),
                 make_float3(dir[3 * size_t(r)], dir[3 * size_t(r) + 1], dir[3 * size_t(r) + 2]));
      if (!kCount && XYZ_ROOT_CLIP) root_clip(s, st);
    }
    if (__ballot(r != 0xFFFFFFFFu) == 0) break;
    bool const act = r != 0xFFFFFFFFu;
#if XYZ_COOP
    bool const fin = coop_step<kCount, kFG>(s, st, act, stk, kHitBlock, keys + (threadIdx.x & ~63u),
                                             reinterpret_cast<uint8_t*>(keys + kHitBlock + (threadIdx.x & ~63u)));
#else
    bool const fin = act && trace_step<kCount>(s, st, stk, kHitBlock);
#endif
    if (act && fin) {
      HitRec const h = trace_result(s, st);
      prim[r] = h.prim;
      if (tout) tout[r] = h.t;
      if (fout) fout[r] = h.f;
      if (gout) gout[r] = h.g;
      if (bout) bout[r] = h.back ? 1 : 0;
      if (kCount) {
        if (counters) {
          counters[r] = st.nn;
          counters[size_t(n) + r] = st.nt;
        }
        sum[0] += st.nn;
        sum[1] += st.nt;
        sum[2] += st.xt;
        sum[3] += st.xb;
      }
      r = 0xFFFFFFFFu;
    }
  }
  if (kCount && totals) {
    // one atomic per wave and counter
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      unsigned long long v = sum[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
      if (__lane_id() == 0) atomicAdd(&totals[k], v);
    }
  }
}

// The kernel instantiation for what the caller asked: counting (per-ray
// counters or launch totals) and barycentrics / side (f, g, back).
void launch_hit_kernel(const DevScene& s, uint32_t grid, size_t lds, hipStream_t stream, uint32_t n,
                       const float* org, const float* dir, int32_t* prim, float* t, float* f,
                       float* g, uint8_t* back, uint32_t* counters, unsigned long long* totals,
                       uint32_t* w) {
  bool const cnt = counters || totals;
  bool const fg = f || g || back;
  auto k = cnt ? (fg ? closest_hit_kernel<true, true> : closest_hit_kernel<true, false>)
               : (fg ? closest_hit_kernel<false, true> : closest_hit_kernel<false, false>);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kHitBlock), lds, stream, s, n, org, dir, prim, t, f, g,
                     back, counters, totals, w);
}

// Clears the closest hit's work counters (one per part of the rays).
__global__ void k_zero_u32(uint32_t* p) {
  if (blockIdx.x == 0 && threadIdx.x < WaveFeeder::kFeedParts) p[threadIdx.x] = 0;
}

int g_hit_cu = 0;

int launch_closest_hit(const DevScene& s, uint32_t n, const float* org, const float* dir,
                       int32_t* prim, float* t, float* f, float* g, uint8_t* back,
                       uint32_t* counters, unsigned long long* totals, hipStream_t stream) {
  if (n == 0) return XYZ_OK;
  if (g_hit_cu == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    g_hit_cu = v;
  }
  size_t const lds = size_t(s.stack_cap) * kHitBlock * sizeof(uint32_t);
  uint32_t const need_blocks = (n + kHitBlock - 1) / kHitBlock;
  uint32_t const grid = std::min<uint32_t>(need_blocks, uint32_t(g_hit_cu) * kHitBlocksPerCu);
  void* work = nullptr;
  hipError_t e = pool_alloc(&work, sizeof(uint32_t) * WaveFeeder::kFeedParts, stream);
  if (e != hipSuccess) return hip_fail(e, "pool_alloc(work counters)");
  uint32_t* w = static_cast<uint32_t*>(work);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, w);
  launch_hit_kernel(s, grid, lds, stream, n, org, dir, prim, t, f, g, back, counters, totals, w);
  e = hipGetLastError();
  (void)hipFreeAsync(work, stream);
  if (e != hipSuccess) return hip_fail(e, "closest_hit launch");
  return XYZ_OK;
}

// ------------------------------------------------- allocation-free variant
// The closest hit's work counters (one per part of the rays) in a fixed device
// array instead of pool memory, for callers that cannot allocate.  Each GPU
// holds its own copy; its address comes from hipGetSymbolAddress once per
// device, and host code uses that address, never the array's name.
constexpr uint32_t kScratchWords = 64u * WaveFeeder::kFeedParts;  // 2 KiB
static __device__ __align__(256) uint32_t scratch[kScratchWords];
// Rays per call: the counters are 32-bit and overshoot their part's end by at
// most one batch of 64 per wave of the grid, so 2^31 rays leave them far from
// wrapping.
constexpr uint64_t kNoAllocMaxRays = 1ull << 31;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// The device's copy of `scratch`, looked up once per device (the device is
// current).
int scratch_address(uint32_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("closest_hit_noalloc: device index out of range");
    return XYZ_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return XYZ_OK;
}

int launch_closest_hit_noalloc(const DevScene& s, uint32_t n, const float* org, const float* dir,
                               int32_t* prim, float* t, float* f, float* g, uint8_t* back,
                               uint32_t* counters, hipStream_t stream) {
  if (uint64_t(n) > kNoAllocMaxRays) {
    set_error("xyz_closest_hit_noalloc: n exceeds the maximum of 2^31 rays per call");
    return XYZ_E_UNSUPPORTED;
  }
  uint32_t* w = nullptr;
  int rc = scratch_address(&w);
  if (rc != XYZ_OK) return rc;
  if (n == 0) return XYZ_OK;
  if (g_hit_cu == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    g_hit_cu = v;
  }
  size_t const lds = size_t(s.stack_cap) * kHitBlock * sizeof(uint32_t);
  uint32_t const need_blocks = (n + kHitBlock - 1) / kHitBlock;
  uint32_t const grid = std::min<uint32_t>(need_blocks, uint32_t(g_hit_cu) * kHitBlocksPerCu);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, w);  // a kernel, not a memset
  launch_hit_kernel(s, grid, lds, stream, n, org, dir, prim, t, f, g, back, counters, nullptr, w);
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "closest_hit_noalloc launch");
  return XYZ_OK;
}

// ------------------------------------------------------------ scene upload
namespace {

template <typename T>
int upload(const std::vector<T>& h, T** d, uint64_t* bytes) {
  *d = nullptr;
  if (h.empty()) return XYZ_OK;
  size_t const nb = h.size() * sizeof(T);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(d), nb);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  e = hipMemcpy(*d, h.data(), nb, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
  *bytes += nb;
  return XYZ_OK;
}

template <typename T>
int upload_field(xyz_scene* sc, const std::vector<T>& h, const T*& field, uint64_t* bytes) {
  T* p = nullptr;
  int const rc = upload(h, &p, bytes);
  if (rc != XYZ_OK) return rc;
  if (p) sc->allocs.push_back(p);
  field = p;
  return XYZ_OK;
}

inline float f3x(const float* p, size_t i, int k) { return p[3 * i + size_t(k)]; }

uint32_t spread10(uint32_t x) {  // 10 bits -> every third bit
  x &= 0x3FFu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// Device triangle layout: each leaf's triangles, in Morton order of their
// centroids within the leaf, grouped in clusters of kClusterSize records.
struct Clustered {
  std::vector<float4> lo, hi;           // per cluster: (box lo - K, P), (box hi + K, Q)
  std::vector<float4> A, B, C;          // per record
  std::vector<int32_t> prim;            // per record
  std::vector<uint32_t> rank, rank_pos; // per record / per leaf rank
};

void cluster_leaves(const FlatTree& T, const float* v0, const std::vector<float>& nrm,
                    const std::vector<float>& e1, const std::vector<float>& e2,
                    std::vector<uint2>* nodes, Clustered* out) {
  size_t n_cl = 0;
  for (const uint2& w : *nodes)
    if ((w.x & 3u) == 3u) n_cl += ((w.x >> 2) + kClusterSize - 1) >> kClusterShift;
  size_t const n_rec = n_cl << kClusterShift;
  out->lo.assign(n_cl, make_float4(0, 0, 0, 0));
  out->hi.assign(n_cl, make_float4(0, 0, 0, 0));
  out->A.assign(n_rec, make_float4(0, 0, 0, 0));
  out->B.assign(n_rec, make_float4(0, 0, 0, 0));
  out->C.assign(n_rec, make_float4(0, 0, 0, 0));
  out->prim.assign(n_rec, -1);
  out->rank.assign(n_rec, 0);
  out->rank_pos.assign(n_rec, 0);
  size_t c0 = 0;
  std::vector<std::pair<uint32_t, uint32_t>> keys;
  for (uint2& w : *nodes) {
    if ((w.x & 3u) != 3u) continue;
    uint32_t const n = w.x >> 2, off = w.y;
    // Morton keys of the centroids (v0 + (e1 + e2) / 3) in the leaf's centroid box
    double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto centroid = [&](size_t i, int k) {
      return double(f3x(v0, i, k)) + (double(e1[3 * i + k]) + double(e2[3 * i + k])) / 3.0;
    };
    for (uint32_t r = 0; r < n; ++r) {
      size_t const i = T.leaf_prims[off + r];
      for (int k = 0; k < 3; ++k) {
        double const c = centroid(i, k);
        if (c < cmin[k]) cmin[k] = c;
        if (c > cmax[k]) cmax[k] = c;
      }
    }
    keys.clear();
    for (uint32_t r = 0; r < n; ++r) {
      size_t const i = T.leaf_prims[off + r];
      uint32_t q[3];
      for (int k = 0; k < 3; ++k) {
        double const ext = cmax[k] - cmin[k];
        double const x = (ext > 0.0) ? (centroid(i, k) - cmin[k]) / ext : 0.0;
        q[k] = (x >= 0.0 && x <= 1.0) ? uint32_t(x * 1023.0) : 0u;  // NaN -> 0
      }
      keys.push_back({spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2), r});
    }
    std::stable_sort(keys.begin(), keys.end());
    size_t const p0 = c0 << kClusterShift;
    for (uint32_t m = 0; m < n; ++m) {
      uint32_t const r = keys[m].second;
      size_t const i = T.leaf_prims[off + r];
      size_t const p = p0 + m;
      out->A[p] = make_float4(f3x(v0, i, 0), f3x(v0, i, 1), f3x(v0, i, 2), nrm[3 * i]);
      out->B[p] = make_float4(e1[3 * i], e1[3 * i + 1], e1[3 * i + 2], nrm[3 * i + 1]);
      out->C[p] = make_float4(e2[3 * i], e2[3 * i + 1], e2[3 * i + 2], nrm[3 * i + 2]);
      out->prim[p] = int32_t(i);
      out->rank[p] = r;
      out->rank_pos[p0 + r] = uint32_t(p);
    }
    uint32_t const nc = (n + kClusterSize - 1) >> kClusterShift;
    for (uint32_t c = 0; c < nc; ++c) {
      uint32_t const m0 = c << kClusterShift;
      uint32_t const cnt = std::min<uint32_t>(kClusterSize, n - m0);
      cluster_bounds(&out->A[p0 + m0], &out->B[p0 + m0], &out->C[p0 + m0], cnt, &out->lo[c0 + c],
                     &out->hi[c0 + c]);
    }
    w.y = uint32_t(c0);
    c0 += nc;
  }
}

}  // namespace

void scene_free(xyz_scene* s) {
  for (void* p : s->allocs) {
    if (p) (void)hipFree(p);
  }
  s->allocs.clear();
}

int scene_create(const xyz_scene_desc* desc, xyz_scene** out) {
  if (!desc || !out) {
    set_error("xyz_scene_create: NULL argument");
    return XYZ_E_INVALID;
  }
  *out = nullptr;
  uint32_t const n = desc->n_tris;
  if (n > 0 && (!desc->v0 || !desc->v1 || !desc->v2)) {
    set_error("xyz_scene_create: NULL vertex arrays");
    return XYZ_E_INVALID;
  }
  if (desc->n_bsdfs > 0 && !desc->bsdfs) {
    set_error("xyz_scene_create: NULL bsdf array");
    return XYZ_E_INVALID;
  }
  uint32_t const nb = desc->n_bsdfs;
  for (uint32_t i = 0; desc->bsdf && i < n; ++i) {
    if (desc->bsdf[i] >= nb) {
      set_error("xyz_scene_create: triangle " + std::to_string(i) + " has bsdf index out of range");
      return XYZ_E_INVALID;
    }
  }
  if (!desc->bsdf && n > 0 && nb == 0) {
    set_error("xyz_scene_create: no bsdfs");
    return XYZ_E_INVALID;
  }
  if (desc->kd_max_depth > 30) {
    // the per-lane LDS traversal stack holds kd_max_depth + 1 entries of 8 bytes
    set_error("xyz_scene_create: kd_max_depth above 30 is not supported on the device");
    return XYZ_E_UNSUPPORTED;
  }
  std::unique_ptr<xyz_scene> sc(new xyz_scene());
  if (desc->kd_builder != XYZ_KD_HOST && desc->kd_builder != XYZ_KD_DEVICE) {
    set_error("xyz_scene_create: kd_builder must be XYZ_KD_HOST or XYZ_KD_DEVICE");
    return XYZ_E_INVALID;
  }
  std::string err = desc->kd_builder == XYZ_KD_DEVICE
                        ? build_flat_tree_device(n, desc->v0, desc->v1, desc->v2, desc->kd_max_bucket,
                                                 desc->kd_max_depth, &sc->tree)
                        : build_flat_tree(n, desc->v0, desc->v1, desc->v2, desc->kd_max_bucket,
                                          desc->kd_max_depth, &sc->tree);
  if (!err.empty()) {
    set_error("xyz_scene_create: " + err);
    return XYZ_E_INVALID;
  }
  FlatTree const& T = sc->tree;
  size_t const n_refs = T.leaf_prims.size();

  // triangle precompute, xyz/geom/triangle.cpp:10-60
  std::vector<float> e1(size_t(n) * 3), e2(size_t(n) * 3), nrm(size_t(n) * 3), ng(size_t(n) * 3);
  for (size_t i = 0; i < n; ++i) {
    float a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = f3x(desc->v0, i, k);
      b[k] = f3x(desc->v1, i, k) - a[k];
      c[k] = f3x(desc->v2, i, k) - a[k];
      e1[3 * i + k] = b[k];
      e2[3 * i + k] = c[k];
    }
    // n = cross(e2, e1), libhi/hi/math/vector3.hpp:126-131
    float const nx = c[1] * b[2] - c[2] * b[1];
    float const ny = c[2] * b[0] - c[0] * b[2];
    float const nz = c[0] * b[1] - c[1] * b[0];
    nrm[3 * i] = nx;
    nrm[3 * i + 1] = ny;
    nrm[3 * i + 2] = nz;
    // ng = normalize(n) = n * (1 / sqrt(dot(n, n)))
    float const l2 = nx * nx + ny * ny + nz * nz;
    float const r = 1.0f / std::sqrt(l2);
    ng[3 * i] = nx * r;
    ng[3 * i + 1] = ny * r;
    ng[3 * i + 2] = nz * r;
  }
  std::vector<uint2> nodes(T.nodes.size() / 2);
  for (size_t i = 0; i < nodes.size(); ++i) nodes[i] = make_uint2(T.nodes[2 * i], T.nodes[2 * i + 1]);
  Clustered CL;
  cluster_leaves(T, desc->v0, nrm, e1, e2, &nodes, &CL);
  // the traversal addresses the node and record arrays with 32-bit byte offsets
  if (CL.A.size() * sizeof(float4) > size_t(0xFFFFFFFFu) ||
      nodes.size() * sizeof(uint2) > size_t(0xFFFFFFFFu)) {
    set_error("xyz_scene_create: the kd tree holds more than 2^28 triangle references or 2^29 nodes");
    return XYZ_E_UNSUPPORTED;
  }

  // per-primitive shading records
  std::vector<float4> png(n), pv0(n), pe1(n), pe2(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t const bs = desc->bsdf ? desc->bsdf[i] : 0u;
    float bsf;
    std::memcpy(&bsf, &bs, 4);
    png[i] = make_float4(ng[3 * i], ng[3 * i + 1], ng[3 * i + 2], bsf);
    // the triangle record's A, B, C (n = cross(e2, e1) in w)
    pv0[i] = make_float4(f3x(desc->v0, i, 0), f3x(desc->v0, i, 1), f3x(desc->v0, i, 2), nrm[3 * i]);
    pe1[i] = make_float4(e1[3 * i], e1[3 * i + 1], e1[3 * i + 2], nrm[3 * i + 1]);
    pe2[i] = make_float4(e2[3 * i], e2[3 * i + 1], e2[3 * i + 2], nrm[3 * i + 2]);
  }
  std::vector<float4> pn0, ps1, ps2;
  if (desc->n0 && desc->n1 && desc->n2) {
    pn0.resize(n);
    ps1.resize(n);
    ps2.resize(n);
    for (size_t i = 0; i < n; ++i) {
      float a[3], b[3], c[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = f3x(desc->n0, i, k);
        b[k] = f3x(desc->n1, i, k) - a[k];
        c[k] = f3x(desc->n2, i, k) - a[k];
      }
      pn0[i] = make_float4(a[0], a[1], a[2], 0.0f);
      ps1[i] = make_float4(b[0], b[1], b[2], 0.0f);
      ps2[i] = make_float4(c[0], c[1], c[2], 0.0f);
    }
  }
  std::vector<float2> puv;
  if (desc->uv) {
    puv.resize(size_t(n) * 3);
    for (size_t i = 0; i < size_t(n) * 3; ++i)
      puv[i] = make_float2(desc->uv[2 * i], desc->uv[2 * i + 1]);
  }

  // materials
  std::vector<DevBsdf> bs(nb);
  for (uint32_t i = 0; i < nb; ++i) {
    xyz_bsdf_desc const& d = desc->bsdfs[i];
    if (d.type > XYZ_LIGHT) {
      set_error("xyz_scene_create: bsdf " + std::to_string(i) + " has unknown type");
      return XYZ_E_INVALID;
    }
    if (d.texture >= 0 && uint32_t(d.texture) >= desc->n_textures) {
      set_error("xyz_scene_create: bsdf " + std::to_string(i) + " texture out of range");
      return XYZ_E_INVALID;
    }
    bs[i].type = d.type;
    bs[i].texture = d.texture;
    for (int k = 0; k < 3; ++k) {
      bs[i].rgb[k] = d.rgb[k];
      bs[i].phi[k] = 0.0f;
    }
    bs[i].shininess = d.shininess;
    bs[i].ior = d.ior;
  }
  // lights and their area CDF, caller order (xyz/core/scene_mqo.cpp:509-518)
  std::vector<float> cdf(1, 0.0f);
  std::vector<uint32_t> lprim;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t const b = desc->bsdf ? desc->bsdf[i] : 0u;
    if (nb > 0 && bs[b].type == XYZ_LIGHT) {
      // area = length(cross(e2, e1)) * 0.5 (xyz/geom/triangle.hpp:57-59)
      float const nx = nrm[3 * size_t(i)], ny = nrm[3 * size_t(i) + 1], nz = nrm[3 * size_t(i) + 2];
      float const area = std::sqrt(nx * nx + ny * ny + nz * nz) * 0.5f;
      lprim.push_back(i);
      cdf.push_back(cdf.back() + area);
    }
  }
  float const light_area = cdf.back();
  for (uint32_t i = 0; i < nb; ++i) {
    if (bs[i].type == XYZ_LIGHT) {
      double const s = 3.14159265358979323846 * double(light_area);
      for (int k = 0; k < 3; ++k) bs[i].phi[k] = float(double(bs[i].rgb[k]) * s);
    }
  }
  // textures
  std::vector<DevTex> tex(desc->n_textures);
  std::vector<float> texels;
  {
    size_t off = 0;
    for (uint32_t t = 0; t < desc->n_textures; ++t) {
      uint32_t const w = desc->tex_wh[2 * t], h = desc->tex_wh[2 * t + 1];
      if (w == 0 || h == 0) {
        set_error("xyz_scene_create: empty texture");
        return XYZ_E_INVALID;
      }
      tex[t].w = w;
      tex[t].h = h;
      tex[t].offset = uint32_t(off);
      tex[t].pad = 0;
      off += size_t(w) * h * 3;
    }
    if (off > 0) {
      if (!desc->texels) {
        set_error("xyz_scene_create: NULL texels");
        return XYZ_E_INVALID;
      }
      texels.assign(desc->texels, desc->texels + off);
    }
    sc->n_texels = off;
  }

  // device upload
  int rc = XYZ_OK;
  uint64_t bytes = 0;
  DevScene& D = sc->dev;
  std::memset(&D, 0, sizeof(D));
#define UP(vec, field)                                 \
  do {                                                 \
    rc = upload_field(sc.get(), vec, D.field, &bytes); \
    if (rc != XYZ_OK) {                                \
      scene_free(sc.get());                            \
      return rc;                                       \
    }                                                  \
  } while (0)
  UP(nodes, nodes);
  UP(CL.lo, clo);
  UP(CL.hi, chi);
  UP(CL.A, triA);
  UP(CL.B, triB);
  UP(CL.C, triC);
  UP(CL.prim, ref_prim);
  UP(CL.rank, rec_rank);
  UP(CL.rank_pos, rank_pos);
  UP(png, prim_ng);
  UP(pn0, prim_n0);
  UP(ps1, prim_s1);
  UP(ps2, prim_s2);
  UP(puv, prim_uv);
  UP(pv0, prim_v0);
  UP(pe1, prim_e1);
  UP(pe2, prim_e2);
  UP(cdf, light_cdf);
  UP(lprim, light_prim);
  UP(bs, bsdfs);
  UP(texels, texels);
  UP(tex, tex);
#undef UP
  D.stack_cap = T.max_depth_param + 1;
  {
    // the root record of root_clip
    float const inf = std::numeric_limits<float>::infinity();
    float4 rl = make_float4(inf, inf, inf, 0.0f), rh = make_float4(-inf, -inf, -inf, 0.0f);
    bool nan = CL.lo.empty();
    for (size_t c = 0; c < CL.lo.size(); ++c) {
      float4 const l = CL.lo[c], h = CL.hi[c];
      nan = nan || std::isnan(l.x) || std::isnan(l.y) || std::isnan(l.z) || std::isnan(l.w) ||
            std::isnan(h.x) || std::isnan(h.y) || std::isnan(h.z) || std::isnan(h.w);
      rl = make_float4(std::min(rl.x, l.x), std::min(rl.y, l.y), std::min(rl.z, l.z), std::max(rl.w, l.w));
      rh = make_float4(std::max(rh.x, h.x), std::max(rh.y, h.y), std::max(rh.z, h.z), std::max(rh.w, h.w));
    }
    if (nan) rl.w = std::numeric_limits<float>::quiet_NaN();
    D.root_lo = rl;
    D.root_hi = rh;
  }
  D.n_lights = uint32_t(lprim.size());
  D.light_area = light_area;
  D.n_bsdfs = nb;
  D.n_tex = desc->n_textures;

  xyz_scene_info& I = sc->info;
  I.n_tris = n;
  I.n_nodes = nodes.size();
  I.n_leaves = T.n_leaves;
  I.n_leaf_refs = n_refs;
  I.depth = T.depth;
  I.max_leaf = T.max_leaf;
  for (int k = 0; k < 3; ++k) {
    I.bmin[k] = T.bmin[k];
    I.bmax[k] = T.bmax[k];
  }
  I.build_seconds = T.build_seconds;
  I.device_bytes = bytes;
  sc->n_tris = n;
  // the host copy of the tree is no longer needed
  std::vector<uint32_t>().swap(sc->tree.nodes);
  std::vector<uint32_t>().swap(sc->tree.leaf_prims);
  *out = sc.release();
  return XYZ_OK;
}

namespace {
// Counts the binary32 values x with 1e-6 <= |x| < kRecipFastMax whose fast
// reciprocal differs from the IEEE quotient 1.0f / x; bit patterns
// [begin, end) of the 2^32.
__global__ void k_check_reciprocal(uint64_t begin, uint64_t end, unsigned long long* bad) {
  uint64_t const stride = uint64_t(gridDim.x) * blockDim.x;
  unsigned long long n = 0;
  for (uint64_t i = begin + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < end; i += stride) {
    float const x = __uint_as_float(uint32_t(i));
    float const ax = fabsf(x);
    if (!(ax >= 1e-6f && ax < kRecipFastMax)) continue;
    if (__float_as_uint(recip_fast(x)) != __float_as_uint(1.0f / x)) ++n;
  }
  if (n != 0) atomicAdd(bad, n);
}
}  // namespace

}  // namespace xyzhip

using namespace xyzhip;

extern "C" {

const char* xyz_last_error(void) { return g_last_error.c_str(); }

const char* xyz_version(void) { return "xyz-hip 0.1 (gfx950)"; }

int xyz_kdtree_build(uint32_t n_tris, const float* v0, const float* v1, const float* v2,
                     uint32_t max_bucket, uint32_t max_depth, xyz_kdtree** out) {
  if (!out) {
    set_error("xyz_kdtree_build: NULL out");
    return XYZ_E_INVALID;
  }
  *out = nullptr;
  std::unique_ptr<xyz_kdtree> t(new xyz_kdtree());
  std::string err = build_flat_tree(n_tris, v0, v1, v2, max_bucket, max_depth, &t->t);
  if (!err.empty()) {
    set_error("xyz_kdtree_build: " + err);
    return XYZ_E_INVALID;
  }
  t->n_tris = n_tris;
  *out = t.release();
  return XYZ_OK;
}

int xyz_kdtree_build_device(uint32_t n_tris, const float* v0, const float* v1, const float* v2,
                            uint32_t max_bucket, uint32_t max_depth, xyz_kdtree** out) {
  if (!out) {
    set_error("xyz_kdtree_build_device: NULL out");
    return XYZ_E_INVALID;
  }
  *out = nullptr;
  std::unique_ptr<xyz_kdtree> t(new xyz_kdtree());
  std::string err = build_flat_tree_device(n_tris, v0, v1, v2, max_bucket, max_depth, &t->t);
  if (!err.empty()) {
    set_error("xyz_kdtree_build_device: " + err);
    return err.rfind("device kd build", 0) == 0 ? XYZ_E_DEVICE : XYZ_E_INVALID;
  }
  t->n_tris = n_tris;
  *out = t.release();
  return XYZ_OK;
}

int xyz_kdtree_info(const xyz_kdtree* tree, xyz_scene_info* info) {
  if (!tree || !info) {
    set_error("xyz_kdtree_info: NULL argument");
    return XYZ_E_INVALID;
  }
  std::memset(info, 0, sizeof(*info));
  info->n_tris = tree->n_tris;
  info->n_nodes = tree->t.nodes.size() / 2;
  info->n_leaves = tree->t.n_leaves;
  info->n_leaf_refs = tree->t.leaf_prims.size();
  info->depth = tree->t.depth;
  info->max_leaf = tree->t.max_leaf;
  for (int k = 0; k < 3; ++k) {
    info->bmin[k] = tree->t.bmin[k];
    info->bmax[k] = tree->t.bmax[k];
  }
  info->build_seconds = tree->t.build_seconds;
  return XYZ_OK;
}

int xyz_kdtree_export(const xyz_kdtree* tree, uint32_t* nodes, uint32_t* leaf_prims) {
  if (!tree) {
    set_error("xyz_kdtree_export: NULL tree");
    return XYZ_E_INVALID;
  }
  if (nodes) std::memcpy(nodes, tree->t.nodes.data(), tree->t.nodes.size() * 4);
  if (leaf_prims) std::memcpy(leaf_prims, tree->t.leaf_prims.data(), tree->t.leaf_prims.size() * 4);
  return XYZ_OK;
}

void xyz_kdtree_destroy(xyz_kdtree* tree) { delete tree; }

int xyz_scene_create(const xyz_scene_desc* desc, xyz_scene** out) { return scene_create(desc, out); }

int xyz_scene_get_info(const xyz_scene* scene, xyz_scene_info* info) {
  if (!scene || !info) {
    set_error("xyz_scene_get_info: NULL argument");
    return XYZ_E_INVALID;
  }
  *info = scene->info;
  return XYZ_OK;
}

void xyz_scene_destroy(xyz_scene* scene) {
  if (!scene) return;
  scene_free(scene);
  delete scene;
}

int xyz_closest_hit(xyz_scene* scene, uint32_t n, const float* org, const float* dir,
                    int32_t* prim, float* t, float* f, float* g, uint8_t* back,
                    uint32_t* counters, void* stream) {
  if (!scene) {
    set_error("xyz_closest_hit: NULL scene");
    return XYZ_E_INVALID;
  }
  if (n > 0 && (!org || !dir || !prim)) {
    set_error("xyz_closest_hit: NULL ray or prim buffer");
    return XYZ_E_INVALID;
  }
  return launch_closest_hit(scene->dev, n, org, dir, prim, t, f, g, back, counters, nullptr,
                            static_cast<hipStream_t>(stream));
}

int xyz_closest_hit_noalloc(xyz_scene* scene, uint32_t n, const float* org, const float* dir,
                            int32_t* prim, float* t, float* f, float* g, uint8_t* back,
                            uint32_t* counters, void* stream) {
  if (!scene) {
    set_error("xyz_closest_hit_noalloc: NULL scene");
    return XYZ_E_INVALID;
  }
  if (n > 0 && (!org || !dir || !prim)) {
    set_error("xyz_closest_hit_noalloc: NULL ray or prim buffer");
    return XYZ_E_INVALID;
  }
  return launch_closest_hit_noalloc(scene->dev, n, org, dir, prim, t, f, g, back, counters,
                                    static_cast<hipStream_t>(stream));
}

int xyz_closest_hit_work(xyz_scene* scene, uint32_t n, const float* org, const float* dir,
                         int32_t* prim, float* t, unsigned long long* work, void* stream) {
  if (!scene) {
    set_error("xyz_closest_hit_work: NULL scene");
    return XYZ_E_INVALID;
  }
  if (!work || (n > 0 && (!org || !dir || !prim))) {
    set_error("xyz_closest_hit_work: NULL ray, prim or work buffer");
    return XYZ_E_INVALID;
  }
  return launch_closest_hit(scene->dev, n, org, dir, prim, t, nullptr, nullptr, nullptr, nullptr,
                            work, static_cast<hipStream_t>(stream));
}

int xyz_check_reciprocal(unsigned long long* mismatches, void* stream) {
  if (!mismatches) {
    set_error("xyz_check_reciprocal: NULL result");
    return XYZ_E_INVALID;
  }
  hipStream_t const st = static_cast<hipStream_t>(stream);
  unsigned long long* bad = nullptr;
  hipError_t e = hipMalloc(&bad, sizeof(unsigned long long));
  if (e != hipSuccess) return hip_fail(e, "xyz_check_reciprocal");
  e = hipMemsetAsync(bad, 0, sizeof(unsigned long long), st);
  for (uint64_t b = 0; e == hipSuccess && b < (1ull << 32); b += (1ull << 30)) {
    hipLaunchKernelGGL(k_check_reciprocal, dim3(4096), dim3(256), 0, st, b, b + (1ull << 30), bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(mismatches, bad, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipError_t const ef = hipFree(bad);
  if (e == hipSuccess) e = ef;
  return e == hipSuccess ? XYZ_OK : hip_fail(e, "xyz_check_reciprocal");
}

}  // extern "C"

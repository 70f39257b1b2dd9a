// Path-tracing entry points: wavefront renderer, path-replay adjoint, and the
// persistent batched renderer with its autotuner.
//
// Wavefront (plain form).  Work item q in [0, W*H*spp) is (sample q / WH, pixel
// q % WH).  Items are processed in waves of up to kWaveCap path slots; per
// bounce:
//   extend : closest hit of every active path               (LDS-stack traversal)
//   shade  : emission + MIS, depth cap, NEE request, RR, BSDF sampling; the
//            survivors are compacted into 8 queues by direction octant with
//            wave64 ballots + prefix popcounts (one atomic per wave and octant)
//   shadow : closest hit of every NEE request; visible -> add contribution
//   prep   : next queues become current (one thread)
// and at the end of the wave the slots' colours are added to the film pixel by
// pixel in sample order (the reference's order, integrator.cpp:119-128).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pt_device.h"
#include "scene_internal.h"

namespace xyzhip {

namespace {

constexpr uint32_t kBlock = 256;
#ifndef XYZ_WAVE_CAP_LOG2
#define XYZ_WAVE_CAP_LOG2 25
#endif
constexpr uint64_t kWaveCap = 1ull << XYZ_WAVE_CAP_LOG2;

// A counter alone in its 256-byte line: device-scope atomics on one line are
// serialised, so counters that many workgroups hit at once do not share one.
struct alignas(256) LineU32 {
  uint32_t v;
  uint32_t pad[63];
};

constexpr uint32_t kShadowAppend = 8;  // append[0..7]: octant queues, append[8]: shadow requests

struct WaveCounters {
  LineU32 append[9];
  LineU32 work_extend;  // persistent extend kernel's work counter
  LineU32 work_shadow;  // persistent shadow kernel's work counter
  uint32_t cur[8];
  unsigned long long queries;
};

struct WaveBufs {
  float4* pos;    // pos.xyz, f_next
  float4* dir;    // dir.xyz, bits(k | specular << 31)
  float4* pw;     // power.xyz
  float4* col;    // colour.xyz
  float4* hit;    // bits(prim | back << 31), t, f, g
  float4* ctot;   // adjoint: the slot's primal colour
  float4* so;     // shadow: origin.xyz, nk2
  float4* sw;     // shadow: w.xyz, nk3
  float4* spe;    // shadow: pe.xyz, bits(slot)
  uint4* smeta;   // shadow: expected prim, type, light bsdf, 0
  float4* smet;   // shadow: fsn, fres, mx, 0
  uint32_t* qa;   // 8 * cap
  uint32_t* qb;   // 8 * cap
  WaveCounters* ctr;
  uint32_t cap;
};

struct GradArgs {
  const float* d_film;
  float* d_tex;
  float* d_emit;
};

// Per-workgroup aggregation of the adjoint's gradient deposits in LDS: the
// emission gradient is a dense array (a handful of lights, hit by every path),
// texel gradients go to a small open-addressing table; one global atomic per
// touched entry is issued when the workgroup ends.  Entries that do not fit go
// straight to global memory.
constexpr uint32_t kTexSlots = 1024;
constexpr uint32_t kTexProbe = 8;
constexpr uint32_t kEmitLds = 3 * 256;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;

struct GradLds {
  uint32_t keys[kTexSlots];
  float vals[kTexSlots * 3];
  float emit[kEmitLds];
};

__device__ __forceinline__ void grad_lds_init(GradLds& L) {
  for (uint32_t i = threadIdx.x; i < kTexSlots; i += blockDim.x) {
    L.keys[i] = kEmptyKey;
    L.vals[3 * i] = L.vals[3 * i + 1] = L.vals[3 * i + 2] = 0.0f;
  }
  for (uint32_t i = threadIdx.x; i < kEmitLds; i += blockDim.x) L.emit[i] = 0.0f;
  __syncthreads();
}

__device__ __forceinline__ void grad_lds_flush(GradLds& L, float* d_tex, float* d_emit,
                                               uint32_t n_emit) {
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kTexSlots; i += blockDim.x) {
    uint32_t const k = L.keys[i];
    if (k == kEmptyKey) continue;
    for (int c = 0; c < 3; ++c) {
      float const v = L.vals[3 * i + c];
      if (v != 0.0f) unsafeAtomicAdd(d_tex + 3 * size_t(k) + c, v);
    }
  }
  uint32_t const m = n_emit < kEmitLds ? n_emit : kEmitLds;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
    float const v = L.emit[i];
    if (v != 0.0f) unsafeAtomicAdd(d_emit + i, v);
  }
}

struct GradReplay {
  const DevTex* tex;
  const DevBsdf* bsdfs;
  float3 w;     // dLoss/d(film sum) of this path's pixel
  float3 ctot;  // the path's total colour from the primal pass
  float* d_tex;
  float* d_emit;
  GradLds* lds;
  __device__ __forceinline__ void albedo(int32_t tex, int32_t texel, float3 alb, float3 prefix) {
    // every later term carries this albedo as a factor: d/dalb = suffix / alb
    uint32_t const key = (tex_offset(tex) / 3) + uint32_t(texel);
    float const sx = ctot.x - prefix.x, sy = ctot.y - prefix.y, sz = ctot.z - prefix.z;
    float const gx = (sx != 0.0f) ? w.x * sx / alb.x : 0.0f;
    float const gy = (sy != 0.0f) ? w.y * sy / alb.y : 0.0f;
    float const gz = (sz != 0.0f) ? w.z * sz / alb.z : 0.0f;
    if (gx == 0.0f && gy == 0.0f && gz == 0.0f) return;
    uint32_t h = (key * 2654435761u) >> 22;  // 10 bits
    for (uint32_t p = 0; p < kTexProbe; ++p) {
      uint32_t const slot = (h + p) & (kTexSlots - 1);
      uint32_t const prev = atomicCAS(&lds->keys[slot], kEmptyKey, key);
      if (prev == kEmptyKey || prev == key) {
        atomicAdd(&lds->vals[3 * slot], gx);
        atomicAdd(&lds->vals[3 * slot + 1], gy);
        atomicAdd(&lds->vals[3 * slot + 2], gz);
        return;
      }
    }
    size_t const base = 3 * size_t(key);
    unsafeAtomicAdd(d_tex + base, gx);
    unsafeAtomicAdd(d_tex + base + 1, gy);
    unsafeAtomicAdd(d_tex + base + 2, gz);
  }
  __device__ __forceinline__ uint32_t tex_offset(int32_t t) const { return tex[t].offset; }
  __device__ __forceinline__ void emission(uint32_t bsdf, float3 add) {
    // Phi = L_e * pi * A, every emission term is linear in its light's L_e
    DevBsdf const& b = bsdfs[bsdf];
    float const g[3] = {(b.rgb[0] != 0.0f && add.x != 0.0f) ? w.x * add.x / b.rgb[0] : 0.0f,
                        (b.rgb[1] != 0.0f && add.y != 0.0f) ? w.y * add.y / b.rgb[1] : 0.0f,
                        (b.rgb[2] != 0.0f && add.z != 0.0f) ? w.z * add.z / b.rgb[2] : 0.0f};
    for (int c = 0; c < 3; ++c) {
      if (g[c] == 0.0f) continue;
      uint32_t const i = 3 * bsdf + uint32_t(c);
      if (i < kEmitLds) atomicAdd(&lds->emit[i], g[c]);
      else unsafeAtomicAdd(d_emit + i, g[c]);
    }
  }
};

__device__ __forceinline__ uint32_t pack_k(uint32_t k, bool spec) { return k | (spec ? 0x80000000u : 0u); }

__device__ __forceinline__ void load_state(const WaveBufs& wb, uint32_t slot, PathState& ps) {
  float4 const p = wb.pos[slot], d = wb.dir[slot], w = wb.pw[slot], c = wb.col[slot];
  ps.pos = ld3(p);
  ps.f_next = p.w;
  ps.dir = ld3(d);
  uint32_t const kb = __float_as_uint(d.w);
  ps.k = kb & 0x7FFFFFFFu;
  ps.specular = (kb >> 31) != 0;
  ps.power = ld3(w);
  ps.color = ld3(c);
}

__device__ __forceinline__ void store_state(const WaveBufs& wb, uint32_t slot, const PathState& ps) {
  wb.pos[slot] = make_float4(ps.pos.x, ps.pos.y, ps.pos.z, ps.f_next);
  wb.dir[slot] = make_float4(ps.dir.x, ps.dir.y, ps.dir.z, __uint_as_float(pack_k(ps.k, ps.specular)));
  wb.pw[slot] = make_float4(ps.power.x, ps.power.y, ps.power.z, 0.0f);
  wb.col[slot] = make_float4(ps.color.x, ps.color.y, ps.color.z, 0.0f);
}

// Total active paths and the queue slot of the j-th one (8 octant regions).
__device__ __forceinline__ uint32_t queue_total(const uint32_t* cnt) {
  uint32_t t = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) t += cnt[o];
  return t;
}
__device__ __forceinline__ uint32_t queue_slot(const uint32_t* cnt, const uint32_t* q, uint32_t cap,
                                               uint32_t j) {
  uint32_t base = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    uint32_t const c = cnt[o];
    if (j < base + c) return q[size_t(o) * cap + (j - base)];
    base += c;
  }
  return 0;
}

// Rays a wave holds at once: 64 while there is work for every wave of the
// grid; when a late bounce leaves few (and expensive) rays, they are spread
// over all waves instead of being packed 64 to a wave.
__device__ __forceinline__ uint32_t wave_ray_limit(uint32_t total) {
  uint32_t const waves = gridDim.x * (blockDim.x / 64u);
  uint32_t const per = (total + waves - 1) / waves;
  return per < 1u ? 1u : (per > 64u ? 64u : per);
}

__global__ void __launch_bounds__(kBlock)
k_camera(DevScene s, DevCamera cam, PTConst c, WaveBufs wb, uint64_t q0, uint32_t n, uint32_t wh,
         uint32_t sample0) {
  uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) {
    wb.ctr->cur[0] = n;
    for (int o = 1; o < 8; ++o) wb.ctr->cur[o] = 0;
    for (int o = 0; o < 9; ++o) wb.ctr->append[o].v = 0;
    wb.ctr->work_extend.v = 0;
    wb.ctr->work_shadow.v = 0;
  }
  if (i >= n) return;
  uint64_t const q = q0 + i;
  uint32_t const pixel = uint32_t(q % wh);
  uint32_t const sample = sample0 + uint32_t(q / wh);
  float4 const u = draw4(c, pixel, sample, 0);
  float3 o, d;
  primary_ray(cam, c, pixel, sample, u.x, u.y, &o, &d);
  wb.pos[i] = make_float4(o.x, o.y, o.z, 0.0f);
  wb.dir[i] = make_float4(d.x, d.y, d.z, __uint_as_float(pack_k(1u, true)));
  wb.pw[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  wb.col[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  wb.qa[i] = i;
}

// Persistent: lanes take active paths from the queues in batches (WaveFeeder)
// and trace them one step at a time, so a wave never waits for its slowest ray.
__global__ void __launch_bounds__(kBlock)
k_extend(DevScene s, WaveBufs wb, const uint32_t* __restrict__ qcur) {
  extern __shared__ uint32_t lds_stack[];
  __shared__ uint32_t cnt[8];
  if (threadIdx.x < 8) cnt[threadIdx.x] = wb.ctr->cur[threadIdx.x];
  __syncthreads();
  __shared__ unsigned long long keys[kBlock];
  uint32_t const total = queue_total(cnt);
  uint32_t const lim = wave_ray_limit(total);
  uint32_t* const stk = lds_stack + threadIdx.x;
  WaveFeeder feed;
  feed.init();
  TraceState st;
  uint32_t slot = 0xFFFFFFFFu;
  bool exhausted = false;
  for (;;) {
    bool const need = (slot == 0xFFFFFFFFu) && !exhausted;
    uint32_t const busy = uint32_t(__popcll(__ballot(slot != 0xFFFFFFFFu)));
    uint32_t const j = feed.refill(need, &wb.ctr->work_extend.v, total, &exhausted, lim - min(lim, busy), lim);
    if (j != 0xFFFFFFFFu) {
      slot = queue_slot(cnt, qcur, wb.cap, j);
      trace_init(st, ld3(wb.pos[slot]), ld3(wb.dir[slot]));
    }
    if (__ballot(slot != 0xFFFFFFFFu) == 0) break;
    bool const act = slot != 0xFFFFFFFFu;
#if XYZ_COOP
    bool const fin = coop_step<false>(s, st, act, stk, kBlock, keys + (threadIdx.x & ~63u));
#else
    bool const fin = act && trace_step<false>(s, st, stk, kBlock);
#endif
    if (act && fin) {
      HitRec const h = trace_result(s, st);
      wb.hit[slot] = make_float4(__int_as_float(h.prim < 0 ? -1 : (h.prim | (h.back ? 0x80000000 : 0))),
                                 h.t, h.f, h.g);
      slot = 0xFFFFFFFFu;
    }
  }
}

__device__ __forceinline__ HitRec unpack_hit(float4 v) {
  HitRec h;
  int32_t const b = __float_as_int(v.x);
  if (b == -1) {
    h.prim = -1;
    h.back = false;
  } else {
    h.prim = b & 0x7FFFFFFF;
    h.back = (uint32_t(b) >> 31) != 0;
  }
  h.t = v.y;
  h.f = v.z;
  h.g = v.w;
  return h;
}

// Workgroup-level append: lanes with `pred` reserve a place in list `key` of
// the workgroup's LDS counters `cnt`; returns the lane's offset in that list.
// One LDS atomic per wave and list.
__device__ __forceinline__ uint32_t lds_wave_append(bool pred, uint32_t key, uint32_t nkeys,
                                                    uint32_t first, uint32_t* cnt) {
  uint32_t const lane = __lane_id();
  unsigned long long const below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t off = 0;
  for (uint32_t o = 0; o < nkeys; ++o) {
    unsigned long long const m = __ballot(pred && key == o);
    if (m == 0) continue;
    uint32_t const leader = __ffsll((long long)m) - 1;
    uint32_t b = 0;
    if (lane == leader) b = atomicAdd(&cnt[first + o], uint32_t(__popcll(m)));
    b = __shfl(b, leader);
    if (pred && key == o) off = b + uint32_t(__popcll(m & below));
  }
  return off;
}

template <bool kReplay>
__global__ void __launch_bounds__(kBlock)
k_shade(DevScene s, PTConst c, WaveBufs wb, const uint32_t* __restrict__ qcur, uint32_t* qnext,
        uint64_t q0, uint32_t wh, uint32_t sample0, GradArgs ga) {
  __shared__ uint32_t cnt[8];
  __shared__ uint32_t bcnt[9], bbase[9];  // this workgroup's appends: 8 octant queues + shadow
  GradLds* glds = nullptr;
  if constexpr (kReplay) {
    __shared__ GradLds glds_mem;  // only the replay instantiation reserves it
    glds = &glds_mem;
    grad_lds_init(glds_mem);
  }
  if (threadIdx.x < 8) cnt[threadIdx.x] = wb.ctr->cur[threadIdx.x];
  __syncthreads();
  uint32_t const total = queue_total(cnt);
  uint32_t const stride = gridDim.x * kBlock;
  // every lane runs the same number of iterations so that ballots are convergent
  for (uint32_t base = blockIdx.x * kBlock; base < total; base += stride) {
    uint32_t const j = base + threadIdx.x;
    bool const active = j < total;
    bool cont = false, has_sr = false;
    uint32_t slot = 0, oct = 0;
    ShadowReq sr;
    if (active) {
      slot = queue_slot(cnt, qcur, wb.cap, j);
      uint64_t const q = q0 + slot;
      uint32_t const pixel = uint32_t(q % wh);
      uint32_t const sample = sample0 + uint32_t(q / wh);
      PathState ps;
      load_state(wb, slot, ps);
      HitRec const h = unpack_hit(wb.hit[slot]);
      float4 const u0 = draw4(c, pixel, sample, 0);
      LightVertex const lv = sample_light(s, u0.z, u0.w);
      if (kReplay) {
        GradReplay g;
        g.tex = s.tex;
        g.bsdfs = s.bsdfs;
        g.w = f3(ga.d_film[3 * size_t(pixel)], ga.d_film[3 * size_t(pixel) + 1],
                 ga.d_film[3 * size_t(pixel) + 2]);
        g.ctot = ld3(wb.ctot[slot]);
        g.d_tex = ga.d_tex;
        g.d_emit = ga.d_emit;
        g.lds = glds;
        cont = shade_vertex(s, c, ps, h, pixel, sample, lv, &sr, &has_sr, g);
      } else {
        NoGrad g;
        cont = shade_vertex(s, c, ps, h, pixel, sample, lv, &sr, &has_sr, g);
      }
      store_state(wb, slot, ps);
      oct = (ps.dir.x < 0.0f ? 1u : 0u) | (ps.dir.y < 0.0f ? 2u : 0u) | (ps.dir.z < 0.0f ? 4u : 0u);
    }
    // appends: LDS counts per workgroup, then one global atomic per list
    if (threadIdx.x < 9) bcnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t const qoff = lds_wave_append(active && cont, oct, 8, 0, bcnt);
    uint32_t const soff = lds_wave_append(active && has_sr, 0, 1, kShadowAppend, bcnt);
    __syncthreads();
    if (threadIdx.x < 9) {
      uint32_t const c = bcnt[threadIdx.x];
      bbase[threadIdx.x] = c ? atomicAdd(&wb.ctr->append[threadIdx.x].v, c) : 0u;
    }
    __syncthreads();
    if (active && cont) qnext[size_t(oct) * wb.cap + bbase[oct] + qoff] = slot;
    uint32_t const r = bbase[kShadowAppend] + soff;
    if (active && has_sr) {
      wb.so[r] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.nk2);
      wb.sw[r] = make_float4(sr.w.x, sr.w.y, sr.w.z, sr.nk3);
      wb.spe[r] = make_float4(sr.pe.x, sr.pe.y, sr.pe.z, __uint_as_float(slot));
      wb.smeta[r] = make_uint4(uint32_t(sr.expected), sr.type, sr.light_bsdf, 0u);
      wb.smet[r] = make_float4(sr.fsn, sr.fres, sr.mx, 0.0f);
    }
  }
  if constexpr (kReplay) grad_lds_flush(*glds, ga.d_tex, ga.d_emit, 3 * s.n_bsdfs);
}

template <bool kReplay>
__global__ void __launch_bounds__(kBlock)
k_shadow(DevScene s, WaveBufs wb, uint64_t q0, uint32_t wh, GradArgs ga) {
  extern __shared__ uint32_t lds_stack[];
  GradLds* glds = nullptr;
  if constexpr (kReplay) {
    __shared__ GradLds glds_mem;
    glds = &glds_mem;
    grad_lds_init(glds_mem);
  }
  __shared__ unsigned long long keys[kBlock];
  uint32_t const total = wb.ctr->append[kShadowAppend].v;
  uint32_t const lim = wave_ray_limit(total);
  uint32_t* const stk = lds_stack + threadIdx.x;
  WaveFeeder feed;
  feed.init();
  TraceState st;
  uint32_t r = 0xFFFFFFFFu;
  bool exhausted = false;
  for (;;) {
    bool const need = (r == 0xFFFFFFFFu) && !exhausted;
    uint32_t const busy = uint32_t(__popcll(__ballot(r != 0xFFFFFFFFu)));
    uint32_t const got = feed.refill(need, &wb.ctr->work_shadow.v, total, &exhausted, lim - min(lim, busy), lim);
    if (got != 0xFFFFFFFFu) {
      r = got;
      trace_init(st, ld3(wb.so[r]), ld3(wb.sw[r]));
    }
    if (__ballot(r != 0xFFFFFFFFu) == 0) break;
    bool const act = r != 0xFFFFFFFFu;
#if XYZ_COOP
    bool const fin = coop_step<false>(s, st, act, stk, kBlock, keys + (threadIdx.x & ~63u));
#else
    bool const fin = act && trace_step<false>(s, st, stk, kBlock);
#endif
    if (!act || !fin) continue;
    HitRec const h = trace_result(s, st);
    uint32_t const rr = r;
    r = 0xFFFFFFFFu;
    uint4 const meta = wb.smeta[rr];
    if (h.prim != int32_t(meta.x)) continue;
    float4 const o = wb.so[rr], w = wb.sw[rr];
    float4 const pe = wb.spe[rr], met = wb.smet[rr];
    ShadowReq sr;
    sr.type = meta.y;
    sr.nk2 = o.w;
    sr.nk3 = w.w;
    sr.pe = ld3(pe);
    sr.fsn = met.x;
    sr.fres = met.y;
    sr.mx = met.z;
    float3 const add = nee_contribution(sr, h.t, s.light_area);
    uint32_t const slot = __float_as_uint(pe.w);
    if constexpr (kReplay) {
      uint64_t const q = q0 + slot;
      uint32_t const pixel = uint32_t(q % wh);
      GradReplay g;
      g.tex = s.tex;
      g.bsdfs = s.bsdfs;
      g.w = f3(ga.d_film[3 * size_t(pixel)], ga.d_film[3 * size_t(pixel) + 1],
               ga.d_film[3 * size_t(pixel) + 2]);
      g.d_tex = ga.d_tex;
      g.d_emit = ga.d_emit;
      g.lds = glds;
      g.emission(meta.z, add);
    }
    float4 const cc = wb.col[slot];
    wb.col[slot] = make_float4(cc.x + add.x, cc.y + add.y, cc.z + add.z, 0.0f);
  }
  if constexpr (kReplay) grad_lds_flush(*glds, ga.d_tex, ga.d_emit, 3 * s.n_bsdfs);
}

__global__ void k_prep(WaveCounters* ctr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t t = ctr->append[kShadowAppend].v;
  for (int o = 0; o < 8; ++o) {
    t += ctr->cur[o];
    ctr->cur[o] = ctr->append[o].v;
    ctr->append[o].v = 0;
  }
  ctr->append[kShadowAppend].v = 0;
  ctr->work_extend.v = 0;
  ctr->work_shadow.v = 0;
  ctr->queries += t;
}

__global__ void k_zero_counters(WaveCounters* ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr->queries = 0;
}

// film[p] += colour of each of p's samples in this wave, in sample order.
__global__ void __launch_bounds__(kBlock)
k_accumulate(float* film, const float4* __restrict__ col, uint64_t q0, uint32_t n, uint32_t wh) {
  uint32_t const p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= wh) return;
  uint64_t q = q0 + (uint64_t(p) + wh - (q0 % wh)) % wh;
  if (q >= q0 + n) return;
  float fx = film[3 * size_t(p)], fy = film[3 * size_t(p) + 1], fz = film[3 * size_t(p) + 2];
  for (; q < q0 + n; q += wh) {
    float4 const c = col[q - q0];
    fx = fx + c.x;
    fy = fy + c.y;
    fz = fz + c.z;
  }
  film[3 * size_t(p)] = fx;
  film[3 * size_t(p) + 1] = fy;
  film[3 * size_t(p) + 2] = fz;
}

__global__ void __launch_bounds__(kBlock) k_copy_col(float4* dst, const float4* src, uint32_t n) {
  uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------------ host side
int g_num_cu = 0;

int num_cu() {
  if (g_num_cu == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    g_num_cu = v;
  }
  return g_num_cu;
}

float hnormalize_scale(const float a[3]) {
  float const l2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  return 1.0f / std::sqrt(l2);
}

// xyz/core/thinlenscamera.cpp:71-90 (basis) with V = 0.0125 / tan(fov_y / 2)
int make_dev_camera(const xyz_camera* cam, DevCamera* dc) {
  if (!cam || cam->width == 0 || cam->height == 0) {
    set_error("camera: zero image size");
    return XYZ_E_INVALID;
  }
  if (!(cam->focal_length > 0.0f) && !(cam->fov_y_deg > 0.0f && cam->fov_y_deg < 180.0f)) {
    set_error("camera: fov_y_deg must be in (0, 180)");
    return XYZ_E_INVALID;
  }
  std::memset(dc, 0, sizeof(*dc));
  float view[3], back[3], right[3], up[3], c[3];
  for (int k = 0; k < 3; ++k) {
    dc->eye[k] = cam->eye[k];
    view[k] = cam->eye[k] - cam->at[k];
  }
  float s = hnormalize_scale(view);
  for (int k = 0; k < 3; ++k) back[k] = view[k] * s;
  // right = normalize(cross(up, back))
  c[0] = cam->up[1] * back[2] - cam->up[2] * back[1];
  c[1] = cam->up[2] * back[0] - cam->up[0] * back[2];
  c[2] = cam->up[0] * back[1] - cam->up[1] * back[0];
  s = hnormalize_scale(c);
  for (int k = 0; k < 3; ++k) right[k] = c[k] * s;
  // up = cross(back, right)
  up[0] = back[1] * right[2] - back[2] * right[1];
  up[1] = back[2] * right[0] - back[0] * right[2];
  up[2] = back[0] * right[1] - back[1] * right[0];
  for (int k = 0; k < 3; ++k) {
    dc->back[k] = back[k];
    dc->right[k] = right[k];
    dc->up[k] = up[k];
  }
  if (!std::isfinite(right[0]) || !std::isfinite(back[0])) {
    set_error("camera: degenerate eye/at/up");
    return XYZ_E_INVALID;
  }
  dc->film = 0.025f;
  dc->aspect = static_cast<float>(int(cam->width)) / int(cam->height);
  if (cam->focal_length > 0.0f) {
    // thin lens: xyz/core/thinlenscamera.cpp:71-77, 192-193
    float const f = cam->focal_length, fnum = cam->f_number;
    float const L = std::sqrt(view[0] * view[0] + view[1] * view[1] + view[2] * view[2]);
    if (!(fnum > 0.0f) || !(L > f) || !std::isfinite(L)) {
      set_error("camera: thin lens needs f_number > 0 and |eye - at| > focal_length");
      return XYZ_E_INVALID;
    }
    dc->thin = 1;
    dc->focus = L;
    dc->v = L * f / (L - f);
    dc->lens_r = f / (2 * fnum);
    dc->focal_plane = dc->film * L / dc->v;
  } else {
    dc->v = float(double(dc->film) * 0.5 / std::tan(double(cam->fov_y_deg) * 3.14159265358979323846 / 360.0));
  }
  dc->width = cam->width;
  dc->height = cam->height;
  return XYZ_OK;
}

int make_pt_const(const xyz_render_params* p, PTConst* c) {
  if (!p) {
    set_error("render: NULL params");
    return XYZ_E_INVALID;
  }
  if (p->max_depth == 0 || p->max_depth > 1000) {
    set_error("render: max_depth must be in [1, 1000]");
    return XYZ_E_INVALID;
  }
  c->kmax = p->max_depth + 1;
  c->rr_start = p->rr_start;
  c->rr_threshold = p->rr_threshold;
  c->seed_lo = uint32_t(p->seed);
  c->seed_hi = uint32_t(p->seed >> 32);
  return XYZ_OK;
}

struct DevMem {
  void* p = nullptr;
  hipStream_t st = nullptr;
  ~DevMem() {
    if (p) (void)hipFreeAsync(p, st);
  }
};

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

int alloc_wave(uint32_t cap, bool replay, hipStream_t st, DevMem* mem, WaveBufs* wb) {
  size_t const f4 = align_up(size_t(cap) * 16);
  size_t const u4 = align_up(size_t(cap) * 16);
  size_t const qsz = align_up(size_t(cap) * 8 * 4);
  size_t const total = f4 * (replay ? 10 : 9) + u4 + 2 * qsz + align_up(sizeof(WaveCounters));
  hipError_t e = pool_alloc(&mem->p, total, st);
  if (e != hipSuccess) return hip_fail(e, "pool_alloc(wave)");
  mem->st = st;
  char* b = static_cast<char*>(mem->p);
  auto take = [&](size_t n) {
    char* r = b;
    b += n;
    return r;
  };
  wb->pos = reinterpret_cast<float4*>(take(f4));
  wb->dir = reinterpret_cast<float4*>(take(f4));
  wb->pw = reinterpret_cast<float4*>(take(f4));
  wb->col = reinterpret_cast<float4*>(take(f4));
  wb->hit = reinterpret_cast<float4*>(take(f4));
  wb->so = reinterpret_cast<float4*>(take(f4));
  wb->sw = reinterpret_cast<float4*>(take(f4));
  wb->spe = reinterpret_cast<float4*>(take(f4));
  wb->smet = reinterpret_cast<float4*>(take(f4));
  wb->ctot = replay ? reinterpret_cast<float4*>(take(f4)) : nullptr;
  wb->smeta = reinterpret_cast<uint4*>(take(u4));
  wb->qa = reinterpret_cast<uint32_t*>(take(qsz));
  wb->qb = reinterpret_cast<uint32_t*>(take(qsz));
  wb->ctr = reinterpret_cast<WaveCounters*>(take(align_up(sizeof(WaveCounters))));
  wb->cap = cap;
  return XYZ_OK;
}

// Internal streams (default priority) that alternate the waves of a render,
// created once per device.
#ifndef XYZ_RENDER_LANES
#define XYZ_RENDER_LANES 2
#endif
constexpr int kLanes = XYZ_RENDER_LANES;
std::mutex g_lane_mu;
std::map<int, std::vector<hipStream_t>> g_lane_streams;

int lane_streams(hipStream_t* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_lane_mu);
  std::vector<hipStream_t>& v = g_lane_streams[dev];
  while (v.size() < size_t(kLanes)) {
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreateWithFlags");
    v.push_back(s);
  }
  for (int i = 0; i < kLanes; ++i) out[i] = v[size_t(i)];
  return XYZ_OK;
}

// Events without timing, destroyed with the set.
struct EventSet {
  std::vector<hipEvent_t> e;
  int create(int n) {
    for (int i = 0; i < n; ++i) {
      hipEvent_t x = nullptr;
      hipError_t const r = hipEventCreateWithFlags(&x, hipEventDisableTiming);
      if (r != hipSuccess) return hip_fail(r, "hipEventCreateWithFlags");
      e.push_back(x);
    }
    return XYZ_OK;
  }
  ~EventSet() {
    for (hipEvent_t x : e) (void)hipEventDestroy(x);
  }
};

// Runs the bounce loop of one wave.  kReplay: adjoint replay pass.
template <bool kReplay>
int run_bounces(const DevScene& s, const PTConst& c, WaveBufs wb, uint64_t q0, uint32_t n,
                uint32_t wh, uint32_t sample0, uint32_t max_depth, GradArgs ga, hipStream_t st) {
  int const ncu = num_cu();
  size_t const lds = size_t(s.stack_cap) * kBlock * sizeof(uint32_t);
  uint32_t const trace_grid = std::max(1u, std::min((n + kBlock - 1) / kBlock, uint32_t(ncu) * 8));
  uint32_t const shade_grid = std::max(1u, std::min((n + kBlock - 1) / kBlock, uint32_t(ncu) * 8));
  uint32_t* qc = wb.qa;
  uint32_t* qn = wb.qb;
  for (uint32_t b = 0; b < max_depth; ++b) {
    hipLaunchKernelGGL(k_extend, dim3(trace_grid), dim3(kBlock), lds, st, s, wb, qc);
    hipLaunchKernelGGL(k_shade<kReplay>, dim3(shade_grid), dim3(kBlock), 0, st, s, c, wb, qc, qn,
                       q0, wh, sample0, ga);
    hipLaunchKernelGGL(k_shadow<kReplay>, dim3(trace_grid), dim3(kBlock), lds, st, s, wb, q0, wh, ga);
    hipLaunchKernelGGL(k_prep, dim3(1), dim3(64), 0, st, wb.ctr);
    std::swap(qc, qn);
  }
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render launch");
  return XYZ_OK;
}

int render_impl(xyz_scene* sc, const xyz_camera* cam, const xyz_render_params* prm, float* film,
                xyz_render_stats* stats, hipStream_t st, bool backward, const float* d_film,
                float* d_tex, float* d_emit) {
  if (!sc) {
    set_error("render: NULL scene");
    return XYZ_E_INVALID;
  }
  DevCamera dc;
  PTConst pc;
  int rc = make_dev_camera(cam, &dc);
  if (rc != XYZ_OK) return rc;
  rc = make_pt_const(prm, &pc);
  if (rc != XYZ_OK) return rc;
  if (!backward && !film) {
    set_error("render: NULL film");
    return XYZ_E_INVALID;
  }
  if (backward && (!d_film || (!d_tex && sc->n_texels > 0) || !d_emit)) {
    set_error("render_backward: NULL gradient buffer");
    return XYZ_E_INVALID;
  }
  uint64_t const wh = uint64_t(cam->width) * cam->height;
  if (wh >= (1ull << 32)) {
    set_error("render: image too large");
    return XYZ_E_UNSUPPORTED;
  }
  uint64_t const total = wh * prm->spp;
  if (uint64_t(prm->sample_offset) + prm->spp >= (1ull << 32)) {
    set_error("render: sample index overflow");
    return XYZ_E_UNSUPPORTED;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (stats) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
  }
  if (total > 0) {
    uint32_t const cap = uint32_t(std::min<uint64_t>(total, kWaveCap));
    uint64_t const n_waves = (total + cap - 1) / cap;
    // consecutive waves alternate between internal streams, so that the tail
    // of one wave (a few long paths) overlaps the full bounces of the next;
    // films still receive the waves in sample order (events below)
    int const lanes = n_waves > 1 ? kLanes : 1;
    hipStream_t ls[kLanes];
    for (int i = 0; i < kLanes; ++i) ls[i] = st;
    if (lanes > 1) {
      rc = lane_streams(ls);
      if (rc != XYZ_OK) return rc;
    }
    DevMem mem[kLanes];
    WaveBufs wb[kLanes];
    for (int i = 0; i < lanes; ++i) {
      rc = alloc_wave(cap, backward, st, &mem[i], &wb[i]);
      if (rc != XYZ_OK) return rc;
      hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, wb[i].ctr);
    }
    EventSet ev;
    if (lanes > 1) {
      rc = ev.create(2 * kLanes + 1);
      if (rc != XYZ_OK) return rc;
      hipEvent_t const fork = ev.e[2 * kLanes];
      hipError_t e = hipEventRecord(fork, st);
      for (int i = 0; i < lanes && e == hipSuccess; ++i) e = hipStreamWaitEvent(ls[i], fork, 0);
      if (e != hipSuccess) return hip_fail(e, "render fork");
    }
    GradArgs none{nullptr, nullptr, nullptr};
    GradArgs ga{d_film, d_tex, d_emit};
    for (uint64_t k = 0; k < n_waves; ++k) {
      int const i = int(k % uint64_t(lanes));
      hipStream_t const s = ls[i];
      uint64_t const q0 = k * cap;
      uint32_t const n = uint32_t(std::min<uint64_t>(cap, total - q0));
      uint32_t const grid = (n + kBlock - 1) / kBlock;
      hipLaunchKernelGGL(k_camera, dim3(grid), dim3(kBlock), 0, s, sc->dev, dc, pc, wb[i], q0, n,
                         uint32_t(wh), prm->sample_offset);
      rc = run_bounces<false>(sc->dev, pc, wb[i], q0, n, uint32_t(wh), prm->sample_offset,
                              prm->max_depth, none, s);
      if (rc != XYZ_OK) return rc;
      if (!backward) {
        // wave k is added to the film after wave k - 1 (sample order per pixel)
        if (lanes > 1 && k > 0) {
          hipError_t const e = hipStreamWaitEvent(s, ev.e[(k - 1) % uint64_t(lanes)], 0);
          if (e != hipSuccess) return hip_fail(e, "render wait");
        }
        hipLaunchKernelGGL(k_accumulate, dim3(uint32_t((wh + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           s, film, wb[i].col, q0, n, uint32_t(wh));
        if (lanes > 1) {
          hipError_t const e = hipEventRecord(ev.e[i], s);
          if (e != hipSuccess) return hip_fail(e, "render record");
        }
      } else {
        // replay: same paths from their seeds, now with the primal colours known
        hipLaunchKernelGGL(k_copy_col, dim3(grid), dim3(kBlock), 0, s, wb[i].ctot, wb[i].col, n);
        hipLaunchKernelGGL(k_camera, dim3(grid), dim3(kBlock), 0, s, sc->dev, dc, pc, wb[i], q0, n,
                           uint32_t(wh), prm->sample_offset);
        rc = run_bounces<true>(sc->dev, pc, wb[i], q0, n, uint32_t(wh), prm->sample_offset,
                               prm->max_depth, ga, s);
        if (rc != XYZ_OK) return rc;
      }
    }
    if (lanes > 1) {  // join: the caller's stream waits for every lane
      for (int i = 0; i < lanes; ++i) {
        hipEvent_t const done = ev.e[kLanes + i];
        hipError_t e = hipEventRecord(done, ls[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, done, 0);
        if (e != hipSuccess) return hip_fail(e, "render join");
      }
    }
    if (stats) {
      unsigned long long q[kLanes] = {};
      hipError_t e = hipSuccess;
      for (int i = 0; i < lanes && e == hipSuccess; ++i)
        e = hipMemcpyAsync(&q[i], &wb[i].ctr->queries, sizeof(q[i]), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hip_fail(e, "render stats");
      stats->queries = 0;
      for (int i = 0; i < lanes; ++i) stats->queries += q[i];
    }
  }
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render");
  if (stats) {
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    stats->seconds = double(ms) * 1e-3;
    stats->paths = total;
    if (total == 0) stats->queries = 0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return XYZ_OK;
}

// ================================================================ batched mode
// One persistent launch renders many jobs.  Work items are (job, tile of
// `tile` pixels) and are taken by whole waves, so the lanes of a wave always
// share one scene (the wave-cooperative traversal tests other lanes' rays).
// Each lane owns one pixel of the tile and traces that pixel's samples in
// sample order, adding them to the film in that order (integrator.cpp:119-128),
// so films are bit-identical to the wavefront renderer's.  A lane is a small
// state machine -- extension trace, shading, NEE shadow trace -- so that every
// lane reaches the cooperative traversal step together.
struct BatchJob {
  DevScene s;
  DevCamera cam;
  PTConst c;
  float* film;
  uint32_t spp, sample0, n_tiles, tile_base;
};

enum : uint32_t { kIdle = 0, kExtend = 1, kShadow = 2 };

__global__ void k_batch(const BatchJob* __restrict__ jobs, uint32_t n_jobs, uint32_t total_tiles,
                        uint32_t tile, uint32_t* work) {
  extern __shared__ uint32_t lds_stack[];
  __shared__ unsigned long long keys_all[1024];
  unsigned long long* const keys = keys_all + (threadIdx.x & ~63u);
  uint32_t* const stk = lds_stack + threadIdx.x;
  uint32_t const lane = __lane_id();
  unsigned long long const below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  // wave-uniform tile state
  uint32_t job = 0, tile_next = 0, tile_end = 0;
  bool have_tile = false;
  // lane state
  uint32_t phase = kIdle, pixel = 0, smp = 0, smp_end = 0;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  PathState ps;
  LightVertex lv;
  ShadowReq sr;
  bool pending_cont = false;
  TraceState st;

  for (;;) {
    // next tile once every lane is idle and the current tile is handed out
    if (__ballot(phase != kIdle) == 0 && tile_next >= tile_end) {
      uint32_t it = 0;
      if (lane == 0) it = atomicAdd(work, 1u);
      it = __shfl(it, 0);
      if (it >= total_tiles) break;
      uint32_t lo = 0, hi = n_jobs;
      while (hi - lo > 1) {
        uint32_t const mid = (lo + hi) >> 1;
        if (jobs[mid].tile_base <= it) lo = mid; else hi = mid;
      }
      job = lo;
      uint32_t const wh = jobs[job].cam.width * jobs[job].cam.height;
      tile_next = (it - jobs[job].tile_base) * tile;
      tile_end = min(tile_next + tile, wh);
      have_tile = true;
    }
    BatchJob const& jb = jobs[job];
    // idle lanes take the tile's next pixels
    bool const need = have_tile && phase == kIdle;
    unsigned long long const m = __ballot(need && tile_next < tile_end);
    if (m != 0) {
      uint32_t const rank = uint32_t(__popcll(m & below));
      uint32_t const avail = tile_end - tile_next;
      if (need && rank < avail) {
        pixel = tile_next + rank;
        float* px = jb.film + 3 * size_t(pixel);
        fx = px[0];
        fy = px[1];
        fz = px[2];
        smp = jb.sample0;
        smp_end = jb.sample0 + jb.spp;
        phase = (smp < smp_end) ? kExtend : kIdle;
        if (phase == kExtend) {
          float4 const u0 = draw4(jb.c, pixel, smp, 0);
          lv = sample_light(jb.s, u0.z, u0.w);
          primary_ray(jb.cam, jb.c, pixel, smp, u0.x, u0.y, &ps.pos, &ps.dir);
          ps.power = f3(1.0f, 1.0f, 1.0f);
          ps.color = f3(0.0f, 0.0f, 0.0f);
          ps.f_next = 0.0f;
          ps.k = 1;
          ps.specular = true;
          trace_init(st, ps.pos, ps.dir);
        }
      }
      uint32_t const want = uint32_t(__popcll(m));
      tile_next += want < avail ? want : avail;
    }
    bool const act = phase != kIdle;
    bool const fin = coop_step<false>(jb.s, st, act, stk, blockDim.x, keys);
    if (!(act && fin)) continue;
    HitRec const h = trace_result(jb.s, st);
    bool path_done = false;
    if (phase == kExtend) {
      NoGrad g;
      bool has_sr = false;
      bool const cont = shade_vertex(jb.s, jb.c, ps, h, pixel, smp, lv, &sr, &has_sr, g);
      if (has_sr) {
        phase = kShadow;
        pending_cont = cont;
        trace_init(st, sr.o, sr.w);
      } else if (cont) {
        trace_init(st, ps.pos, ps.dir);
      } else {
        path_done = true;
      }
    } else {
      if (h.prim == sr.expected) ps.color = add3(ps.color, nee_contribution(sr, h.t, jb.s.light_area));
      if (pending_cont) {
        phase = kExtend;
        trace_init(st, ps.pos, ps.dir);
      } else {
        path_done = true;
      }
    }
    if (path_done) {
      fx = fx + ps.color.x;
      fy = fy + ps.color.y;
      fz = fz + ps.color.z;
      ++smp;
      if (smp < smp_end) {
        float4 const u0 = draw4(jb.c, pixel, smp, 0);
        lv = sample_light(jb.s, u0.z, u0.w);
        primary_ray(jb.cam, jb.c, pixel, smp, u0.x, u0.y, &ps.pos, &ps.dir);
        ps.power = f3(1.0f, 1.0f, 1.0f);
        ps.color = f3(0.0f, 0.0f, 0.0f);
        ps.f_next = 0.0f;
        ps.k = 1;
        ps.specular = true;
        phase = kExtend;
        trace_init(st, ps.pos, ps.dir);
      } else {
        float* px = jb.film + 3 * size_t(pixel);
        px[0] = fx;
        px[1] = fy;
        px[2] = fz;
        phase = kIdle;
      }
    }
  }
}

__global__ void k_reset_work(uint32_t* work) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *work = 0;
}

__global__ void __launch_bounds__(kBlock) k_fill_zero(float* p, size_t n) {
  size_t const i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

struct BatchPlan {
  std::vector<BatchJob> jobs;
  uint32_t total_tiles = 0;
  uint32_t max_stack = 1;
  uint64_t paths = 0;
};

int plan_batch(uint32_t n, xyz_scene* const* scenes, const xyz_camera* cams,
               const xyz_render_params* prm, float* const* films, uint32_t tile, BatchPlan* plan) {
  plan->jobs.resize(n);
  // zeroed so that the bytes of the job table (the graph cache key) are defined
  std::memset(static_cast<void*>(plan->jobs.data()), 0, n * sizeof(BatchJob));
  plan->total_tiles = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!scenes[i] || !films[i]) {
      set_error("render_batch: NULL scene or film in job " + std::to_string(i));
      return XYZ_E_INVALID;
    }
    BatchJob& j = plan->jobs[i];
    int rc = make_dev_camera(&cams[i], &j.cam);
    if (rc != XYZ_OK) return rc;
    rc = make_pt_const(&prm[i], &j.c);
    if (rc != XYZ_OK) return rc;
    j.s = scenes[i]->dev;
    j.film = films[i];
    j.spp = prm[i].spp;
    j.sample0 = prm[i].sample_offset;
    uint64_t const wh = uint64_t(cams[i].width) * cams[i].height;
    j.n_tiles = uint32_t((wh + tile - 1) / tile);
    j.tile_base = plan->total_tiles;
    plan->total_tiles += j.n_tiles;
    plan->max_stack = std::max(plan->max_stack, j.s.stack_cap);
    plan->paths += wh * prm[i].spp;
  }
  return XYZ_OK;
}

xyz_batch_tuning default_tuning() {
  xyz_batch_tuning t;
  t.tile_pixels = 256;
  t.blocks_per_cu = 4;
  t.block_threads = 128;
  t.use_graph = 1;
  return t;
}

struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  void* jobs_dev = nullptr;
  uint32_t* work = nullptr;
};

std::mutex g_graph_mu;
std::map<std::string, GraphEntry> g_graphs;
hipStream_t g_capture_stream = nullptr;

int launch_batch_raw(const BatchPlan& plan, const xyz_batch_tuning& t, void* jobs_dev,
                     uint32_t* work, hipStream_t st) {
  uint32_t const threads = t.block_threads;
  size_t const lds = size_t(plan.max_stack) * threads * sizeof(uint32_t);
  uint32_t const grid = std::max(1u, uint32_t(num_cu()) * t.blocks_per_cu);
  hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(64), 0, st, work);
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(threads), lds, st,
                     static_cast<const BatchJob*>(jobs_dev), uint32_t(plan.jobs.size()),
                     plan.total_tiles, t.tile_pixels, work);
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render_batch launch");
  return XYZ_OK;
}

int check_tuning(const xyz_batch_tuning& t) {
  if (t.tile_pixels == 0 || t.block_threads == 0 || t.block_threads > 1024 ||
      (t.block_threads % 64) != 0 || t.blocks_per_cu == 0 || t.blocks_per_cu > 64) {
    set_error("render_batch: invalid tuning");
    return XYZ_E_INVALID;
  }
  return XYZ_OK;
}

int render_batch_impl(uint32_t n, xyz_scene* const* scenes, const xyz_camera* cams,
                      const xyz_render_params* prm, float* const* films,
                      const xyz_batch_tuning* tuning, xyz_render_stats* stats, hipStream_t st) {
  if (n == 0) return XYZ_OK;
  if (!scenes || !cams || !prm || !films) {
    set_error("render_batch: NULL argument");
    return XYZ_E_INVALID;
  }
  xyz_batch_tuning const t = tuning ? *tuning : default_tuning();
  int rc = check_tuning(t);
  if (rc != XYZ_OK) return rc;
  BatchPlan plan;
  rc = plan_batch(n, scenes, cams, prm, films, t.tile_pixels, &plan);
  if (rc != XYZ_OK) return rc;
  if (size_t(plan.max_stack) * t.block_threads * sizeof(uint32_t) > 65536) {
    set_error("render_batch: LDS stack exceeds 64 KiB (lower block_threads)");
    return XYZ_E_UNSUPPORTED;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (stats) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
  }
  if (t.use_graph) {
    // key: the job table bytes and the tuning; the graph replays one launch
    std::string key(reinterpret_cast<const char*>(plan.jobs.data()), plan.jobs.size() * sizeof(BatchJob));
    key.append(reinterpret_cast<const char*>(&t), sizeof(t));
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find(key);
    if (it == g_graphs.end()) {
      GraphEntry ge;
      size_t const jb = plan.jobs.size() * sizeof(BatchJob);
      hipError_t e = hipMalloc(&ge.jobs_dev, jb);
      if (e == hipSuccess) e = hipMemcpy(ge.jobs_dev, plan.jobs.data(), jb, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&ge.work), sizeof(uint32_t));
      if (e != hipSuccess) return hip_fail(e, "render_batch graph buffers");
      if (!g_capture_stream) {
        e = hipStreamCreate(&g_capture_stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
      }
      e = hipStreamBeginCapture(g_capture_stream, hipStreamCaptureModeRelaxed);
      if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
      rc = launch_batch_raw(plan, t, ge.jobs_dev, ge.work, g_capture_stream);
      e = hipStreamEndCapture(g_capture_stream, &ge.graph);
      if (rc != XYZ_OK) return rc;
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0);
      if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
      if (g_graphs.size() > 64) {
        for (auto& kv : g_graphs) {
          (void)hipGraphExecDestroy(kv.second.exec);
          (void)hipGraphDestroy(kv.second.graph);
          (void)hipFree(kv.second.jobs_dev);
          (void)hipFree(kv.second.work);
        }
        g_graphs.clear();
      }
      it = g_graphs.emplace(key, ge).first;
    }
    if (stats) (void)hipEventRecord(e0, st);
    hipError_t const e = hipGraphLaunch(it->second.exec, st);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  } else {
    void* jobs_dev = nullptr;
    uint32_t* work = nullptr;
    size_t const jb = plan.jobs.size() * sizeof(BatchJob);
    hipError_t e = pool_alloc(&jobs_dev, jb + 256, st);
    if (e != hipSuccess) return hip_fail(e, "pool_alloc(batch)");
    work = reinterpret_cast<uint32_t*>(static_cast<char*>(jobs_dev) + align_up(jb));
    e = hipMemcpyAsync(jobs_dev, plan.jobs.data(), jb, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(batch)");
    // the host table must outlive the copy
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if (stats) (void)hipEventRecord(e0, st);
    rc = launch_batch_raw(plan, t, jobs_dev, work, st);
    (void)hipFreeAsync(jobs_dev, st);
    if (rc != XYZ_OK) return rc;
  }
  if (stats) {
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    stats->seconds = double(ms) * 1e-3;
    stats->paths = plan.paths;
    stats->queries = 0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return XYZ_OK;
}

int zero_films(uint32_t n, const xyz_camera* cams, float* const* films, hipStream_t st) {
  for (uint32_t i = 0; i < n; ++i) {
    size_t const m = size_t(cams[i].width) * cams[i].height * 3;
    hipLaunchKernelGGL(k_fill_zero, dim3(uint32_t((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                       films[i], m);
  }
  hipError_t const e = hipGetLastError();
  return e == hipSuccess ? XYZ_OK : hip_fail(e, "zero films");
}

// Project-written autotuner: times every candidate (tile size, persistent
// workgroups per CU, workgroup size) once after a warm-up and keeps the fastest.
int autotune_impl(uint32_t n, xyz_scene* const* scenes, const xyz_camera* cams,
                  const xyz_render_params* prm, float* const* films, xyz_batch_tuning* best,
                  hipStream_t st) {
  if (!best) {
    set_error("batch_autotune: NULL best");
    return XYZ_E_INVALID;
  }
  *best = default_tuning();
  if (n == 0) return XYZ_OK;
  // candidates are timed on a reduced sample count of the same jobs
  std::vector<xyz_render_params> trial(prm, prm + n);
  for (auto& p : trial) p.spp = std::max(1u, std::min(p.spp, 4u));
  prm = trial.data();
  const uint32_t tiles[] = {64, 256, 1024};
  const uint32_t bpc[] = {2, 4, 8};
  const uint32_t thr[] = {64, 128, 256};
  double best_s = 1e30;
  bool first = true;
  for (uint32_t tp : tiles) {
    for (uint32_t bc : bpc) {
      for (uint32_t th : thr) {
        xyz_batch_tuning t;
        t.tile_pixels = tp;
        t.blocks_per_cu = bc;
        t.block_threads = th;
        t.use_graph = 0;
        uint32_t max_stack = 1;
        for (uint32_t i = 0; i < n; ++i) max_stack = std::max(max_stack, scenes[i]->dev.stack_cap);
        if (size_t(max_stack) * th * sizeof(uint32_t) > 65536) continue;
        if (first) {  // warm-up
          int rc = render_batch_impl(n, scenes, cams, prm, films, &t, nullptr, st);
          if (rc != XYZ_OK) return rc;
          first = false;
        }
        xyz_render_stats s;
        int rc = render_batch_impl(n, scenes, cams, prm, films, &t, &s, st);
        if (rc != XYZ_OK) return rc;
        if (s.seconds < best_s) {
          best_s = s.seconds;
          *best = t;
        }
      }
    }
  }
  best->use_graph = 1;
  return zero_films(n, cams, films, st);
}

}  // namespace

}  // namespace xyzhip

using namespace xyzhip;

extern "C" {

int xyz_render(xyz_scene* scene, const xyz_camera* camera, const xyz_render_params* params,
               float* film, xyz_render_stats* stats, void* stream) {
  return render_impl(scene, camera, params, film, stats, static_cast<hipStream_t>(stream), false,
                     nullptr, nullptr, nullptr);
}

int xyz_render_backward(xyz_scene* scene, const xyz_camera* camera, const xyz_render_params* params,
                        const float* d_film, float* d_texels, float* d_emission, void* stream) {
  return render_impl(scene, camera, params, nullptr, nullptr, static_cast<hipStream_t>(stream),
                     true, d_film, d_texels, d_emission);
}

int xyz_render_batch(uint32_t n, xyz_scene* const* scenes, const xyz_camera* cameras,
                     const xyz_render_params* params, float* const* films,
                     const xyz_batch_tuning* tuning, xyz_render_stats* stats, void* stream) {
  return render_batch_impl(n, scenes, cameras, params, films, tuning, stats,
                           static_cast<hipStream_t>(stream));
}

int xyz_batch_autotune(uint32_t n, xyz_scene* const* scenes, const xyz_camera* cameras,
                       const xyz_render_params* params, float* const* films,
                       xyz_batch_tuning* best, void* stream) {
  return autotune_impl(n, scenes, cameras, params, films, best, static_cast<hipStream_t>(stream));
}

}  // extern "C"

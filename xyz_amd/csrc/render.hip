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

#include "bdpt_device.h"
#include "pt_device.h"
#include "scene_internal.h"

namespace xyzhip {

namespace {

constexpr uint32_t kBlock = 256;
// Path slots per wave: 2^26 (about 17 GB of wave buffers per internal stream);
// the 1920x1080 256-spp frame measured 498-505 Msamples/s at 2^25, 516-532 at
// 2^26 and 526-529 at 2^27.
#ifndef XYZ_WAVE_CAP_LOG2
#define XYZ_WAVE_CAP_LOG2 26
#endif
constexpr uint64_t kWaveCap = 1ull << XYZ_WAVE_CAP_LOG2;
// Bytes of adjoint vertex tape per internal stream (the adjoint's waves are
// cut to fit).
constexpr uint64_t kTapeBudget = 6ull << 30;

// A counter alone in its 256-byte line: device-scope atomics on one line are
// serialised, so counters that many workgroups hit at once do not share one.
struct alignas(256) LineU32 {
  uint32_t v;
  uint32_t pad[63];
};

constexpr uint32_t kMaxWaveJobs = 64;  // jobs one wave may hold
constexpr uint32_t kFeedParts = 8;     // parts of the trace kernels' work, one per XCD
constexpr uint32_t kSegPad = 256;      // queue segments start at multiples of this position

// One render job -- scene, camera, integrator constants, film -- on the device.
// Its work items are the global items [base, base + wh * spp); item base + q
// is sample sample0 + q / wh of pixel q % wh.
// Workgroups per CU the guided shading kernel is compiled for (4: 4 waves per
// SIMD, at most 128 VGPRs).
#ifndef XYZ_GUIDED_SHADE_BLOCKS
#define XYZ_GUIDED_SHADE_BLOCKS 4
#endif
// Waves per SIMD the extension and shadow kernels are compiled for (0: the
// compiler's choice).
#ifndef XYZ_TRACE_WAVES_PER_EU
#define XYZ_TRACE_WAVES_PER_EU 0
#endif
// Workgroups per CU the plain and replay shading kernels are compiled for
// (1: the compiler's register choice).
#ifndef XYZ_SHADE_BLOCKS
#define XYZ_SHADE_BLOCKS 1
#endif

struct RJob {
  DevScene s;
  DevCamera cam;
  PTConst c;
  DevImportons imp;  // importance-driven path tracing only
  float* film;
  unsigned long long base;
  uint32_t wh, sample0;
};

// The jobs of one wave: local job l is global job jlo + l and owns the wave's
// path slots [start[l], start[l + 1]); slot i holds global item q0 + i.
struct WaveJobs {
  unsigned long long q0;
  uint32_t jlo, nj;
  uint32_t start[kMaxWaveJobs + 1];
};

// Queue segments of a wave.  Extension queues: segment g = 8 l + o holds the
// active paths of local job l whose direction lies in octant o, at entries
// [8 start[l] + o m_l, ...) of a queue array (m_l = start[l + 1] - start[l]).
// Shadow requests of job l sit at entries [start[l], ...) of the request
// arrays.  Persistent kernels address a queue set through virtual positions:
// the segments end to end, each padded to a multiple of kSegPad, so that an
// aligned batch of positions always lies in one segment, i.e. one job.
struct WaveCounters {
  LineU32 append[kMaxWaveJobs * 9];     // [9 l + o]: o < 8 next octant queue, o = 8 shadow requests
  LineU32 work_extend[kFeedParts];      // persistent extend kernel's work counters, one per part
  LineU32 work_shadow[kFeedParts];      // persistent shadow kernel's work counters
  uint32_t cur[kMaxWaveJobs * 8];       // current extension queue sizes
  uint32_t vext[kMaxWaveJobs * 8 + 1];  // padded prefix of cur
  uint32_t n_active;                    // sum of cur
  unsigned long long queries;
};

struct WaveBufs {
  // path state, one 64-byte record per slot at ps[4 * slot]: (pos, f_next),
  // (dir, bits(k | specular << 31)), (power, 0), (colour, 0).  Shading reads
  // and rewrites whole records of slots spread over the octant queues; as one
  // record a slot's state is one whole 64-byte segment instead of four
  // partial ones.
  float4* ps;
  // the extension hits of the current bounce, by virtual queue position (the
  // extension and shading kernels both walk the queues by position, so both
  // access it in whole segments): bits(prim | back << 31), t, f, g
  float4* hit;
  float4* ctot;   // adjoint: emission per unit throughput of the light the path ends on, bits(vertex count)
  // shadow requests, one 64-byte record per request at rqs[4 * rq]:
  // (origin, nk2), (w, nk3), (pe, bits(slot)), bits(expected prim,
  // type | extra << 7 | vertex << 8, light bsdf), t_target.  The shadow
  // kernel reads a request before and after its query; as one record that
  // is one segment.  smet[rq]: fsn, fres, mx, histogram density, written and
  // read only with `extra` (a METAL request or a guided one; otherwise
  // 0, 0, 0, 1).
  float4* rqs;
  float4* smet;
  uint32_t* qa;   // 8 * cap
  uint32_t* qb;   // 8 * cap
  WaveCounters* ctr;
  uint32_t cap;
  // adjoint: per (vertex vj, slot), record r = vj * cap + slot, tape_depth
  // vertices.  tape_ab[2r], tape_ab[2r + 1]: (throughput before the albedo,
  // bits(texel key)), (albedo, scale after the NEE step, negative when the
  // vertex made a NEE request), written together once per vertex; tape_c[r]:
  // (NEE term per unit throughput after the albedo, 0), written by the shadow
  // kernel for every request (0 when occluded).  The slots a shading wave
  // holds are spread over the octant queues, so each record is one 32-byte
  // store: partial cache lines written back to HBM cost a read each.
  float4* tape_ab;
  float4* tape_c;
  uint32_t tape_depth;
};

struct GradArgs {
  const float* d_film;
  float* d_tex;
  float* d_emit;
};

// Per-workgroup aggregation of the adjoint's gradient deposits in LDS.  The
// emission gradient is a dense array (a handful of lights, reached by every
// path); texel gradients go to a small open-addressing table.  One global
// atomic per touched entry is issued when the workgroup ends; entries that do
// not fit go straight to global memory.
constexpr uint32_t kTexSlots = 1024;
constexpr uint32_t kTexProbe = 8;
constexpr uint32_t kEmitLds = 3 * 256;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;

struct EmitLds {
  float emit[kEmitLds];
};
struct TexLds {
  uint32_t keys[kTexSlots];
  float vals[kTexSlots * 3];
};

__device__ __forceinline__ void emit_lds_init(EmitLds& L) {
  for (uint32_t i = threadIdx.x; i < kEmitLds; i += blockDim.x) L.emit[i] = 0.0f;
  __syncthreads();
}
__device__ __forceinline__ void emit_lds_flush(EmitLds& L, float* d_emit, uint32_t n_emit) {
  __syncthreads();
  uint32_t const m = n_emit < kEmitLds ? n_emit : kEmitLds;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
    float const v = L.emit[i];
    if (v != 0.0f) unsafeAtomicAdd(d_emit + i, v);
  }
}
__device__ __forceinline__ void emit_add(EmitLds* L, float* d_emit, uint32_t i, float v) {
  if (v == 0.0f) return;
  if (i < kEmitLds) atomicAdd(&L->emit[i], v);
  else unsafeAtomicAdd(d_emit + i, v);
}

__device__ __forceinline__ void tex_lds_init(TexLds& L) {
  for (uint32_t i = threadIdx.x; i < kTexSlots; i += blockDim.x) {
    L.keys[i] = kEmptyKey;
    L.vals[3 * i] = L.vals[3 * i + 1] = L.vals[3 * i + 2] = 0.0f;
  }
  __syncthreads();
}
__device__ __forceinline__ void tex_lds_flush(TexLds& L, float* d_tex) {
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kTexSlots; i += blockDim.x) {
    uint32_t const k = L.keys[i];
    if (k == kEmptyKey) continue;
    for (int c = 0; c < 3; ++c) {
      float const v = L.vals[3 * i + c];
      if (v != 0.0f) unsafeAtomicAdd(d_tex + 3 * size_t(k) + c, v);
    }
  }
}
// Adds g to texel `key` (the texel's index in the concatenated textures).
__device__ __forceinline__ void tex_add(TexLds* L, float* d_tex, uint32_t key, float3 g) {
  if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f) return;
  uint32_t const h = (key * 2654435761u) >> 22;  // 10 bits
  for (uint32_t p = 0; p < kTexProbe; ++p) {
    uint32_t const slot = (h + p) & (kTexSlots - 1);
    uint32_t const prev = atomicCAS(&L->keys[slot], kEmptyKey, key);
    if (prev == kEmptyKey || prev == key) {
      atomicAdd(&L->vals[3 * slot], g.x);
      atomicAdd(&L->vals[3 * slot + 1], g.y);
      atomicAdd(&L->vals[3 * slot + 2], g.z);
      return;
    }
  }
  size_t const base = 3 * size_t(key);
  unsafeAtomicAdd(d_tex + base, g.x);
  unsafeAtomicAdd(d_tex + base + 1, g.y);
  unsafeAtomicAdd(d_tex + base + 2, g.z);
}

// Gradient hooks of the adjoint's replay (see NoGrad).  Each shading vertex
// keeps its record in registers until finish() writes it to the slot's tape;
// a path that ends on a light writes that light's emission per unit
// throughput to ctot.  The L_e derivative of
// an emission term is the term with pi A_lights in place of Phi
// (Phi = L_e pi A_lights) and is deposited at once.
struct GradReplay {
  const DevTex* tex;
  float3 w;       // dLoss/d(film sum) of this path's pixel
  double unit;    // pi A_lights
  float* d_emit;
  EmitLds* lds;
  float4* ctot;
  float4* tape_ab;
  uint32_t cap, tape_depth;
  uint32_t slot;
  uint32_t vj;    // the current vertex
  float sigma;    // its throughput scale after the NEE step
  bool recorded;  // vertex() ran for this shading call
  float4 ra;      // the record's first half: throughput, bits(texel key)
  float3 ralb;    // its albedo
  __device__ __forceinline__ void bind(const WaveBufs& wb, uint32_t s) {
    ctot = wb.ctot;
    tape_ab = wb.tape_ab;
    cap = wb.cap;
    tape_depth = wb.tape_depth;
    slot = s;
    vj = 0;
    sigma = 1.0f;
    recorded = false;
  }
  __device__ __forceinline__ void emission(uint32_t bsdf, float3 power, double wfac,
                                           const DevBsdf& b, uint32_t k) {
    ctot[slot] = make_float4(float(double(b.phi[0]) * wfac), float(double(b.phi[1]) * wfac),
                             float(double(b.phi[2]) * wfac), __uint_as_float(k - 1u));
    double const u = unit * wfac;
    emit_add(lds, d_emit, 3 * bsdf, w.x * float(double(power.x) * u));
    emit_add(lds, d_emit, 3 * bsdf + 1, w.y * float(double(power.y) * u));
    emit_add(lds, d_emit, 3 * bsdf + 2, w.z * float(double(power.z) * u));
  }
  __device__ __forceinline__ void vertex(uint32_t j, float3 power, float3 alb, int32_t texture,
                                         int32_t texel) {
    vj = j;
    sigma = 1.0f;
    if (j >= tape_depth) return;  // the depth cap keeps j < tape_depth
    uint32_t const key = (texture >= 0 && texel >= 0)
                             ? tex[texture].offset / 3u + uint32_t(texel) : kEmptyKey;
    ra = make_float4(power.x, power.y, power.z, __uint_as_float(key));
    ralb = alb;
    recorded = true;
  }
  __device__ __forceinline__ void scale(float sc) { sigma = sigma * sc; }
  // After the shading call: the vertex's record, flagged with whether it made
  // a NEE request, and the path's vertex count.
  __device__ __forceinline__ void finish(bool nee) {
    if (!recorded) return;
    size_t const r = size_t(vj) * cap + slot;
    tape_ab[2 * r] = ra;
    tape_ab[2 * r + 1] = make_float4(ralb.x, ralb.y, ralb.z, nee ? -sigma : sigma);
    ctot[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(vj + 1u));
  }
};

__device__ __forceinline__ uint32_t pack_k(uint32_t k, bool spec) { return k | (spec ? 0x80000000u : 0u); }

__device__ __forceinline__ void load_state(const WaveBufs& wb, uint32_t slot, PathState& ps) {
  float4 const* const r = wb.ps + 4 * size_t(slot);
  float4 const p = r[0], d = r[1], w = r[2], c = r[3];
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
  float4* const r = wb.ps + 4 * size_t(slot);
  r[0] = make_float4(ps.pos.x, ps.pos.y, ps.pos.z, ps.f_next);
  r[1] = make_float4(ps.dir.x, ps.dir.y, ps.dir.z, __uint_as_float(pack_k(ps.k, ps.specular)));
  r[2] = make_float4(ps.power.x, ps.power.y, ps.power.z, 0.0f);
  r[3] = make_float4(ps.color.x, ps.color.y, ps.color.z, 0.0f);
}

__device__ __forceinline__ uint32_t pad_seg(uint32_t c) { return (c + kSegPad - 1u) & ~(kSegPad - 1u); }

// Queue entry of position r of extension segment g.
__device__ __forceinline__ uint32_t ext_entry(const WaveJobs& wj, uint32_t g, uint32_t r) {
  uint32_t const l = g >> 3, o = g & 7u;
  uint32_t const s0 = wj.start[l];
  return 8u * s0 + o * (wj.start[l + 1] - s0) + r;
}

// The segment holding virtual position v: the last g with vst[g] <= v, for
// non-decreasing starts vst[0..nseg] with vst[0] = 0 <= v < vst[nseg].
__device__ __forceinline__ uint32_t seg_of(const uint32_t* vst, uint32_t nseg, uint32_t v) {
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1u) {
    uint32_t const mid = (lo + hi) >> 1;
    if (vst[mid] <= v) lo = mid; else hi = mid;
  }
  return lo;
}

// Local job of wave slot i.
__device__ __forceinline__ uint32_t job_of_slot(const WaveJobs& wj, uint32_t i) {
  uint32_t lo = 0, hi = wj.nj;
  while (hi - lo > 1u) {
    uint32_t const mid = (lo + hi) >> 1;
    if (wj.start[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// Rays a wave holds at once: 64 while there is work for every wave of the
// grid; when a late bounce leaves few (and expensive) rays, they are spread
// over all waves instead of being packed 64 to a wave.  A power of two, so
// that an aligned batch of this size never crosses a queue segment.
__device__ __forceinline__ uint32_t wave_ray_limit(uint32_t total) {
  uint32_t const waves = gridDim.x * (blockDim.x / 64u);
  uint32_t const per = (total + waves - 1u) / waves;
  uint32_t lim = 1u;
  while (lim < per && lim < 64u) lim <<= 1;
  return lim;
}

// Per-wave source of queue positions for the persistent trace kernels.
// Batches of `lim` consecutive virtual positions come from a global counter,
// one atomic per batch; padding positions are skipped.  A wave traces the
// rays of one job at a time: a batch of another job is handed out only once
// all of the wave's lanes are idle.  All lanes call take() together.
struct SegFeeder {
  uint32_t base, count, next, seg;
  uint32_t part, parts_left;
  bool done;
  // The positions [0, vtotal) are split into kFeedParts contiguous parts on
  // kSegPad boundaries (so no claim of lim <= 64 positions crosses one), with
  // a counter each; the wave starts on part `first` (its block's XCD under
  // round-robin dispatch) and moves to the next part when one is used up.
  __device__ __forceinline__ void init(uint32_t first) {
    base = count = next = seg = 0;
    part = first % kFeedParts;
    parts_left = kFeedParts;
    done = false;
  }
  // Lanes with `need` receive the entry r of segment *lane_seg (returned), or
  // UINT32_MAX.  cnt[g]: real size of segment g; segs_per_job segments per
  // job; *job: the wave's job, *switched: set when it changed.
  __device__ __forceinline__ uint32_t take(bool need, LineU32* counters, uint32_t vtotal,
                                           uint32_t lim, const uint32_t* vst, uint32_t nseg,
                                           const uint32_t* cnt, uint32_t segs_per_job,
                                           uint32_t* job, bool* switched, uint32_t* lane_seg) {
    uint32_t const lane = __lane_id();
    unsigned long long const below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t busy = uint32_t(__popcll(__ballot(!need)));
    uint32_t room = lim > busy ? lim - busy : 0u;
    uint32_t got = 0xFFFFFFFFu;
    for (;;) {
      unsigned long long const m = __ballot(need);
      if (m == 0 || room == 0 || done) break;
      if (next == count) {
        if (parts_left == 0) {
          done = true;
          break;
        }
        uint32_t const pt = __builtin_amdgcn_readfirstlane(part);
        uint32_t const units = vtotal / kSegPad;
        uint32_t const lo = (units * pt / kFeedParts) * kSegPad;
        uint32_t const hi = (units * (pt + 1u) / kFeedParts) * kSegPad;
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&counters[pt].v, lim);
        b = __builtin_amdgcn_readfirstlane(__shfl(b, 0)) + lo;
        if (b >= hi) {
          part = (pt + 1u) % kFeedParts;
          parts_left = __builtin_amdgcn_readfirstlane(parts_left) - 1u;
          count = next = 0;
          continue;
        }
        uint32_t const g = seg_of(vst, nseg, b);
        uint32_t const r0 = b - vst[g];
        uint32_t const c = cnt[g];
        base = r0;
        seg = g;
        next = 0;
        count = (r0 < c) ? ((c - r0 < lim) ? c - r0 : lim) : 0u;
        continue;
      }
      uint32_t const j = seg / segs_per_job;
      if (j != *job) {
        if (busy > 0) break;  // the current job's rays first
        *job = j;
        *switched = true;
      }
      uint32_t const avail = count - next;
      uint32_t const rank = uint32_t(__popcll(m & below));
      uint32_t take_n = uint32_t(__popcll(m));
      take_n = take_n < avail ? take_n : avail;
      take_n = take_n < room ? take_n : room;
      if (need && rank < take_n) {
        got = base + next + rank;
        *lane_seg = seg;
        need = false;
      }
      next += take_n;
      room -= take_n;
      busy += take_n;
    }
    return got;
  }
};

// Exclusive prefix sum over a 512-thread workgroup (sc: 512 LDS words, which
// hold the exclusive prefix on return); returns the total.
__device__ __forceinline__ uint32_t block_scan512(uint32_t v, uint32_t* sc) {
  uint32_t const t = threadIdx.x;
  sc[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < 512u; off <<= 1) {
    uint32_t const a = (t >= off) ? sc[t - off] : 0u;
    __syncthreads();
    sc[t] += a;
    __syncthreads();
  }
  uint32_t const incl = sc[t];
  uint32_t const total = sc[511];
  __syncthreads();
  sc[t] = incl - v;
  __syncthreads();
  return total;
}

// Device copies of the job and wave tables, written through kernel arguments
// (stream-ordered; no host buffer has to outlive a copy).
__global__ void k_put_job(RJob j, RJob* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = j;
}
__global__ void k_put_wave(WaveJobs w, WaveJobs* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = w;
}

// Queue counters of a fresh wave: every path of job l in its octant-0 queue.
__global__ void __launch_bounds__(512) k_init_wave(const WaveJobs* __restrict__ wjp,
                                                   WaveCounters* ctr) {
  __shared__ uint32_t sc[512];
  WaveJobs const& wj = *wjp;
  uint32_t const k = threadIdx.x, nseg = wj.nj * 8u;
  uint32_t c = 0;
  if (k < nseg) {
    uint32_t const l = k >> 3;
    c = ((k & 7u) == 0u) ? wj.start[l + 1] - wj.start[l] : 0u;
    ctr->cur[k] = c;
  }
  for (uint32_t a = k; a < wj.nj * 9u; a += 512u) ctr->append[a].v = 0u;
  uint32_t const tot = block_scan512(k < nseg ? pad_seg(c) : 0u, sc);
  if (k < nseg) ctr->vext[k] = sc[k];
  if (k == 0) {
    ctr->vext[nseg] = tot;
    ctr->n_active = wj.start[wj.nj];
    for (uint32_t p = 0; p < kFeedParts; ++p) {
      ctr->work_extend[p].v = 0u;
      ctr->work_shadow[p].v = 0u;
    }
  }
}

__global__ void __launch_bounds__(kBlock)
k_camera(const RJob* __restrict__ jobs, const WaveJobs* __restrict__ wjp, WaveBufs wb, uint32_t n) {
  uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  WaveJobs const& wj = *wjp;
  uint32_t const l = job_of_slot(wj, i);
  RJob const& J = jobs[wj.jlo + l];
  unsigned long long const item = wj.q0 + i - J.base;
  uint32_t const pixel = uint32_t(item % J.wh);
  uint32_t const sample = J.sample0 + uint32_t(item / J.wh);
  float4 const u = draw4(J.c, pixel, sample, 0);
  float3 o, d;
  primary_ray(J.cam, J.c, pixel, sample, u.x, u.y, &o, &d);
  float4* const r = wb.ps + 4 * size_t(i);
  r[0] = make_float4(o.x, o.y, o.z, 0.0f);
  r[1] = make_float4(d.x, d.y, d.z, __uint_as_float(pack_k(1u, true)));
  r[2] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  r[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (wb.ctot) wb.ctot[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // no vertex, no emission
  wb.qa[8u * wj.start[l] + (i - wj.start[l])] = i;
}

// Persistent: lanes take active paths from the queues (SegFeeder) and trace
// them one step at a time, so a wave never waits for its slowest ray.
__global__ void __launch_bounds__(kBlock)
#if XYZ_TRACE_WAVES_PER_EU
__attribute__((amdgpu_waves_per_eu(XYZ_TRACE_WAVES_PER_EU, XYZ_TRACE_WAVES_PER_EU)))
#endif
k_extend(const RJob* __restrict__ jobs, const WaveJobs* __restrict__ wjp, WaveBufs wb,
         const uint32_t* __restrict__ qcur) {
  extern __shared__ uint32_t lds_stack[];
  __shared__ uint32_t vst[kMaxWaveJobs * 8 + 1];
  __shared__ uint32_t cnt[kMaxWaveJobs * 8];
  __shared__ unsigned long long keys[2 * kBlock];  // keys, then batch_owner's tables
  WaveJobs const& wj = *wjp;
  uint32_t const nseg = wj.nj * 8u;
  for (uint32_t k = threadIdx.x; k <= nseg; k += kBlock) {
    vst[k] = wb.ctr->vext[k];
    if (k < nseg) cnt[k] = wb.ctr->cur[k];
  }
  __syncthreads();
  uint32_t const vtotal = vst[nseg];
  uint32_t const lim = wave_ray_limit(wb.ctr->n_active);
  uint32_t* const stk = lds_stack + threadIdx.x;
  SegFeeder feed;
  feed.init(blockIdx.x);
  uint32_t job = 0;
  const DevScene* sp = &jobs[wj.jlo].s;  // the wave's job (uniform)
  TraceState st;
  uint32_t slot = 0xFFFFFFFFu, hv = 0;
  for (;;) {
    bool switched = false;
    uint32_t lseg = 0;
    uint32_t const r = feed.take(slot == 0xFFFFFFFFu, wb.ctr->work_extend, vtotal, lim, vst,
                                 nseg, cnt, 8u, &job, &switched, &lseg);
    if (switched) sp = &jobs[wj.jlo + __builtin_amdgcn_readfirstlane(job)].s;
    DevScene const& s = *sp;
    if (r != 0xFFFFFFFFu) {
      hv = vst[lseg] + r;  // the query's virtual position
      slot = qcur[ext_entry(wj, lseg, r)];
      trace_init(st, ld3(wb.ps[4 * size_t(slot)]), ld3(wb.ps[4 * size_t(slot) + 1]));
    }
    if (__ballot(slot != 0xFFFFFFFFu) == 0) break;
    bool const act = slot != 0xFFFFFFFFu;
    bool const fin = coop_step<false>(s, st, act, stk, kBlock, keys + (threadIdx.x & ~63u),
                                            reinterpret_cast<uint8_t*>(keys + kBlock + (threadIdx.x & ~63u)));
    if (act && fin) {
      HitRec const h = trace_result(s, st);
      wb.hit[hv] = make_float4(__int_as_float(h.prim < 0 ? -1 : (h.prim | (h.back ? 0x80000000 : 0))),
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

// Shading of the active paths.  A workgroup takes 256 consecutive virtual
// positions at a time, which lie in one queue segment (one job), so the
// workgroup's appends go to that job's queues: counted in LDS, then one global
// atomic per queue.
template <bool kReplay, bool kGuide>
__global__ void __launch_bounds__(kBlock, kGuide ? XYZ_GUIDED_SHADE_BLOCKS : XYZ_SHADE_BLOCKS)
k_shade(const RJob* __restrict__ jobs, const WaveJobs* __restrict__ wjp, WaveBufs wb,
        const uint32_t* __restrict__ qcur, uint32_t* qnext, GradArgs ga) {
  __shared__ uint32_t vst[kMaxWaveJobs * 8 + 1];
  __shared__ uint32_t bcnt[9], bbase[9];  // this workgroup's appends: 8 octant queues + shadow
  EmitLds* glds = nullptr;
  if constexpr (kReplay) {
    __shared__ EmitLds glds_mem;  // only the replay instantiation reserves it
    glds = &glds_mem;
    emit_lds_init(glds_mem);
  }
  WaveJobs const& wj = *wjp;
  uint32_t const nseg = wj.nj * 8u;
  for (uint32_t k = threadIdx.x; k <= nseg; k += kBlock) vst[k] = wb.ctr->vext[k];
  __syncthreads();
  uint32_t const vtotal = vst[nseg];
  // every thread of a workgroup runs the same iterations (barriers below)
  for (uint32_t vb = blockIdx.x * kBlock; vb < vtotal; vb += gridDim.x * kBlock) {
    uint32_t const g = __builtin_amdgcn_readfirstlane(seg_of(vst, nseg, vb));
    uint32_t const l = g >> 3;
    uint32_t const r = vb - vst[g] + threadIdx.x;
    bool const active = r < wb.ctr->cur[g];
    RJob const& J = jobs[wj.jlo + l];
    bool cont = false, has_sr = false;
    uint32_t slot = 0, oct = 0;
    ShadowReq sr;
    if constexpr (kGuide) {
      // importance-driven: the wave queries the importon grid for its guided
      // lanes together, between the two shading steps
      PathState ps{};
      HitRec h{};
      LightVertex lv{};
      VertexCtx x{};
      uint32_t pixel = 0, sample = 0;
      bool go = false;
      NoGrad ng;
      if (active) {
        slot = qcur[ext_entry(wj, g, r)];
        unsigned long long const item = wj.q0 + slot - J.base;
        pixel = uint32_t(item % J.wh);
        sample = J.sample0 + uint32_t(item / J.wh);
        load_state(wb, slot, ps);
        h = unpack_hit(wb.hit[vb + threadIdx.x]);
        float4 const u0 = draw4(J.c, pixel, sample, 0);
        lv = sample_light(J.s, u0.z, u0.w);
        go = shade_begin<NoGrad, true>(J.s, J.c, ps, h, pixel, sample, x, ng);
      }
      uint32_t hc[kImpN];
      wave_importance_counts(J.imp, go && x.guided, ps.pos, x.B, hc);
      if (active) {
        if (go) cont = shade_end<NoGrad, true>(J.s, J.c, ps, h, x, lv, &sr, &has_sr, ng, hc);
        store_state(wb, slot, ps);
        oct = (ps.dir.x < 0.0f ? 1u : 0u) | (ps.dir.y < 0.0f ? 2u : 0u) |
              (ps.dir.z < 0.0f ? 4u : 0u);
      }
    } else if (active) {
      slot = qcur[ext_entry(wj, g, r)];
      unsigned long long const item = wj.q0 + slot - J.base;
      uint32_t const pixel = uint32_t(item % J.wh);
      uint32_t const sample = J.sample0 + uint32_t(item / J.wh);
      PathState ps;
      load_state(wb, slot, ps);
      HitRec const h = unpack_hit(wb.hit[vb + threadIdx.x]);
      float4 const u0 = draw4(J.c, pixel, sample, 0);
      LightVertex const lv = sample_light(J.s, u0.z, u0.w);
      if (kReplay) {
        GradReplay gr;
        gr.tex = J.s.tex;
        gr.w = f3(ga.d_film[3 * size_t(pixel)], ga.d_film[3 * size_t(pixel) + 1],
                  ga.d_film[3 * size_t(pixel) + 2]);
        gr.unit = kPi * double(J.s.light_area);
        gr.d_emit = ga.d_emit;
        gr.lds = glds;
        gr.bind(wb, slot);
        cont = shade_vertex(J.s, J.c, ps, h, pixel, sample, lv, &sr, &has_sr, gr);
        gr.finish(has_sr);
      } else {
        NoGrad ng;
        cont = shade_vertex(J.s, J.c, ps, h, pixel, sample, lv, &sr, &has_sr, ng);
      }
      store_state(wb, slot, ps);
      oct = (ps.dir.x < 0.0f ? 1u : 0u) | (ps.dir.y < 0.0f ? 2u : 0u) | (ps.dir.z < 0.0f ? 4u : 0u);
    }
    // appends: LDS counts per workgroup, then one global atomic per list
    if (threadIdx.x < 9) bcnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t const qoff = lds_wave_append(active && cont, oct, 8, 0, bcnt);
    uint32_t const soff = lds_wave_append(active && has_sr, 0, 1, 8, bcnt);
    __syncthreads();
    if (threadIdx.x < 9) {
      uint32_t const c = bcnt[threadIdx.x];
      bbase[threadIdx.x] = c ? atomicAdd(&wb.ctr->append[9u * l + threadIdx.x].v, c) : 0u;
    }
    __syncthreads();
    if (active && cont) qnext[ext_entry(wj, 8u * l + oct, bbase[oct] + qoff)] = slot;
    uint32_t const rq = wj.start[l] + bbase[8] + soff;
    if (active && has_sr) {
      float4* const q = wb.rqs + 4 * size_t(rq);
      bool const extra = (sr.type == XYZ_METAL) || (sr.gd != 1.0f);
      q[0] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.nk2);
      q[1] = make_float4(sr.w.x, sr.w.y, sr.w.z, sr.nk3);
      q[2] = make_float4(sr.pe.x, sr.pe.y, sr.pe.z, __uint_as_float(slot));
      q[3] = make_float4(__uint_as_float(uint32_t(sr.expected)),
                         __uint_as_float(sr.type | (extra ? 0x80u : 0u) | (sr.vj << 8)),
                         __uint_as_float(sr.light_bsdf), sr.t_target);
      if (extra) wb.smet[rq] = make_float4(sr.fsn, sr.fres, sr.mx, sr.gd);
    }
  }
  if constexpr (kReplay) emit_lds_flush(*glds, ga.d_emit, 3 * jobs[wj.jlo].s.n_bsdfs);
}

template <bool kReplay>
__global__ void __launch_bounds__(kBlock)
#if XYZ_TRACE_WAVES_PER_EU
__attribute__((amdgpu_waves_per_eu(XYZ_TRACE_WAVES_PER_EU, XYZ_TRACE_WAVES_PER_EU)))
#endif
k_shadow(const RJob* __restrict__ jobs, const WaveJobs* __restrict__ wjp, WaveBufs wb, GradArgs ga) {
  extern __shared__ uint32_t lds_stack[];
  __shared__ uint32_t vst[kMaxWaveJobs + 1];
  __shared__ uint32_t cnt[kMaxWaveJobs];
  EmitLds* glds = nullptr;
  if constexpr (kReplay) {
    __shared__ EmitLds glds_mem;
    glds = &glds_mem;
    emit_lds_init(glds_mem);
  }
  __shared__ unsigned long long keys[2 * kBlock];  // keys, then batch_owner's tables
  WaveJobs const& wj = *wjp;
  if (threadIdx.x < 64) {  // padded prefix of the jobs' request counts
    uint32_t const k = threadIdx.x;
    uint32_t const c = (k < wj.nj) ? wb.ctr->append[9u * k + 8u].v : 0u;
    uint32_t excl, total;
    wave_prefix(pad_seg(c), &excl, &total);
    if (k < wj.nj) {
      cnt[k] = c;
      vst[k] = excl;
    }
    if (k == 0) vst[wj.nj] = total;
  }
  __syncthreads();
  uint32_t const nseg = wj.nj;
  uint32_t const vtotal = vst[nseg];
  uint32_t n_req = 0;
  for (uint32_t k = 0; k < nseg; ++k) n_req += cnt[k];
  uint32_t const lim = wave_ray_limit(n_req);
  uint32_t* const stk = lds_stack + threadIdx.x;
  SegFeeder feed;
  feed.init(blockIdx.x);
  uint32_t job = 0;
  const DevScene* sp = &jobs[wj.jlo].s;  // the wave's job (uniform)
  TraceState st;
  uint32_t rq = 0xFFFFFFFFu;
  for (;;) {
    bool switched = false;
    uint32_t lseg = 0;
    uint32_t const r = feed.take(rq == 0xFFFFFFFFu, wb.ctr->work_shadow, vtotal, lim, vst, nseg,
                                 cnt, 1u, &job, &switched, &lseg);
    if (switched) sp = &jobs[wj.jlo + __builtin_amdgcn_readfirstlane(job)].s;
    DevScene const& s = *sp;
    if (r != 0xFFFFFFFFu) {
      rq = wj.start[lseg] + r;
      float4 const* const q = wb.rqs + 4 * size_t(rq);
      trace_init(st, ld3(q[0]), ld3(q[1]));
      // The query only asks whether its closest hit is the shading point's
      // triangle, which it can find only at that triangle's distance T (the
      // request's t_target; tri_test on the same record values gives the same
      // bits).  With the root interval [t_min, T] in place of [t_min, +inf]
      // the traversal visits the same leaves up to the first whose interval
      // reaches past T, each interval cut at T.  A result of the full query at
      // or before T is the least key of a subset that still holds it, so it is
      // kept; a result past T comes from a leaf whose interval reaches past T,
      // and no later leaf can accept T.  So the capped query finds the
      // triangle exactly when the full query does, at t = T.
      st.t_cap = st.t_max = q[3].w;
    }
    if (__ballot(rq != 0xFFFFFFFFu) == 0) break;
    bool const act = rq != 0xFFFFFFFFu;
    bool const fin = coop_step<false, false>(s, st, act, stk, kBlock, keys + (threadIdx.x & ~63u),
                                            reinterpret_cast<uint8_t*>(keys + kBlock + (threadIdx.x & ~63u)));
    if (!act || !fin) continue;
    HitRec const h = trace_result(s, st);
    uint32_t const rr = rq;
    rq = 0xFFFFFFFFu;
    float4 const* const q = wb.rqs + 4 * size_t(rr);
    float4 const m4 = q[3];
    uint3 const meta = make_uint3(__float_as_uint(m4.x), __float_as_uint(m4.y), __float_as_uint(m4.z));
    if (h.prim != int32_t(meta.x)) {
      if constexpr (kReplay) {
        // occluded: the vertex's NEE term is 0
        uint32_t const vj = meta.y >> 8;
        if (vj < wb.tape_depth)
          wb.tape_c[size_t(vj) * wb.cap + __float_as_uint(q[2].w)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
      continue;
    }
    float4 const o = q[0], w = q[1], pe = q[2];
    float4 const met = (meta.y & 0x80u) ? wb.smet[rr] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    ShadowReq sr;
    sr.type = meta.y & 0x7Fu;
    sr.nk2 = o.w;
    sr.nk3 = w.w;
    sr.pe = ld3(pe);
    sr.fsn = met.x;
    sr.fres = met.y;
    sr.mx = met.z;
    sr.gd = met.w;
    double wfac = 0.0;
    float3 const add = nee_contribution(sr, h.t, s.light_area, kReplay ? &wfac : nullptr);
    uint32_t const slot = __float_as_uint(pe.w);
    if constexpr (kReplay) {
      // the vertex's NEE term per unit throughput, and the sampled light's
      // L_e derivative (pi A_lights in place of Phi)
      RJob const& J = jobs[wj.jlo + job];
      unsigned long long const item = wj.q0 + slot - J.base;
      uint32_t const pixel = uint32_t(item % J.wh);
      uint32_t const vj = meta.y >> 8;
      if (vj < wb.tape_depth) {
        size_t const ti = size_t(vj) * wb.cap + slot;
        float4 const ta = wb.tape_ab[2 * ti], tb = wb.tape_ab[2 * ti + 1];
        float3 const pw = mul3(ld3(ta), ld3(tb));  // the throughput after the albedo
        auto const& lb = gp(s.bsdfs)[meta.z];
        wb.tape_c[ti] = make_float4(float(double(lb.phi[0]) * wfac), float(double(lb.phi[1]) * wfac),
                                    float(double(lb.phi[2]) * wfac), 0.0f);
        double const u = kPi * double(s.light_area) * wfac;
        emit_add(glds, ga.d_emit, 3 * meta.z, ga.d_film[3 * size_t(pixel)] * float(double(pw.x) * u));
        emit_add(glds, ga.d_emit, 3 * meta.z + 1, ga.d_film[3 * size_t(pixel) + 1] * float(double(pw.y) * u));
        emit_add(glds, ga.d_emit, 3 * meta.z + 2, ga.d_film[3 * size_t(pixel) + 2] * float(double(pw.z) * u));
      }
    }
    float4* const cr = wb.ps + 4 * size_t(slot) + 3;
    float4 const cc = *cr;
    *cr = make_float4(cc.x + add.x, cc.y + add.y, cc.z + add.z, 0.0f);
  }
  if constexpr (kReplay) emit_lds_flush(*glds, ga.d_emit, 3 * jobs[wj.jlo].s.n_bsdfs);
}

// The adjoint's reverse sweep over each slot's recorded path (one thread per
// slot): with R = the emission of the light the path ended on (per unit
// throughput), for vertex j from the last to the first
//   g = Y_j + sigma_j R,   dS/d(albedo_j) += beta_j g,   R = albedo_j g,
// where S is the pixel's film sum; the texel gradient is w * beta_j * g with w
// = dLoss/dS.  No factor is divided out, so albedo channels at 0 get their
// derivative too.
__global__ void __launch_bounds__(kBlock)
k_adjoint_sweep(const RJob* __restrict__ jobs, const WaveJobs* __restrict__ wjp, WaveBufs wb,
                GradArgs ga, uint32_t n) {
  __shared__ TexLds lds;
  tex_lds_init(lds);
  uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    WaveJobs const& wj = *wjp;
    RJob const& J = jobs[wj.jlo + job_of_slot(wj, i)];
    unsigned long long const item = wj.q0 + i - J.base;
    uint32_t const pixel = uint32_t(item % J.wh);
    float3 const w = f3(ga.d_film[3 * size_t(pixel)], ga.d_film[3 * size_t(pixel) + 1],
                        ga.d_film[3 * size_t(pixel) + 2]);
    float4 const e = wb.ctot[i];
    uint32_t cnt = __float_as_uint(e.w);
    cnt = cnt < wb.tape_depth ? cnt : wb.tape_depth;
    float3 R = ld3(e);
    for (uint32_t j = cnt; j-- > 0;) {
      size_t const ti = size_t(j) * wb.cap + i;
      float4 const a = wb.tape_ab[2 * ti], b = wb.tape_ab[2 * ti + 1];
      // the NEE term only where the vertex made a request (sign of the scale)
      bool const nee = (__float_as_uint(b.w) >> 31) != 0u;
      float3 const c = nee ? ld3(wb.tape_c[ti]) : f3(0.0f, 0.0f, 0.0f);
      float3 const g = add3(c, scl3(R, fabsf(b.w)));
      uint32_t const key = __float_as_uint(a.w);
      if (key != kEmptyKey) tex_add(&lds, ga.d_tex, key, mul3(w, mul3(ld3(a), g)));
      R = mul3(ld3(b), g);
    }
  }
  tex_lds_flush(lds, ga.d_tex);
}

// End of a bounce: the next queues become current; their padded prefix for
// the next bounce's virtual positions; counted queries.
__global__ void __launch_bounds__(512) k_prep(const WaveJobs* __restrict__ wjp, WaveCounters* ctr) {
  __shared__ uint32_t sc[512];
  __shared__ unsigned long long qsum;
  __shared__ uint32_t nsum;
  WaveJobs const& wj = *wjp;
  uint32_t const k = threadIdx.x, nseg = wj.nj * 8u;
  if (k == 0) {
    qsum = 0;
    nsum = 0;
  }
  __syncthreads();
  uint32_t done = 0, nx = 0;
  if (k < nseg) {
    uint32_t const a = 9u * (k >> 3) + (k & 7u);
    done = ctr->cur[k];
    nx = ctr->append[a].v;
    ctr->cur[k] = nx;
    ctr->append[a].v = 0u;
  }
  if (k < wj.nj) {
    done += ctr->append[9u * k + 8u].v;
    ctr->append[9u * k + 8u].v = 0u;
  }
  if (done) atomicAdd(&qsum, static_cast<unsigned long long>(done));
  if (nx) atomicAdd(&nsum, nx);
  uint32_t const tot = block_scan512(k < nseg ? pad_seg(nx) : 0u, sc);
  if (k < nseg) ctr->vext[k] = sc[k];
  if (k == 0) {
    ctr->vext[nseg] = tot;
    ctr->n_active = nsum;
    ctr->queries += qsum;
    for (uint32_t p = 0; p < kFeedParts; ++p) {
      ctr->work_extend[p].v = 0u;
      ctr->work_shadow[p].v = 0u;
    }
  }
}

__global__ void k_zero_counters(WaveCounters* ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr->queries = 0;
}

// film[p] += colour of each of p's samples in this wave, in sample order;
// blockIdx.y = local job.
__global__ void __launch_bounds__(kBlock)
k_accumulate(const RJob* __restrict__ jobs, const WaveJobs* __restrict__ wjp,
             const float4* __restrict__ ps) {
  WaveJobs const& wj = *wjp;
  uint32_t const l = blockIdx.y;
  if (l >= wj.nj) return;
  RJob const& J = jobs[wj.jlo + l];
  uint32_t const p = blockIdx.x * kBlock + threadIdx.x;
  uint32_t const wh = J.wh;
  if (p >= wh) return;
  uint32_t const s0 = wj.start[l];
  uint32_t const n = wj.start[l + 1] - s0;
  unsigned long long const q0 = wj.q0 + s0 - J.base;  // job item of slot s0
  unsigned long long q = q0 + (uint64_t(p) + wh - (q0 % wh)) % wh;
  if (q >= q0 + n) return;
  float* const film = J.film;
  float fx = film[3 * size_t(p)], fy = film[3 * size_t(p) + 1], fz = film[3 * size_t(p) + 2];
  for (; q < q0 + n; q += wh) {
    float4 const c = ps[4 * size_t(s0 + (q - q0)) + 3];
    fx = fx + c.x;
    fy = fy + c.y;
    fz = fz + c.z;
  }
  film[3 * size_t(p)] = fx;
  film[3 * size_t(p) + 1] = fy;
  film[3 * size_t(p) + 2] = fz;
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

// Workspace of one call, from pool_alloc or big_alloc, released on `st`.
struct DevMem {
  void* p = nullptr;
  hipStream_t st = nullptr;
  ~DevMem() { big_free(p, st); }
};

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Bytes of one lane's wave buffers; tape_depth > 0: with the adjoint's
// buffers (ctot and the vertex tape).
// Virtual queue positions beyond a wave's slots: each of at most 8 segments
// per job is padded to a multiple of kSegPad.
constexpr size_t kHitPad = size_t(8) * kMaxWaveJobs * kSegPad;

size_t wave_bytes(uint32_t cap, uint32_t tape_depth) {
  bool const replay = tape_depth > 0;
  size_t const f4 = align_up(size_t(cap) * 16);
  size_t const hb = align_up((size_t(cap) + kHitPad) * 16);
  size_t const qsz = align_up(size_t(cap) * 8 * 4);
  size_t const tape = replay ? align_up(size_t(cap) * tape_depth * 16) : 0;
  // path state 4, requests 4, smet 1, ctot 1 (replay); hits; queues; tape ab, c
  return f4 * (replay ? 10 : 9) + hb + 2 * qsz + align_up(sizeof(WaveCounters)) + 3 * tape;
}

// Lays one lane's wave buffers out in `mem` (wave_bytes(cap, tape_depth) bytes).
void carve_wave(void* mem, uint32_t cap, uint32_t tape_depth, WaveBufs* wb) {
  bool const replay = tape_depth > 0;
  size_t const f4 = align_up(size_t(cap) * 16);
  size_t const hb = align_up((size_t(cap) + kHitPad) * 16);
  size_t const qsz = align_up(size_t(cap) * 8 * 4);
  size_t const tape = replay ? align_up(size_t(cap) * tape_depth * 16) : 0;
  char* b = static_cast<char*>(mem);
  auto take = [&](size_t n) {
    char* r = b;
    b += n;
    return r;
  };
  wb->ps = reinterpret_cast<float4*>(take(4 * f4));
  wb->hit = reinterpret_cast<float4*>(take(hb));
  wb->rqs = reinterpret_cast<float4*>(take(4 * f4));
  wb->smet = reinterpret_cast<float4*>(take(f4));
  wb->ctot = replay ? reinterpret_cast<float4*>(take(f4)) : nullptr;

  wb->qa = reinterpret_cast<uint32_t*>(take(qsz));
  wb->qb = reinterpret_cast<uint32_t*>(take(qsz));
  wb->ctr = reinterpret_cast<WaveCounters*>(take(align_up(sizeof(WaveCounters))));
  wb->cap = cap;
  wb->tape_ab = replay ? reinterpret_cast<float4*>(take(2 * tape)) : nullptr;
  wb->tape_c = replay ? reinterpret_cast<float4*>(take(tape)) : nullptr;
  wb->tape_depth = tape_depth;
}

// One lane's wave buffers (big_alloc).
int alloc_wave(uint32_t cap, uint32_t tape_depth, hipStream_t st, DevMem* mem, WaveBufs* wb) {
  hipError_t const e = big_alloc(&mem->p, wave_bytes(cap, tape_depth), st);
  if (e != hipSuccess) return hip_fail(e, "big_alloc(wave)");
  mem->st = st;
  carve_wave(mem->p, cap, tape_depth, wb);
  return XYZ_OK;
}

// Internal streams (default priority) that alternate the waves of a render,
// created once per device.
#ifndef XYZ_RENDER_LANES
#define XYZ_RENDER_LANES 2
#endif
constexpr int kLanes = XYZ_RENDER_LANES;
// Three lanes ended a batched render with a host fault whose cause is not
// known; only the settings run by the test suite are accepted.
static_assert(kLanes >= 1 && kLanes <= 2, "XYZ_RENDER_LANES must be 1 or 2");
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

// After a fork onto internal streams: records an event on each lane and
// makes the caller's stream wait for it.  join() does so and reports errors;
// an armed guard that goes out of scope does so on the way out (error paths),
// so that memory released on the caller's stream afterwards is not in use.
struct LaneJoin {
  hipStream_t st = nullptr;
  const hipStream_t* ls = nullptr;
  const hipEvent_t* done = nullptr;
  int n = 0;
  void arm(hipStream_t s, const hipStream_t* lanes, const hipEvent_t* ev, int count) {
    st = s;
    ls = lanes;
    done = ev;
    n = count;
  }
  hipError_t join() {
    hipError_t e = hipSuccess;
    for (int i = 0; i < n; ++i) {
      hipError_t r = hipEventRecord(done[i], ls[i]);
      if (r == hipSuccess) r = hipStreamWaitEvent(st, done[i], 0);
      if (r != hipSuccess && e == hipSuccess) e = r;
    }
    n = 0;
    return e;
  }
  ~LaneJoin() {
    if (n > 0) (void)join();
  }
};

// Persistent workgroups per CU of the extension and shadow kernels.
#ifndef XYZ_TRACE_BLOCKS_PER_CU
#define XYZ_TRACE_BLOCKS_PER_CU 8
#endif

// Runs the bounce loop of one wave.  kReplay: adjoint replay pass.
template <bool kReplay>
int run_bounces(const RJob* jobs, const WaveJobs* wj, WaveBufs wb, uint32_t n, uint32_t depth,
                uint32_t stack_cap, GradArgs ga, hipStream_t st, bool guided = false) {
  int const ncu = num_cu();
  size_t const lds = size_t(stack_cap) * kBlock * sizeof(uint32_t);
  uint32_t const trace_grid =
      std::max(1u, std::min((n + kBlock - 1) / kBlock, uint32_t(ncu) * XYZ_TRACE_BLOCKS_PER_CU));
  uint32_t const shade_grid = std::max(1u, std::min((n + kBlock - 1) / kBlock, uint32_t(ncu) * 8));
  uint32_t* qc = wb.qa;
  uint32_t* qn = wb.qb;
  for (uint32_t b = 0; b < depth; ++b) {
    hipLaunchKernelGGL(k_extend, dim3(trace_grid), dim3(kBlock), lds, st, jobs, wj, wb, qc);
    if constexpr (!kReplay) {
      if (guided)
        hipLaunchKernelGGL((k_shade<false, true>), dim3(shade_grid), dim3(kBlock), 0, st, jobs, wj,
                           wb, qc, qn, ga);
      else
        hipLaunchKernelGGL((k_shade<false, false>), dim3(shade_grid), dim3(kBlock), 0, st, jobs,
                           wj, wb, qc, qn, ga);
    } else {
      hipLaunchKernelGGL((k_shade<true, false>), dim3(shade_grid), dim3(kBlock), 0, st, jobs, wj,
                         wb, qc, qn, ga);
    }
    hipLaunchKernelGGL(k_shadow<kReplay>, dim3(trace_grid), dim3(kBlock), lds, st, jobs, wj, wb, ga);
    hipLaunchKernelGGL(k_prep, dim3(1), dim3(512), 0, st, wj, wb.ctr);
    std::swap(qc, qn);
  }
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render launch");
  return XYZ_OK;
}

// One job of a render call, before planning.
struct JobSpec {
  RJob j;
  unsigned long long items;
  uint32_t max_depth;
};

// The waves of a render: the jobs' work items laid end to end and cut into
// waves of up to `cap` path slots and kMaxWaveJobs jobs.
struct RenderPlan {
  std::vector<JobSpec> jobs;
  std::vector<WaveJobs> waves;
  std::vector<uint32_t> wdepth, wstack, wmaxwh;
  uint32_t cap = 0, tape_depth = 0;
  int lanes = 1;
  bool backward = false;
};

void plan_render(std::vector<JobSpec> const& js, bool backward, RenderPlan* p) {
  p->jobs.clear();
  for (JobSpec const& x : js)
    if (x.items > 0) p->jobs.push_back(x);
  p->backward = backward;
  std::vector<JobSpec>& jobs = p->jobs;
  if (jobs.empty()) return;
  unsigned long long total = 0;
  for (JobSpec& x : jobs) {
    x.j.base = total;
    total += x.items;
  }
  // the adjoint's vertex tape: 48 bytes per slot and shading vertex, at most
  // kTapeBudget bytes per lane
  unsigned long long slot_cap = kWaveCap;
  if (backward) {
    for (JobSpec const& x : jobs) p->tape_depth = std::max(p->tape_depth, x.max_depth);
    slot_cap = std::max<unsigned long long>(
        kBlock, std::min<unsigned long long>(kWaveCap, kTapeBudget / (48ull * p->tape_depth)));
  }
  uint32_t const cap = uint32_t(std::min<unsigned long long>(total, slot_cap));
  p->cap = cap;
  unsigned long long q = 0;
  size_t j = 0;
  while (q < total) {
    WaveJobs w;
    std::memset(&w, 0, sizeof(w));
    w.q0 = q;
    w.jlo = uint32_t(j);
    uint32_t slots = 0, depth = 0, stack = 1, maxwh = 0;
    while (slots < cap && j < jobs.size() && w.nj < kMaxWaveJobs) {
      unsigned long long const end = jobs[j].j.base + jobs[j].items;
      uint32_t const take = uint32_t(std::min<unsigned long long>(cap - slots, end - q));
      w.start[w.nj++] = slots;
      slots += take;
      q += take;
      depth = std::max(depth, jobs[j].max_depth);
      stack = std::max(stack, jobs[j].j.s.stack_cap);
      maxwh = std::max(maxwh, jobs[j].j.wh);
      if (q == end) ++j; else break;
    }
    w.start[w.nj] = slots;
    p->waves.push_back(w);
    p->wdepth.push_back(depth);
    p->wstack.push_back(stack);
    p->wmaxwh.push_back(maxwh);
  }
  p->lanes = p->waves.size() > 1 ? kLanes : 1;
}

// Bytes of a render's device tables (jobs, then waves).
size_t tables_bytes(const RenderPlan& p) {
  return align_up(p.jobs.size() * sizeof(RJob)) + p.waves.size() * sizeof(WaveJobs);
}

// Issues a planned render on `st` (and the internal lane streams, joined back
// into `st` before the return): tables at `tables`, lane i's wave buffers wb[i].
int enqueue_render(const RenderPlan& p, void* tables, WaveBufs* wb, hipStream_t st, GradArgs ga,
                   bool guided, EventSet* ext_ev = nullptr) {
  RJob* const jobs_dev = static_cast<RJob*>(tables);
  WaveJobs* const waves_dev =
      reinterpret_cast<WaveJobs*>(static_cast<char*>(tables) + align_up(p.jobs.size() * sizeof(RJob)));
  for (size_t i = 0; i < p.jobs.size(); ++i)
    hipLaunchKernelGGL(k_put_job, dim3(1), dim3(1), 0, st, p.jobs[i].j, jobs_dev + i);
  for (size_t k = 0; k < p.waves.size(); ++k)
    hipLaunchKernelGGL(k_put_wave, dim3(1), dim3(1), 0, st, p.waves[k], waves_dev + k);
  int const lanes = p.lanes;
  for (int i = 0; i < lanes; ++i)
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, wb[i].ctr);
  hipStream_t ls[kLanes];
  for (int i = 0; i < kLanes; ++i) ls[i] = st;
  int rc = XYZ_OK;
  hipError_t e = hipSuccess;
  // events of the fork, the per-wave film order and the join; a caller that
  // captures the render into a graph keeps them until the capture has ended
  EventSet own_ev;
  EventSet& ev = ext_ev ? *ext_ev : own_ev;
  // once the lanes are forked, every return joins them back into `st`
  LaneJoin join;
  if (lanes > 1) {
    rc = lane_streams(ls);
    if (rc != XYZ_OK) return rc;
    if (ev.e.size() < size_t(2 * kLanes + 1)) rc = ev.create(2 * kLanes + 1);
    if (rc != XYZ_OK) return rc;
    hipEvent_t const fork = ev.e[2 * kLanes];
    e = hipEventRecord(fork, st);
    if (e != hipSuccess) return hip_fail(e, "render fork");
    join.arm(st, ls, &ev.e[kLanes], lanes);
    for (int i = 0; i < lanes && e == hipSuccess; ++i) e = hipStreamWaitEvent(ls[i], fork, 0);
    if (e != hipSuccess) return hip_fail(e, "render fork");
  }
  GradArgs none{nullptr, nullptr, nullptr};
  for (size_t k = 0; k < p.waves.size(); ++k) {
    int const i = int(k % size_t(lanes));
    hipStream_t const s = ls[i];
    WaveJobs const* const wj = waves_dev + k;
    uint32_t const n = p.waves[k].start[p.waves[k].nj];
    uint32_t const grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_init_wave, dim3(1), dim3(512), 0, s, wj, wb[i].ctr);
    hipLaunchKernelGGL(k_camera, dim3(grid), dim3(kBlock), 0, s, jobs_dev, wj, wb[i], n);
    if (p.backward) {
      // the paths are traced again from their seeds, each slot recording its
      // vertices, then swept in reverse
      rc = run_bounces<true>(jobs_dev, wj, wb[i], n, p.wdepth[k], p.wstack[k], ga, s);
      if (rc != XYZ_OK) return rc;
      hipLaunchKernelGGL(k_adjoint_sweep, dim3(grid), dim3(kBlock), 0, s, jobs_dev, wj, wb[i], ga, n);
      continue;
    }
    rc = run_bounces<false>(jobs_dev, wj, wb[i], n, p.wdepth[k], p.wstack[k], none, s, guided);
    if (rc != XYZ_OK) return rc;
    // wave k is added to the films after wave k - 1 (sample order per pixel)
    if (lanes > 1 && k > 0) {
      e = hipStreamWaitEvent(s, ev.e[(k - 1) % size_t(lanes)], 0);
      if (e != hipSuccess) return hip_fail(e, "render wait");
    }
    hipLaunchKernelGGL(k_accumulate, dim3((p.wmaxwh[k] + kBlock - 1) / kBlock, p.waves[k].nj),
                       dim3(kBlock), 0, s, jobs_dev, wj, wb[i].ps);
    if (lanes > 1) {
      e = hipEventRecord(ev.e[i], s);
      if (e != hipSuccess) return hip_fail(e, "render record");
    }
  }
  if (lanes > 1) {  // join: the caller's stream waits for every lane
    e = join.join();
    if (e != hipSuccess) return hip_fail(e, "render join");
  }
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render");
  return XYZ_OK;
}

// Closest-hit queries counted by a finished render's lanes (synchronises `st`).
int render_queries(const WaveBufs* wb, int lanes, hipStream_t st, unsigned long long* queries) {
  unsigned long long q[kLanes] = {};
  hipError_t e = hipSuccess;
  for (int i = 0; i < lanes && e == hipSuccess; ++i)
    e = hipMemcpyAsync(&q[i], &wb[i].ctr->queries, sizeof(q[i]), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "render stats");
  *queries = 0;
  for (int i = 0; i < lanes; ++i) *queries += q[i];
  return XYZ_OK;
}

// Renders jobs with the wavefront kernels (see RenderPlan); consecutive waves
// alternate between internal streams, so that the tail of one wave (a few
// long paths) overlaps the full bounces of the next.  Each job's film
// receives its waves in sample order (an event per wave).  Workspace comes
// from the stream-ordered pool.  backward: the adjoint of the single job (ga).
// queries: total closest-hit queries (may be NULL; then nothing waits on the
// device).
int render_jobs(std::vector<JobSpec> js, hipStream_t st, bool backward, GradArgs ga,
                unsigned long long* queries, bool guided = false) {
  if (queries) *queries = 0;
  RenderPlan p;
  plan_render(js, backward, &p);
  if (p.jobs.empty()) return XYZ_OK;
  DevMem tables;
  hipError_t e = pool_alloc(&tables.p, tables_bytes(p), st);
  if (e != hipSuccess) return hip_fail(e, "pool_alloc(render tables)");
  tables.st = st;
  DevMem mem[kLanes];
  WaveBufs wb[kLanes];
  for (int i = 0; i < p.lanes; ++i) {
    int const rc = alloc_wave(p.cap, p.tape_depth, st, &mem[i], &wb[i]);
    if (rc != XYZ_OK) return rc;
  }
  int rc = enqueue_render(p, tables.p, wb, st, ga, guided);
  if (rc != XYZ_OK) return rc;
  if (queries) return render_queries(wb, p.lanes, st, queries);
  return XYZ_OK;
}

// imp: the importon map of importance-driven path tracing, or NULL.
int render_impl(xyz_scene* sc, const xyz_camera* cam, const xyz_render_params* prm, float* film,
                xyz_render_stats* stats, hipStream_t st, bool backward, const float* d_film,
                float* d_tex, float* d_emit, const xyz_importon_map* imp = nullptr) {
  if (!sc) {
    set_error("render: NULL scene");
    return XYZ_E_INVALID;
  }
  JobSpec js;
  std::memset(&js, 0, sizeof(js));
  int rc = make_dev_camera(cam, &js.j.cam);
  if (rc != XYZ_OK) return rc;
  rc = make_pt_const(prm, &js.j.c);
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
  if (uint64_t(prm->sample_offset) + prm->spp >= (1ull << 32)) {
    set_error("render: sample index overflow");
    return XYZ_E_UNSUPPORTED;
  }
  js.j.s = sc->dev;
  js.j.film = film;
  js.j.wh = uint32_t(wh);
  js.j.sample0 = prm->sample_offset;
  js.items = wh * prm->spp;
  js.max_depth = prm->max_depth;
  if (imp) {
    if (imp->scene != sc) {
      set_error("render_idpt: the importon map belongs to another scene");
      return XYZ_E_INVALID;
    }
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != imp->device) {
      set_error("render_idpt: the importon map lives on another device");
      return XYZ_E_INVALID;
    }
    js.j.imp = imp->dev;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (stats) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
  }
  unsigned long long q = 0;
  rc = render_jobs({js}, st, backward, GradArgs{d_film, d_tex, d_emit}, stats ? &q : nullptr,
                   imp != nullptr);
  if (rc != XYZ_OK) return rc;
  if (stats) {
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    stats->seconds = double(ms) * 1e-3;
    stats->paths = js.items;
    stats->queries = q;
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
  __shared__ uint8_t owners_all[1024];
  unsigned long long* const keys = keys_all + (threadIdx.x & ~63u);
  uint8_t* const owners = owners_all + (threadIdx.x & ~63u);
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
    bool const fin = coop_step<false, true, 1>(jb.s, st, act, stk, blockDim.x, keys, owners);
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
  t.engine = 1;
  return t;
}

// A captured batch and the device memory its nodes address.  Replays of one
// entry share that memory, so each replay waits for the previous one (`done`)
// on whatever stream it is issued.
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipEvent_t done = nullptr;
  // engine 0: the job table and the work counter
  void* jobs_dev = nullptr;
  uint32_t* work = nullptr;
  // engine 1: the render tables and each lane's wave buffers
  void* tables = nullptr;
  void* wave_mem[kLanes] = {};
  WaveBufs wb[kLanes];
  int lanes = 0;
  uint64_t last_use = 0;
};

std::mutex g_graph_mu;
std::map<std::string, GraphEntry> g_graphs;
uint64_t g_graph_clock = 0;
hipStream_t g_capture_stream = nullptr;
// Cached graphs at most (a shared-wave entry holds its wave buffers, up to
// ~34 GB for 2^26-slot waves on two lanes).
constexpr size_t kMaxCachedGraphs = 2;

void free_graph_entry(GraphEntry& ge) {
  if (ge.done) (void)hipEventSynchronize(ge.done);  // its last replay has finished
  if (ge.exec) (void)hipGraphExecDestroy(ge.exec);
  if (ge.graph) (void)hipGraphDestroy(ge.graph);
  if (ge.done) (void)hipEventDestroy(ge.done);
  if (ge.jobs_dev) (void)hipFree(ge.jobs_dev);
  if (ge.work) (void)hipFree(ge.work);
  if (ge.tables) (void)hipFree(ge.tables);
  for (void* m : ge.wave_mem)
    if (m) (void)hipFree(m);
  ge = GraphEntry();
}

// Makes room for one more entry (call with g_graph_mu held): the least recently
// used entries go, each after its last replay.
void evict_graphs() {
  while (g_graphs.size() >= kMaxCachedGraphs) {
    auto victim = g_graphs.begin();
    for (auto it = g_graphs.begin(); it != g_graphs.end(); ++it)
      if (it->second.last_use < victim->second.last_use) victim = it;
    free_graph_entry(victim->second);
    g_graphs.erase(victim);
  }
}

// Replays a cached entry on `st` after its previous replay.
int replay_graph(GraphEntry& ge, hipStream_t st) {
  hipError_t e = hipStreamWaitEvent(st, ge.done, 0);
  if (e == hipSuccess) e = hipGraphLaunch(ge.exec, st);
  if (e == hipSuccess) e = hipEventRecord(ge.done, st);
  ge.last_use = ++g_graph_clock;
  return e == hipSuccess ? XYZ_OK : hip_fail(e, "graph replay");
}

int capture_stream(hipStream_t* out) {
  if (!g_capture_stream) {
    hipError_t const e = hipStreamCreateWithFlags(&g_capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
  }
  *out = g_capture_stream;
  return XYZ_OK;
}

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
  if (t.engine > 1) {
    set_error("render_batch: unknown engine");
    return XYZ_E_INVALID;
  }
  if (t.engine == 1) return XYZ_OK;
  if (t.tile_pixels == 0 || t.block_threads == 0 || t.block_threads > 1024 ||
      (t.block_threads % 64) != 0 || t.blocks_per_cu == 0 || t.blocks_per_cu > 64) {
    set_error("render_batch: invalid tuning");
    return XYZ_E_INVALID;
  }
  return XYZ_OK;
}

// Engine 1: all jobs through the wavefront kernels in shared waves.
int render_batch_wavefront(uint32_t n, xyz_scene* const* scenes, const xyz_camera* cams,
                           const xyz_render_params* prm, float* const* films,
                           xyz_render_stats* stats, hipStream_t st, bool use_graph) {
  std::vector<JobSpec> js(n);
  unsigned long long paths = 0;
  for (uint32_t i = 0; i < n; ++i) {
    JobSpec& x = js[i];
    std::memset(&x, 0, sizeof(x));
    if (!scenes[i] || !films[i]) {
      set_error("render_batch: NULL scene or film in job " + std::to_string(i));
      return XYZ_E_INVALID;
    }
    int rc = make_dev_camera(&cams[i], &x.j.cam);
    if (rc != XYZ_OK) return rc;
    rc = make_pt_const(&prm[i], &x.j.c);
    if (rc != XYZ_OK) return rc;
    uint64_t const wh = uint64_t(cams[i].width) * cams[i].height;
    if (wh >= (1ull << 32) || uint64_t(prm[i].sample_offset) + prm[i].spp >= (1ull << 32)) {
      set_error("render_batch: image too large or sample index overflow in job " + std::to_string(i));
      return XYZ_E_UNSUPPORTED;
    }
    x.j.s = scenes[i]->dev;
    x.j.film = films[i];
    x.j.wh = uint32_t(wh);
    x.j.sample0 = prm[i].sample_offset;
    x.items = wh * prm[i].spp;
    x.max_depth = prm[i].max_depth;
    paths += x.items;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (stats) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
  }
  unsigned long long q = 0;
  int rc = XYZ_OK;
  if (use_graph) {
    // one capture per job table: the render's tables and wave buffers are
    // allocated before the capture and stay with the cached graph; a repeated
    // batch is one hipGraphLaunch
    std::string key("engine1:");
    for (JobSpec const& x : js) {
      key.append(reinterpret_cast<const char*>(&x.j), sizeof(x.j));
      key.append(reinterpret_cast<const char*>(&x.items), sizeof(x.items));
      key.append(reinterpret_cast<const char*>(&x.max_depth), sizeof(x.max_depth));
    }
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find(key);
    if (it == g_graphs.end()) {
      evict_graphs();
      RenderPlan p;
      plan_render(js, false, &p);
      GraphEntry ge;
      ge.lanes = p.lanes;
      hipStream_t cs = nullptr;
      rc = capture_stream(&cs);
      if (rc != XYZ_OK) return rc;
      hipError_t e = hipMalloc(&ge.tables, std::max<size_t>(tables_bytes(p), 256));
      for (int i = 0; i < p.lanes && e == hipSuccess; ++i) {
        e = hipMalloc(&ge.wave_mem[i], wave_bytes(p.cap, 0));
        if (e == hipSuccess) carve_wave(ge.wave_mem[i], p.cap, 0, &ge.wb[i]);
      }
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ge.done, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ge.done, cs);
      // the work of earlier calls on the caller's stream is ordered by the
      // replay; the capture stream itself has nothing in flight here
      if (e == hipSuccess) e = hipStreamSynchronize(cs);
      if (e != hipSuccess) {
        free_graph_entry(ge);
        return hip_fail(e, "render_batch graph set-up");
      }
      {
        EventSet ev;  // kept until the capture has ended
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess)
          rc = enqueue_render(p, ge.tables, ge.wb, cs, GradArgs{nullptr, nullptr, nullptr}, false, &ev);
        hipError_t const ee = e == hipSuccess ? hipStreamEndCapture(cs, &ge.graph) : e;
        if (e == hipSuccess) e = ee;
      }
      if (rc == XYZ_OK && e == hipSuccess) e = hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0);
      if (rc != XYZ_OK || e != hipSuccess) {
        free_graph_entry(ge);
        return rc != XYZ_OK ? rc : hip_fail(e, "render_batch capture");
      }
      it = g_graphs.emplace(key, ge).first;
    }
    rc = replay_graph(it->second, st);
    if (rc == XYZ_OK && stats) rc = render_queries(it->second.wb, it->second.lanes, st, &q);
  } else {
    rc = render_jobs(std::move(js), st, false, GradArgs{nullptr, nullptr, nullptr},
                     stats ? &q : nullptr);
  }
  if (rc != XYZ_OK) return rc;
  if (stats) {
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    stats->seconds = double(ms) * 1e-3;
    stats->paths = paths;
    stats->queries = q;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
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
  if (t.engine == 1)
    return render_batch_wavefront(n, scenes, cams, prm, films, stats, st, t.use_graph != 0);
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
    key.insert(0, "engine0:");
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find(key);
    if (it == g_graphs.end()) {
      evict_graphs();
      GraphEntry ge;
      size_t const jb = plan.jobs.size() * sizeof(BatchJob);
      hipStream_t cs = nullptr;
      rc = capture_stream(&cs);
      if (rc != XYZ_OK) return rc;
      hipError_t e = hipMalloc(&ge.jobs_dev, jb);
      if (e == hipSuccess) e = hipMemcpy(ge.jobs_dev, plan.jobs.data(), jb, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&ge.work), sizeof(uint32_t));
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ge.done, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ge.done, cs);
      if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
      if (e != hipSuccess) {
        free_graph_entry(ge);
        return hip_fail(e, "render_batch graph set-up");
      }
      rc = launch_batch_raw(plan, t, ge.jobs_dev, ge.work, cs);
      e = hipStreamEndCapture(cs, &ge.graph);
      if (rc == XYZ_OK && e == hipSuccess) e = hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0);
      if (rc != XYZ_OK || e != hipSuccess) {
        free_graph_entry(ge);
        return rc != XYZ_OK ? rc : hip_fail(e, "render_batch capture");
      }
      it = g_graphs.emplace(key, ge).first;
    }
    if (stats) (void)hipEventRecord(e0, st);
    rc = replay_graph(it->second, st);
    if (rc != XYZ_OK) return rc;
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
  const xyz_render_params* const full = prm;
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
        t.engine = 0;
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
  // the best engine-0 tuning against engine 1, on more samples: engine 1 gains
  // from large waves, which a 4-sample trial does not form
  std::vector<xyz_render_params> trial16(full, full + n);
  for (auto& p : trial16) p.spp = std::max(1u, std::min(p.spp, 16u));
  xyz_batch_tuning wave = *best;
  wave.engine = 1;
  double secs[2] = {0.0, 0.0};
  for (int k = 0; k < 2; ++k) {
    xyz_batch_tuning const t = k == 0 ? *best : wave;
    int rc = render_batch_impl(n, scenes, cams, trial16.data(), films, &t, nullptr, st);  // warm-up
    if (rc != XYZ_OK) return rc;
    xyz_render_stats s;
    rc = render_batch_impl(n, scenes, cams, trial16.data(), films, &t, &s, st);
    if (rc != XYZ_OK) return rc;
    secs[k] = s.seconds;
  }
  if (secs[1] < secs[0]) best->engine = 1;
  best->use_graph = 1;
  return zero_films(n, cams, films, st);
}

// ================================================================ BDPT
// Work items are (sample, pixel) pairs, pixel-major within a sample so that
// neighbouring threads trace neighbouring primary rays.  A chunk of samples
// writes each item's own colour to a buffer, then one thread per pixel adds
// them to the film in sample order (the order of BidirectionalPathTracer::run,
// integrator.cpp:1059-1081), so own colours are summed deterministically;
// light-tracing contributions land on other pixels through float atomics.
#ifndef XYZ_BDPT_WAVES_PER_EU
#define XYZ_BDPT_WAVES_PER_EU 8
#endif
template <bool kCount>
__global__ void __launch_bounds__(kBlock)
#if XYZ_BDPT_WAVES_PER_EU > 0
    __attribute__((amdgpu_waves_per_eu(XYZ_BDPT_WAVES_PER_EU)))
#endif
    k_bdpt(DevScene s, DevCamera cam, BdptConst bc,
                                                 float4* colours, float* splat, uint32_t n,
                                                 uint32_t sample0, uint64_t items,
                                                 unsigned long long* queries) {
  extern __shared__ uint32_t lds_stack[];
  uint32_t* const stk = lds_stack + threadIdx.x;
  uint32_t q = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < items;
       i += uint64_t(gridDim.x) * blockDim.x) {
    uint32_t const smp = sample0 + uint32_t(i / n);
    uint32_t const p = uint32_t(i % n);
    uint32_t const px = p % cam.width, py = p / cam.width;
    float3 const c = bdpt_sample<kCount>(s, cam, bc, px, py, smp, stk, blockDim.x, splat, &q);
    colours[i] = make_float4(c.x, c.y, c.z, 0.0f);
  }
  if (kCount && q) atomicAdd(queries, static_cast<unsigned long long>(q));
}

// Queued form of the same estimator, in three passes per chunk of items:
//   1. k_bdpt_paths: a thread per item builds the two subpaths (the wave's
//      lanes step their subpaths together and trace each round's rays with
//      the cooperative traversal), stores their vertices, and writes the
//      visibility rays of all its connections to a queue (a count pass over
//      the connections first, then one atomic per wave for the wave's
//      entries);
//   2. the cooperative closest-hit kernel over the queue;
//   3. k_bdpt_connect: a thread per item reloads its subpaths and evaluates
//      its connections in the reference's order, reading the queue's hits in
//      the order the rays were written.
// The connections' queries are the bulk of the estimator's work; traced by one
// thread each they diverge with the scene's size.  xyz_render_bdpt takes the
// queued form for scenes of at least XYZ_BDPT_QUEUED_MIN_REFS kd leaf
// references (measured crossover between 2.2e5 and 4.6e5 references on the
// batch workload's scenes at 256x256, 16 spp), unless the caller names the
// form (params->engine).
#ifndef XYZ_BDPT_QUEUED_MIN_REFS
#define XYZ_BDPT_QUEUED_MIN_REFS 300000
#endif

// Most connection queries of one sample: kmax - 2 for light tracing and
// sum over t = 2 .. kmax - 1 of kmax - t for the connections.
__host__ __device__ __forceinline__ uint32_t bdpt_max_queries(uint32_t kmax) {
  return kmax < 2 ? 0u : (kmax - 2) * (kmax + 1) / 2;
}

template <bool kCount>
__global__ void __launch_bounds__(kBlock)
#if XYZ_BDPT_WAVES_PER_EU > 0
    __attribute__((amdgpu_waves_per_eu(XYZ_BDPT_WAVES_PER_EU)))
#endif
    k_bdpt_paths(DevScene s, DevCamera cam, BdptConst bc, float4* __restrict__ verts,
                 uint32_t* __restrict__ lens, uint32_t* __restrict__ qoff, float* __restrict__ qorg,
                 float* __restrict__ qdir, uint32_t* __restrict__ qcount, uint32_t qcap,
                 uint32_t n, uint32_t sample0, uint32_t items, unsigned long long* queries) {
  extern __shared__ uint32_t lds_stack[];
  __shared__ unsigned long long keys[2 * kBlock];  // coop_step's tables, per wave
  uint32_t* const stk = lds_stack + threadIdx.x;
  uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
  bool const act = i < items;
  uint32_t const kmax = bc.c.kmax;
  uint32_t nq = 0, q = 0;
  BVert L[kBdptMaxDepth + 3], E[kBdptMaxDepth + 3];
  {
    uint32_t const smp = sample0 + (act ? i / n : 0u);
    uint32_t const p = act ? i % n : 0u;
    bdpt_build_wave<kCount>(s, cam, bc, act, p % cam.width, p / cam.width, smp, stk, kBlock,
                            keys + (threadIdx.x & ~63u),
                            reinterpret_cast<uint8_t*>(keys + kBlock + (threadIdx.x & ~63u)), L, E, &q);
  }
  if (act) {
    BdptCountQueries c{0};
    (void)bdpt_connect(s, cam, bc, L, E, nullptr, c);
    nq = c.n;
  }
  // this item's queue entries: consecutive, after the wave's earlier items
  uint32_t excl, total;
  wave_prefix(nq, &excl, &total);
  uint32_t base = 0;
  if (__lane_id() == 0 && total > 0) base = atomicAdd(qcount, total);
  base = uint32_t(__builtin_amdgcn_readfirstlane(base));
  if (act) {
    uint32_t const off = base + excl;
    qoff[i] = off;
    size_t const fs = items;  // distance between an entry's fields
    uint32_t const entries = kmax + 2;
    uint32_t const nl = bdpt_path_len(L, kmax, 1), ne = bdpt_path_len(E, kmax, 2);
    lens[i] = nl | (ne << 16);
    for (uint32_t k = 0; k <= nl; ++k) bvert_store(verts + size_t(k) * 4 * fs + i, fs, L[k]);
    for (uint32_t k = 0; k <= ne; ++k)
      bvert_store(verts + size_t(entries + k) * 4 * fs + i, fs, E[k]);
    BdptRecordQueries r{qorg, qdir, off, qcap};
    (void)bdpt_connect(s, cam, bc, L, E, nullptr, r);
  }
  if (kCount) {
    unsigned long long v = q + nq;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (__lane_id() == 0 && v) atomicAdd(queries, v);
  }
}

__global__ void __launch_bounds__(kBlock)
#if XYZ_BDPT_WAVES_PER_EU > 0
    __attribute__((amdgpu_waves_per_eu(XYZ_BDPT_WAVES_PER_EU)))
#endif
    k_bdpt_connect(DevScene s, DevCamera cam, BdptConst bc, const float4* __restrict__ verts,
                   const uint32_t* __restrict__ lens, const uint32_t* __restrict__ qoff,
                   const int32_t* __restrict__ qprim, const float* __restrict__ qt,
                   float4* __restrict__ colours, float* splat, uint32_t items) {
  uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= items) return;
  uint32_t const kmax = bc.c.kmax;
  BVert L[kBdptMaxDepth + 3], E[kBdptMaxDepth + 3];
  for (uint32_t k = 0; k < kmax + 2; ++k) {
    bvert_clear(L[k]);
    bvert_clear(E[k]);
  }
  size_t const fs = items;
  uint32_t const entries = kmax + 2;
  uint32_t const ln = lens[i];
  uint32_t const nl = ln & 0xFFFFu, ne = ln >> 16;
  for (uint32_t k = 0; k <= nl; ++k) bvert_load(verts + size_t(k) * 4 * fs + i, fs, L[k]);
  for (uint32_t k = 0; k <= ne; ++k) bvert_load(verts + size_t(entries + k) * 4 * fs + i, fs, E[k]);
  BdptReplayQueries r{qprim, qt, qoff[i]};
  float3 const c = bdpt_connect(s, cam, bc, L, E, splat, r);
  colours[i] = make_float4(c.x, c.y, c.z, 0.0f);
}

__global__ void __launch_bounds__(kBlock) k_bdpt_sum(float* film, const float4* colours,
                                                     uint32_t n, uint32_t count) {
  uint32_t const p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  float* const dst = film + 3 * size_t(p);
  float x = dst[0], y = dst[1], z = dst[2];
  for (uint32_t k = 0; k < count; ++k) {
    float4 const c = colours[size_t(k) * n + p];
    x = x + c.x;
    y = y + c.y;
    z = z + c.z;
  }
  dst[0] = x;
  dst[1] = y;
  dst[2] = z;
}

__global__ void __launch_bounds__(kBlock) k_add_into(float* dst, const float* src, size_t n) {
  size_t const i = size_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) dst[i] = dst[i] + src[i];
}

__global__ void k_zero_u64(unsigned long long* p) {
  if (threadIdx.x == 0) *p = 0ull;
}
__global__ void k_zero_u32x(uint32_t* p) {
  if (threadIdx.x == 0) *p = 0u;
}

int render_bdpt_impl(xyz_scene* sc, const xyz_camera* cam, const xyz_render_params* prm,
                     float* film, float* splat, xyz_render_stats* stats, hipStream_t st) {
  if (!sc || !film) {
    set_error("render_bdpt: NULL scene or film");
    return XYZ_E_INVALID;
  }
  DevCamera dc;
  int rc = make_dev_camera(cam, &dc);
  if (rc != XYZ_OK) return rc;
  BdptConst bc;
  rc = make_pt_const(prm, &bc.c);
  if (rc != XYZ_OK) return rc;
  if (prm->max_depth > kBdptMaxDepth) {
    set_error("render_bdpt: max_depth must be <= 16");
    return XYZ_E_UNSUPPORTED;
  }
  if (prm->engine > 2) {
    set_error("render_bdpt: engine must be 0 (by scene size), 1 (per-thread) or 2 (queued)");
    return XYZ_E_INVALID;
  }
  bool const queued =
      prm->engine == 2 || (prm->engine == 0 && sc->info.n_leaf_refs >= XYZ_BDPT_QUEUED_MIN_REFS);
  uint64_t const wh = uint64_t(cam->width) * cam->height;
  if (wh >= (1ull << 31)) {
    set_error("render_bdpt: image too large");
    return XYZ_E_UNSUPPORTED;
  }
  if (uint64_t(prm->sample_offset) + prm->spp >= (1ull << 32)) {
    set_error("render_bdpt: sample index overflow");
    return XYZ_E_UNSUPPORTED;
  }
  // GetConstFactor() = (f / F)^2 / A (thinlenscamera.hpp:83-85); a pinhole given
  // by its field of view takes its lens-film distance V as f
  float const f = (cam->focal_length > 0.0f) ? cam->focal_length : dc.v;
  float const r = f / dc.film;
  bc.C = (r * r) / dc.aspect;
  bc.inv_area = sc->dev.n_lights ? 1.0f / sc->dev.light_area : 0.0f;
  if (prm->spp == 0) return XYZ_OK;
  DevMem own_splat, qbuf;
  own_splat.st = qbuf.st = st;
  float* sp = splat;
  hipError_t e = hipSuccess;
  if (!sp) {
    e = pool_alloc(&own_splat.p, wh * 3 * sizeof(float), st);
    if (e != hipSuccess) return hip_fail(e, "render_bdpt: splat buffer");
    sp = static_cast<float*>(own_splat.p);
    hipLaunchKernelGGL(k_fill_zero, dim3(uint32_t((wh * 3 + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       st, sp, size_t(wh * 3));
  }
  unsigned long long* qd = nullptr;
  if (stats) {
    e = pool_alloc(&qbuf.p, sizeof(unsigned long long), st);
    if (e != hipSuccess) return hip_fail(e, "render_bdpt: counter");
    qd = static_cast<unsigned long long*>(qbuf.p);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, qd);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (stats) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
  }
  // samples in chunks of at most 2^24 (sample, pixel) items (256 MiB of colours);
  // the queued form also keeps each item's subpaths and its worst-case queue
  // entries, within about 6 GiB
  uint32_t const n = uint32_t(wh);
  uint64_t chunk_items = 1ull << 24;
  uint32_t const kmax = bc.c.kmax;
  uint32_t const wmax = bdpt_max_queries(kmax);
  size_t const item_bytes = size_t(kmax + 2) * 2 * 4 * sizeof(float4) + 2 * sizeof(uint32_t) +
                            size_t(wmax) * (6 * sizeof(float) + sizeof(int32_t) + sizeof(float));
  if (queued) chunk_items = std::min<uint64_t>(chunk_items, (6ull << 30) / item_bytes);
  uint32_t const per_chunk = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(prm->spp, chunk_items / wh)));
  if (queued && uint64_t(per_chunk) * wh * wmax >= (1ull << 32)) {
    set_error("render_bdpt: image too large for the queued connections");
    return XYZ_E_UNSUPPORTED;
  }
  DevMem colbuf;
  colbuf.st = st;
  e = big_alloc(&colbuf.p, size_t(per_chunk) * wh * sizeof(float4), st);
  if (e != hipSuccess) return hip_fail(e, "render_bdpt: colour buffer");
  float4* const cols = static_cast<float4*>(colbuf.p);
  size_t const lds = size_t(sc->dev.stack_cap) * kBlock * sizeof(uint32_t);
  size_t const cap = queued ? size_t(per_chunk) * wh : 0;  // items per chunk
  size_t const qcap = cap * wmax;                           // queue entries per chunk
  // the queued form's buffers, carved from one block: subpath vertices, path
  // lengths, queue offsets, query origins, directions, hit ids and distances,
  // and the queue counter
  size_t const qn = std::max<size_t>(1, qcap);
  size_t const qsizes[8] = {cap * size_t(kmax + 2) * 2 * 4 * sizeof(float4), cap * sizeof(uint32_t),
                            cap * sizeof(uint32_t), qn * 3 * sizeof(float), qn * 3 * sizeof(float),
                            qn * sizeof(int32_t), qn * sizeof(float), sizeof(uint32_t)};
  char* qparts[8] = {};
  DevMem qmem;
  qmem.st = st;
  if (queued) {
    size_t total = 0;
    for (size_t b : qsizes) total += align_up(b);
    e = big_alloc(&qmem.p, total, st);
    if (e != hipSuccess) return hip_fail(e, "render_bdpt: path and query buffers");
    char* at = static_cast<char*>(qmem.p);
    for (int i = 0; i < 8; ++i) {
      qparts[i] = at;
      at += align_up(qsizes[i]);
    }
  }
  float4* const verts = reinterpret_cast<float4*>(qparts[0]);
  uint32_t* const lens = reinterpret_cast<uint32_t*>(qparts[1]);
  uint32_t* const qoff = reinterpret_cast<uint32_t*>(qparts[2]);
  float* const qorg = reinterpret_cast<float*>(qparts[3]);
  float* const qdir = reinterpret_cast<float*>(qparts[4]);
  int32_t* const qprim = reinterpret_cast<int32_t*>(qparts[5]);
  float* const qt = reinterpret_cast<float*>(qparts[6]);
  uint32_t* const qcount = reinterpret_cast<uint32_t*>(qparts[7]);
  for (uint32_t done = 0; done < prm->spp; done += per_chunk) {
    uint32_t const count = std::min(per_chunk, prm->spp - done);
    uint64_t const items = uint64_t(count) * wh;
    if (queued) {
      uint32_t const grid = uint32_t((items + kBlock - 1) / kBlock);
      hipLaunchKernelGGL(k_zero_u32x, dim3(1), dim3(64), 0, st, qcount);
      if (stats) {
        hipLaunchKernelGGL(k_bdpt_paths<true>, dim3(grid), dim3(kBlock), lds, st, sc->dev, dc, bc,
                           verts, lens, qoff, qorg, qdir, qcount, uint32_t(qcap), n,
                           prm->sample_offset + done, uint32_t(items), qd);
      } else {
        hipLaunchKernelGGL(k_bdpt_paths<false>, dim3(grid), dim3(kBlock), lds, st, sc->dev, dc, bc,
                           verts, lens, qoff, qorg, qdir, qcount, uint32_t(qcap), n,
                           prm->sample_offset + done, uint32_t(items), qd);
      }
      uint32_t nq = 0;
      e = hipMemcpyAsync(&nq, qcount, sizeof(nq), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hip_fail(e, "render_bdpt: query count");
      if (nq > qcap) {
        set_error("render_bdpt: connection queue overflow");
        return XYZ_E_DEVICE;
      }
      rc = launch_closest_hit(sc->dev, nq, qorg, qdir, qprim, qt, nullptr, nullptr, nullptr, nullptr,
                              nullptr, st);
      if (rc != XYZ_OK) return rc;
      hipLaunchKernelGGL(k_bdpt_connect, dim3(grid), dim3(kBlock), 0, st, sc->dev, dc, bc, verts, lens,
                         qoff, qprim, qt, cols, sp, uint32_t(items));
    } else {
      uint32_t const grid = uint32_t(std::min<uint64_t>((items + kBlock - 1) / kBlock, 1u << 20));
      if (stats) {
        hipLaunchKernelGGL(k_bdpt<true>, dim3(grid), dim3(kBlock), lds, st, sc->dev, dc, bc, cols, sp,
                           n, prm->sample_offset + done, items, qd);
      } else {
        hipLaunchKernelGGL(k_bdpt<false>, dim3(grid), dim3(kBlock), lds, st, sc->dev, dc, bc, cols,
                           sp, n, prm->sample_offset + done, items, qd);
      }
    }
    hipLaunchKernelGGL(k_bdpt_sum, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, film, cols,
                       n, count);
  }
  if (!splat) {
    hipLaunchKernelGGL(k_add_into, dim3(uint32_t((wh * 3 + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       st, film, sp, size_t(wh * 3));
  }
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "render_bdpt launch");
  if (stats) {
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    unsigned long long q = 0;
    e = hipMemcpyAsync(&q, qd, sizeof(q), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "render_bdpt stats");
    stats->seconds = double(ms) * 1e-3;
    stats->paths = wh * prm->spp;
    stats->queries = q;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return XYZ_OK;
}

// Self-check of the Metal lobe's pow: out[i] = pow_det(x[i], y[i]).
__global__ void k_eval_pow(uint32_t n, const float* __restrict__ x, const float* __restrict__ y,
                           float* __restrict__ out) {
  uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = pow_det(x[i], y[i]);
}

}  // namespace

}  // namespace xyzhip

using namespace xyzhip;

extern "C" {

int xyz_eval_pow(uint32_t n, const float* x, const float* y, float* out, void* stream) {
  if (n > 0 && (!x || !y || !out)) {
    set_error("xyz_eval_pow: NULL buffer");
    return XYZ_E_INVALID;
  }
  if (n == 0) return XYZ_OK;
  hipLaunchKernelGGL(k_eval_pow, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                     n, x, y, out);
  hipError_t const e = hipGetLastError();
  return e == hipSuccess ? XYZ_OK : hip_fail(e, "xyz_eval_pow");
}

int xyz_render(xyz_scene* scene, const xyz_camera* camera, const xyz_render_params* params,
               float* film, xyz_render_stats* stats, void* stream) {
  return render_impl(scene, camera, params, film, stats, static_cast<hipStream_t>(stream), false,
                     nullptr, nullptr, nullptr);
}

int xyz_render_idpt(xyz_scene* scene, const xyz_importon_map* map, const xyz_camera* camera,
                    const xyz_render_params* params, float* film, xyz_render_stats* stats,
                    void* stream) {
  if (!map) {
    set_error("render_idpt: NULL importon map");
    return XYZ_E_INVALID;
  }
  return render_impl(scene, camera, params, film, stats, static_cast<hipStream_t>(stream), false,
                     nullptr, nullptr, nullptr, map);
}

int xyz_render_bdpt(xyz_scene* scene, const xyz_camera* camera, const xyz_render_params* params,
                    float* film, float* splat_film, xyz_render_stats* stats, void* stream) {
  return render_bdpt_impl(scene, camera, params, film, splat_film, stats,
                          static_cast<hipStream_t>(stream));
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

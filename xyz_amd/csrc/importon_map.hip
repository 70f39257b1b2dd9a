// Importon map of importance-driven path tracing: light paths traced on the
// device (Scene::CreateLightImportanceMap, xyz/core/scene_photonmap.cpp:272-334,
// with Bsdf::BoundsImportance), a uniform grid over the importons for the
// fixed-radius query, and the histogram query itself as a kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bdpt_device.h"
#include "pt_device.h"
#include "scene_internal.h"

namespace xyzhip {
namespace {

constexpr uint32_t kEmitBlock = 256;
constexpr uint32_t kImportonStream = 4u;  // Philox stream of the light paths
constexpr float kSearchRadius = 0.1f;      // CONFIG_SEARCH_RADIUS (xyz/core/config.hpp:56)
constexpr uint64_t kMaxCells = 1ull << 22;
constexpr uint32_t kMaxEmit = 1u << 28;  // light paths per map

// Bsdf::BoundsImportance: *dir holds the incoming direction and becomes the
// outgoing one.  u.x: Russian roulette / Fresnel choice, u.y: glass's cosine
// test, u.y u.z: a direction (Lambert, Metal).
__device__ __forceinline__ bool bounds_importance(const DevBsdf& b, bool back, const Basis& B,
                                                  float4 u, float3* dir, bool* specular) {
  switch (b.type) {
    case XYZ_LAMBERT: {  // lambert.cpp:16-40 (its cosine ratio takes the shading normal twice)
      float const cs = -dot3(*dir, B.ns);
      if (cs <= 0.0f) return false;
      float const cg = -dot3(*dir, B.ns);
      float const f = cs / cg;
      if (u.x >= b.rgb[1] * f) return false;
      *specular = false;
      bd_diffused(u.y, u.z, B.t, B.ng, B.b, dir);
      return true;
    }
    case XYZ_MIRROR: {  // mirror.cpp:18-48
      float const cs = -dot3(*dir, B.ns);
      if (cs <= 0.0f) return false;
      float const cg = -dot3(*dir, B.ng);
      *dir = add3(*dir, scl3l(2.0f * cs, B.ns));
      float const co = dot3(*dir, B.ng);
      if (co <= 0.0f) return false;
      if (u.x >= co / cg) return false;
      *specular = true;
      return true;
    }
    case XYZ_GLASS: {  // glass.cpp:16-91
      float const cs = -dot3(*dir, B.ns);
      if (cs <= 0.0f) return false;
      float const cg = -dot3(*dir, B.ng);
      float const ior_sq = glass_ior_sq(b, back);
      float const nk = cs;
      float const bb = ior_sq + nk * nk - 1.0f;
      *specular = true;
      if (bb > 0.0f) {
        float const g = sqrtf(bb);
        float const gpc = g + nk, gmc = g - nk;
        float const alpha = gmc / gpc;
        float const beta = (nk * gpc - 1.0f) / (nk * gmc + 1.0f);
        float const fr = 0.5f * alpha * alpha * (beta * beta + 1.0f);
        if (u.x < fr) {
          *dir = add3(*dir, scl3l(2.0f * nk, B.ns));
          float const co = dot3(*dir, B.ng);
          if (co <= 0.0f) return false;
          if (u.y >= co / cg) return false;
        } else {
          float const sq = sqrtf(ior_sq);
          float3 const v = sub3(*dir, scl3l(gmc, B.ns));
          *dir = f3(v.x / sq, v.y / sq, v.z / sq);
          float const co = -dot3(*dir, B.ng);
          if (co <= 0.0f) return false;
          if (u.y >= co / cg) return false;
        }
      } else {
        *dir = add3(*dir, scl3l(2.0f * nk, B.ns));
        float const co = dot3(*dir, B.ng);
        if (co <= 0.0f) return false;
        if (u.y >= co / cg) return false;
      }
      return true;
    }
    case XYZ_METAL: {  // metal.cpp:22-66
      float const nk1 = -dot3(*dir, B.ns);
      if (nk1 <= 0.0f) return false;
      float const cg = -dot3(*dir, B.ng);
      float3 hv;
      bd_glossy_half(u.y, u.z, b.shininess, B.t, B.ns, B.b, &hv);
      float const hk = -dot3(hv, *dir);
      *dir = add3(*dir, scl3l(2.0f * hk, hv));
      float const nk2 = dot3(*dir, B.ns);
      if (nk2 <= 0.0f) return false;
      float const co = dot3(B.ng, *dir);
      if (co <= 0.0f) return false;
      float const f = fake_fresnel(hk) / hmaxf(nk1, nk2) * nk2 * (co / cg);
      if (u.x >= b.rgb[1] * f) return false;
      *specular = false;
      return true;
    }
    default:
      return false;  // light.cpp:11-16
  }
}

// One thread per light path.  Count pass (kWrite false): importons per path.
// Write pass: path i's importons at offsets[i], in vertex order, at most
// counts[i] of them (the count pass's result).
template <bool kWrite>
__global__ void __launch_bounds__(kEmitBlock)
    k_emit(DevScene s, PTConst c, uint32_t n_emit, uint32_t* counts, const uint32_t* offsets,
           float4* pos_out, float4* dir_out) {
  extern __shared__ uint32_t lds_stack[];
  uint32_t* const stk = lds_stack + threadIdx.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_emit;
       i += gridDim.x * blockDim.x) {
    uint32_t cnt = 0;
    uint32_t const base = kWrite ? offsets[i] : 0u;
    uint32_t const lim = kWrite ? counts[i] : 0xFFFFFFFFu;
    float4 const d0 = draw4(c, i, 0u, 0u, kImportonStream);
    LightVertex const lv = sample_light(s, d0.x, d0.y);
    if (lv.prim >= 0) {
      // SampleLightPosition's basis and ComputeDiffusedVector about the light's ng
      float3 t, b;
      basis_about(lv.ng, &t, &b);
      float3 dir;
      bd_diffused(d0.z, d0.w, t, lv.ng, b, &dir);
      float3 pos = lv.pos;
      bool specular = false;
      for (uint32_t depth = 0; depth < c.kmax; ++depth) {
        HitRec const h = trace_lane(s, pos, dir, stk, blockDim.x);
        if (h.prim < 0) break;
        uint32_t bsdf_id;
        Basis B = shading_basis(s, h.prim, h.f, h.g, h.back, &bsdf_id);
        auto const& bs = gp(s.bsdfs)[bsdf_id];
        pos = add3(pos, scl3(dir, h.t));
        if (bs.type == XYZ_LAMBERT) {
          basis_about(B.ng, &B.t, &B.b);  // Triangle::GeometricBasis
          if (depth > 0 && specular) {
            if (kWrite && cnt < lim) {
              pos_out[base + cnt] = make_float4(pos.x, pos.y, pos.z, 0.0f);
              dir_out[base + cnt] = make_float4(dir.x, dir.y, dir.z, 0.0f);
            }
            ++cnt;
          }
        }
        float4 const u = draw4(c, i, 0u, depth + 1u, kImportonStream);
        if (!bounds_importance(bs, h.back, B, u, &dir, &specular)) break;
      }
    }
    if (!kWrite) counts[i] = cnt;
  }
}

// cdf[17] of the importance histogram at each vertex (basis: position, shading
// normal, tangent, binormal).
__global__ void __launch_bounds__(kEmitBlock)
    k_importance_query(DevImportons m, uint32_t n, const float* __restrict__ basis, float* cdf) {
  uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* q = basis + 12 * size_t(i);
  Basis B;
  float3 const p = f3(q[0], q[1], q[2]);
  B.ns = f3(q[3], q[4], q[5]);
  B.t = f3(q[6], q[7], q[8]);
  B.b = f3(q[9], q[10], q[11]);
  B.ng = B.ns;
  uint32_t c[kImpN];
  importance_counts(m, p, B, c);
  float* out = cdf + 17 * size_t(i);
  out[0] = 0.0f;
  uint32_t pre = 0;
#pragma unroll
  for (uint32_t k = 1; k <= kImpN; ++k) {
    pre += c[k - 1];
    out[k] = float(kImpN * k + pre);
  }
}

void map_free(xyz_importon_map* m) {
  for (void* p : m->allocs)
    if (p) (void)hipFree(p);
  m->allocs.clear();
}

int dev_alloc(xyz_importon_map* m, size_t bytes, void** p) {
  *p = nullptr;
  hipError_t const e = hipMalloc(p, std::max<size_t>(bytes, 16));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(importon map)");
  m->allocs.push_back(*p);
  return XYZ_OK;
}

// Uniform grid over m->host: cells at least the search radius wide (at most
// kMaxCells of them), importons sorted by cell with a stable counting sort,
// uploaded with the cell table.
int build_grid(xyz_importon_map* m) {
  size_t const n = m->host.size() / 6;
  DevImportons& d = m->dev;
  std::memset(&d, 0, sizeof(d));
  d.r2 = kSearchRadius * kSearchRadius;
  d.reach = kSearchRadius * 1.0625f;
  d.n = uint32_t(n);
  float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      float const v = m->host[6 * i + size_t(a)];
      if (!std::isfinite(v)) {
        set_error("importon map: non-finite importon position");
        return XYZ_E_INVALID;
      }
      if (i == 0 || v < lo[a]) lo[a] = v;
      if (i == 0 || v > hi[a]) hi[a] = v;
    }
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(hi[a] - lo[a])) {
      set_error("importon map: importon coordinates span more than the float range");
      return XYZ_E_INVALID;
    }
  float cell = kSearchRadius;
  uint64_t dims[3];
  for (;;) {
    float const inv = 1.0f / cell;
    uint64_t total = 1;
    for (int a = 0; a < 3; ++a) {
      dims[a] = uint64_t(std::floor(double(hi[a] - lo[a]) * double(inv))) + 1;
      total *= dims[a];
    }
    if (total <= kMaxCells) break;
    cell *= 1.25f;
  }
  d.inv_cell = 1.0f / cell;
  for (int a = 0; a < 3; ++a) {
    d.lo[a] = lo[a];
    d.dim[a] = uint32_t(dims[a]);
  }
  size_t const ncell = size_t(dims[0] * dims[1] * dims[2]);
  std::vector<uint32_t> cell_of(n), start(ncell + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    uint32_t ci[3];
    for (int a = 0; a < 3; ++a)
      ci[a] = uint32_t(imp_cell(m->host[6 * i + size_t(a)], d.lo[a], d.inv_cell, d.dim[a]));
    cell_of[i] = (ci[2] * d.dim[1] + ci[1]) * d.dim[0] + ci[0];
    ++start[cell_of[i] + 1];
  }
  for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  std::vector<float> sp(4 * n), sd(4 * n);
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (size_t i = 0; i < n; ++i) {
    size_t const k = fill[cell_of[i]]++;
    for (int a = 0; a < 3; ++a) {
      sp[4 * k + size_t(a)] = m->host[6 * i + size_t(a)];
      sd[4 * k + size_t(a)] = m->host[6 * i + 3 + size_t(a)];
    }
    sp[4 * k + 3] = sd[4 * k + 3] = 0.0f;
  }
  void *pp, *pd, *pc;
  int rc = dev_alloc(m, sp.size() * sizeof(float), &pp);
  if (rc == XYZ_OK) rc = dev_alloc(m, sd.size() * sizeof(float), &pd);
  if (rc == XYZ_OK) rc = dev_alloc(m, start.size() * sizeof(uint32_t), &pc);
  if (rc != XYZ_OK) return rc;
  hipError_t e = hipSuccess;
  if (n > 0) e = hipMemcpy(pp, sp.data(), sp.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && n > 0)
    e = hipMemcpy(pd, sd.data(), sd.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(pc, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy(importon grid)");
  d.pos = static_cast<const float4*>(pp);
  d.dir = static_cast<const float4*>(pd);
  d.cell_start = static_cast<const uint32_t*>(pc);
  return XYZ_OK;
}

// Device scratch freed at scope exit.
struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
  int alloc(size_t bytes, void** out) {
    *out = nullptr;
    hipError_t const e = hipMalloc(out, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(importon emission)");
    p.push_back(*out);
    return XYZ_OK;
  }
};

int emit(xyz_scene* sc, uint32_t n_emit, uint32_t max_depth, uint64_t seed, xyz_importon_map* m) {
  m->host.clear();
  if (n_emit == 0) return XYZ_OK;
  PTConst c;
  std::memset(&c, 0, sizeof(c));
  c.kmax = max_depth + 1;
  c.seed_lo = uint32_t(seed);
  c.seed_hi = uint32_t(seed >> 32);
  Scratch sc_mem;
  void *counts, *offsets;
  int rc = sc_mem.alloc(size_t(n_emit) * sizeof(uint32_t), &counts);
  if (rc == XYZ_OK) rc = sc_mem.alloc(size_t(n_emit) * sizeof(uint32_t), &offsets);
  if (rc != XYZ_OK) return rc;
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      ncu <= 0)
    ncu = 256;
  uint32_t const grid =
      std::max(1u, std::min((n_emit + kEmitBlock - 1) / kEmitBlock, uint32_t(ncu) * 8));
  size_t const lds = size_t(std::max(1u, sc->dev.stack_cap)) * kEmitBlock * sizeof(uint32_t);
  hipLaunchKernelGGL(k_emit<false>, dim3(grid), dim3(kEmitBlock), lds, nullptr, sc->dev, c, n_emit,
                     static_cast<uint32_t*>(counts), nullptr, nullptr, nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "importon count launch");
  std::vector<uint32_t> cnt(n_emit), off(n_emit);
  e = hipMemcpy(cnt.data(), counts, cnt.size() * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "importon counts");
  uint64_t total = 0;
  for (uint32_t i = 0; i < n_emit; ++i) {
    off[i] = uint32_t(total);
    total += cnt[i];
  }
  if (total >= (1ull << 31)) {
    set_error("importon map: too many importons");
    return XYZ_E_UNSUPPORTED;
  }
  if (total == 0) return XYZ_OK;
  void *pos, *dir;
  rc = sc_mem.alloc(size_t(total) * sizeof(float4), &pos);
  if (rc == XYZ_OK) rc = sc_mem.alloc(size_t(total) * sizeof(float4), &dir);
  if (rc != XYZ_OK) return rc;
  e = hipMemcpy(offsets, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "importon offsets");
  hipLaunchKernelGGL(k_emit<true>, dim3(grid), dim3(kEmitBlock), lds, nullptr, sc->dev, c, n_emit,
                     static_cast<uint32_t*>(counts), static_cast<const uint32_t*>(offsets),
                     static_cast<float4*>(pos), static_cast<float4*>(dir));
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "importon write launch");
  std::vector<float4> hp(total), hd(total);
  e = hipMemcpy(hp.data(), pos, hp.size() * sizeof(float4), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hd.data(), dir, hd.size() * sizeof(float4), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "importon download");
  m->host.resize(6 * size_t(total));
  for (size_t i = 0; i < total; ++i) {
    float* q = &m->host[6 * i];
    q[0] = hp[i].x; q[1] = hp[i].y; q[2] = hp[i].z;
    q[3] = hd[i].x; q[4] = hd[i].y; q[5] = hd[i].z;
  }
  return XYZ_OK;
}

int finish_map(xyz_scene* sc, xyz_importon_map* m, xyz_importon_map** out,
               std::chrono::steady_clock::time_point t0) {
  int const rc = build_grid(m);
  if (rc != XYZ_OK) {
    map_free(m);
    delete m;
    return rc;
  }
  m->scene = sc;
  (void)hipGetDevice(&m->device);
  m->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = m;
  return XYZ_OK;
}

}  // namespace
}  // namespace xyzhip

using namespace xyzhip;

extern "C" {

int xyz_importon_map_create(xyz_scene* scene, uint32_t n_emit, uint32_t max_depth, uint64_t seed,
                            xyz_importon_map** out) {
  if (!scene || !out) {
    set_error("xyz_importon_map_create: NULL argument");
    return XYZ_E_INVALID;
  }
  *out = nullptr;
  if (max_depth < 1 || max_depth > 255) {
    set_error("xyz_importon_map_create: max_depth must be in [1, 255]");
    return XYZ_E_INVALID;
  }
  if (n_emit > kMaxEmit) {
    set_error("xyz_importon_map_create: n_emit above 2^28");
    return XYZ_E_UNSUPPORTED;
  }
  auto const t0 = std::chrono::steady_clock::now();
  auto* m = new xyz_importon_map();
  m->n_emit = n_emit;
  int const rc = emit(scene, n_emit, max_depth, seed, m);
  if (rc != XYZ_OK) {
    delete m;
    return rc;
  }
  return finish_map(scene, m, out, t0);
}

int xyz_importon_map_from_host(xyz_scene* scene, uint64_t n, const float* importons,
                               xyz_importon_map** out) {
  if (!scene || !out || (n > 0 && !importons)) {
    set_error("xyz_importon_map_from_host: NULL argument");
    return XYZ_E_INVALID;
  }
  *out = nullptr;
  if (n >= (1ull << 31)) {
    set_error("xyz_importon_map_from_host: too many importons");
    return XYZ_E_UNSUPPORTED;
  }
  for (uint64_t i = 0; i < 6 * n; ++i)
    if (!std::isfinite(importons[i])) {
      set_error("xyz_importon_map_from_host: non-finite importon");
      return XYZ_E_INVALID;
    }
  auto const t0 = std::chrono::steady_clock::now();
  auto* m = new xyz_importon_map();
  m->host.assign(importons, importons + 6 * n);
  return finish_map(scene, m, out, t0);
}

int xyz_importon_map_info(const xyz_importon_map* map, xyz_importon_info* info) {
  if (!map || !info) {
    set_error("xyz_importon_map_info: NULL argument");
    return XYZ_E_INVALID;
  }
  std::memset(info, 0, sizeof(*info));
  info->n_importons = map->host.size() / 6;
  info->n_emit = map->n_emit;
  for (int a = 0; a < 3; ++a) info->grid[a] = map->dev.dim[a];
  info->cell = 1.0f / map->dev.inv_cell;
  info->radius = kSearchRadius;
  info->seconds = map->seconds;
  return XYZ_OK;
}

int xyz_importon_map_export(const xyz_importon_map* map, float* importons) {
  if (!map || (!importons && !map->host.empty())) {
    set_error("xyz_importon_map_export: NULL argument");
    return XYZ_E_INVALID;
  }
  if (!map->host.empty()) std::memcpy(importons, map->host.data(), map->host.size() * sizeof(float));
  return XYZ_OK;
}

int xyz_importance_query(const xyz_importon_map* map, uint32_t n, const float* basis, float* cdf,
                         void* stream) {
  if (!map || (n > 0 && (!basis || !cdf))) {
    set_error("xyz_importance_query: NULL argument");
    return XYZ_E_INVALID;
  }
  if (n == 0) return XYZ_OK;
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != map->device) {
    set_error("xyz_importance_query: the importon map lives on another device");
    return XYZ_E_INVALID;
  }
  hipLaunchKernelGGL(k_importance_query, dim3((n + kEmitBlock - 1) / kEmitBlock), dim3(kEmitBlock),
                     0, static_cast<hipStream_t>(stream), map->dev, n, basis, cdf);
  hipError_t const e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "importance query launch");
  return XYZ_OK;
}

void xyz_importon_map_destroy(xyz_importon_map* map) {
  if (!map) return;
  map_free(map);
  delete map;
}

}  // extern "C"

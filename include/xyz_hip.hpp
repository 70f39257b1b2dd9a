// xyz_hip.hpp -- header-only C++ adapter over the C ABI (xyz_hip.h) that keeps
// the shapes of the reference renderer's interfaces for the hot path:
//
//   xyz::Intersection            (reference xyz/geom/intersection.hpp:8-41)
//   xyz::KDTree::Build           (reference xyz/geom/kdtree.hpp:31)
//   xyz::Scene::FindIntersection (reference xyz/core/scene.hpp:236-238,277-295)
//   xyz::PathTracer::run body    (reference xyz/integrator/integrator.cpp:119-128)
//   ImageFilm::SaveAsPfm         (reference xyz/core/imagefilm.cpp:153-206)
//
// The reference's code calls these one ray or one pixel at a time; over a GPU
// boundary the useful unit is a batch, so the batched forms are primary and the
// single-ray form exists for drop-in compatibility only.  Errors become
// std::runtime_error carrying xyz_last_error().
#ifndef XYZ_HIP_HPP_
#define XYZ_HIP_HPP_

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "xyz_hip.h"

namespace xyz_hip {

inline void check(int rc) {
  if (rc != XYZ_OK) throw std::runtime_error(std::string("xyz_hip: ") + xyz_last_error());
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Ray interval and hit record, as xyz::Intersection.
struct Intersection {
  float t_min = 1e-5f;
  float t_max = INFINITY;
  float f = 0.0f, g = 0.0f;
  bool is_back_side = false;
};

// Owning device buffer.
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  void resize(size_t n) {
    if (n <= n_) return;
    release();
    hip_check(hipMalloc(reinterpret_cast<void**>(&p_), (n ? n : 1) * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void upload(const T* h, size_t n, hipStream_t s) {
    resize(n);
    hip_check(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  }
  void download(T* h, size_t n, hipStream_t s) const {
    hip_check(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  }
  T* get() const { return p_; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Host kd-tree (KDTree::Build with the reference's split rules), flattened.
class KDTree {
 public:
  KDTree(const std::vector<float>& v0, const std::vector<float>& v1, const std::vector<float>& v2,
         uint32_t max_bucket = 16, uint32_t max_depth = 16) {
    check(xyz_kdtree_build(uint32_t(v0.size() / 3), v0.data(), v1.data(), v2.data(), max_bucket,
                           max_depth, &t_));
  }
  ~KDTree() { xyz_kdtree_destroy(t_); }
  KDTree(const KDTree&) = delete;
  KDTree& operator=(const KDTree&) = delete;
  xyz_scene_info Info() const {
    xyz_scene_info i;
    check(xyz_kdtree_info(t_, &i));
    return i;
  }

 private:
  xyz_kdtree* t_ = nullptr;
};

// Device-resident scene: Scene::FindIntersection over batches and the
// path-tracing pass.
class Scene {
 public:
  explicit Scene(const xyz_scene_desc& d) { check(xyz_scene_create(&d, &s_)); }
  ~Scene() {
    xyz_scene_destroy(s_);
    if (own_stream_) (void)hipStreamDestroy(stream_);
  }
  Scene(const Scene&) = delete;
  Scene& operator=(const Scene&) = delete;

  xyz_scene* handle() const { return s_; }

  // Batched FindIntersection for rays in HOST memory.  prim[i] is the caller
  // index of the hit triangle or -1 (the reference's nullptr); params[i] gets
  // t_max / f / g / is_back_side of the hit.
  void FindIntersection(size_t n, const float* org, const float* dir, int32_t* prim,
                        Intersection* params) {
    hipStream_t s = stream();
    d_org_.upload(org, 3 * n, s);
    d_dir_.upload(dir, 3 * n, s);
    d_prim_.resize(n);
    d_t_.resize(n);
    d_f_.resize(n);
    d_g_.resize(n);
    d_back_.resize(n);
    check(xyz_closest_hit(s_, uint32_t(n), d_org_.get(), d_dir_.get(), d_prim_.get(), d_t_.get(),
                          d_f_.get(), d_g_.get(), d_back_.get(), nullptr, s));
    std::vector<float> t(n), f(n), g(n);
    std::vector<uint8_t> b(n);
    d_prim_.download(prim, n, s);
    d_t_.download(t.data(), n, s);
    d_f_.download(f.data(), n, s);
    d_g_.download(g.data(), n, s);
    d_back_.download(b.data(), n, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (size_t i = 0; i < n; ++i) {
      params[i].t_max = t[i];
      params[i].f = f[i];
      params[i].g = g[i];
      params[i].is_back_side = b[i] != 0;
    }
  }

  // One-ray form of Scene::FindIntersection (drop-in shape; use the batch).
  int32_t FindIntersection(const float o[3], const float d[3], Intersection* p) {
    int32_t prim = -1;
    FindIntersection(1, o, d, &prim, p);
    return prim;
  }

  // One progressive pass of PathTracer::run: adds the sum of samples
  // [sample_offset, sample_offset + spp) of every pixel to `film_sum`
  // (host, width*height*3, row 0 = bottom).
  void Render(const xyz_camera& cam, const xyz_render_params& prm, std::vector<float>* film_sum) {
    size_t const n = size_t(cam.width) * cam.height * 3;
    film_sum->resize(n, 0.0f);
    hipStream_t s = stream();
    d_film_.upload(film_sum->data(), n, s);
    check(xyz_render(s_, &cam, &prm, d_film_.get(), nullptr, s));
    d_film_.download(film_sum->data(), n, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  }

 private:
  hipStream_t stream() {
    if (!stream_) {
      hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
      own_stream_ = true;
    }
    return stream_;
  }
  xyz_scene* s_ = nullptr;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  DeviceBuffer<float> d_org_, d_dir_, d_t_, d_f_, d_g_, d_film_;
  DeviceBuffer<int32_t> d_prim_;
  DeviceBuffer<uint8_t> d_back_;
};

// ImageFilm::SaveAsPfm: "PF\nW H\n-1\n", little-endian RGB floats, rows bottom
// up, value = sum / spp, non-finite pixels written as 0.
inline bool SaveAsPfm(const char* path, uint32_t w, uint32_t h, const std::vector<float>& film_sum,
                      float spp) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  std::fprintf(f, "PF\n%u %u\n-1\n", w, h);
  float const s = 1.0f / spp;
  for (size_t p = 0; p < size_t(w) * h; ++p) {
    float rgb[3] = {film_sum[3 * p] * s, film_sum[3 * p + 1] * s, film_sum[3 * p + 2] * s};
    if (!std::isfinite(rgb[0]) || !std::isfinite(rgb[1]) || !std::isfinite(rgb[2]))
      rgb[0] = rgb[1] = rgb[2] = 0.0f;
    std::fwrite(rgb, sizeof(float), 3, f);
  }
  return std::fclose(f) == 0;
}

}  // namespace xyz_hip

#endif  // XYZ_HIP_HPP_

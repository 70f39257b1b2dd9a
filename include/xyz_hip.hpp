// xyz_hip.hpp -- header-only C++ adapter over the C ABI (xyz_hip.h) that keeps
// the shapes of the reference renderer's interfaces for the hot path:
//
//   xyz::Intersection            (reference xyz/geom/intersection.hpp:8-41)
//   xyz::KDTree::Build           (reference xyz/geom/kdtree.hpp:31)
//   xyz::Scene::FindIntersection (reference xyz/core/scene.hpp:236-238,277-295)
//   xyz::PathTracer::run body    (reference xyz/integrator/integrator.cpp:119-128)
//   ImageFilm::SaveAsPfm/PfmRgb/Float/Hdr (reference xyz/core/imagefilm.cpp:108-241,
//                                 xyz/core/imagefilm_hdr.cpp)
//
// The reference's code calls these one ray or one pixel at a time; over a GPU
// boundary the useful unit is a batch, so the batched forms are primary and the
// single-ray form exists for drop-in compatibility only.  Errors become
// std::runtime_error carrying xyz_last_error().
#ifndef XYZ_HIP_HPP_
#define XYZ_HIP_HPP_

#include <hip/hip_runtime_api.h>

#include <algorithm>
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
    xyz_importon_map_destroy(imap_);
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

  // Scene::CreateLightImportanceMap (reference CONFIG_EMITED_SIZE = 100000
  // light paths of at most max_depth + 1 segments); replaces an earlier map.
  size_t CreateLightImportanceMap(uint32_t n_emit = 100000, uint32_t max_depth = 15,
                                  uint64_t seed = 0) {
    xyz_importon_map* m = nullptr;
    check(xyz_importon_map_create(s_, n_emit, max_depth, seed, &m));
    xyz_importon_map_destroy(imap_);
    imap_ = m;
    xyz_importon_info info;
    check(xyz_importon_map_info(imap_, &info));
    return size_t(info.n_importons);
  }

  // One progressive pass of ImportanceDrivenPathTracer::run (after
  // CreateLightImportanceMap); film contract of Render.
  void RenderImportanceDriven(const xyz_camera& cam, const xyz_render_params& prm,
                              std::vector<float>* film_sum) {
    if (!imap_) throw std::runtime_error("RenderImportanceDriven: no importon map");
    size_t const n = size_t(cam.width) * cam.height * 3;
    film_sum->resize(n, 0.0f);
    hipStream_t s = stream();
    d_film_.upload(film_sum->data(), n, s);
    check(xyz_render_idpt(s_, imap_, &cam, &prm, d_film_.get(), nullptr, s));
    d_film_.download(film_sum->data(), n, s);
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  }

  // One progressive pass of BidirectionalPathTracer::run: own colours and
  // light-tracing contributions added to the same film.
  void RenderBidirectional(const xyz_camera& cam, const xyz_render_params& prm,
                           std::vector<float>* film_sum) {
    size_t const n = size_t(cam.width) * cam.height * 3;
    film_sum->resize(n, 0.0f);
    hipStream_t s = stream();
    d_film_.upload(film_sum->data(), n, s);
    check(xyz_render_bdpt(s_, &cam, &prm, d_film_.get(), nullptr, nullptr, s));
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
  xyz_importon_map* imap_ = nullptr;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  DeviceBuffer<float> d_org_, d_dir_, d_t_, d_f_, d_g_, d_film_;
  DeviceBuffer<int32_t> d_prim_;
  DeviceBuffer<uint8_t> d_back_;
};

// ---- film files (xyz/core/imagefilm.cpp:108-206, xyz/core/imagefilm_hdr.cpp) --
// A film holds per-pixel sums, row 0 = bottom; every writer divides by the
// sample count first (value = sum * (1 / spp)).  Like ImageFilm, the .float
// and .hdr writers treat the film as CIE XYZ and convert it to RGB with the
// CCIR 601-1 matrix; the PFM writers write the values as they are.  For
// byte-identical files build with -ffp-contract=off.

// hi::CCIR601_1_XYZ2RGB (libhi/hi/image/color_space.hpp:70-79), in float.
inline void Ccir601XyzToRgb(const float xyz[3], float rgb[3]) {
  rgb[0] = 1.9104f * xyz[0] - 0.5338f * xyz[1] - 0.2891f * xyz[2];
  rgb[1] = -0.9844f * xyz[0] + 1.9985f * xyz[1] - 0.0279f * xyz[2];
  rgb[2] = 0.0585f * xyz[0] - 0.1187f * xyz[1] + 0.9017f * xyz[2];
}

// ImageFilm::SaveAsPfm: "PF\nW H\n-1\n", little-endian floats, rows bottom up,
// non-finite pixels written as 0.
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

// ImageFilm::SaveAsPfmRgb: the PFM layout, values unchecked.
inline bool SaveAsPfmRgb(const char* path, uint32_t w, uint32_t h,
                         const std::vector<float>& film_sum, float spp) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  std::fprintf(f, "PF\n%u %u\n-1\n", w, h);
  float const s = 1.0f / spp;
  for (size_t p = 0; p < size_t(w) * h; ++p) {
    float const v[3] = {film_sum[3 * p] * s, film_sum[3 * p + 1] * s, film_sum[3 * p + 2] * s};
    std::fwrite(v, sizeof(float), 3, f);
  }
  return std::fclose(f) == 0;
}

// ImageFilm::SaveAsFloat: headerless RGB floats, rows bottom up, negative
// components clamped to 0.
inline bool SaveAsFloat(const char* path, uint32_t w, uint32_t h,
                        const std::vector<float>& film_sum, float spp) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  float const s = 1.0f / spp;
  for (size_t p = 0; p < size_t(w) * h; ++p) {
    float const xyz[3] = {film_sum[3 * p] * s, film_sum[3 * p + 1] * s, film_sum[3 * p + 2] * s};
    float rgb[3];
    Ccir601XyzToRgb(xyz, rgb);
    for (float& c : rgb)
      if (c < 0) c = 0;
    std::fwrite(rgb, sizeof(float), 3, f);
  }
  return std::fclose(f) == 0;
}

namespace detail {

// RGBE pixel (xyz/core/imagefilm_hdr.cpp:12-28): the largest component d maps
// to a mantissa byte by frexp(d) * 255.9999 / d.
inline void RgbeEncode(const float rgb[3], uint8_t out[4]) {
  float const d = std::max(std::max(rgb[0], rgb[1]), rgb[2]);
  if (d <= 1e-32) {
    out[0] = out[1] = out[2] = out[3] = 0;
    return;
  }
  int e = 0;
  float const f = float(double(std::frexp(d, &e)) * 255.9999 / double(d));
  // a negative component (XYZ -> RGB can give one) is undefined in the
  // reference's cast; here it truncates to int and wraps modulo 256
  out[0] = uint8_t(int(rgb[0] * f));
  out[1] = uint8_t(int(rgb[1] * f));
  out[2] = uint8_t(int(rgb[2] * f));
  out[3] = uint8_t(e + 128);
}

// One scanline, run-length encoded per component (xyz/core/imagefilm_hdr.cpp:
// 44-109): lines shorter than 8 or longer than 32767 pixels are written flat;
// otherwise "2 2 hi lo", then for each component a sequence of literal packets
// (count <= 128, then the bytes) and run packets (128 + count <= 127, then the
// byte); runs shorter than 4 stay literal.
inline void RgbeWriteLine(const std::vector<uint8_t>& line, std::string* out) {
  int const size = int(line.size() / 4);
  if (size < 8 || size > 0x7fff) {
    out->append(reinterpret_cast<const char*>(line.data()), line.size());
    return;
  }
  out->push_back(char(2));
  out->push_back(char(2));
  out->push_back(char((size >> 8) & 0xff));
  out->push_back(char(size & 0xff));
  auto at = [&](int i, int c) { return line[size_t(4 * i + c)]; };
  for (int c = 0; c < 4; ++c) {
    for (int j = 0, count = 0; j < size; j += count) {
      // the next run of at least 4 equal bytes, from j on
      int begin = j;
      for (; begin < size; begin += count) {
        for (count = 1; count < 127 && begin + count < size && at(begin + count, c) == at(begin, c);
             ++count) {
        }
        if (count >= 4) break;
      }
      // a short run (2 or 3 equal bytes) right before it goes out as a run
      if (begin - j > 1 && begin - j < 4) {
        int k = j + 1;
        while (at(k++, c) == at(j, c)) {
          if (k == begin) {
            out->push_back(char(begin - j + 128));
            out->push_back(char(at(j, c)));
            j = begin;
            break;
          }
        }
      }
      // literals up to the run
      while (j < begin) {
        int n = begin - j;
        if (n > 128) n = 128;
        out->push_back(char(n));
        while (n--) out->push_back(char(at(j++, c)));
      }
      if (count >= 4) {
        out->push_back(char(count + 128));
        out->push_back(char(at(begin, c)));
      } else {
        count = 0;
      }
    }
  }
}

}  // namespace detail

// ImageFilm::SaveAsHdr (xyz/core/imagefilm_hdr.cpp:114-187): Radiance RGBE,
// "-Y H +X W", top row first, XYZ -> RGB, non-finite pixels as 0.
inline bool SaveAsHdr(const char* path, uint32_t w, uint32_t h, const std::vector<float>& film_sum,
                      float spp) {
  std::string out =
      "#?RADIANCE\n"
      "# Made with Mikado Renderer\n"
      "FORMAT=32-bit_rle_rgbe\n"
      "EXPOSURE=1.0000000000000\n"
      "\n";
  out += "-Y " + std::to_string(h) + " +X " + std::to_string(w) + "\n";
  float const s = 1.0f / spp;
  std::vector<uint8_t> line(size_t(w) * 4);
  for (uint32_t y = 0; y < h; ++y) {
    size_t const row = size_t(h - 1 - y) * w;
    for (uint32_t x = 0; x < w; ++x) {
      size_t const p = row + x;
      float const xyz[3] = {film_sum[3 * p] * s, film_sum[3 * p + 1] * s, film_sum[3 * p + 2] * s};
      float rgb[3];
      Ccir601XyzToRgb(xyz, rgb);
      if (!std::isfinite(rgb[0]) || !std::isfinite(rgb[1]) || !std::isfinite(rgb[2]))
        rgb[0] = rgb[1] = rgb[2] = 0.0f;
      detail::RgbeEncode(rgb, &line[4 * size_t(x)]);
    }
    detail::RgbeWriteLine(line, &out);
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  size_t const wrote = std::fwrite(out.data(), 1, out.size(), f);
  return std::fclose(f) == 0 && wrote == out.size();
}

}  // namespace xyz_hip

#endif  // XYZ_HIP_HPP_

// Exercises include/xyz_hip.hpp the way reference host code would: build a
// kd-tree from caller-ordered triangles, run FindIntersection on a batch and one
// progressive path-tracing pass, write a PFM.  Mode "host" touches no GPU.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "xyz_hip.hpp"

int main(int argc, char** argv) {
  // "films <prefix>": render on the device and write the film files
  bool const films = argc > 2 && std::strcmp(argv[1], "films") == 0;
  bool const gpu = films || (argc > 1 && std::strcmp(argv[1], "gpu") == 0);
  // two triangles of a unit quad at z = 1 facing -z, one light quad above
  std::vector<float> v0 = {0, 0, 1, 0, 0, 1, 0.2f, 2, 0.2f, 0.2f, 2, 0.2f};
  // light winding gives cross(e2, e1) = -y: it faces down toward the quad
  std::vector<float> v1 = {1, 0, 1, 1, 1, 1, 0.8f, 2, 0.8f, 0.2f, 2, 0.8f};
  std::vector<float> v2 = {1, 1, 1, 0, 1, 1, 0.8f, 2, 0.2f, 0.8f, 2, 0.8f};
  xyz_hip::KDTree tree(v0, v1, v2);
  xyz_scene_info info = tree.Info();
  std::printf("kdtree: %llu nodes, %llu refs\n", (unsigned long long)info.n_nodes,
              (unsigned long long)info.n_leaf_refs);
  if (!gpu) return info.n_leaf_refs == 4 ? 0 : 1;

  xyz_bsdf_desc bs[2] = {{XYZ_LAMBERT, {0.5f, 0.5f, 0.5f}, 0, 1.5f, -1},
                         {XYZ_LIGHT, {10, 10, 10}, 0, 1.5f, -1}};
  uint32_t tri_bsdf[4] = {0, 0, 1, 1};
  xyz_scene_desc d = {};
  d.n_tris = 4;
  d.v0 = v0.data();
  d.v1 = v1.data();
  d.v2 = v2.data();
  d.bsdf = tri_bsdf;
  d.n_bsdfs = 2;
  d.bsdfs = bs;
  xyz_hip::Scene scene(d);
  float o[6] = {0.25f, 0.5f, 0, 0.75f, 0.5f, 0};
  float dir[6] = {0, 0, 1, 0, 0, 1};
  int32_t prim[2];
  xyz_hip::Intersection hit[2];
  scene.FindIntersection(2, o, dir, prim, hit);
  std::printf("hits: %d t=%g, %d t=%g\n", prim[0], hit[0].t_max, prim[1], hit[1].t_max);
  if (prim[0] != 1 || prim[1] != 0 || hit[0].t_max != 1.0f) return 2;
  if (films) {
    xyz_camera fc = {{0.5f, 0.5f, -2}, {0.5f, 0.5f, 1}, {0, 1, 0}, 40.0f, 48, 40};
    xyz_render_params fp = {};
    fp.spp = 8;
    fp.max_depth = 4;
    fp.rr_threshold = 0.2f;
    fp.seed = 3;
    std::vector<float> f;
    scene.Render(fc, fp, &f);
    std::string const pre = argv[2];
    bool ok = xyz_hip::SaveAsPfm((pre + ".pfm").c_str(), 48, 40, f, 8.0f);
    ok = ok && xyz_hip::SaveAsPfmRgb((pre + ".rgb.pfm").c_str(), 48, 40, f, 8.0f);
    ok = ok && xyz_hip::SaveAsFloat((pre + ".float").c_str(), 48, 40, f, 8.0f);
    ok = ok && xyz_hip::SaveAsHdr((pre + ".hdr").c_str(), 48, 40, f, 8.0f);
    std::printf("films written: %d\n", int(ok));
    return ok ? 0 : 6;
  }
  xyz_camera cam = {{0.5f, 0.5f, -2}, {0.5f, 0.5f, 1}, {0, 1, 0}, 40.0f, 16, 16};
  xyz_render_params prm = {};
  prm.spp = 4;
  prm.max_depth = 4;
  prm.rr_threshold = 0.2f;
  std::vector<float> film;
  scene.Render(cam, prm, &film);
  double sum = 0;
  for (float v : film) sum += v;
  std::printf("film sum %g\n", sum);
  if (!(sum > 0 && xyz_hip::SaveAsPfm("/tmp/xyz_adapter_example.pfm", 16, 16, film, 4.0f)))
    return 3;
  // no specular surface: no importons, and the importance-driven pass is the
  // path-tracing pass bit for bit
  size_t const n_imp = scene.CreateLightImportanceMap(1000, 4, 0);
  std::vector<float> idpt;
  scene.RenderImportanceDriven(cam, prm, &idpt);
  bool const same = std::memcmp(idpt.data(), film.data(), film.size() * sizeof(float)) == 0;
  std::printf("importons %zu, idpt film equal to pt: %d\n", n_imp, int(same));
  if (n_imp != 0 || !same) return 4;
  std::vector<float> bdpt;
  scene.RenderBidirectional(cam, prm, &bdpt);
  double bsum = 0;
  for (float v : bdpt) bsum += v;
  std::printf("bdpt film sum %g\n", bsum);
  return bsum > 0 ? 0 : 5;
}

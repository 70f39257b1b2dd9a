// Writes a film of per-pixel sums with the adapter's ImageFilm writers.
// usage: film_writer <raw float32 film (h*w*3)> <w> <h> <spp> <out prefix>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "xyz_hip.hpp"

int main(int argc, char** argv) {
  if (argc != 6) return 2;
  uint32_t const w = uint32_t(std::atoi(argv[2])), h = uint32_t(std::atoi(argv[3]));
  float const spp = float(std::atof(argv[4]));
  std::vector<float> film(size_t(w) * h * 3);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f || std::fread(film.data(), sizeof(float), film.size(), f) != film.size()) return 3;
  std::fclose(f);
  std::string const p = argv[5];
  bool ok = xyz_hip::SaveAsPfm((p + ".pfm").c_str(), w, h, film, spp);
  ok = ok && xyz_hip::SaveAsPfmRgb((p + ".rgb.pfm").c_str(), w, h, film, spp);
  ok = ok && xyz_hip::SaveAsFloat((p + ".float").c_str(), w, h, film, spp);
  ok = ok && xyz_hip::SaveAsHdr((p + ".hdr").c_str(), w, h, film, spp);
  return ok ? 0 : 1;
}

// Metasequoia (.mqo, "Format Text Ver 1.0") scene loader feeding
// xyz_scene_create: the reference's Scene::Load (xyz/core/scene_mqo.cpp:547-777)
// with its parsing rules, its material -> BSDF mapping, centimetre -> metre
// scaling, angle-weighted smooth normals and the camera of the Scene chunk.
//
// The parser works on the same stream operations as the reference (token
// reads, getline up to delimiters, brace counting), so a file is read the same
// way, including its quirks: a face that is not a triangle becomes the
// degenerate triangle (0, 0, 0) with material 0; a face line reads exactly
// three attribute tags; a material's tex()/aplane()/bump() consumes the file
// up to the next ')'.  Where the reference's behaviour is undefined (an index
// out of range, a file that ends without "Eof" or mid-chunk, an empty line
// inside a skipped block) the loader reports an error or stops at the end of
// the text.  Triangles keep file order: the reference's final Morton sort
// (scene_mqo.cpp:676-692) is not reproduced, so primitive ids are file order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "spectra.h"
#include "xyz_internal.h"

struct xyz_mqo {
  std::vector<float> v0, v1, v2, n0, n1, n2, uv;
  std::vector<uint32_t> bsdf;
  std::vector<xyz_bsdf_desc> bsdfs;
  xyz_mqo_info info;
};

namespace xyzhip {
namespace {

struct V3 {
  float x = 0, y = 0, z = 0;
};
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// libhi/hi/math/vector3.hpp:121-131 operation order
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length3(V3 a) { return std::sqrt(dot3(a, a)); }
// hi::normalize: a * rsqrt(dot(a, a)) with rsqrt(x) = 1 / sqrt(x)
inline V3 normalize3(V3 a) { return a * (1.0f / std::sqrt(dot3(a, a))); }

void trim(std::string& s) {  // hi::trim (libhi/hi/util/string.hpp:22-34)
  size_t const b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) {
    s.clear();
    return;
  }
  size_t const e = s.find_last_not_of(" \t\r\n");
  s = s.substr(b, e - b + 1);
}

// getline that drops the '\r' of a CRLF line, as a text-mode stream does
bool getline_text(std::istream& in, std::string& s, char delim = '\n') {
  if (!std::getline(in, s, delim)) return false;
  if (delim == '\n' && !s.empty() && s.back() == '\r') s.pop_back();
  return true;
}

void getstring(std::istream& in, std::string& s) {  // scene_mqo.cpp:49-52
  std::getline(in, s, '"');
  std::getline(in, s, '"');
}

// Reads lines until the braces opened on them are closed (scene_mqo.cpp:54-67).
bool skip_block(std::istream& in) {
  int indent = 0;
  std::string line;
  do {
    if (!getline_text(in, line)) return false;
    trim(line);
    if (line.empty()) continue;
    if (line.back() == '{') ++indent;
    else if (line.back() == '}') --indent;
  } while (indent > 0);
  return true;
}

// spd2xyz: binary32 dot products of the spectrum with the colour matching
// functions, summed in wavelength order (libhi/hi/image/spectrum.cpp:478-485,554-559).
void spd2xyz(const double* spd, float xyz[3]) {
  for (int k = 0; k < 3; ++k) {
    float sum = 0.0f;
    for (int i = 0; i < kSpectrumSize; ++i) {
      float const a = float(spd[i]);
      float const b = float(kCie1931Cmf[k][i] / kCmfScale);
      sum += a * b;
    }
    xyz[k] = sum;
  }
}

void clamp01(float xyz[3]) {
  for (int k = 0; k < 3; ++k) {
    if (xyz[k] < 0.0f) xyz[k] = 0.0f;
    else if (xyz[k] > 1.0f) xyz[k] = 1.0f;
  }
}

struct Material {  // MoqMaterial defaults, scene_mqo.cpp:9-36
  std::string name, tex, aplane, bump;
  size_t shader = 3;
  float col[4] = {1, 1, 1, 1};
  float dif = 0.8f, amb = 0.6f, emi = 0, spc = 0, power = 5;
};

// Albedo of a Lambert or Metal material: a Macbeth patch named in the material
// name (first match in table order), else the gamma-expanded colour
// pow(col * spc, 2.2) through the CCIR 601-1 RGB -> XYZ matrix
// (libhi/hi/image/color_space.hpp:58-68); clamped to [0, 1].
void material_xyz(const Material& m, float xyz[3]) {
  for (const NamedSpectrum& s : kMacbeth) {
    if (m.name.find(s.key) != std::string::npos) {
      spd2xyz(s.v, xyz);
      clamp01(xyz);
      return;
    }
  }
  float const r = std::pow(m.col[0] * m.spc, 2.2f);
  float const g = std::pow(m.col[1] * m.spc, 2.2f);
  float const b = std::pow(m.col[2] * m.spc, 2.2f);
  xyz[0] = 0.6069f * r + 0.1735f * g + 0.2003f * b;
  xyz[1] = 0.2989f * r + 0.5866f * g + 0.1145f * b;
  xyz[2] = 0.0000f * r + 0.0661f * g + 1.1162f * b;
  clamp01(xyz);
}

xyz_bsdf_desc make_bsdf(uint32_t type, const float rgb[3], float shininess) {
  xyz_bsdf_desc d;
  std::memset(&d, 0, sizeof(d));
  d.type = type;
  for (int k = 0; k < 3; ++k) d.rgb[k] = rgb ? rgb[k] : 0.0f;
  d.shininess = shininess;
  d.ior = 1.5f;
  d.texture = -1;
  return d;
}

// The material chunk (scene_mqo.cpp:69-236).
void load_materials(std::istream& in, size_t count, std::vector<xyz_bsdf_desc>* out) {
  std::string line, tag, data;
  for (size_t i = 0; i < count; ++i) {
    Material mat;
    getline_text(in, line);
    std::istringstream lin(line);
    getstring(lin, mat.name);
    while (!lin.eof()) {
      std::getline(lin, tag, '(');
      trim(tag);
      if (tag.empty()) break;
      if (tag == "shader") {
        std::getline(lin, data, ')');
        std::istringstream(data) >> mat.shader;
      } else if (tag == "col") {
        std::getline(lin, data, ')');
        std::istringstream din(data);
        din >> mat.col[0] >> mat.col[1] >> mat.col[2] >> mat.col[3];
      } else if (tag == "dif") {
        std::getline(lin, data, ')');
        std::istringstream(data) >> mat.dif;
      } else if (tag == "amb") {
        std::getline(lin, data, ')');
        std::istringstream(data) >> mat.amb;
      } else if (tag == "emi") {
        std::getline(lin, data, ')');
        std::istringstream(data) >> mat.emi;
      } else if (tag == "spc") {
        std::getline(lin, data, ')');
        std::istringstream(data) >> mat.spc;
      } else if (tag == "power") {
        std::getline(lin, data, ')');
        std::istringstream(data) >> mat.power;
      } else if (tag == "tex" || tag == "aplane" || tag == "bump") {
        std::string& s = tag == "tex" ? mat.tex : (tag == "aplane" ? mat.aplane : mat.bump);
        getstring(lin, s);
        std::getline(in, data, ')');  // reads the file stream, as the reference does
      }
    }
    float xyz[3];
    if (mat.emi == 1.0f) {
      out->push_back(make_bsdf(XYZ_LIGHT, nullptr, 0.0f));
    } else if (mat.col[3] == 1.0f) {
      if (mat.spc == 1.0f && mat.power > 0.0f) {
        if (mat.power == 100.0f) {
          out->push_back(make_bsdf(XYZ_MIRROR, nullptr, 0.0f));
        } else {
          material_xyz(mat, xyz);
          out->push_back(make_bsdf(XYZ_METAL, xyz, mat.power));
        }
      } else {
        material_xyz(mat, xyz);
        out->push_back(make_bsdf(XYZ_LAMBERT, xyz, 0.0f));
      }
    } else {
      out->push_back(make_bsdf(XYZ_GLASS, nullptr, 0.0f));
    }
  }
}

struct Face {  // MoqFace defaults, scene_mqo.cpp:38-47
  int v0 = 0, v1 = 0, v2 = 0, m = 0;
  float s0 = 0, t0 = 0, s1 = 1, t1 = 0, s2 = 0, t2 = 1;
};

// One object chunk (scene_mqo.cpp:238-530).  Returns false with *err set on an
// error.
bool load_object(std::istream& in, std::vector<V3>& vertices, xyz_mqo* M, std::string* err) {
  size_t const vn = vertices.size();
  std::string line, tag;
  int vertex_count = 0, triangle_count = 0;
  int visible = 0, shading = 1;
  std::vector<Face> data;
  for (;;) {
    if (!(in >> line)) {
      *err = "object chunk not closed";
      return false;
    }
    if (line == "visible") {
      in >> visible;
    } else if (line == "shading") {
      in >> shading;
    } else if (line == "vertex") {
      if (!visible) {
        if (!skip_block(in)) break;
        continue;
      }
      in >> vertex_count;
      if (vertex_count < 0) vertex_count = 0;
      getline_text(in, line);  // {
      for (int i = 0; i < vertex_count; ++i) {
        getline_text(in, line);
        std::istringstream sin(line);
        V3 p;
        sin >> p.x >> p.y >> p.z;
        vertices.push_back(p * 0.01f);  // cm -> m
      }
      getline_text(in, line);  // }
    } else if (line == "face") {
      if (!visible) {
        if (!skip_block(in)) break;
        continue;
      }
      in >> triangle_count;
      if (triangle_count < 0) triangle_count = 0;
      data.resize(size_t(triangle_count));
      getline_text(in, line);  // {
      for (int i = 0; i < triangle_count; ++i) {
        getline_text(in, line);
        std::replace(line.begin(), line.end(), '(', ' ');
        std::replace(line.begin(), line.end(), ')', ' ');
        std::istringstream sin(line);
        size_t n = 0;
        sin >> n;
        if (n == 3) {
          for (int j = 0; j < 3; ++j) {
            sin >> tag;
            if (tag == "V") {
              sin >> data[i].v0 >> data[i].v1 >> data[i].v2;
            } else if (tag == "M") {
              sin >> data[i].m;
            } else if (tag == "UV") {
              sin >> data[i].s0 >> data[i].t0 >> data[i].s1 >> data[i].t1 >> data[i].s2 >> data[i].t2;
            }
          }
        }
      }
      getline_text(in, line);  // }
    } else if (line == "}") {
      break;
    } else {
      if (!skip_block(in)) break;
    }
  }
  M->info.n_objects++;
  if (!visible) return true;

  size_t const nb = M->bsdfs.size();
  for (int i = 0; i < triangle_count; ++i) {
    const Face& f = data[i];
    int const idx[3] = {f.v0, f.v1, f.v2};
    for (int k = 0; k < 3; ++k) {
      if (idx[k] < 0 || idx[k] >= vertex_count || vn + size_t(idx[k]) >= vertices.size()) {
        *err = "face " + std::to_string(i) + " has a vertex index out of range";
        return false;
      }
    }
    if (f.m < 0 || size_t(f.m) >= nb) {
      *err = "face " + std::to_string(i) + " has a material index out of range";
      return false;
    }
  }
  // angle-weighted vertex normals (scene_mqo.cpp:446-482)
  std::vector<V3> normals;
  if (shading) {
    normals.assign(size_t(vertex_count), V3{});
    for (int i = 0; i < triangle_count; ++i) {
      const Face& f = data[i];
      V3 const p0 = vertices[vn + f.v0], p1 = vertices[vn + f.v1], p2 = vertices[vn + f.v2];
      V3 const ea = p1 - p2, eb = p2 - p0, ec = p0 - p1;
      float const a = length3(ea), b = length3(eb), c = length3(ec);
      // std::max(0.0, 1 - M_1_PI * acos(float)) evaluated in double, stored as float
      auto weight = [](float cosv) {
        double const w = 1.0 - 0.318309886183790671538 * double(std::acos(cosv));
        return float((0.0 < w) ? w : 0.0);
      };
      float const aa = weight(dot3(eb, ec) / (b * c));
      float const ab = weight(dot3(ec, ea) / (c * a));
      float const ac = weight(dot3(ea, eb) / (a * b));
      V3 const n = normalize3(cross3(neg(eb), ec));  // NormalVector(ec, -eb) = cross(-eb, ec)
      normals[f.v0] = normals[f.v0] + n * aa;
      normals[f.v1] = normals[f.v1] + n * ab;
      normals[f.v2] = normals[f.v2] + n * ac;
    }
    for (V3& n : normals) n = normalize3(n);
  }
  float object_area = 0.0f;  // summed per object, then into the scene total
  for (int i = 0; i < triangle_count; ++i) {
    const Face& f = data[i];
    V3 const p[3] = {vertices[vn + f.v0], vertices[vn + f.v1], vertices[vn + f.v2]};
    V3 const e1 = p[1] - p[0], e2 = p[2] - p[0];
    V3 nn[3];
    if (shading) {
      nn[0] = normals[f.v0];
      nn[1] = normals[f.v1];
      nn[2] = normals[f.v2];
    } else {  // flat: every corner carries ng = normalize(cross(e2, e1)) (triangle.cpp:10-33)
      nn[0] = nn[1] = nn[2] = normalize3(cross3(e2, e1));
    }
    std::vector<float>* vd[3] = {&M->v0, &M->v1, &M->v2};
    std::vector<float>* nd[3] = {&M->n0, &M->n1, &M->n2};
    for (int c = 0; c < 3; ++c) {
      vd[c]->insert(vd[c]->end(), {p[c].x, p[c].y, p[c].z});
      nd[c]->insert(nd[c]->end(), {nn[c].x, nn[c].y, nn[c].z});
    }
    float const uv[6] = {f.s0, f.t0, f.s1, f.t1, f.s2, f.t2};
    M->uv.insert(M->uv.end(), uv, uv + 6);
    M->bsdf.push_back(uint32_t(f.m));
    // Triangle::area = length(cross(e2, e1)) * 0.5 (triangle.hpp:57-59)
    float const area = length3(cross3(e2, e1)) * 0.5f;
    object_area += area;
    if (M->bsdfs[f.m].type == XYZ_LIGHT) {
      M->info.n_lights++;
      M->info.light_area += area;
    }
  }
  M->info.surface_area += object_area;
  if (shading) M->info.n_smooth_objects++;
  return true;
}

// The Scene chunk's camera (scene_mqo.cpp:568-618).
bool load_camera(std::istream& in, xyz_mqo_info* info) {
  float pos[3] = {0, 0, 400}, look[3] = {0, 100, 0};
  float head = 0, pitch = 0;
  std::string line, tag;
  int indent = 0;
  do {
    if (!getline_text(in, line)) return false;
    trim(line);
    if (line.empty()) continue;
    if (line.back() == '{') {
      ++indent;
    } else if (line.back() == '}') {
      --indent;
    } else {
      std::istringstream din(line);
      din >> tag;
      if (tag == "pos") din >> pos[0] >> pos[1] >> pos[2];
      else if (tag == "lookat") din >> look[0] >> look[1] >> look[2];
      else if (tag == "head") din >> head;
      else if (tag == "pich") din >> pitch;
    }
  } while (indent > 0);
  {  // pitch (Y-Z)
    float const s = std::sin(pitch), c = std::cos(pitch);
    float const v[3] = {pos[0], pos[1], pos[2]};
    pos[1] = c * v[1] + s * v[2];
    pos[2] = s * v[1] - c * v[2];
  }
  {  // head (Z-X), head -= pi in double, stored as float
    head = float(double(head) - 3.14159265358979323846);
    float const s = std::sin(head), c = std::cos(head);
    float const v[3] = {pos[0], pos[1], pos[2]};
    pos[2] = c * v[2] + s * v[0];
    pos[0] = s * v[2] - c * v[0];
  }
  for (int k = 0; k < 3; ++k) {
    pos[k] *= 0.01f;
    look[k] *= 0.01f;
    info->eye[k] = look[k] + pos[k];
    info->at[k] = look[k];
  }
  info->has_camera = 1;
  return true;
}

}  // namespace

int mqo_load(const char* text, uint64_t len, xyz_mqo** out) {
  if (!out || (!text && len > 0)) {
    set_error("xyz_mqo_load: NULL argument");
    return XYZ_E_INVALID;
  }
  *out = nullptr;
  std::unique_ptr<xyz_mqo> M(new xyz_mqo());
  std::memset(&M->info, 0, sizeof(M->info));
  // the reference camera's defaults (xyz/core/thinlenscamera.cpp:21-28)
  float const eye0[3] = {0.0f, 1.0f, 4.0f}, at0[3] = {0.0f, 1.0f, 0.0f};
  for (int k = 0; k < 3; ++k) {
    M->info.eye[k] = eye0[k];
    M->info.at[k] = at0[k];
  }
  spd2xyz(kMercuryArc, M->info.light_power);
  std::istringstream in(std::string(text ? text : "", size_t(len)));
  std::string line, tag;
  getline_text(in, line);
  if (line != "Metasequoia Document") {
    set_error("xyz_mqo_load: not a Metasequoia document");
    return XYZ_E_INVALID;
  }
  getline_text(in, line);
  if (line != "Format Text Ver 1.0") {
    set_error("xyz_mqo_load: unsupported format (want \"Format Text Ver 1.0\")");
    return XYZ_E_INVALID;
  }
  std::vector<V3> vertices;
  std::string err;
  while (in >> tag) {
    if (tag == "Scene") {
      if (!load_camera(in, &M->info)) break;
    }
    if (tag == "Material") {
      size_t count = 0;
      in >> count;
      if (count > 0) {
        getline_text(in, line);  // {
        load_materials(in, count, &M->bsdfs);
        getline_text(in, line);  // }
      } else {
        float xyz[3];
        spd2xyz(kMacbeth[18].v, xyz);  // mcc-white
        clamp01(xyz);
        M->bsdfs.push_back(make_bsdf(XYZ_LAMBERT, xyz, 0.0f));
        if (!skip_block(in)) break;
      }
    } else if (tag == "Object") {
      getline_text(in, line);  // "name" {
      if (!load_object(in, vertices, M.get(), &err)) {
        set_error("xyz_mqo_load: " + err);
        return XYZ_E_INVALID;
      }
      getline_text(in, line);
    } else if (tag == "Eof") {
      break;
    }
  }
  M->info.n_tris = uint32_t(M->bsdf.size());
  M->info.n_bsdfs = uint32_t(M->bsdfs.size());
  if (M->info.n_tris > 0 && M->bsdfs.empty()) {
    set_error("xyz_mqo_load: faces without a material chunk");
    return XYZ_E_INVALID;
  }
  // every light emits the reference's global power: L_e = Phi / (pi * light area)
  // (xyz/core/scene_pathtracing.cpp:89-90)
  if (M->info.light_area > 0.0f) {
    for (xyz_bsdf_desc& d : M->bsdfs) {
      if (d.type != XYZ_LIGHT) continue;
      for (int k = 0; k < 3; ++k)
        d.rgb[k] = float(double(M->info.light_power[k]) /
                         (3.14159265358979323846 * double(M->info.light_area)));
    }
  }
  *out = M.release();
  return XYZ_OK;
}

}  // namespace xyzhip

extern "C" {

int xyz_mqo_load(const char* text, uint64_t len, xyz_mqo** out) {
  return xyzhip::mqo_load(text, len, out);
}

int xyz_mqo_load_file(const char* path, xyz_mqo** out) {
  if (!path || !out) {
    xyzhip::set_error("xyz_mqo_load_file: NULL argument");
    return XYZ_E_INVALID;
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    xyzhip::set_error(std::string("xyz_mqo_load_file: cannot open ") + path);
    return XYZ_E_INVALID;
  }
  std::ostringstream ss;
  ss << f.rdbuf();
  std::string const s = ss.str();
  return xyzhip::mqo_load(s.data(), s.size(), out);
}

int xyz_mqo_describe(const xyz_mqo* m, xyz_scene_desc* desc, xyz_mqo_info* info) {
  if (!m) {
    xyzhip::set_error("xyz_mqo_describe: NULL scene");
    return XYZ_E_INVALID;
  }
  if (info) *info = m->info;
  if (desc) {
    std::memset(desc, 0, sizeof(*desc));
    desc->n_tris = m->info.n_tris;
    bool const any = m->info.n_tris > 0;
    desc->v0 = any ? m->v0.data() : nullptr;
    desc->v1 = any ? m->v1.data() : nullptr;
    desc->v2 = any ? m->v2.data() : nullptr;
    desc->n0 = any ? m->n0.data() : nullptr;
    desc->n1 = any ? m->n1.data() : nullptr;
    desc->n2 = any ? m->n2.data() : nullptr;
    desc->uv = any ? m->uv.data() : nullptr;
    desc->bsdf = any ? m->bsdf.data() : nullptr;
    desc->n_bsdfs = m->info.n_bsdfs;
    desc->bsdfs = m->bsdfs.empty() ? nullptr : m->bsdfs.data();
  }
  return XYZ_OK;
}

void xyz_mqo_destroy(xyz_mqo* m) { delete m; }

}  // extern "C"

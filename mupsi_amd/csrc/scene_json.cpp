// scene_json.cpp — the reference's scene.json schema, OBJ meshes and HDR images,
// parsed without nlohmann/Eigen/OpenCV, lowered to a mupsi_scene_desc and uploaded.
//
// Schema keys, defaults and quirks follow Controller::load
// (src/controller/controller.cpp:40-264):
//   - "transform": {"position", "scale"} (the README's "translate" is ignored);
//   - "emission" and "int_medium" are read for spheres only; int_medium is
//     presence-only and attaches the GP medium, else the sdf_medium, else the
//     external_gp_medium (whose grid files resolve against the scene file's
//     directory and, as GridMeanFunction, read as zeros where absent);
//   - a mesh "texture" applies only to a lambertian bsdf;
//   - an unknown bsdf name falls back to lambertian, an unknown primitive type
//     is skipped, an unknown mean_type or a missing "primitives" array fails.
// OBJ parsing follows Mesh::fetchFrom (src/primitives/mesh.cpp:20-136): v, vn,
// vt, f with v, v/vt, v//vn, v/vt/vn; negative indices; non-triangles skipped.
// Images: PFM (exact float), Radiance RGBE .hdr; rows stored top-down like
// the OpenCV imread the reference uses.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mupsi.h"
#include "scene_host.h"

namespace {

// ---------------------------------------------------------------- JSON
struct Json {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  bool has(const std::string& k) const { return kind == Obj && obj.count(k) != 0; }
  const Json* get(const std::string& k) const {
    if (kind != Obj) return nullptr;
    auto it = obj.find(k);
    return it == obj.end() ? nullptr : &it->second;
  }
};

struct Parser {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end) {
      if (std::isspace((unsigned char)*p)) {
        ++p;
      } else if (p + 1 < end && p[0] == '/' && p[1] == '/') {  // tolerate // comments
        while (p < end && *p != '\n') ++p;
      } else {
        break;
      }
    }
  }
  bool fail(const char* m) {
    if (err.empty()) err = m;
    return false;
  }
  bool lit(const char* s, size_t n) {
    if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
    p += n;
    return true;
  }
  bool string(std::string& out) {
    if (p >= end || *p != '"') return fail("expected string");
    ++p;
    out.clear();
    while (p < end && *p != '"') {
      if (*p == '\\') {
        ++p;
        if (p >= end) return fail("bad escape");
        char c = *p++;
        switch (c) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (end - p < 4) return fail("bad \\u escape");
            unsigned v = (unsigned)std::strtoul(std::string(p, 4).c_str(), nullptr, 16);
            p += 4;
            if (v < 0x80) {
              out += (char)v;
            } else if (v < 0x800) {
              out += (char)(0xC0 | (v >> 6));
              out += (char)(0x80 | (v & 0x3F));
            } else {
              out += (char)(0xE0 | (v >> 12));
              out += (char)(0x80 | ((v >> 6) & 0x3F));
              out += (char)(0x80 | (v & 0x3F));
            }
            break;
          }
          default: out += c;
        }
      } else {
        out += *p++;
      }
    }
    if (p >= end) return fail("unterminated string");
    ++p;
    return true;
  }
  bool value(Json& v, int depth) {
    if (depth > 64) return fail("nesting too deep");
    ws();
    if (p >= end) return fail("unexpected end of input");
    char c = *p;
    if (c == '{') {
      ++p;
      v.kind = Json::Obj;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      while (true) {
        ws();
        std::string key;
        if (!string(key)) return false;
        ws();
        if (p >= end || *p != ':') return fail("expected ':'");
        ++p;
        Json child;
        if (!value(child, depth + 1)) return false;
        v.obj[key] = std::move(child);
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p;
      v.kind = Json::Arr;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      while (true) {
        Json child;
        if (!value(child, depth + 1)) return false;
        v.arr.push_back(std::move(child));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = Json::Str;
      return string(v.str);
    }
    if (lit("true", 4)) {
      v.kind = Json::Bool;
      v.b = true;
      return true;
    }
    if (lit("false", 5)) {
      v.kind = Json::Bool;
      v.b = false;
      return true;
    }
    if (lit("null", 4)) {
      v.kind = Json::Null;
      return true;
    }
    if (c == '-' || std::isdigit((unsigned char)c)) {
      char* q = nullptr;
      std::string tok;
      const char* s = p;
      while (s < end && (std::isdigit((unsigned char)*s) || *s == '-' || *s == '+' || *s == '.' ||
                         *s == 'e' || *s == 'E'))
        ++s;
      tok.assign(p, s);
      v.num = std::strtod(tok.c_str(), &q);
      if (q == tok.c_str()) return fail("bad number");
      v.kind = Json::Num;
      p += (q - tok.c_str());
      return true;
    }
    return fail("unexpected character");
  }
};

bool parse_json(const std::string& text, Json& out, std::string& err) {
  Parser ps{text.data(), text.data() + text.size(), {}};
  if (!ps.value(out, 0)) {
    err = ps.err;
    return false;
  }
  ps.ws();
  if (ps.p != ps.end) {
    err = "trailing characters after JSON value";
    return false;
  }
  return true;
}

// ---------------------------------------------------------------- schema helpers
struct SchemaError {
  std::string msg;
};

float num_f(const Json* j, float def) {
  if (!j) return def;
  if (j->kind == Json::Num) return (float)j->num;
  if (j->kind == Json::Bool) return j->b ? 1.0f : 0.0f;
  throw SchemaError{"expected a number"};
}
int num_i(const Json* j, int def) {
  if (!j) return def;
  if (j->kind == Json::Num) return (int)j->num;
  throw SchemaError{"expected an integer"};
}
std::string str_v(const Json* j, const std::string& def) {
  if (!j) return def;
  if (j->kind == Json::Str) return j->str;
  throw SchemaError{"expected a string"};
}
void vec3(const Json* j, float out[3]) {  // jsonToVec3 (controller.cpp:32-36)
  if (!j || j->kind != Json::Arr || j->arr.size() < 3) throw SchemaError{"expected [x, y, z]"};
  for (int a = 0; a < 3; ++a) out[a] = num_f(&j->arr[a], 0.0f);
}

// ---------------------------------------------------------------- OBJ
struct ObjMesh {
  std::vector<float> v, vn, vt;
  std::vector<int32_t> f, ft;
};

bool load_obj(const std::string& path, ObjMesh& m, std::string& err) {
  std::ifstream file(path);
  if (!file.is_open()) {
    err = "cannot open OBJ file " + path;
    return false;
  }
  std::string line;
  auto norm_idx = [](int idx, size_t count) -> int32_t {
    if (idx > 0) return idx - 1;
    if (idx < 0) return (int32_t)count + idx;
    return 0;
  };
  while (std::getline(file, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    std::istringstream iss(line);
    std::string type;
    iss >> type;
    if (type == "v") {
      float x = 0, y = 0, z = 0;
      iss >> x >> y >> z;
      m.v.insert(m.v.end(), {x, y, z});
    } else if (type == "vn") {
      float x = 0, y = 0, z = 0;
      iss >> x >> y >> z;
      m.vn.insert(m.vn.end(), {x, y, z});
    } else if (type == "vt") {
      float u = 0, v = 0;
      iss >> u >> v;
      m.vt.insert(m.vt.end(), {u, v});
    } else if (type == "f") {
      std::vector<int32_t> vi, ti;
      std::string spec;
      while (iss >> spec) {
        int a = 0, b = 0;
        try {
          size_t s1 = spec.find('/');
          if (s1 == std::string::npos) {
            a = std::stoi(spec);
          } else {
            a = std::stoi(spec.substr(0, s1));
            size_t s2 = spec.find('/', s1 + 1);
            if (s2 == std::string::npos) {
              b = std::stoi(spec.substr(s1 + 1));
            } else if (s1 + 1 != s2) {
              b = std::stoi(spec.substr(s1 + 1, s2 - s1 - 1));
            }
          }
        } catch (...) {
          err = "malformed face in " + path;
          return false;
        }
        vi.push_back(norm_idx(a, m.v.size() / 3));
        ti.push_back(norm_idx(b, m.vt.size() / 2));
      }
      if (vi.size() != 3) continue;  // < 3 or > 3 vertices: skipped with a warning
      m.f.insert(m.f.end(), vi.begin(), vi.end());
      m.ft.insert(m.ft.end(), ti.begin(), ti.end());
    }
  }
  if (m.v.empty() || m.f.empty()) {
    err = "no valid geometry in " + path;
    return false;
  }
  return true;
}

// ---------------------------------------------------------------- images
struct Image {
  int w = 0, h = 0;
  std::vector<float> rgb;  // row 0 = top
};

bool load_pfm(const std::string& path, Image& img, std::string& err) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) {
    err = "cannot open " + path;
    return false;
  }
  char magic[3] = {0};
  int w = 0, h = 0;
  double scale = 0;
  if (std::fscanf(fp, "%2s %d %d %lf", magic, &w, &h, &scale) != 4 || std::fgetc(fp) == EOF ||
      w <= 0 || h <= 0) {
    std::fclose(fp);
    err = "bad PFM header in " + path;
    return false;
  }
  const bool color = std::strcmp(magic, "PF") == 0;
  if (!color && std::strcmp(magic, "Pf") != 0) {
    std::fclose(fp);
    err = "not a PFM file: " + path;
    return false;
  }
  const int ch = color ? 3 : 1;
  std::vector<float> raw((size_t)w * h * ch);
  const size_t got = std::fread(raw.data(), 4, raw.size(), fp);
  std::fclose(fp);
  if (got != raw.size()) {
    err = "truncated PFM " + path;
    return false;
  }
  const bool big_endian = scale > 0;
  if (big_endian) {
    for (float& f : raw) {
      uint32_t u;
      std::memcpy(&u, &f, 4);
      u = __builtin_bswap32(u);
      std::memcpy(&f, &u, 4);
    }
  }
  img.w = w;
  img.h = h;
  img.rgb.resize((size_t)w * h * 3);
  for (int y = 0; y < h; ++y) {  // PFM rows run bottom-up
    const float* src = raw.data() + (size_t)(h - 1 - y) * w * ch;
    for (int x = 0; x < w; ++x)
      for (int c = 0; c < 3; ++c) img.rgb[((size_t)y * w + x) * 3 + c] = src[x * ch + (color ? c : 0)];
  }
  return true;
}

// Radiance RGBE (.hdr), flat or new-style RLE scanlines.
bool load_hdr(const std::string& path, Image& img, std::string& err) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) {
    err = "cannot open " + path;
    return false;
  }
  char line[512];
  bool fmt_ok = false;
  while (std::fgets(line, sizeof line, fp)) {
    if (line[0] == '\n') break;
    if (std::strstr(line, "FORMAT=32-bit_rle_rgbe")) fmt_ok = true;
  }
  int w = 0, h = 0;
  if (!fmt_ok || !std::fgets(line, sizeof line, fp) || std::sscanf(line, "-Y %d +X %d", &h, &w) != 2 ||
      w <= 0 || h <= 0) {
    std::fclose(fp);
    err = "unsupported HDR header in " + path;
    return false;
  }
  img.w = w;
  img.h = h;
  img.rgb.resize((size_t)w * h * 3);
  std::vector<uint8_t> scan((size_t)w * 4);
  for (int y = 0; y < h; ++y) {
    int c0 = std::fgetc(fp), c1 = std::fgetc(fp), c2 = std::fgetc(fp), c3 = std::fgetc(fp);
    if (c3 == EOF) {
      std::fclose(fp);
      err = "truncated HDR " + path;
      return false;
    }
    if (c0 == 2 && c1 == 2 && ((c2 << 8) | c3) == w && w >= 8 && w < 32768) {
      for (int ch = 0; ch < 4; ++ch) {
        int x = 0;
        while (x < w) {
          int cnt = std::fgetc(fp);
          if (cnt == EOF) break;
          if (cnt > 128) {
            cnt -= 128;
            int val = std::fgetc(fp);
            for (int i = 0; i < cnt && x < w; ++i) scan[(size_t)x++ * 4 + ch] = (uint8_t)val;
          } else {
            for (int i = 0; i < cnt && x < w; ++i) scan[(size_t)x++ * 4 + ch] = (uint8_t)std::fgetc(fp);
          }
        }
      }
    } else {
      scan[0] = (uint8_t)c0;
      scan[1] = (uint8_t)c1;
      scan[2] = (uint8_t)c2;
      scan[3] = (uint8_t)c3;
      if (w > 1 && std::fread(scan.data() + 4, 4, (size_t)w - 1, fp) != (size_t)w - 1) {
        std::fclose(fp);
        err = "truncated HDR " + path;
        return false;
      }
    }
    for (int x = 0; x < w; ++x) {
      const uint8_t* px = scan.data() + (size_t)x * 4;
      const float f = px[3] ? std::ldexp(1.0f, (int)px[3] - (128 + 8)) : 0.0f;
      for (int c = 0; c < 3; ++c) img.rgb[((size_t)y * w + x) * 3 + c] = px[c] * f;
    }
  }
  std::fclose(fp);
  return true;
}

bool load_image(const std::string& path, Image& img, std::string& err) {
  const size_t dot = path.find_last_of('.');
  std::string ext = dot == std::string::npos ? "" : path.substr(dot);
  for (char& c : ext) c = (char)std::tolower((unsigned char)c);
  if (ext == ".pfm") return load_pfm(path, img, err);
  if (ext == ".hdr") return load_hdr(path, img, err);
  err = "unsupported image format (PFM and Radiance HDR are read): " + path;
  return false;
}

// Storage that keeps every array the desc points at alive until upload.
struct DescStore {
  std::vector<std::unique_ptr<ObjMesh>> meshes;
  std::vector<std::unique_ptr<Image>> images;
  std::vector<mupsi_primitive_desc> prims;
  std::vector<std::unique_ptr<std::vector<float>>> grids;
};

// GridMeanFunction(filepath, res, ...) (meanfunction.cpp:41-54): res^3 floats,
// zero-filled first, so a missing or short file leaves zeros.
std::vector<float>* read_grid(DescStore& store, const std::string& path, int res) {
  auto g = std::make_unique<std::vector<float>>((size_t)res * res * res, 0.0f);
  if (FILE* fp = std::fopen(path.c_str(), "rb")) {
    const size_t got = std::fread(g->data(), sizeof(float), g->size(), fp);
    (void)got;
    std::fclose(fp);
  }
  store.grids.push_back(std::move(g));
  return store.grids.back().get();
}

void fill_mesh(const ObjMesh& m, mupsi_mesh_desc& d) {
  std::memset(&d, 0, sizeof d);
  d.vertices = m.v.data();
  d.n_vertices = (int32_t)(m.v.size() / 3);
  d.normals = m.vn.empty() ? nullptr : m.vn.data();
  d.n_normals = (int32_t)(m.vn.size() / 3);
  d.texcoords = m.vt.empty() ? nullptr : m.vt.data();
  d.n_texcoords = (int32_t)(m.vt.size() / 2);
  d.faces = m.f.data();
  d.face_tex = m.vt.empty() ? nullptr : m.ft.data();
  d.n_faces = (int32_t)(m.f.size() / 3);
  d.scale = 1.0f;
}

void read_transform(const Json* t, mupsi_mesh_desc& d) {
  if (!t) return;
  if (t->has("position")) vec3(t->get("position"), d.position);
  d.scale = num_f(t->get("scale"), 1.0f);
}

std::string resolve(const std::string& base_dir, const std::string& p) {
  (void)base_dir;
  return p;  // the reference opens paths as given (relative to the working directory)
}

}  // namespace

extern "C" mupsi_status mupsi_scene_load_json(mupsi_ctx* ctx, const char* path, mupsi_scene** out,
                                              int32_t* spp_out) {
  std::string& g_json_error = mupsi::g_last_error;
  if (!ctx || !path || !out) return MUPSI_ERR_INVALID;
  *out = nullptr;
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) {
    g_json_error = std::string("cannot open ") + path;
    return MUPSI_ERR_IO;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  Json j;
  std::string err;
  if (!parse_json(ss.str(), j, err)) {
    g_json_error = std::string(path) + ": " + err;
    return MUPSI_ERR_PARSE;
  }
  if (j.kind != Json::Obj) {
    g_json_error = "scene.json must hold an object";
    return MUPSI_ERR_PARSE;
  }
  const std::string dir = "";
  DescStore store;
  mupsi_scene_desc d;
  std::memset(&d, 0, sizeof d);
  try {
    d.max_bounce = num_i(j.get("max_bounce"), 3);   // PathTracerSettings default
    if (spp_out) *spp_out = num_i(j.get("spp"), 1);  // Controller default spp_ = 1
    d.width = num_i(j.get("width"), 256);
    d.height = num_i(j.get("height"), 256);
    const Json* cam = j.get("camera");
    if (!cam || cam->kind != Json::Obj) throw SchemaError{"missing camera"};
    vec3(cam->get("pos"), d.cam_pos);
    vec3(cam->get("dir"), d.cam_dir);
    d.cam_up[0] = 0;
    d.cam_up[1] = 1;
    d.cam_up[2] = 0;
    if (cam->has("up")) vec3(cam->get("up"), d.cam_up);
    d.fov = num_f(cam->get("fov"), 45.0f);
    d.gp_mode = str_v(j.get("gp_mode"), "single_realization") == "renewal_plus"
                    ? MUPSI_GP_RENEWAL_PLUS
                    : MUPSI_GP_SINGLE_REALIZATION;
    // mean function of a gp_medium / sdf_medium block (controller.cpp:85-109,140-168)
    const auto parse_mean = [&](const Json* blk, const char* what) -> mupsi_status {
      const std::string mt = str_v(blk->get("mean_type"), "sphere");
      if (mt == "sphere") {
        d.mean_type = MUPSI_MEAN_SPHERE;
        vec3(blk->get("mean_center"), d.mean_center);
        d.mean_radius = num_f(blk->get("mean_radius"), 70.0f);
      } else if (mt == "mesh") {
        d.mean_type = MUPSI_MEAN_MESH;
        auto m = std::make_unique<ObjMesh>();
        if (!load_obj(resolve(dir, str_v(blk->get("file"), "")), *m, err)) {
          g_json_error = std::string(what) + " mesh: " + err;
          return MUPSI_ERR_IO;
        }
        fill_mesh(*m, d.mean_mesh);
        read_transform(blk->get("transform"), d.mean_mesh);
        store.meshes.push_back(std::move(m));
      } else {
        g_json_error = "unknown mean_type: " + mt;
        return MUPSI_ERR_PARSE;
      }
      return MUPSI_OK;
    };
    const Json* gpm = j.get("gp_medium");
    if (!gpm && j.has("sdf_medium")) {
      d.has_gp_medium = 1;
      d.medium_kind = MUPSI_MEDIUM_SDF;
      const mupsi_status ms = parse_mean(j.get("sdf_medium"), "sdf_medium");
      if (ms != MUPSI_OK) return ms;
    } else if (!gpm && j.has("external_gp_medium")) {
      // ExternalGPMedium (controller.cpp:170-189): grid files relative to the
      // scene file's directory
      const Json* e = j.get("external_gp_medium");
      const std::string sp(path);
      const size_t slash = sp.find_last_of("/\\");
      const std::string cfg_dir = slash == std::string::npos ? "" : sp.substr(0, slash + 1);
      d.has_gp_medium = 1;
      d.medium_kind = MUPSI_MEDIUM_GRID;
      mupsi_grid_medium_desc& g = d.grid;
      g.resolution = num_i(e->get("resolution"), 48);
      if (g.resolution < 2 || g.resolution > 1024) throw SchemaError{"external_gp_medium.resolution outside [2, 1024]"};
      vec3(e->get("bounds_min"), g.bounds_min);
      vec3(e->get("bounds_max"), g.bounds_max);
      g.noise_scale = num_f(e->get("noise_scale"), 1.0f);
      g.mean = read_grid(store, cfg_dir + str_v(e->get("mean_file"), ""), g.resolution)->data();
      // the reference builds a variance grid whenever configDir + variance_file is
      // non-empty; a missing file reads as zeros (no noise), as no grid does
      const std::string vf = str_v(e->get("variance_file"), "");
      g.variance = vf.empty() ? nullptr : read_grid(store, cfg_dir + vf, g.resolution)->data();
    }
    if (gpm) {
      d.has_gp_medium = 1;
      d.medium_kind = MUPSI_MEDIUM_GP;
      const mupsi_status ms = parse_mean(gpm, "gp_medium");
      if (ms != MUPSI_OK) return ms;
      const Json* k = j.get("kernel");
      d.kernel.sigma = num_f(k ? k->get("sigma") : nullptr, 1.0f);
      d.kernel.kernel_radius = num_f(k ? k->get("length_scale") : nullptr, 1.0f);
      d.kernel.aniso[0] = d.kernel.aniso[1] = d.kernel.aniso[2] = 1.0f;
      if (k && k->has("aniso")) vec3(k->get("aniso"), d.kernel.aniso);
      d.kernel.points_per_cell = num_i(k ? k->get("points_per_cell") : nullptr, 3);
    }
    if (j.has("skydrome")) {
      auto img = std::make_unique<Image>();
      if (!load_image(resolve(dir, str_v(j.get("skydrome"), "")), *img, err)) {
        g_json_error = "skydrome: " + err;
        return MUPSI_ERR_IO;
      }
      d.sky.rgb = img->rgb.data();
      d.sky.width = img->w;
      d.sky.height = img->h;
      store.images.push_back(std::move(img));
    }
    const Json* prims = j.get("primitives");
    if (!prims || prims->kind != Json::Arr) {
      g_json_error = "no primitives array in config";
      return MUPSI_ERR_PARSE;
    }
    for (const Json& pj : prims->arr) {
      mupsi_primitive_desc pd;
      std::memset(&pd, 0, sizeof pd);
      const std::string type = str_v(pj.get("type"), "sphere");
      const std::string bs = str_v(pj.get("bsdf"), "lambertian");
      pd.bsdf = bs == "null" ? MUPSI_BSDF_NULL
                             : (bs == "specular" ? MUPSI_BSDF_SPECULAR : MUPSI_BSDF_LAMBERTIAN);
      if (type == "sphere") {
        pd.type = MUPSI_PRIM_SPHERE;
        vec3(pj.get("center"), pd.center);
        pd.radius = num_f(pj.get("radius"), 100.0f);
        if (pj.has("emission")) {
          pd.has_emission = 1;
          vec3(pj.get("emission"), pd.emission);
        }
        pd.int_medium = pj.has("int_medium") ? 1 : 0;
      } else if (type == "mesh") {
        pd.type = MUPSI_PRIM_MESH;
        auto m = std::make_unique<ObjMesh>();
        if (!load_obj(resolve(dir, str_v(pj.get("file"), "")), *m, err)) {
          g_json_error = "mesh: " + err;
          return MUPSI_ERR_IO;
        }
        fill_mesh(*m, pd.mesh);
        read_transform(pj.get("transform"), pd.mesh);
        store.meshes.push_back(std::move(m));
        // "emission" on a mesh is ignored by the reference (controller.cpp:225-228);
        // the scene key "extensions": {"mesh_emission": true} honours it
        const Json* ext = j.get("extensions");
        const Json* me = ext ? ext->get("mesh_emission") : nullptr;
        if (pj.has("emission") && me && me->kind == Json::Bool && me->b) {
          pd.has_emission = 1;
          vec3(pj.get("emission"), pd.emission);
        }
        if (pj.has("texture") && bs == "lambertian") {
          auto img = std::make_unique<Image>();
          if (!load_image(resolve(dir, str_v(pj.get("texture"), "")), *img, err)) {
            g_json_error = "texture: " + err;
            return MUPSI_ERR_IO;
          }
          pd.albedo.rgb = img->rgb.data();
          pd.albedo.width = img->w;
          pd.albedo.height = img->h;
          store.images.push_back(std::move(img));
        }
      } else {
        continue;  // unknown primitive type: skipped (controller.cpp:257-259)
      }
      store.prims.push_back(pd);
    }
  } catch (const SchemaError& e) {
    g_json_error = std::string(path) + ": " + e.msg;
    return MUPSI_ERR_PARSE;
  }
  d.primitives = store.prims.empty() ? nullptr : store.prims.data();
  d.n_primitives = (int32_t)store.prims.size();
  return mupsi_scene_upload(ctx, &d, out);
}

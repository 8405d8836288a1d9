// Host scene loader: .nls (JSON) + OBJ + CSV spectra, and flattening to the
// C ABI.  See scene.hpp for the mapping to the reference loader.
#include "scene.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

#include "json.hpp"

namespace lamh {

// ---------------------------------------------------------------------------
// spectra
// ---------------------------------------------------------------------------

static bool read_file(const std::string &path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// std::stof semantics for one CSV cell: leading blanks skipped, the longest
// valid prefix is converted, an empty cell is an error.
static bool cell_to_float(const std::string &cell, float &v) {
  const char *b = cell.c_str();
  char *e = nullptr;
  v = std::strtof(b, &e);
  return e != b;
}

static std::vector<std::string> split_lines(const std::string &text) {
  std::vector<std::string> lines;
  std::string cur;
  for (char c : text) {
    if (c == '\n') { lines.push_back(cur); cur.clear(); }
    else if (c != '\r') cur.push_back(c);
  }
  if (!cur.empty()) lines.push_back(cur);
  return lines;
}

static bool blank(const std::string &s) {
  for (char c : s) if (c != ' ' && c != '\t') return false;
  return true;
}

// sampledλ(fpath): "λ,value" rows (cg.hpp:208-221)
bool read_sampled_csv(const std::string &path, Sampled &out, std::string &err) {
  std::string text;
  if (!read_file(path, text)) { err = "could not open spectrum file '" + path + "'"; return false; }
  out.l.clear(); out.v.clear();
  for (const std::string &line : split_lines(text)) {
    if (blank(line)) continue;
    size_t c = line.find(',');
    std::string a = line.substr(0, c), b = (c == std::string::npos) ? "" : line.substr(c + 1);
    size_t c2 = b.find(',');
    if (c2 != std::string::npos) b = b.substr(0, c2);
    float l, v;
    if (!cell_to_float(a, l) || !cell_to_float(b, v)) {
      err = "malformed spectrum row '" + line + "' in '" + path + "'";
      return false;
    }
    out.l.push_back(l);
    out.v.push_back(v);
  }
  return true;
}

// sampledλ(fpath, minλ, stepλ, i): value column i, λ = min + n*step
// (cg.hpp:222-236)
bool read_sampled_csv_columns(const std::string &path, float minl, float stepl,
                              int column, Sampled &out, std::string &err) {
  std::string text;
  if (!read_file(path, text)) { err = "could not open table '" + path + "'"; return false; }
  out.l.clear(); out.v.clear();
  uint64_t n = 0;
  for (const std::string &line : split_lines(text)) {
    if (blank(line)) continue;
    out.l.push_back(minl + float(n) * stepl);
    n++;
    std::string rest = line, cell;
    for (int k = 0; k <= column; k++) {
      size_t c = rest.find(',');
      cell = rest.substr(0, c);
      rest = (c == std::string::npos) ? "" : rest.substr(c + 1);
    }
    float v;
    if (!cell_to_float(cell, v)) { err = "malformed row in '" + path + "'"; return false; }
    out.v.push_back(v);
  }
  return true;
}

static inline float fmin_ref(float a, float b) { return a <= b ? a : b; }  // numli.hpp:63
static inline float fmax_ref(float a, float b) { return a >= b ? a : b; }  // numli.hpp:64

// coefficientλ(sampledλ const&): trapezoid integral of the samples over
// [inf+iΔ, inf+(i+1)Δ] divided by Δ (cg.hpp:252-272)
Spectrum box_filter(const Grid &g, const Sampled &s) {
  Spectrum out(g, 0.f);
  const float D = g.delta();
  size_t j = 0;
  const size_t L = s.l.size();
  for (uint32_t i = 0; i < g.n; i++) {
    const float l0 = g.inf + float(i) * D;
    const float l1 = l0 + D;
    float sum = 0.f;
    while (j + 1 < L && s.l[j + 1] <= l0) j++;
    size_t k = j;
    while (k + 1 < L && s.l[k] < l1) {
      float seg0 = fmax_ref(l0, s.l[k]);
      float seg1 = fmin_ref(l1, s.l[k + 1]);
      float seg = seg1 - seg0;
      if (seg > 0) sum += .5f * (s.v[k] + s.v[k + 1]) * seg;
      k++;
    }
    out.v[i] = sum / D;
  }
  return out;
}

// Transpose of box_filter with respect to the sample values: the same
// segment walk, each trapezoid .5*(v[k]+v[k+1])*seg/D sending seg/(2D) of
// the coefficient's gradient to both of its end samples.
void box_filter_transpose(const Grid &g, const std::vector<float> &l, const float *dcoef, float *dv) {
  const float D = g.delta();
  const size_t L = l.size();
  for (size_t k = 0; k < L; k++) dv[k] = 0.f;
  size_t j = 0;
  for (uint32_t i = 0; i < g.n; i++) {
    const float l0 = g.inf + float(i) * D;
    const float l1 = l0 + D;
    while (j + 1 < L && l[j + 1] <= l0) j++;
    size_t k = j;
    while (k + 1 < L && l[k] < l1) {
      float seg = fmin_ref(l1, l[k + 1]) - fmax_ref(l0, l[k]);
      if (seg > 0) {
        float w = .5f * seg / D * dcoef[i];
        dv[k] += w;
        dv[k + 1] += w;
      }
      k++;
    }
  }
}

// operator| (cg.hpp:283-288)
float spectrum_dot(const Grid &g, const Spectrum &a, const Spectrum &b) {
  float sum = 0.f;
  for (uint32_t i = 0; i < g.n; i++) sum += a.v[i] * b.v[i];
  return sum * g.delta();
}

// operator()(float) (cg.hpp:289-297)
float spectrum_eval(const Grid &g, const Spectrum &s, float l) {
  if (l <= g.inf) return s.v[0];
  if (l >= g.sup) return s.v[g.n - 1];
  float x = (l - g.inf) / g.delta();
  uint32_t i = std::min<uint32_t>(static_cast<uint32_t>(x), g.n - 2);
  float t = x - float(i);
  return (1.f - t) * s.v[i] + t * s.v[i + 1];
}

// toString with "{:.2f}" per coefficient (cg.hpp:321-326)
std::string spectrum_to_string(const Spectrum &s) {
  std::string out;
  char buf[64];
  for (float x : s.v) {
    std::snprintf(buf, sizeof buf, "%.2f", double(x));
    out += buf;
  }
  return out;
}

static Spectrum scaled(const Spectrum &s, float x) {
  Spectrum o = s;
  for (float &v : o.v) v = v * x;
  return o;
}
static Spectrum added(const Spectrum &a, const Spectrum &b) {
  Spectrum o = a;
  for (size_t i = 0; i < o.v.size(); i++) o.v[i] = a.v[i] + b.v[i];
  return o;
}

std::string default_data_dir() {
  // <package>/data next to the shared library (lambda_amd/liblambda_amd.so)
  Dl_info info;
  if (dladdr(reinterpret_cast<void *>(&default_data_dir), &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t s = p.rfind('/');
    if (s != std::string::npos) return p.substr(0, s) + "/data";
  }
  return "lambda_amd/data";
}

// CMF globals (cg.hpp:329-337)
bool load_color_tables(Scene &sc, const std::string &data_dir, std::string &err) {
  const std::string d = data_dir.empty() ? default_data_dir() : data_dir;
  const char *xyz[3] = {"xyz_x.csv", "xyz_y.csv", "xyz_z.csv"};
  for (int c = 0; c < 3; c++) {
    Sampled s;
    if (!read_sampled_csv(d + "/" + xyz[c], s, err)) return false;
    sc.cmf[c] = box_filter(sc.grid, s);
  }
  for (int c = 0; c < 3; c++) {
    Sampled s;
    if (!read_sampled_csv_columns(d + "/cie1931-basis-bt709-380+5+780.csv", 380.f, 5.f, c, s, err))
      return false;
    sc.bt709[c] = box_filter(sc.grid, s);
  }
  return true;
}

// ---------------------------------------------------------------------------
// OBJ
// ---------------------------------------------------------------------------

// Decimal parse with the method the reference's OBJ reader uses (fast_obj
// 1.3, lib/fast_obj.h:587-668): integer and fraction digits accumulated in
// double, scaled by an exact power of ten, one final rounding to float.
// Using the same method keeps vertex bits identical to the reference's.
static const char *obj_float(const char *p, float &out) {
  while (*p == ' ' || *p == '\t') p++;
  double sign = 1.0;
  if (*p == '+') p++;
  else if (*p == '-') { sign = -1.0; p++; }
  double num = 0.0;
  while (*p >= '0' && *p <= '9') num = 10.0 * num + double(*p++ - '0');
  if (*p == '.') p++;
  double fra = 0.0, div = 1.0;
  while (*p >= '0' && *p <= '9') { fra = 10.0 * fra + double(*p++ - '0'); div *= 10.0; }
  num += fra / div;
  if (*p == 'e' || *p == 'E') {
    p++;
    bool neg = false;
    if (*p == '+') p++;
    else if (*p == '-') { neg = true; p++; }
    unsigned e = 0;
    while (*p >= '0' && *p <= '9') e = 10 * e + unsigned(*p++ - '0');
    if (e >= 20) num *= 0.0;
    else {
      double pw = 1.0;
      for (unsigned k = 0; k < e; k++) pw *= 10.0;  // exact for e < 23
      num = neg ? num * (1.0 / pw) : num * pw;
    }
  }
  out = float(sign * num);
  return p;
}

static const char *obj_int(const char *p, long &out) {
  long sign = 1, v = 0;
  if (*p == '-') { sign = -1; p++; }
  else if (*p == '+') p++;
  while (*p >= '0' && *p <= '9') v = 10 * v + (*p++ - '0');
  out = sign * v;
  return p;
}

bool load_obj(const std::string &path, Mesh &m, std::string &err) {
  std::string text;
  if (!read_file(path, text)) { err = "could not open mesh '" + path + "'"; return false; }
  m.V.assign(1, V3(0.f, 0.f, 0.f));
  m.N.assign(1, V3(0.f, 0.f, 1.f));
  m.FV.clear(); m.FN.clear(); m.GN.clear();
  size_t n_tex = 1;
  for (const std::string &line : split_lines(text)) {
    const char *p = line.c_str();
    while (*p == ' ' || *p == '\t') p++;
    if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
      V3 v;
      p = obj_float(p + 2, v.x); p = obj_float(p, v.y); p = obj_float(p, v.z);
      m.V.push_back(v);
    } else if (p[0] == 'v' && p[1] == 'n') {
      V3 v;
      p = obj_float(p + 2, v.x); p = obj_float(p, v.y); p = obj_float(p, v.z);
      m.N.push_back(v);
    } else if (p[0] == 'v' && p[1] == 't') {
      n_tex++;
    } else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
      p += 2;
      std::vector<uint32_t> fv, fn;
      while (true) {
        while (*p == ' ' || *p == '\t') p++;
        if (*p == 0 || *p == '#') break;
        long v = 0, t = 0, n = 0;
        p = obj_int(p, v);
        if (*p == '/') {
          p++;
          if (*p != '/') p = obj_int(p, t);
          if (*p == '/') { p++; p = obj_int(p, n); }
        }
        (void)t;
        uint32_t vi, ni;
        if (v < 0) vi = uint32_t(long(m.V.size()) + v);
        else if (v > 0) vi = uint32_t(v);
        else break;
        if (n < 0) ni = uint32_t(long(m.N.size()) + n);
        else ni = uint32_t(n);
        fv.push_back(vi);
        fn.push_back(ni);
        while (*p && *p != ' ' && *p != '\t') p++;
      }
      if (fv.size() < 3) continue;
      for (size_t k = 1; k + 1 < fv.size(); k++) {  // fan
        uint32_t a[3] = {0, uint32_t(k), uint32_t(k + 1)};
        for (int q = 0; q < 3; q++) {
          m.FV.push_back(fv[a[q]]);
          m.FN.push_back(fn[a[q]]);
        }
      }
    }
  }
  for (size_t i = 0; i < m.FV.size(); i++) {
    if (m.FV[i] >= m.V.size() || m.FN[i] >= m.N.size()) {
      err = "face index out of range in '" + path + "'";
      return false;
    }
  }
  // geometric normals, cg.hpp:2263-2266
  const size_t nf = m.n_faces();
  m.GN.resize(nf);
  for (size_t f = 0; f < nf; f++) {
    V3 v0 = m.V[m.FV[3 * f]], v1 = m.V[m.FV[3 * f + 1]], v2 = m.V[m.FV[3 * f + 2]];
    V3 e0 = v1 - v0, e2 = v2 - v0;
    m.GN[f] = normalized(cross(e0, e2));
  }
  return true;
}

// ---------------------------------------------------------------------------
// .nls
// ---------------------------------------------------------------------------

namespace {

struct Loader {
  Scene &sc;
  std::string err;
  explicit Loader(Scene &s) : sc(s) {}

  static float f(const Json &j) { return float(j.num()); }
  static V3 v3(const Json &j) { return V3(f(j[0]), f(j[1]), f(j[2])); }

  // loadTransform (cg.hpp:2003-2019)
  Xform transform(const Json &j) {
    V3 s(1.f, 1.f, 1.f), t(0.f, 0.f, 0.f), axis(0.f, 0.f, 1.f);
    float deg = 0.f;
    if (j.has("scale")) { try { s = v3(j.at("scale")); } catch (...) { s = V3(1.f, 1.f, 1.f); } }
    if (j.has("translate")) { try { t = v3(j.at("translate")); } catch (...) { t = V3(0.f, 0.f, 0.f); } }
    if (j.has("rotate")) {
      try {
        const Json &r = j.at("rotate");
        axis = v3(r.at("axis"));
        deg = f(r.at("degrees"));
      } catch (...) { axis = V3(0.f, 0.f, 1.f); deg = 0.f; }
    }
    return xf_then(xf_then(xf_scale(s), xf_rotate(axis, deg)), xf_translate(t));
  }

  // loadSpectrum (cg.hpp:2020-2034)
  Spectrum spectrum(const Json &j) {
    const Json &val = j.at("val");
    Spectrum s;
    if (val.is_string()) {
      Sampled smp;
      std::string e;
      if (!read_sampled_csv(val.str(), smp, e)) throw JsonError(e);
      s = box_filter(sc.grid, smp);
    } else if (val.is_number()) {
      s = Spectrum(sc.grid, f(val));
    } else {
      V3 c = v3(val);
      // linRGB2coefλ (cg.hpp:342-343)
      s = added(added(scaled(sc.bt709[0], c.x), scaled(sc.bt709[1], c.y)), scaled(sc.bt709[2], c.z));
    }
    float mul = 1.f;
    if (j.has("mult")) { try { mul = f(j.at("mult")); } catch (...) {} }
    return scaled(s, mul);
  }

  // loadCamera (cg.hpp:2039-2063) + camera::init (cg.hpp:1158-1164)
  void camera(const Json &j) {
    Camera &c = sc.cam;
    V3 pos = v3(j.at("pos")), look = v3(j.at("look_at")), up = v3(j.at("up"));
    uint64_t width = uint64_t(j.at("width").num());
    float fov = f(j.at("fov"));
    float ar = 1.77778f, dof = 0.f, fd = 1.f;
    if (j.has("ar")) ar = f(j.at("ar"));
    if (j.has("dof")) dof = f(j.at("dof"));
    if (j.has("f")) fd = f(j.at("f"));
    c.fov = fov; c.dof = dof; c.D = fd;
    c.width = width;
    c.height = uint64_t(std::ceil(float(width) / ar));
    V3 z = normalized(pos - look);
    V3 x = cross(up, z);
    V3 y = cross(z, x);
    c.x = x; c.y = y; c.z = z; c.pos = pos;
    const float pi = 3.14159265358979323846f;
    c.delta = 2 * c.D * std::tan(.5f * (fov * pi / 180.f)) / float(c.height);
    c.h = c.delta * float(c.height);
    c.w = c.delta * float(c.width);
  }

  uint32_t material_index(const std::string &name) {
    for (size_t i = 0; i < sc.materials.size(); i++)
      if (sc.materials[i].name == name) return uint32_t(i);
    throw JsonError("unknown material '" + name + "'");
  }

  // loadMaterials (cg.hpp:2359-2444)
  void materials(const Json &j) {
    for (size_t i = 0; i < j.size(); i++) {
      const Json &jm = j[i];
      Material m;
      m.name = jm.at("name").str();
      const std::string type = jm.at("type").str();
      if (type == "lambertian") {
        m.type = LAM_MAT_LAMBERTIAN;
        m.s[0] = spectrum(jm.at("albedo"));
      } else if (type == "blinn") {
        m.type = LAM_MAT_BLINN;
        const char *keys[5] = {"Kd", "Ks", "Kt", "Le", "ior"};
        for (int k = 0; k < 5; k++) {
          try { m.s[k] = spectrum(jm.at(keys[k])); }
          catch (...) { m.s[k] = Spectrum(sc.grid, k == 4 ? 1.54f : 0.f); }
        }
        m.alpha = 1024.f;
        if (jm.has("glossiness")) { try { m.alpha = f(jm.at("glossiness")); } catch (...) {} }
        // blinn::init (cg.hpp:931-940)
        float diff = std::sqrt(spectrum_dot(sc.grid, m.s[0], m.s[0]));
        m.w_r = std::sqrt(spectrum_dot(sc.grid, m.s[1], m.s[1]));
        m.w_t = std::sqrt(spectrum_dot(sc.grid, m.s[2], m.s[2]));
        float tot = diff + m.w_r + m.w_t;
        if (!(tot > 0.f)) throw JsonError("blinn material '" + m.name + "' has no lobe (Kd, Ks, Kt all zero)");
        m.p_d = diff / tot;
        m.p_spec = (m.w_r + m.w_t) / tot;
      } else if (type == "thinfilm") {
        m.type = LAM_MAT_THINFILM;
        m.s[0] = spectrum(jm.at("film_ior"));
        try { m.s[1] = spectrum(jm.at("base_ior")); } catch (...) { m.s[1] = Spectrum(sc.grid, 1.54f); }
        m.alpha = 1024.f;
        m.d = 400.f;
        if (jm.has("glossiness")) { try { m.alpha = f(jm.at("glossiness")); } catch (...) {} }
        if (jm.has("thickness")) { try { m.d = f(jm.at("thickness")); } catch (...) {} }
      } else if (type == "microfacet") {
        m.type = LAM_MAT_MICROFACET;
      } else {
        continue;  // MaterialType::NONE is skipped (cg.hpp:2409)
      }
      sc.materials.push_back(m);
    }
  }

  int mesh_index(const std::string &path) {
    for (size_t i = 0; i < sc.meshes.size(); i++)
      if (sc.meshes[i].name == path) return int(i);
    Mesh m;
    m.name = path;
    std::string e;
    if (!load_obj(path, m, e)) throw JsonError(e);
    sc.meshes.push_back(std::move(m));
    return int(sc.meshes.size() - 1);
  }

  // loadObjects (cg.hpp:2307-2353); groups ignore their own transform
  void objects(const Json &j) {
    for (size_t i = 0; i < j.size(); i++) {
      const Json &jo = j[i];
      Object o;
      o.name = jo.at("name").str();
      const std::string type = jo.at("type").str();
      if (type == "group") { objects(jo.at("children")); continue; }
      if (type == "sphere") o.type = LAM_OBJ_SPHERE;
      else if (type == "plane") o.type = LAM_OBJ_PLANE;
      else if (type == "trimesh") o.type = LAM_OBJ_TRIMESH;
      else throw JsonError("unknown object type '" + type + "'");
      o.T = transform(jo.at("transform"));
      o.mat = material_index(jo.at("material").str());
      if (o.type == LAM_OBJ_TRIMESH) o.mesh = uint32_t(mesh_index(jo.at("source").str()));
      o.obj = uint32_t(sc.objects.size());
      sc.objects.push_back(o);
    }
  }

  // emitter auto-material, named by the radiance's toString (cg.hpp:2089-2095, 2113-2119)
  uint32_t emitter_material(const Spectrum &rad, bool directional, float spec) {
    const std::string name = "emitter_" + spectrum_to_string(rad);
    for (size_t i = 0; i < sc.materials.size(); i++)
      if (sc.materials[i].name == name) return uint32_t(i);
    Material m;
    m.name = name;
    m.type = directional ? LAM_MAT_DIREMITTER : LAM_MAT_EMITTER;
    m.s[0] = rad;
    m.specular = spec;
    sc.materials.push_back(m);
    return uint32_t(sc.materials.size() - 1);
  }

  // loadLights (cg.hpp:2124-2186)
  void lights(const Json &j) {
    for (size_t i = 0; i < j.size(); i++) {
      const Json &jl = j[i];
      Light l;
      l.name = jl.at("name").str();
      const std::string type = jl.at("type").str();
      if (type == "ambient") {
        l.type = LAM_LIGHT_AMBIENT;
        l.radiance = spectrum(jl.at("radiance"));
      } else if (type == "point") {
        l.type = LAM_LIGHT_POINT;
        l.radiance = spectrum(jl.at("radiant_intensity"));
        l.pos = v3(jl.at("pos"));
      } else if (type == "direction") {
        l.type = LAM_LIGHT_DIR;
        l.radiance = spectrum(jl.at("radiant_intensity"));
        l.dir = v3(jl.at("dir"));
      } else if (type == "sphere") {
        l.type = LAM_LIGHT_SPHERE;
        l.radiance = spectrum(jl.at("radiance"));
        Object s;
        s.type = LAM_OBJ_SPHERE;
        s.T = transform(jl.at("transform"));
        float sx = l2(s.T.column(0)), sy = l2(s.T.column(1)), sz = l2(s.T.column(2));
        if (std::fabs(sx - sy) >= 0.01f || std::fabs(sx - sz) >= 0.01f)
          throw JsonError("sphere light '" + l.name + "' needs a uniform scale");
        s.mat = emitter_material(l.radiance, false, 0.f);
        s.name = l.name;
        l.size = sx;
        s.obj = uint32_t(sc.objects.size());
        l.object = s.obj;
        sc.objects.push_back(s);
      } else if (type == "quaddir") {
        l.type = LAM_LIGHT_QUADDIR;
        l.radiance = spectrum(jl.at("radiance"));
        Object p;
        p.type = LAM_OBJ_PLANE;
        p.T = transform(jl.at("transform"));
        float spec = 1024.f;
        if (jl.has("specular")) { try { spec = f(jl.at("specular")); } catch (...) {} }
        p.mat = emitter_material(l.radiance, true, spec);
        p.name = l.name;
        l.spec = spec;
        l.inv_area = 1.f / (l2(p.T.column(0)) * l2(p.T.column(1)));
        p.obj = uint32_t(sc.objects.size());
        l.object = p.obj;
        sc.objects.push_back(p);
      } else {
        throw JsonError("unknown light type '" + type + "'");
      }
      sc.lights.push_back(l);
    }
  }
};

}  // namespace

int load_nls_text(const std::string &text, Scene &sc, std::string &err,
                  const std::string &data_dir) {
  Json j;
  try {
    j = Json::parse(text);
  } catch (const std::exception &e) {
    err = std::string("scene JSON: ") + e.what();
    return LAM_E_PARSE;
  }
  sc = Scene();
  try {
    if (j.has("spectral_grid")) {
      const Json &g = j.at("spectral_grid");
      sc.grid.n = uint32_t(g.at("n").num());
      sc.grid.inf = float(g.at("inf").num());
      sc.grid.sup = float(g.at("sup").num());
      if (sc.grid.n < 2 || !(sc.grid.sup > sc.grid.inf)) throw JsonError("bad spectral_grid");
    }
    if (j.has("hero_samples")) {
      sc.hero = uint32_t(j.at("hero_samples").num());
      if (sc.hero != 4 && sc.hero != 16) throw JsonError("hero_samples must be 4 or 16");
    }
  } catch (const std::exception &e) {
    err = e.what();
    return LAM_E_PARSE;
  }
  if (!load_color_tables(sc, data_dir, err)) return LAM_E_IO;
  Loader L(sc);
  try {
    L.camera(j.at("camera"));
    L.materials(j.at("materials"));
    L.objects(j.at("objects"));
    L.lights(j.at("lights"));
  } catch (const std::exception &e) {
    err = e.what();
    return LAM_E_PARSE;
  }
  if (sc.lights.empty()) {  // the reference divides by zero (cg.hpp:1554)
    err = "scene has no lights";
    return LAM_E_PARSE;
  }
  return LAM_OK;
}

int load_nls(const std::string &path, Scene &sc, std::string &err,
             const std::string &data_dir) {
  std::string text;
  if (!read_file(path, text)) {
    err = "could not open scene file '" + path + "'";
    return LAM_E_IO;
  }
  return load_nls_text(text, sc, err, data_dir);
}

// ---------------------------------------------------------------------------
// flattening
// ---------------------------------------------------------------------------

void FlatScene::refresh() {
  desc.n_spectra = uint32_t(spectra.size() / (desc.grid.n ? desc.grid.n : 1));
  desc.spectra = spectra.data();
  desc.cmf = cmf.data();
  desc.n_objects = uint32_t(objects.size());
  desc.objects = objects.data();
  desc.n_materials = uint32_t(materials.size());
  desc.materials = materials.data();
  desc.n_lights = uint32_t(lights.size());
  desc.lights = lights.data();
  desc.n_tris = uint32_t(tri_obj.size());
  desc.tri_v = tri_v.data();
  desc.tri_n = tri_n.data();
  desc.tri_gn = tri_gn.data();
  desc.tri_obj = tri_obj.data();
  desc.tri_prim = tri_prim.data();
}

void flatten(const Scene &sc, FlatScene &out) {
  out = FlatScene();
  out.desc.grid.n = sc.grid.n;
  out.desc.grid.inf = sc.grid.inf;
  out.desc.grid.sup = sc.grid.sup;
  // spectra table with exact de-duplication
  std::map<std::vector<float>, int32_t> seen;
  auto spec_index = [&](const Spectrum &s) -> int32_t {
    if (s.empty()) return -1;
    auto it = seen.find(s.v);
    if (it != seen.end()) return it->second;
    int32_t idx = int32_t(out.spectra.size() / sc.grid.n);
    out.spectra.insert(out.spectra.end(), s.v.begin(), s.v.end());
    seen[s.v] = idx;
    return idx;
  };
  for (int c = 0; c < 3; c++) out.cmf.insert(out.cmf.end(), sc.cmf[c].v.begin(), sc.cmf[c].v.end());
  for (const Material &m : sc.materials) {
    lam_material lm{};
    lm.type = m.type;
    for (int k = 0; k < 5; k++) lm.spec[k] = spec_index(m.s[k]);
    lm.alpha = m.alpha; lm.d = m.d; lm.p_d = m.p_d; lm.p_spec = m.p_spec;
    lm.w_r = m.w_r; lm.w_t = m.w_t; lm.specular = m.specular;
    out.materials.push_back(lm);
  }
  for (const Light &l : sc.lights) {
    lam_light ll{};
    ll.type = l.type;
    ll.object = l.object;
    ll.spec = spec_index(l.radiance);
    ll.size = l.size; ll.inv_area = l.inv_area; ll.specular = l.spec;
    for (int k = 0; k < 3; k++) { ll.pos[k] = l.pos[k]; ll.dir[k] = l.dir[k]; }
    out.lights.push_back(ll);
  }
  for (const Object &o : sc.objects) {
    lam_object lo{};
    lo.type = o.type;
    lo.material = o.mat;
    lo.mesh = o.mesh;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 4; c++) {
        lo.T.M[r * 4 + c] = o.T.M(r, c);
        lo.T.Minv[r * 4 + c] = o.T.Minv(r, c);
      }
    out.objects.push_back(lo);
    if (o.type != LAM_OBJ_TRIMESH) continue;
    const Mesh &m = sc.meshes[o.mesh];
    const size_t nf = m.n_faces();
    for (size_t f = 0; f < nf; f++) {
      for (int k = 0; k < 3; k++) {
        V3 v = o.T.pntToWorld(m.V[m.FV[3 * f + k]]);
        V3 n = o.T.normalToWorld(m.N[m.FN[3 * f + k]]);
        for (int q = 0; q < 3; q++) { out.tri_v.push_back(v[q]); out.tri_n.push_back(n[q]); }
      }
      V3 gn = normalized(o.T.normalToWorld(m.GN[f]));
      for (int q = 0; q < 3; q++) out.tri_gn.push_back(gn[q]);
      out.tri_obj.push_back(o.obj);
      out.tri_prim.push_back(uint32_t(f));
    }
  }
  const Camera &c = sc.cam;
  lam_camera &lc = out.desc.camera;
  for (int k = 0; k < 3; k++) {
    lc.pos[k] = c.pos[k]; lc.x[k] = c.x[k]; lc.y[k] = c.y[k]; lc.z[k] = c.z[k];
  }
  lc.D = c.D; lc.delta = c.delta; lc.w = c.w; lc.h = c.h; lc.dof = c.dof; lc.fov = c.fov;
  lc.width = uint32_t(c.width);
  lc.height = uint32_t(c.height);
  out.refresh();
}

}  // namespace lamh

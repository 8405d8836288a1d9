// Host scene model and .nls loader.
//
// Mirrors nl::cg::scene (lib/numli/cg.hpp:1167-1180) and the loader in
// nl::cg::load (cg.hpp:1991-2474): same JSON schema, same defaults, same
// object/material/light ordering, same emitter auto-materials.  Spectra
// are coefficient spectra on a runtime grid (the reference's fixed
// coefficientλ<80> on [400,720] is the default grid).
//
// Differences, all deliberate:
//  * load failures return an error code with a message instead of
//    crashing (missing OBJ: the reference dereferences NULL, cg.hpp:2213;
//    unknown material name: the reference indexes past the list);
//  * OBJ faces with more than 3 vertices are fan-triangulated;
//  * optional top-level keys "spectral_grid": {"n","inf","sup"} and
//    "hero_samples" select a non-reference spectral grid / hero count.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "lambda_amd.h"
#include "linalg.hpp"

namespace lamh {

struct Grid {
  uint32_t n = 80;
  float inf = 400.f, sup = 720.f;
  float delta() const { return (sup - inf) / float(n - 1); }  // cg.hpp:246
};

// coefficientλ on a runtime grid (cg.hpp:241-327)
struct Spectrum {
  std::vector<float> v;
  Spectrum() = default;
  Spectrum(const Grid &g, float x) : v(g.n, x) {}
  bool empty() const { return v.empty(); }
};

// sampledλ (cg.hpp:204-237)
struct Sampled {
  std::vector<float> l, v;
};

struct Camera {
  V3 pos, x, y, z;
  float fov = 0, dof = 0, D = 1, delta = 0, h = 0, w = 0;
  uint64_t width = 0, height = 0;
};

struct Material {
  uint32_t type = LAM_MAT_NONE;
  std::string name;
  Spectrum s[5];
  float alpha = 0, d = 0, p_d = 0, p_spec = 0, w_r = 0, w_t = 0, specular = 0;
};

struct Object {
  uint32_t type = 0;
  std::string name;
  Xform T;
  uint32_t mat = 0, obj = 0, mesh = LAM_NONE;
};

struct Light {
  uint32_t type = 0;
  std::string name;
  Spectrum radiance;
  float size = 0, inv_area = 0, spec = 0;
  V3 pos, dir;
  uint32_t object = LAM_NONE;
};

// trimeshdata (cg.hpp:766-820).  Index 0 of V and N is the dummy entry
// fast_obj inserts ((0,0,0) and (0,0,1)), so OBJ indices map directly and a
// face without `vn` uses the (0,0,1) normal as in the reference.
struct Mesh {
  std::string name;  // source path (de-dup key, cg.hpp:2295)
  std::vector<V3> V, N, GN;
  std::vector<uint32_t> FV, FN;  // 3 per face
  size_t n_faces() const { return FV.size() / 3; }
};

struct Scene {
  Grid grid;
  uint32_t hero = 16;
  Camera cam;
  std::vector<Material> materials;
  std::vector<Object> objects;
  std::vector<Light> lights;
  std::vector<Mesh> meshes;
  Spectrum cmf[3];    // CMF_X/Y/Z (cg.hpp:329-331)
  Spectrum bt709[3];  // bt_709_r/g/b (cg.hpp:332-337)
};

// ---- spectra ------------------------------------------------------------
bool read_sampled_csv(const std::string &path, Sampled &out, std::string &err);
bool read_sampled_csv_columns(const std::string &path, float minl, float stepl,
                              int column, Sampled &out, std::string &err);
Spectrum box_filter(const Grid &g, const Sampled &s);          // cg.hpp:252-272
// dv[k] = Σ_i dcoef[i] ∂box_filter(g, {l, v})[i]/∂v[k] (box_filter is linear in v)
void box_filter_transpose(const Grid &g, const std::vector<float> &l, const float *dcoef, float *dv);
float spectrum_dot(const Grid &g, const Spectrum &a, const Spectrum &b);  // cg.hpp:283-288
float spectrum_eval(const Grid &g, const Spectrum &s, float l);  // cg.hpp:289-297
std::string spectrum_to_string(const Spectrum &s);               // cg.hpp:321-326

// Directory holding xyz_{x,y,z}.csv and the BT.709 basis table.
std::string default_data_dir();
bool load_color_tables(Scene &sc, const std::string &data_dir, std::string &err);

// ---- loader ---------------------------------------------------------------
bool load_obj(const std::string &path, Mesh &m, std::string &err);
// loadNLS (cg.hpp:2455-2470).  Returns LAM_OK or LAM_E_IO / LAM_E_PARSE.
int load_nls(const std::string &path, Scene &sc, std::string &err,
             const std::string &data_dir = "");
int load_nls_text(const std::string &text, Scene &sc, std::string &err,
                  const std::string &data_dir = "");

// ---- flattening to the C ABI ---------------------------------------------
struct FlatScene {
  lam_scene_desc desc{};
  std::vector<float> spectra, cmf, tri_v, tri_n, tri_gn;
  std::vector<uint32_t> tri_obj, tri_prim;
  std::vector<lam_object> objects;
  std::vector<lam_material> materials;
  std::vector<lam_light> lights;
  void refresh();  // point desc at the vectors
};
void flatten(const Scene &sc, FlatScene &out);

}  // namespace lamh

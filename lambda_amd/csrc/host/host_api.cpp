// lambda_host.h implementation: loader handle, tone map, PNG, BVH checks.
#include "lambda_host.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bvh8.hpp"
#include "scene.hpp"

struct lamh_scene {
  lamh::Scene scene;
  lamh::FlatScene flat;
};

namespace {

void set_err(char *err, size_t len, const std::string &msg) {
  if (!err || len == 0) return;
  std::snprintf(err, len, "%s", msg.c_str());
}

int finish_load(int rc, lamh_scene *s, lamh_scene **out, char *err, size_t errlen, const std::string &msg) {
  if (rc != LAM_OK) {
    set_err(err, errlen, msg);
    delete s;
    return rc;
  }
  lamh::flatten(s->scene, s->flat);
  *out = s;
  return LAM_OK;
}

// ---- PNG (RGB8, zlib stored blocks) ----------------------------------------
uint32_t crc_table[256];
bool crc_ready = false;
void crc_init() {
  for (uint32_t n = 0; n < 256; n++) {
    uint32_t c = n;
    for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
    crc_table[n] = c;
  }
  crc_ready = true;
}
uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
  if (!crc_ready) crc_init();
  for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c;
}
void put32(std::vector<uint8_t> &v, uint32_t x) {
  v.push_back(uint8_t(x >> 24)); v.push_back(uint8_t(x >> 16));
  v.push_back(uint8_t(x >> 8)); v.push_back(uint8_t(x));
}
void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
  put32(out, uint32_t(data.size()));
  std::vector<uint8_t> td(type, type + 4);
  td.insert(td.end(), data.begin(), data.end());
  out.insert(out.end(), td.begin(), td.end());
  put32(out, crc32(td.data(), td.size()) ^ 0xFFFFFFFFu);
}

// linRGB2sRGB + sRGB2rgb24 (cg.hpp:149-162, numli.hpp:105-106)
uint8_t to_byte(float cl) {
  const float s = cl > 0.0031308f ? 1.055f * std::pow(cl, 1.f / 2.4f) - 0.055f : 12.92f * cl;
  const float v = 255 * s + 0.5f;
  if (!(v > 0.f)) return 0;
  if (v >= 255.f) return 255;
  return uint8_t(int(v));
}

}  // namespace

extern "C" {

int lamh_load_nls(const char *path, const char *data_dir, lamh_scene **out, char *err, size_t errlen) {
  if (!path || !out) return LAM_E_INVALID;
  *out = nullptr;
  lamh_scene *s = new lamh_scene();
  std::string msg;
  const int rc = lamh::load_nls(path, s->scene, msg, data_dir ? data_dir : "");
  return finish_load(rc, s, out, err, errlen, msg);
}

int lamh_load_nls_text(const char *text, const char *data_dir, lamh_scene **out, char *err, size_t errlen) {
  if (!text || !out) return LAM_E_INVALID;
  *out = nullptr;
  lamh_scene *s = new lamh_scene();
  std::string msg;
  const int rc = lamh::load_nls_text(text, s->scene, msg, data_dir ? data_dir : "");
  return finish_load(rc, s, out, err, errlen, msg);
}

void lamh_scene_free(lamh_scene *s) { delete s; }

const lam_scene_desc *lamh_scene_desc(const lamh_scene *s) { return s ? &s->flat.desc : nullptr; }

uint32_t lamh_scene_hero(const lamh_scene *s) { return s ? s->scene.hero : 16; }

const char *lamh_object_name(const lamh_scene *s, uint32_t i) {
  if (!s || i >= s->scene.objects.size()) return nullptr;
  return s->scene.objects[i].name.c_str();
}
const char *lamh_material_name(const lamh_scene *s, uint32_t i) {
  if (!s || i >= s->scene.materials.size()) return nullptr;
  return s->scene.materials[i].name.c_str();
}

int lamh_spectrum_grad_to_samples(const lam_grid *grid, const float *sample_l, size_t n_samples, float mult,
                                   const float *dcoef, float *dsamples) {
  if (!grid || !sample_l || !dcoef || !dsamples || n_samples == 0 || grid->n < 2) return LAM_E_INVALID;
  for (size_t k = 1; k < n_samples; k++)
    if (!(sample_l[k] > sample_l[k - 1])) return LAM_E_INVALID;
  lamh::Grid g;
  g.n = grid->n;
  g.inf = grid->inf;
  g.sup = grid->sup;
  std::vector<float> l(sample_l, sample_l + n_samples);
  lamh::box_filter_transpose(g, l, dcoef, dsamples);
  for (size_t k = 0; k < n_samples; k++) dsamples[k] *= mult;
  return LAM_OK;
}

int lamh_tone_map(const float *in, uint32_t w, uint32_t h, float y_mid, float *out) {
  if (!in || !out) return LAM_E_INVALID;
  const size_t N = size_t(w) * h;
  if (N == 0) return LAM_OK;
  float maxY = 0.f, sumlogY = 0.f;
  const float eps = FLT_EPSILON;
  for (size_t i = 0; i < N; i++) {
    const float Y = 0.2126f * in[3 * i] + 0.7152f * in[3 * i + 1] + 0.0722f * in[3 * i + 2];
    if (Y > maxY) maxY = Y;
    sumlogY += std::log(Y >= eps ? Y : eps);
  }
  sumlogY /= float(N);
  const float Ybar = std::exp(sumlogY);
  const float exposure = y_mid / Ybar;
  for (size_t i = 0; i < 3 * N; i++) {
    const float x = 10 * (in[i] * exposure);
    // koiFilmic = filmicToneMap(10L, .3, .14, .3, 2.8, 10) (cg.hpp:181-194)
    out[i] = (x * (0.3f * x + 0.14f)) / (x * (0.3f * x + 2.8f) + 10.f);
  }
  return LAM_OK;
}

int lamh_write_png(const char *path, const float *rgb, uint32_t w, uint32_t h) {
  if (!path || !rgb || w == 0 || h == 0) return LAM_E_INVALID;
  std::vector<uint8_t> px(size_t(w) * h * 3);
  for (size_t i = 0; i < px.size(); i++) px[i] = to_byte(rgb[i]);
  return lamh_write_png8(path, px.data(), w, h);
}

int lamh_write_png8(const char *path, const uint8_t *rgb8, uint32_t w, uint32_t h) {
  if (!path || !rgb8 || w == 0 || h == 0) return LAM_E_INVALID;
  std::vector<uint8_t> raw;
  raw.reserve(size_t(h) * (size_t(w) * 3 + 1));
  for (uint32_t y = 0; y < h; y++) {
    raw.push_back(0);  // filter: none
    raw.insert(raw.end(), rgb8 + size_t(y) * w * 3, rgb8 + size_t(y + 1) * w * 3);
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  uint32_t a = 1, b = 0;
  for (uint8_t v : raw) { a = (a + v) % 65521u; b = (b + a) % 65521u; }
  size_t pos = 0;
  do {
    const size_t n = std::min<size_t>(65535, raw.size() - pos);
    const bool last = pos + n == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back(uint8_t(n)); z.push_back(uint8_t(n >> 8));
    z.push_back(uint8_t(~n)); z.push_back(uint8_t((~n) >> 8));
    z.insert(z.end(), raw.begin() + long(pos), raw.begin() + long(pos + n));
    pos += n;
  } while (pos < raw.size());
  put32(z, (b << 16) | a);
  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  std::vector<uint8_t> ihdr;
  put32(ihdr, w); put32(ihdr, h);
  ihdr.push_back(8); ihdr.push_back(2); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
  chunk(png, "IHDR", ihdr);
  chunk(png, "IDAT", z);
  chunk(png, "IEND", {});
  FILE *f = std::fopen(path, "wb");
  if (!f) return LAM_E_IO;
  const size_t wr = std::fwrite(png.data(), 1, png.size(), f);
  std::fclose(f);
  return wr == png.size() ? LAM_OK : LAM_E_IO;
}

int lamh_image_name(const char *scene_path, uint32_t spp, uint32_t bounces, float p, const char *suffix, char *out,
                    size_t outlen) {
  if (!scene_path || !out) return LAM_E_INVALID;
  const std::string f = scene_path;
  if (f.size() < 10) return LAM_E_INVALID;
  char pb[32];
  std::snprintf(pb, sizeof pb, "%.2f", double(p));
  const std::string name = "render" + f.substr(6, f.size() - 10) + "-" + std::to_string(spp) + "spp-" +
                           std::to_string(bounces) + "b-" + pb + "p" + (suffix ? suffix : "") + ".png";
  if (name.size() + 1 > outlen) return LAM_E_LIMIT;
  std::memcpy(out, name.c_str(), name.size() + 1);
  return LAM_OK;
}

int lamh_bvh8_check(const float *tri_v, uint32_t n, lam_scene_stats *stats, int *conservative) {
  if ((!tri_v && n) || !stats || !conservative) return LAM_E_INVALID;
  lamh::Bvh8 b;
  if (!lamh::build_bvh8(tri_v, n, b)) return LAM_E_LIMIT;
  *stats = lam_scene_stats{};
  stats->n_nodes = uint32_t(b.nodes.size());
  stats->n_leaves = b.n_leaves;
  stats->n_tris = n;
  stats->depth = b.depth;
  stats->stack_bound = b.stack_bound;
  stats->sah_cost = b.sah_cost;
  *conservative = 1;
  // exact boxes bottom-up (children have larger indices than parents)
  std::vector<float> lo(b.nodes.size() * 3, FLT_MAX), hi(b.nodes.size() * 3, -FLT_MAX);
  std::vector<uint8_t> seen(n, 0);
  for (size_t i = b.nodes.size(); i-- > 0;) {
    const lamh::Node8 &nd = b.nodes[i];
    for (int k = 0; k < nd.nchild; k++) {
      float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
      const uint32_t c = nd.child[k];
      if (c & lamh::kLeafBit) {
        const uint32_t first = c & lamh::kLeafFirstMask, cnt = ((c >> lamh::kLeafCountShift) & 7u) + 1;
        for (uint32_t t = first; t < first + cnt; t++) {
          if (t >= n) return LAM_E_BVH;
          seen[t]++;
          const float *v = tri_v + size_t(b.order[t]) * 9;
          for (int q = 0; q < 3; q++)
            for (int a = 0; a < 3; a++) {
              clo[a] = std::min(clo[a], v[3 * q + a]);
              chi[a] = std::max(chi[a], v[3 * q + a]);
            }
        }
      } else {
        for (int a = 0; a < 3; a++) { clo[a] = lo[3 * c + a]; chi[a] = hi[3 * c + a]; }
      }
      float dlo[3], dhi[3];
      lamh::decode_child_box(nd, k, dlo, dhi);
      for (int a = 0; a < 3; a++) {
        if (dlo[a] > clo[a] || dhi[a] < chi[a]) *conservative = 0;
        lo[3 * i + a] = std::min(lo[3 * i + a], clo[a]);
        hi[3 * i + a] = std::max(hi[3 * i + a], chi[a]);
      }
    }
  }
  for (uint32_t t = 0; t < n; t++)
    if (seen[t] != 1) *conservative = 0;  // every triangle in exactly one leaf
  return LAM_OK;
}

}  // extern "C"

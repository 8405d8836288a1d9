// Renderer over the C ABI (see renderer.hpp).
#include "renderer.hpp"

#include <stdexcept>

#include "lambda_host.h"

namespace lamh {

Renderer::~Renderer() {
  if (ctx_) lam_destroy(ctx_);
}

void Renderer::upload(const Scene &scene) {
  if (!ctx_) {
    const int rc = lam_create(&ctx_, DEVICE);
    if (rc != LAM_OK) throw std::runtime_error("lam_create failed (" + std::to_string(rc) + ")");
  }
  if (uploaded_ == &scene) return;
  flatten(scene, flat_);
  if (lam_scene_upload(ctx_, &flat_.desc) != LAM_OK) throw std::runtime_error(lam_last_error(ctx_));
  uploaded_ = &scene;
}

lam_render_params Renderer::params(const Scene &scene) const {
  lam_render_params p{};
  p.spp = uint32_t(SPP);
  p.max_scatterings = uint32_t(MAX_SCATTERINGS);
  p.sample_p = SAMPLE_P;
  p.hero = scene.hero;
  p.seed = SEED;
  p.flags = FLAGS;
  return p;
}

// Renderer::render (src/renderer.cpp:14-61)
void Renderer::render(const Scene &scene, rendering &buffer) {
  upload(scene);
  buffer.img.init(scene.cam.width, scene.cam.height);
  const lam_render_params p = params(scene);
  if (lam_render(ctx_, &p, buffer.img.data.data(), nullptr) != LAM_OK)
    throw std::runtime_error(lam_last_error(ctx_));
}

std::vector<float> Renderer::tracePath(const Scene &scene, uint32_t pixel, uint32_t sample) {
  upload(scene);
  const lam_render_params p = params(scene);
  std::vector<float> out(p.hero);
  if (lam_trace_paths(ctx_, &p, &pixel, &sample, 1, out.data()) != LAM_OK)
    throw std::runtime_error(lam_last_error(ctx_));
  return out;
}

// Renderer::toneMap (src/renderer.cpp:149-176)
void Renderer::toneMap(const rendering &in, rendering &tm, float Y_MID) const {
  tm.img.init(in.img.width, in.img.height);
  lamh_tone_map(in.img.data.data(), uint32_t(in.img.width), uint32_t(in.img.height), Y_MID, tm.img.data.data());
}

// Renderer::saveImage (src/renderer.cpp:179-191)
void Renderer::saveImage(const rendering &buffer, std::string fpath, std::string suffix) const {
  char name[4096];
  if (lamh_image_name(fpath.c_str(), uint32_t(SPP), uint32_t(MAX_SCATTERINGS), SAMPLE_P, suffix.c_str(), name,
                      sizeof name) != LAM_OK)
    throw std::runtime_error("cannot derive an image name from '" + fpath + "'");
  if (lamh_write_png(name, buffer.img.data.data(), uint32_t(buffer.img.width), uint32_t(buffer.img.height)) !=
      LAM_OK)
    throw std::runtime_error(std::string("could not write ") + name);
}

}  // namespace lamh

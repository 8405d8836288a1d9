// Renderer: the reference's include/renderer.hpp interface over the C ABI.
//
// `render`, `toneMap` and `saveImage` keep their names, argument meaning
// and defaults (include/renderer.hpp:32-52).  `render` flattens the scene,
// uploads it to the GPU once (re-uploading only when a different scene is
// passed) and calls lam_render; failures throw std::runtime_error with the
// library's message, where the reference would abort or trap.
// `tracePath` takes (pixel, sample) instead of (λ, ray, RNG&): paths are
// keyed by a counter-based RNG, so a path is named by its pixel and sample
// index (lam_trace_paths).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "lambda_amd.h"
#include "scene.hpp"

namespace lamh {

struct rgb {
  float r, g, b;
};

struct image {
  size_t width = 0, height = 0;
  std::vector<float> data;  // width*height*3, top row first
  void init(size_t w, size_t h) {
    width = w;
    height = h;
    data.assign(w * h * 3, 0.f);
  }
};

struct rendering {
  image img;
};

class Renderer {
 public:
  uint64_t MAX_SCATTERINGS = 4;
  uint64_t SPP = 4;
  float SAMPLE_P = 0.5;
  uint64_t SEED = 0;
  int DEVICE = 0;
  uint32_t FLAGS = 0;

  Renderer() = default;
  Renderer(const Renderer &) = delete;
  Renderer &operator=(const Renderer &) = delete;
  ~Renderer();

  void render(const Scene &scene, rendering &buffer);
  void toneMap(const rendering &buffer, rendering &tm_buffer, float Y_MID = 0.12f) const;
  std::vector<float> tracePath(const Scene &scene, uint32_t pixel, uint32_t sample);
  void saveImage(const rendering &buffer, std::string fpath, std::string suffix = "") const;

 private:
  lam_ctx *ctx_ = nullptr;
  const Scene *uploaded_ = nullptr;
  FlatScene flat_;
  void upload(const Scene &scene);
  lam_render_params params(const Scene &scene) const;
};

}  // namespace lamh

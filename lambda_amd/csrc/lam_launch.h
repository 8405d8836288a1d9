// Internal interface between the C-ABI runtime (lam_api.cpp) and the
// kernels (lam_kernels.hip, lam_wavefront.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "lam_device.h"

namespace lam {

struct PathArgs {
  uint64_t seed;
  uint32_t max_scat;
  float sample_p;
};

hipError_t launch_trace_rays(const DevScene &sc, const lam_ray *rays, size_t n, lam_hit *hits, uint32_t *occl,
                             uint32_t *counts, uint32_t *err, hipStream_t st);
hipError_t launch_trace_paths(const DevScene &sc, const PathArgs &pa, int hero, const uint32_t *pixel,
                              const uint32_t *sample, size_t n, float *out, uint32_t *err, hipStream_t st);
hipError_t launch_render_mega(const DevScene &sc, const PathArgs &pa, int hero, uint32_t sample, uint32_t npix,
                              float *film, uint32_t *err, hipStream_t st);
hipError_t launch_resolve(const DevScene &sc, const float *film, uint32_t w, uint32_t h, float inv_n, float *rgb,
                          float *xyz, hipStream_t st);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t st);
hipError_t launch_math(int fn, const float *a, const float *b, size_t n, float *out, hipStream_t st);

}  // namespace lam

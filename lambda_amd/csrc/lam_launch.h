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

// ---- wavefront integrator (lam_wavefront.hip) ------------------------------
// Path state and queues for `cap` path slots; β, L and the pending NEE term
// are stored [k][slot] (H rows of cap floats).
struct WaveState {
  float *beta, *L, *pend;
  void *ray[2];   // extend queues, 32 B per ray (o, tmax | d, slot)
  Hit *hits;      // 16 B per queue entry
  void *sray;     // shadow queue, 32 B per ray (o, tmax | d, skip)
  uint32_t *spath;
  size_t cap;
};
struct WaveLaunch {
  uint32_t trace_blocks;    // persistent extend / shadow / finish grid (64-lane blocks)
  uint32_t shade_blocks;    // shade grid (256-lane blocks)
  uint32_t tail_threshold;  // queue length at or below which k_finish takes over
  uint32_t max_iters;       // 0: pure wavefront; else cap on bounce iterations per pass
};
// bounce iterations enqueued per pass
inline uint32_t wave_iterations(uint32_t max_scat, const WaveLaunch &wl) {
  const uint32_t full = max_scat + 2;
  return wl.max_iters && wl.max_iters < full ? wl.max_iters : full;
}
// HIP events around every finish, extend and shadow launch of one pass
struct WaveProfile {
  hipEvent_t ev[4];
  hipEvent_t *fin0, *ext0, *ext1, *shd0, *shd1;
};
constexpr uint32_t kCtrlWords = 16;
constexpr uint32_t kCtrlOverflow = 15;

hipError_t enqueue_pass(int hero, const DevScene &sc, const PathArgs &pa, const uint32_t *pixel,
                        const uint32_t *sample, uint32_t pixel_offset, const WaveState &ws, uint32_t n,
                        uint32_t *ctrl, float *film, float *out, const WaveLaunch &wl, hipStream_t st,
                        WaveProfile *prof);
hipError_t enqueue_pass_set(uint32_t *ctrl, uint32_t pass, hipStream_t st);

// ---- path-replay adjoint (lam_adjoint.hip) --------------------------------------
constexpr uint32_t kMaxDiffParams = 8192;  // gradient slots held in LDS per block
hipError_t launch_adjoint(int hero, const DevScene &sc, const PathArgs &pa, const int32_t *mat_d,
                          const int32_t *mat_ior, const int32_t *mat_albedo, uint32_t n_params, uint32_t sample,
                          uint32_t npix, float inv_n, const float *dloss, float *grad, uint32_t *ctr, uint32_t *err,
                          uint32_t blocks, hipStream_t st);
hipError_t enqueue_pass_next(uint32_t *ctrl, hipStream_t st);

}  // namespace lam

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

// Closest hit (occl null) or occlusion of n rays with the persistent queue
// traversal on min(blocks, n/64) blocks: *fetch must be zero at launch,
// spill holds (kStack - kLdsStack) * 64 * blocks words.
hipError_t launch_trace_rays(const DevScene &sc, const lam_ray *rays, size_t n, lam_hit *hits, uint32_t *occl,
                             uint32_t *counts, uint32_t *err, uint32_t *fetch, uint32_t *spill, uint32_t blocks,
                             hipStream_t st);
// One thread per path on min(blocks, n/64) blocks (grid-stride); spill
// holds (kStack - kLdsStack) * 64 * blocks words (traversal-stack spill).
hipError_t launch_trace_paths(const DevScene &sc, const PathArgs &pa, int hero, const uint32_t *pixel,
                              const uint32_t *sample, size_t n, float *out, uint32_t *err, uint32_t *spill,
                              uint32_t blocks, hipStream_t st);
hipError_t launch_render_mega(const DevScene &sc, const PathArgs &pa, int hero, uint32_t sample, uint32_t npix,
                              float *film, uint32_t *err, uint32_t *spill, uint32_t blocks, hipStream_t st);
// tone map + 8-bit sRGB of n linear-RGB pixels; partial: 1024 doubles,
// exposure: 1 float of device scratch
hipError_t launch_tonemap8(const float *rgb, uint32_t n, float y_mid, double *partial, float *exposure,
                           uint8_t *out, hipStream_t st);
hipError_t launch_resolve(const DevScene &sc, const float *film, uint32_t w, uint32_t h, float inv_n, float *rgb,
                          float *xyz, hipStream_t st);
// zeroes bytes / 16 float4 words at p (16-byte aligned; a tail below 16 bytes is left)
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
  // traversal-stack spill columns of the extend / shadow grid: entry k >=
  // kLdsStack of lane g at spill[(k - kLdsStack) * spill_lanes + g]
  uint32_t *spill;
  uint32_t spill_lanes;
  // rays traced by the lane's passes (nullable): kRaySlots slots, 128 B
  // apart, each [0] extend, [1] shadow; their sum is the count
  unsigned long long *rays;
};
// counter slots of WaveState::rays: the finish kernel's waves add to slot
// (block % kRaySlots), so that its end-of-kernel atomics spread over many
// cache lines instead of serialising on one address
constexpr uint32_t kRaySlots = 64;
constexpr uint32_t kRaySlotWords = 16;
struct WaveLaunch {
  uint32_t trace_blocks;    // persistent extend / shadow grid (64-lane blocks, split stacks)
  uint32_t shade_blocks;    // shade grid (256-lane blocks)
  uint32_t tail_threshold;  // queue length at or below which k_finish takes over
  uint32_t max_iters;       // 0: pure wavefront; else cap on bounce iterations per pass
  uint32_t guide_blocks;    // guiding-network grid (256-lane blocks, 51 KB LDS each)
  uint32_t finish_blocks;   // persistent finish grid (64-lane blocks, split stacks)
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
  // (nullable) fetch counts of the extend and shadow kernels: extend nodes,
  // triangles, rays, then shadow nodes, triangles, rays
  unsigned long long *counts;
};
constexpr uint32_t kCtrlWords = 16;
constexpr uint32_t kCtrlOverflow = 15;

// ---- path guiding (lam_guide.hip; the guided shade kernel in lam_wavefront.hip) ----------
// Device view of the guiding network for the shade kernel: inputs of a vertex
// are its position in the scene box, its outgoing direction and its hero
// wavelength, each mapped to [0,1].
struct GuideDev {
  const uint4 *frag;      // packed bf16 fragments + fp32 biases
  uint4 *mix;             // network outputs per extend-queue entry: 40 fp16 (5 uint4)
  float alpha;            // probability of the guide branch at supported vertices
  uint32_t mat_mask;      // guided material types (bit 1 << type)
  uint32_t lobes;         // LAM_GUIDE_* lobes the mixture stands in for
  uint32_t max_vertex;    // guided vertices: bounce index < max_vertex
  f3 lo, inv_ext;         // scene box
  float4 *rec;            // training records (4 float4 each); null when not training
  uint32_t *rec_count;    // records appended (may exceed rec_cap; the excess is dropped)
  uint32_t rec_cap;
  uint32_t *rec_of_slot;  // record written by each path slot at the current vertex
};
// Training state: fp32 master parameters and Adam moments, the packed
// network (updated in place), scal[0] = Σw and scal[1] = loss of the last
// step's batch, and a workspace of guide_train_bytes(batch)
struct GuideTrain {
  float *th, *m, *v, *scal, *ws;
  void *pack;
  uint32_t batch, step;
  float lr, b1, b2, eps;
  uint64_t seed;
};
size_t guide_train_bytes(uint32_t batch);
hipError_t guide_pack(const float *th, void *pack, hipStream_t st);
hipError_t guide_infer(const void *pack, const float *in, uint32_t n, float *raw, uint32_t blocks, hipStream_t st);
// loss and gradient of host rows (batch x 10: in[6], ω[3], w) without an
// update: grad (kGParams floats) and raw (optional, batch x 40) on the device
hipError_t guide_loss_grad_rows(GuideTrain &gt, const float *rows, uint32_t batch, float *grad, float *raw,
                                hipStream_t st);
hipError_t guide_train_rows(GuideTrain &gt, const float *rows, uint32_t batch, uint32_t steps, hipStream_t st);
hipError_t guide_train_records(GuideTrain &gt, const float4 *rec, const uint32_t *rec_count, uint32_t rec_cap,
                               uint32_t steps, hipStream_t st);
// training weights of the records of the pass just traced (needs its final L)
hipError_t guide_weights(int hero, const WaveState &ws, const GuideDev &gd, hipStream_t st);

// gd: null for the plain render loop; train: write training records.
// With film and out both null the pass ends before its film kernel
// (enqueue_film then splats it, e.g. after waiting for another stream).
// m: samples per pass (without lists): n = m x pixels, slot q npix + i is
// pixel pixel_offset + i at sample ctrl[CTRL_PASS] + q.
hipError_t enqueue_pass(int hero, const DevScene &sc, const PathArgs &pa, const uint32_t *pixel,
                        const uint32_t *sample, uint32_t pixel_offset, const WaveState &ws, uint32_t n,
                        uint32_t *ctrl, float *film, float *out, const WaveLaunch &wl, hipStream_t st,
                        WaveProfile *prof, const GuideDev *gd = nullptr, bool train = false, uint32_t m = 1);
hipError_t enqueue_pass_set(uint32_t *ctrl, uint32_t pass, hipStream_t st);
// the film kernel of a pass of n paths (m samples of n/m pixels from
// pixel_offset), each pixel's samples added in sample order
hipError_t enqueue_film(int hero, const DevScene &sc, const PathArgs &pa, uint32_t pixel_offset,
                        const WaveState &ws, uint32_t n, const uint32_t *ctrl, float *film, hipStream_t st,
                        uint32_t m = 1);

// ---- BVH8 build on the GPU (lam_bvh.hip) -----------------------------------------
struct GpuBvhStats {
  uint32_t n_nodes, depth, n_leaves, stack_bound;
};
// device scratch for gpu_build_bvh8 over n triangles
size_t gpu_bvh_scratch_bytes(uint32_t n);
// tri/shd: the triangle and shading records (3 float4 each) in input order;
// nodes: room for n records; tri_out/shd_out: the records in leaf order.
// Synchronises the stream (the collapse reads each level's node count).
hipError_t gpu_build_bvh8(const float4 *tri, const float4 *shd, uint32_t n, void *scratch, Node8d *nodes,
                          float4 *tri_out, float4 *shd_out, GpuBvhStats *st, hipStream_t stream);
// the build's stable LSD radix sort of (key, value) pairs by the low `bits`
// bits of the keys: scratch of radix_sort_scratch(n) bytes; kin/vin are
// left unchanged, the sorted pairs go to kout/vout
size_t radix_sort_scratch(uint32_t n);
hipError_t radix_sort_pairs(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, uint32_t n,
                            uint32_t bits, void *scratch, hipStream_t st);

// ---- path-replay adjoint (lam_adjoint.hip) --------------------------------------
constexpr uint32_t kMaxDiffParams = 8192;  // gradient slots held in LDS per block
// replay state of one pass lane: f = 5 hero x cap floats (adjoint weights,
// pending sampled factor u/v, pending NEE term u/v), u = 2 cap words (their
// materials)
struct ReplayBufs {
  float *f;
  uint32_t *u;
};
// The replay pass of the n slots of the forward pass just traced on the
// same lane (m samples of n/m pixels from pixel_offset, ws.L = each path's
// total radiance).  Its kernels add their blocks' gradient sums into the
// lane's partial rows `part` (wl.shade_blocks rows of n_params floats).
hipError_t enqueue_replay(int hero, const DevScene &sc, const PathArgs &pa, uint32_t pixel_offset, uint32_t m,
                          const WaveState &ws, const ReplayBufs &rb, const int32_t *mat_d, const int32_t *mat_ior,
                          const int32_t *mat_albedo, uint32_t n_params, float *part, uint32_t n, uint32_t *ctrl,
                          const float *dloss, float inv_n, const WaveLaunch &wl, hipStream_t st);
// grad += the sum of `rows` partial rows
hipError_t enqueue_sum_rows(const float *part, uint32_t rows, uint32_t n_params, float *grad, hipStream_t st);
// one bounce iteration's start in a pass without a tail: counters reset and
// the closest hits of extend queue q (the render's k_iter_begin and k_extend)
hipError_t enqueue_iter_extend(const DevScene &sc, const WaveState &ws, uint32_t *ctrl, uint32_t q,
                               const WaveLaunch &wl, hipStream_t st);
hipError_t enqueue_pass_next(uint32_t *ctrl, uint32_t stride, hipStream_t st);

}  // namespace lam

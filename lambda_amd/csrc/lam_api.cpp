// C-ABI runtime: context, scene upload (BVH8 build + HBM layout), render,
// path and ray test hooks, timing.  See include/lambda_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host/bvh8.hpp"
#include "lam_guide.h"
#include "lam_launch.h"
#include "lambda_amd.h"

static constexpr int kMaxLanes = 4;
// launches of the batch trace kernel that one timed call may queue
static constexpr uint32_t kTraceCounters = 4096;

// ---- device scratch of lam_extend_dev_noalloc ------------------------------
// A 256-byte header (word 0: the ray-queue counter, word 4: the traversal-
// stack overflow flag) followed by the traversal-stack spill columns of a
// grid of up to kNoAllocBlocks 64-lane blocks, (kStack - kLdsStack) words per
// lane: 64 + 104 * 2048 * 64 words = 54,526,208 bytes.  Every GPU the library
// runs on holds one copy until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host it names a stand-in, not device
// memory): it uses the per-device address from hipGetSymbolAddress.
static constexpr uint32_t kNoAllocBlocks = 2048;
static constexpr size_t kNoAllocHeaderWords = 64;
static constexpr size_t kNoAllocWords =
    kNoAllocHeaderWords + size_t(lam::kStack - lam::kLdsStack) * kNoAllocBlocks * 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {
// per-device address of `scratch`, resolved once per device by lam_create
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_rc[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// with `device` current; returns the device address or nullptr (*rc set)
uint32_t *scratch_address(int device, hipError_t *rc) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    *rc = hipErrorInvalidDevice;
    return nullptr;
  }
  if (!g_scratch_done[device]) {
    void *p = nullptr;
    g_scratch_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_rc[device] == hipSuccess ? static_cast<uint32_t *>(p) : nullptr;
    g_scratch_done[device] = true;
  }
  *rc = g_scratch_rc[device];
  return g_scratch_ptr[device];
}
}  // namespace

struct lam_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool has_scene = false;
  lam::DevScene sc{};
  lam_scene_stats stats{};
  std::vector<void *> scene_bufs;
  uint32_t *d_err = nullptr;  // 16 B: [0] traversal-stack overflow flag
  // grow-only scratch for the test hooks and the film
  void *scratch[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t scratch_size[6] = {0, 0, 0, 0, 0, 0};
  lam_timing timing{};
  hipEvent_t ev[2] = {nullptr, nullptr};
  // wavefront path state, queues, control block and stream of each pipeline
  // lane (lane 0 runs on `stream`; the others only trace when passes overlap)
  struct Lane {
    lam::WaveState ws{};
    int hero = 0;
    std::vector<void *> bufs;
    uint32_t *ctrl = nullptr;
    hipStream_t stream = nullptr;  // lanes 1.. (lane 0 uses ctx->stream)
    hipEvent_t ev_film = nullptr;  // recorded after the lane's latest film kernel
    hipEvent_t ev_join = nullptr;
    void *mix = nullptr;  // guiding-network outputs per queue entry
    size_t mix_cap = 0;
    uint32_t *spill = nullptr;  // traversal-stack spill columns (WaveState::spill)
    uint32_t spill_lanes = 0;
    unsigned long long *rays = nullptr;  // ray counters of the lane's passes (WaveState::rays)
    // adjoint replay state (lam::ReplayBufs) for rep_cap slots of rep_hero wavelengths
    void *rep_mem = nullptr;
    size_t rep_cap = 0;
    int rep_hero = 0;
    lam::ReplayBufs rep{};
  } lane[kMaxLanes];
  // block counts are set on the first render (they depend on the CU count);
  // the finish kernel takes over below 131,072 paths or after 12 bounces
  // the film (scratch[3]) of the last render: pixels, hero count and the
  // samples per pixel it holds (0: none to accumulate into)
  size_t film_npix = 0;
  uint32_t film_hero = 0;
  uint64_t film_spp = 0;
  size_t img_npix = 0;       // pixels of the last rendered linear-RGB image (scratch[4])
  void *tone_buf = nullptr;  // tone map: 1024 per-block log sums (double) + the exposure
  unsigned long long *fetch_counts = nullptr;  // LAM_FLAG_PROFILE_FETCH: 6 counters (WaveProfile::counts)
  lam::WaveLaunch wl{0, 0, 131072u, 12u, 0, 0};
  int cus = 0;
  // this device's copy of `scratch` (lam_extend_dev_noalloc), set by lam_create
  uint32_t *noalloc = nullptr;
  hipError_t noalloc_rc = hipSuccess;
  uint32_t trace_per_cu = 16;  // extend / shadow blocks per CU in render passes
  uint32_t batch_per_cu = 24;  // ... for ray batches outside a render
  uint32_t batch_blocks = 0;
  uint32_t finish_per_cu = 10;  // finish-kernel blocks (64 lanes) per CU
  uint32_t shade_per_cu = 3;    // shade-kernel blocks (256 lanes) per CU
  uint32_t *trace_ctr = nullptr;  // fetch counters of the batch trace kernel (one per launch)
  uint32_t lanes = 3;     // passes in flight
  uint32_t tail_div = 16;  // the finish kernel's threshold is at most 1/tail_div of a chunk
  uint32_t samples_per_pass = 32;  // at most this many samples of every pixel traced together in one pass
  std::vector<hipEvent_t> prof_ev;  // profile mode: 4 + 4 per iteration
  float box_lo[3] = {0.f, 0.f, 0.f}, box_hi[3] = {1.f, 1.f, 1.f};  // scene box (guiding inputs)
  // guiding network: fp32 master parameters, Adam moments, gradient,
  // bf16-rounded copy (one allocation), packed fragments, scalars
  bool guide_ready = false;
  float *g_th = nullptr, *g_m = nullptr, *g_v = nullptr, *g_grad = nullptr, *g_scal = nullptr;
  void *g_pack = nullptr;
  uint32_t g_step = 0;
  float *g_ws = nullptr;  // training batch workspace
  size_t g_ws_bytes = 0;
  // guided-render training records
  float4 *g_rec = nullptr;
  uint32_t g_rec_cap = 0;
  uint32_t *g_rec_count = nullptr, *g_rec_of_slot = nullptr;
  size_t g_rec_slot_cap = 0;
};

// Path slots per wavefront pass.  The frame is processed in chunks of this
// many paths; 4 Mi slots with 16 hero wavelengths is ~0.9 GB of path state.
static constexpr size_t kMaxWaveCap = size_t(1) << 22;
// path slots of one wavefront pass of several samples per pixel
// (samples_per_pass): up to ~11 GB of path state per lane at 4 hero
// wavelengths (HBM is 288 GB)
static constexpr size_t kMaxPassSlots = size_t(1) << 26;

namespace {

int fail(lam_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPC(call)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail(ctx, LAM_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

int ensure_scratch(lam_ctx *ctx, int slot, size_t bytes) {
  if (ctx->scratch_size[slot] >= bytes) return LAM_OK;
  if (ctx->scratch[slot]) HIPC(hipFree(ctx->scratch[slot]));
  ctx->scratch[slot] = nullptr;
  ctx->scratch_size[slot] = 0;
  HIPC(hipMalloc(&ctx->scratch[slot], bytes));
  ctx->scratch_size[slot] = bytes;
  return LAM_OK;
}

void free_scene(lam_ctx *ctx) {
  for (void *p : ctx->scene_bufs) (void)hipFree(p);
  ctx->scene_bufs.clear();
  ctx->has_scene = false;
  ctx->sc = lam::DevScene{};
}

template <typename T>
int upload(lam_ctx *ctx, const T *host, size_t count, const T **dev) {
  *dev = nullptr;
  if (count == 0) return LAM_OK;
  void *p = nullptr;
  HIPC(hipMalloc(&p, count * sizeof(T)));
  ctx->scene_bufs.push_back(p);
  HIPC(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  *dev = static_cast<const T *>(p);
  ctx->stats.device_bytes += count * sizeof(T);
  return LAM_OK;
}

int check_err_flag(lam_ctx *ctx) {
  uint32_t flag = 0;
  HIPC(hipMemcpy(&flag, ctx->d_err, 4, hipMemcpyDeviceToHost));
  if (flag) return fail(ctx, LAM_E_BVH, "BVH traversal stack overflow");
  return LAM_OK;
}

int clear_err_flag(lam_ctx *ctx) {
  HIPC(lam::launch_zero(ctx->d_err, 16, ctx->stream));
  return LAM_OK;
}

double event_ms(lam_ctx *ctx) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) != hipSuccess) return 0.0;
  return double(ms);
}

lam::PathArgs path_args(const lam_render_params *p) {
  lam::PathArgs pa;
  pa.seed = p->seed;
  pa.max_scat = p->max_scatterings;
  pa.sample_p = p->sample_p;
  return pa;
}

// samples: the call renders p->spp samples per pixel (lam_trace_paths
// traces the (pixel, sample) pairs it is given instead)
int check_params(lam_ctx *ctx, const lam_render_params *p, bool samples = true) {
  if (!p) return fail(ctx, LAM_E_INVALID, "params is NULL");
  if (p->hero != 4 && p->hero != 16) return fail(ctx, LAM_E_INVALID, "hero must be 4 or 16");
  // Renderer::render divides the film by SPP·HERO_SAMPLES (renderer.cpp:51-52):
  // no samples would be 0/0 in every pixel
  if (samples && p->spp == 0) return fail(ctx, LAM_E_INVALID, "spp must be at least 1");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (ctx->sc.n_lights == 0) return fail(ctx, LAM_E_INVALID, "scene has no lights");
  return LAM_OK;
}

// BVH8 and leaf-ordered triangle/shading records built on the device from
// the records in input order; fails with LAM_E_BVH if the tree needs a
// deeper traversal stack than the kernels hold.
int build_bvh_on_device(lam_ctx *ctx, const std::vector<float> &tri, const std::vector<float> &shd, uint32_t nt,
                        lam::GpuBvhStats *st) {
  const size_t rec = size_t(nt) * 48;
  void *tin = nullptr, *sin = nullptr, *scratch = nullptr, *nodes = nullptr;
  auto cleanup = [&]() {
    for (void *p : {tin, sin, scratch, nodes})
      if (p) (void)hipFree(p);
  };
  auto dev = [&](hipError_t e, const char *what) {
    if (e == hipSuccess) return LAM_OK;
    cleanup();
    return fail(ctx, LAM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  };
  int rc;
  if ((rc = dev(hipMalloc(&tin, rec), "hipMalloc"))) return rc;
  if ((rc = dev(hipMalloc(&sin, rec), "hipMalloc"))) return rc;
  if ((rc = dev(hipMalloc(&scratch, lam::gpu_bvh_scratch_bytes(nt)), "hipMalloc"))) return rc;
  if ((rc = dev(hipMalloc(&nodes, size_t(nt) * sizeof(lam::Node8d)), "hipMalloc"))) return rc;
  void *tout = nullptr, *sout = nullptr;
  if ((rc = dev(hipMalloc(&tout, rec), "hipMalloc"))) return rc;
  ctx->scene_bufs.push_back(tout);
  if ((rc = dev(hipMalloc(&sout, rec), "hipMalloc"))) return rc;
  ctx->scene_bufs.push_back(sout);
  if ((rc = dev(hipMemcpy(tin, tri.data(), rec, hipMemcpyHostToDevice), "hipMemcpy"))) return rc;
  if ((rc = dev(hipMemcpy(sin, shd.data(), rec, hipMemcpyHostToDevice), "hipMemcpy"))) return rc;
  if ((rc = dev(lam::gpu_build_bvh8(static_cast<const float4 *>(tin), static_cast<const float4 *>(sin), nt, scratch,
                                    static_cast<lam::Node8d *>(nodes), static_cast<float4 *>(tout),
                                    static_cast<float4 *>(sout), st, ctx->stream),
                "BVH build")))
    return rc;
  if (st->stack_bound > uint32_t(lam::kStack)) {
    cleanup();
    return fail(ctx, LAM_E_BVH, "BVH needs a traversal stack of " + std::to_string(st->stack_bound) +
                                    " entries (limit " + std::to_string(lam::kStack) + ")");
  }
  // the nodes into a buffer of their own size
  void *exact = nullptr;
  const size_t nbytes = size_t(st->n_nodes) * sizeof(lam::Node8d);
  if ((rc = dev(hipMalloc(&exact, nbytes), "hipMalloc"))) return rc;
  ctx->scene_bufs.push_back(exact);
  if ((rc = dev(hipMemcpy(exact, nodes, nbytes, hipMemcpyDeviceToDevice), "hipMemcpy"))) return rc;
  cleanup();
  ctx->sc.nodes = static_cast<const lam::Node8d *>(exact);
  ctx->sc.tris = static_cast<const float4 *>(tout);
  ctx->sc.shade = static_cast<const float4 *>(sout);
  ctx->stats.device_bytes += 2 * rec + nbytes;
  return LAM_OK;
}

// Film for a render of the current camera: zeroed, or with
// LAM_FLAG_ACCUMULATE the previous render's film (same pixels and hero
// count).  The film counts as empty until finish_film records the render.
int begin_film(lam_ctx *ctx, const lam_render_params *p, size_t npix, float **film, uint64_t *base_spp) {
  if (p->flags & LAM_FLAG_ACCUMULATE) {
    if (ctx->film_spp == 0 || ctx->film_npix != npix || ctx->film_hero != p->hero)
      return fail(ctx, LAM_E_INVALID, "no film of this camera and hero count to accumulate into");
    *base_spp = ctx->film_spp;
  } else {
    int rc;
    if ((rc = ensure_scratch(ctx, 3, npix * 32))) return rc;  // film: sum + compensation
    HIPC(lam::launch_zero(ctx->scratch[3], npix * 32, ctx->stream));
    *base_spp = 0;
  }
  ctx->film_spp = 0;
  *film = static_cast<float *>(ctx->scratch[3]);
  return LAM_OK;
}

void finish_film(lam_ctx *ctx, const lam_render_params *p, size_t npix, uint64_t base_spp) {
  ctx->img_npix = npix;
  ctx->film_npix = npix;
  ctx->film_hero = p->hero;
  ctx->film_spp = base_spp + p->spp;
}

constexpr size_t kRayBytes = size_t(lam::kRaySlots) * lam::kRaySlotWords * 8;

// the rays counted by the lanes' passes: copied on the library stream, added
// to the timing after it synchronises
int read_rays_async(lam_ctx *ctx, std::vector<unsigned long long> &rays) {
  rays.assign(size_t(kMaxLanes) * lam::kRaySlots * lam::kRaySlotWords, 0ull);
  for (int l = 0; l < kMaxLanes; l++)
    if (ctx->lane[l].rays)
      HIPC(hipMemcpyAsync(rays.data() + size_t(l) * lam::kRaySlots * lam::kRaySlotWords, ctx->lane[l].rays,
                          kRayBytes, hipMemcpyDeviceToHost, ctx->stream));
  return LAM_OK;
}
void add_rays(lam_ctx *ctx, const std::vector<unsigned long long> &rays) {
  for (size_t i = 0; i + 1 < rays.size(); i += lam::kRaySlotWords) {
    ctx->timing.extend_rays += rays[i];
    ctx->timing.shadow_rays += rays[i + 1];
  }
}

void free_wave(lam_ctx *ctx, int l) {
  lam_ctx::Lane &ln = ctx->lane[l];
  for (void *p : ln.bufs) (void)hipFree(p);
  ln.bufs.clear();
  ln.ws = lam::WaveState{};
  ln.hero = 0;
}

// Launch shape (from the CU count), lane 0's control block and the fetch
// counters of the batch trace kernel, on first use.
int ensure_launch(lam_ctx *ctx) {
  if (!ctx->lane[0].ctrl) {
    HIPC(hipMalloc(reinterpret_cast<void **>(&ctx->lane[0].ctrl), lam::kCtrlWords * 4));
    HIPC(hipMalloc(reinterpret_cast<void **>(&ctx->trace_ctr), kTraceCounters * 4));
    int cus = 0;
    HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ctx->cus = cus > 0 ? cus : 256;
    const uint32_t c = uint32_t(ctx->cus);
    // extend/shadow: 6 KB of LDS stack per 64-lane block (the rest of the
    // stack spills to HBM); trace_per_cu blocks per CU in render passes (16:
    // the passes of the other lanes run beside them; a full 24 made the 1080p
    // render 6% slower), batch_per_cu for a lone ray batch (24, the most the
    // LDS stacks leave room for beside other work);
    // finish: 6 KB of LDS stack per 64-lane block, 10 per CU;
    // shade: 256-lane blocks, grid-stride, one occupancy round (its launch
    // bounds give 3 waves per SIMD = 3 blocks per CU)
    ctx->wl.trace_blocks = c * ctx->trace_per_cu;
    ctx->batch_blocks = c * ctx->batch_per_cu;
    ctx->wl.finish_blocks = c * ctx->finish_per_cu;
    ctx->wl.shade_blocks = c * ctx->shade_per_cu;
    // guiding network: 51 KB of LDS per 256-lane block, 3 blocks per CU
    ctx->wl.guide_blocks = c * 3u;
  }
  return LAM_OK;
}

// Traversal-stack spill columns of lane l for its largest grid: the trace
// grid (extend, shadow), the batch grid (lane 0: lone ray batches) or the
// finish grid (finish, and on lane 0 the adjoint, megakernel and path-trace
// hooks, capped at it).
int ensure_spill(lam_ctx *ctx, int l) {
  lam_ctx::Lane &ln = ctx->lane[l];
  const uint32_t glanes =
      std::max(std::max(ctx->wl.trace_blocks, ctx->wl.finish_blocks), l == 0 ? ctx->batch_blocks : 0u) * 64u;
  if (ln.spill_lanes < glanes) {
    if (ln.spill) HIPC(hipFree(ln.spill));
    ln.spill = nullptr;
    ln.spill_lanes = 0;
    HIPC(hipMalloc(reinterpret_cast<void **>(&ln.spill), size_t(lam::kStack - lam::kLdsStack) * glanes * 4));
    ln.spill_lanes = glanes;
  }
  ln.ws.spill = ln.spill;
  ln.ws.spill_lanes = ln.spill_lanes;
  return LAM_OK;
}

// Launch shape, control block, spill columns and path state of lane l for
// chunks of up to cap paths.  A lane's stream and events are created on its
// first use.
int ensure_wave(lam_ctx *ctx, size_t cap, int hero, int l = 0) {
  int rc;
  if ((rc = ensure_launch(ctx))) return rc;
  lam_ctx::Lane &ln = ctx->lane[l];
  if (!ln.ev_film) {
    if (l > 0) {
      HIPC(hipMalloc(reinterpret_cast<void **>(&ln.ctrl), lam::kCtrlWords * 4));
      HIPC(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
    }
    HIPC(hipEventCreateWithFlags(&ln.ev_join, hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&ln.ev_film, hipEventDisableTiming));
  }
  if (!ln.rays) {
    HIPC(hipMalloc(reinterpret_cast<void **>(&ln.rays), kRayBytes));
    HIPC(lam::launch_zero(ln.rays, kRayBytes, ctx->stream));
  }
  if ((rc = ensure_spill(ctx, l))) return rc;
  if (ln.ws.cap >= cap && ln.hero >= hero) {
    ln.ws.rays = ln.rays;
    return LAM_OK;
  }
  free_wave(ctx, l);
  cap = std::max(cap, size_t(64));
  auto alloc = [&](size_t bytes, void **p) -> int {
    HIPC(hipMalloc(p, bytes));
    ln.bufs.push_back(*p);
    return LAM_OK;
  };
  void *p;
  const size_t hf = size_t(hero) * cap * 4;
  if ((rc = alloc(hf, &p))) return rc;
  ln.ws.beta = static_cast<float *>(p);
  if ((rc = alloc(hf, &p))) return rc;
  ln.ws.L = static_cast<float *>(p);
  if ((rc = alloc(hf, &p))) return rc;
  ln.ws.pend = static_cast<float *>(p);
  if ((rc = alloc(cap * 32, &ln.ws.ray[0]))) return rc;
  if ((rc = alloc(cap * 32, &ln.ws.ray[1]))) return rc;
  if ((rc = alloc(cap * 16, &p))) return rc;
  ln.ws.hits = static_cast<lam::Hit *>(p);
  if ((rc = alloc(cap * 32, &ln.ws.sray))) return rc;
  if ((rc = alloc(cap * 4, &p))) return rc;
  ln.ws.spath = static_cast<uint32_t *>(p);
  ln.ws.cap = cap;
  ln.ws.spill = ln.spill;
  ln.ws.spill_lanes = ln.spill_lanes;
  ln.ws.rays = ln.rays;
  ln.hero = hero;
  return LAM_OK;
}

// guiding-network output buffer of lane l (80 B per path slot)
// the adjoint's replay state for the lane's path slots (its rows are
// ws.cap apart, so it follows the wave state's capacity)
int ensure_replay(lam_ctx *ctx, int l, int hero) {
  lam_ctx::Lane &ln = ctx->lane[l];
  const size_t cap = ln.ws.cap;
  if (ln.rep_mem && ln.rep_cap >= cap && ln.rep_hero >= hero) return LAM_OK;
  if (ln.rep_mem) HIPC(hipFree(ln.rep_mem));
  ln.rep_mem = nullptr;
  ln.rep_cap = 0;
  const size_t fb = (5 * size_t(hero) * cap * 4 + 255) / 256 * 256;
  HIPC(hipMalloc(&ln.rep_mem, fb + 2 * cap * 4));
  char *b = static_cast<char *>(ln.rep_mem);
  ln.rep.f = reinterpret_cast<float *>(b);
  ln.rep.u = reinterpret_cast<uint32_t *>(b + fb);
  ln.rep_cap = cap;
  ln.rep_hero = hero;
  return LAM_OK;
}

// Shape of the wavefront passes of a frame of npix pixels: a frame of at
// most kMaxPassSlots pixels is one chunk, each pass tracing m samples of
// every pixel (longer queues, one tail per m samples); a larger frame goes
// in chunks of kMaxWaveCap pixels, one sample per pass.  m is at most
// mmax = min(samples_per_pass, kMaxPassSlots / npix), and the spp samples
// go in g >= 2 passes, g rounded up to a multiple of the lanes (g =
// ceil(spp / mmax) before rounding), m = ceil(spp / g): every lane traces
// its passes beside the others', whose tails and film kernels then overlap
// (1080p, 64 spp: three passes of 22, 21-22 samples, 2-3% faster than two
// of 32 on two lanes, 10% faster than four of 16 on three lanes).
size_t pass_pixels(size_t npix) { return npix <= kMaxPassSlots ? npix : kMaxWaveCap; }
static uint32_t pass_samples_max(const lam_ctx *ctx, size_t npix) {
  return uint32_t(std::max<size_t>(1, std::min<size_t>(size_t(ctx->samples_per_pass), kMaxPassSlots / npix)));
}
uint32_t pass_samples(const lam_ctx *ctx, size_t npix, uint32_t spp, uint32_t lanes) {
  if (npix > kMaxPassSlots || spp <= 1) return 1u;
  const uint32_t mmax = pass_samples_max(ctx, npix);
  const uint32_t l = std::max(1u, std::min(lanes, spp));
  uint32_t g = std::max(2u, (spp + mmax - 1) / mmax);
  g = (g + l - 1) / l * l;
  return (spp + g - 1) / g;
}
// an upper bound of pass_samples(ctx, npix, s, ·) for every s <= spp (g >= 2
// and g >= ceil(s / mmax)), for sizing the path state
uint32_t pass_samples_bound(const lam_ctx *ctx, size_t npix, uint32_t spp) {
  if (npix > kMaxPassSlots) return 1u;
  return std::max(1u, std::min(pass_samples_max(ctx, npix), (spp + 1) / 2));
}

// samples per training pass of a guided render: at least four training
// passes (the network is updated after each), at most 8 samples each
uint32_t train_pass_samples(const lam_ctx *ctx, size_t npix, uint32_t spp, uint32_t train_passes) {
  const uint32_t tp = std::min(train_passes, spp);
  return std::max(1u, std::min({8u, tp / 4u, pass_samples_bound(ctx, npix, spp)}));
}

int ensure_mix(lam_ctx *ctx, int l) {
  lam_ctx::Lane &ln = ctx->lane[l];
  if (ln.mix_cap >= ln.ws.cap) return LAM_OK;
  if (ln.mix) HIPC(hipFree(ln.mix));
  ln.mix = nullptr;
  ln.mix_cap = 0;
  HIPC(hipMalloc(&ln.mix, ln.ws.cap * 80));
  ln.mix_cap = ln.ws.cap;
  return LAM_OK;
}

lam::WaveLaunch launch_for(const lam_ctx *ctx, const lam_render_params *p) {
  lam::WaveLaunch wl = ctx->wl;
  if (p->flags & LAM_FLAG_NO_TAIL) {
    wl.tail_threshold = 0;
    wl.max_iters = 0;
  }
  return wl;
}

int check_ctrl_overflow(lam_ctx *ctx) {
  for (const lam_ctx::Lane &ln : ctx->lane) {
    if (!ln.ctrl) continue;
    uint32_t flag = 0;
    HIPC(hipMemcpy(&flag, ln.ctrl + lam::kCtrlOverflow, 4, hipMemcpyDeviceToHost));
    if (flag) return fail(ctx, LAM_E_BVH, "BVH traversal stack overflow");
  }
  return LAM_OK;
}

void free_guide(lam_ctx *ctx) {
  void *bufs[] = {ctx->g_th, ctx->g_pack, ctx->g_ws, ctx->g_rec, ctx->g_rec_count, ctx->g_rec_of_slot};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  for (lam_ctx::Lane &ln : ctx->lane)
    if (ln.mix) (void)hipFree(ln.mix);
  ctx->g_th = ctx->g_m = ctx->g_v = ctx->g_grad = ctx->g_scal = nullptr;
  ctx->g_pack = nullptr;
  ctx->g_ws = nullptr;
  ctx->g_ws_bytes = 0;
  for (lam_ctx::Lane &ln : ctx->lane) {
    ln.mix = nullptr;
    ln.mix_cap = 0;
  }
  ctx->g_rec = nullptr;
  ctx->g_rec_cap = 0;
  ctx->g_rec_count = ctx->g_rec_of_slot = nullptr;
  ctx->g_rec_slot_cap = 0;
  ctx->guide_ready = false;
}

int ensure_guide(lam_ctx *ctx) {
  if (ctx->g_th) return LAM_OK;
  const size_t np = (size_t(lam::kGParams) + 63) / 64 * 64;
  void *p = nullptr;
  HIPC(hipMalloc(&p, np * 4 * 4 + 64));
  float *f = static_cast<float *>(p);
  ctx->g_th = f;
  ctx->g_m = f + np;
  ctx->g_v = f + 2 * np;
  ctx->g_grad = f + 3 * np;
  ctx->g_scal = f + 4 * np;
  HIPC(hipMalloc(&ctx->g_pack, lam::kGPackBytes));
  return LAM_OK;
}

int ensure_guide_ws(lam_ctx *ctx, uint32_t batch) {
  const size_t need = lam::guide_train_bytes(batch);
  if (ctx->g_ws_bytes >= need) return LAM_OK;
  if (ctx->g_ws) HIPC(hipFree(ctx->g_ws));
  ctx->g_ws = nullptr;
  ctx->g_ws_bytes = 0;
  HIPC(hipMalloc(reinterpret_cast<void **>(&ctx->g_ws), need));
  ctx->g_ws_bytes = need;
  return LAM_OK;
}

// He-uniform hidden layers, a 10x smaller output layer, zero biases (seeded)
void guide_init_params(uint64_t seed, std::vector<float> &th) {
  th.assign(lam::kGParams, 0.f);
  uint64_t st = seed ^ 0x6a09e667f3bcc909ull;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * 0x1.0p-53;
  };
  const int W[5] = {lam::kGW1, lam::kGW2, lam::kGW3, lam::kGW4, lam::kGW5};
  const int fan_in[5] = {lam::kGFeat, lam::kGHid, lam::kGHid, lam::kGHid, lam::kGHid};
  const int count[5] = {lam::kGHid * lam::kGFeat, lam::kGHid * lam::kGHid, lam::kGHid * lam::kGHid,
                        lam::kGHid * lam::kGHid, lam::kGOut * lam::kGHid};
  for (int l = 0; l < 5; l++) {
    const double a = std::sqrt(6.0 / fan_in[l]) * (l == 4 ? 0.1 : 1.0);
    for (int i = 0; i < count[l]; i++) th[size_t(W[l]) + i] = float((2.0 * next() - 1.0) * a);
  }
}

int guide_upload_params(lam_ctx *ctx, const float *th) {
  int rc;
  if ((rc = ensure_guide(ctx))) return rc;
  const size_t np = (size_t(lam::kGParams) + 63) / 64 * 64;
  HIPC(hipMemcpyAsync(ctx->g_th, th, size_t(lam::kGParams) * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(lam::launch_zero(ctx->g_m, np * 8, ctx->stream));  // m and v
  HIPC(lam::guide_pack(ctx->g_th, ctx->g_pack, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->g_step = 0;
  ctx->guide_ready = true;
  return LAM_OK;
}

lam::GuideTrain guide_train_state(lam_ctx *ctx, uint32_t batch, float lr, uint64_t seed) {
  lam::GuideTrain gt{};
  gt.th = ctx->g_th;
  gt.m = ctx->g_m;
  gt.v = ctx->g_v;
  gt.scal = ctx->g_scal;
  gt.ws = ctx->g_ws;
  gt.pack = ctx->g_pack;
  gt.batch = batch;
  gt.step = ctx->g_step;
  gt.lr = lr;
  gt.b1 = 0.9f;
  gt.b2 = 0.99f;
  gt.eps = 1e-8f;
  gt.seed = seed;
  return gt;
}

// Guided-render settings for render_wavefront
struct GuideRun {
  const lam_guide_params *gp;
  lam::GuideDev gd_render, gd_train;
};

// Wavefront render of every pixel for passes spp_offset .. spp_offset+spp-1.
//
// Consecutive passes go round-robin to `lanes` lanes, each with its own
// path state, control block and stream, so that the tail of one pass (a few
// paths with many bounces, finished by k_finish on a handful of CUs) shares
// the GPU with the full wavefronts of the next.  Each lane's pass is
// captured once per chunk as a hipGraph (its first node advances the lane's
// pass index by the number of lanes) and replayed.  The film kernels stay in
// pass order: each runs after its own pass and after the previous pass's
// film on the previous lane, so the Kahan film receives the same sums in the
// same order as a one-lane render and the image is bit-identical to one.
//
// LAM_FLAG_NO_PIPE keeps every pass on lane 0.  LAM_FLAG_PROFILE (one lane,
// no graph) times every extend and shadow launch with events;
// LAM_FLAG_NO_GRAPH launches without a graph.
int render_wavefront(lam_ctx *ctx, const lam_render_params *p, float *film, size_t npix,
                     const GuideRun *gr = nullptr) {
  // pixels per chunk and samples per pass (pass_samples); a guided render's
  // training passes keep one sample each
  const size_t cpix = pass_pixels(npix);
  const size_t cap = cpix * pass_samples_bound(ctx, npix, p->spp);
  const bool profile = p->flags & LAM_FLAG_PROFILE;
  const bool graph = !(p->flags & (LAM_FLAG_NO_GRAPH | LAM_FLAG_PROFILE));
  // every configured lane gets its state on the first render, also when
  // this one has fewer passes than lanes, so that later renders allocate
  // nothing
  const uint32_t lanes_cfg = (p->flags & (LAM_FLAG_NO_PIPE | LAM_FLAG_PROFILE)) ? 1u : ctx->lanes;
  const uint32_t lanes = std::min(lanes_cfg, p->spp);
  int rc;
  for (uint32_t l = 0; l < lanes_cfg; l++) {
    if ((rc = ensure_wave(ctx, cap, int(p->hero), int(l)))) return rc;
    if (gr && (rc = ensure_mix(ctx, int(l)))) return rc;
  }
  hipStream_t st[kMaxLanes];
  lam::WaveState ws[kMaxLanes];
  uint32_t *ctrl[kMaxLanes];
  lam::GuideDev gdl[kMaxLanes];
  for (int l = 0; l < kMaxLanes; l++) {
    lam_ctx::Lane &ln = ctx->lane[l];
    st[l] = l == 0 ? ctx->stream : ln.stream;
    ws[l] = ln.ws;
    ctrl[l] = ln.ctrl;
    // every control block (also of lanes this render leaves idle) starts at
    // zero, so that the counters read after the render are this render's
    if (ctrl[l]) HIPC(lam::launch_zero(ctrl[l], lam::kCtrlWords * 4, st[0]));
    if (ln.rays) HIPC(lam::launch_zero(ln.rays, kRayBytes, st[0]));
    if (gr) {
      gdl[l] = gr->gd_render;
      gdl[l].mix = static_cast<uint4 *>(ln.mix);
    }
  }
  const lam::PathArgs pa = path_args(p);
  const lam::WaveLaunch wl_all = launch_for(ctx, p);
  const uint32_t iters = lam::wave_iterations(p->max_scatterings, wl_all);
  lam::WaveProfile prof{};
  if (profile) {
    const size_t need = 4 + 5 * size_t(iters);
    while (ctx->prof_ev.size() < need) {
      hipEvent_t e;
      HIPC(hipEventCreate(&e));
      ctx->prof_ev.push_back(e);
    }
    for (int i = 0; i < 4; i++) prof.ev[i] = ctx->prof_ev[i];
    prof.fin0 = ctx->prof_ev.data() + 4;
    prof.ext0 = prof.fin0 + iters;
    prof.ext1 = prof.ext0 + iters;
    prof.shd0 = prof.ext1 + iters;
    prof.shd1 = prof.shd0 + iters;
    if (p->flags & LAM_FLAG_PROFILE_FETCH) {
      if (!ctx->fetch_counts) HIPC(hipMalloc(reinterpret_cast<void **>(&ctx->fetch_counts), 64));
      HIPC(lam::launch_zero(ctx->fetch_counts, 64, st[0]));
      prof.counts = ctx->fetch_counts;
    }
  }
  for (size_t off = 0; off < npix; off += cpix) {
    const uint32_t n = uint32_t(std::min(cpix, npix - off));
    // the finish kernel takes over below 1/tail_div of the chunk's paths
    // (and at most tail_threshold): guided passes then shade most vertices
    // with the learned proposal, and small frames keep a full wavefront
    // longer
    lam::WaveLaunch wl = wl_all;
    wl.tail_threshold = std::min(wl.tail_threshold, n / ctx->tail_div);
    uint32_t first = 0;  // passes of this chunk already traced
    if (gr && off == 0) {
      // training passes (lane 0): trace with records, weight them, then Adam
      const lam_guide_params *gp = gr->gp;
      const uint32_t tp = std::min(gp->train_passes, p->spp);
      lam::GuideTrain gt = guide_train_state(ctx, gp->batch, gp->lr, gp->seed);
      lam::GuideDev gtr = gr->gd_train;
      gtr.mix = gdl[0].mix;
      ctx->timing.guide_records = 0;
      const auto t0 = std::chrono::steady_clock::now();
      // passes of mt samples of every pixel, each followed by mt x
      // steps_per_pass Adam steps on its records
      const uint32_t mt = train_pass_samples(ctx, npix, p->spp, gp->train_passes);
      for (; first < tp; first += std::min(mt, tp - first)) {
        const uint32_t mg = std::min(mt, tp - first);
        HIPC(lam::launch_zero(gtr.rec_count, 16, st[0]));
        HIPC(lam::enqueue_pass_set(ctrl[0], p->spp_offset + first, st[0]));
        HIPC(lam::enqueue_pass(int(p->hero), ctx->sc, pa, nullptr, nullptr, uint32_t(off), ws[0], n * mg, ctrl[0],
                               film, nullptr, wl, st[0], nullptr, &gtr, true, mg));
        HIPC(lam::guide_weights(int(p->hero), ws[0], gtr, st[0]));
        HIPC(lam::guide_train_records(gt, gtr.rec, gtr.rec_count, gtr.rec_cap, gp->steps_per_pass * mg, st[0]));
        uint32_t cnt = 0;
        float sc[2] = {0.f, 0.f};
        HIPC(hipMemcpyAsync(&cnt, gtr.rec_count, 4, hipMemcpyDeviceToHost, st[0]));
        HIPC(hipMemcpyAsync(sc, ctx->g_scal, 8, hipMemcpyDeviceToHost, st[0]));
        HIPC(hipStreamSynchronize(st[0]));
        ctx->timing.guide_records += std::min(cnt, gtr.rec_cap);
        if (first == 0) ctx->timing.guide_loss_first = sc[1];
        ctx->timing.guide_loss_last = sc[1];
      }
      ctx->g_step = gt.step;
      HIPC(hipStreamSynchronize(st[0]));
      ctx->timing.guide_train_ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (first >= p->spp) continue;
    // passes of m samples of every pixel of the chunk (the last may hold fewer)
    const uint32_t left = p->spp - first;
    const uint32_t m = pass_samples(ctx, npix, left, lanes_cfg);
    const uint32_t ng = (left + m - 1) / m, nfull = left / m;
    if (off == 0) {
      ctx->timing.pass_samples = m;
      ctx->timing.passes = ng;
    }
    const uint32_t nl = std::min(lanes, ng);
    const uint32_t npass = n * m;
    wl.tail_threshold = std::min(wl_all.tail_threshold, npass / ctx->tail_div);
    // with several lanes the film kernel is enqueued on its own, after the wait
    float *pass_film = nl > 1 ? nullptr : film;
    if (nl > 1) {  // the other lanes start after everything enqueued on lane 0 so far
      HIPC(hipEventRecord(ctx->lane[0].ev_join, st[0]));
      for (uint32_t l = 1; l < nl; l++) HIPC(hipStreamWaitEvent(st[l], ctx->lane[0].ev_join, 0));
    }
    hipGraphExec_t ge[kMaxLanes] = {};
    hipError_t e = hipSuccess;
    if (graph) {
      for (uint32_t l = 0; l < nl && e == hipSuccess; l++) {
        hipGraph_t g = nullptr;
        if ((e = hipStreamBeginCapture(st[l], hipStreamCaptureModeThreadLocal)) != hipSuccess) break;
        hipError_t e1 = lam::enqueue_pass_next(ctrl[l], nl * m, st[l]);
        hipError_t e2 = lam::enqueue_pass(int(p->hero), ctx->sc, pa, nullptr, nullptr, uint32_t(off), ws[l], npass,
                                          ctrl[l], pass_film, nullptr, wl, st[l], nullptr, gr ? &gdl[l] : nullptr,
                                          false, m);
        hipError_t e3 = hipStreamEndCapture(st[l], &g);
        e = e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3;
        if (e == hipSuccess) e = hipGraphInstantiate(&ge[l], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        // the graph's first node advances the first sample by nl m: start one
        // step back (modulo 2^32)
        if (e == hipSuccess) e = lam::enqueue_pass_set(ctrl[l], p->spp_offset + first + (l - nl) * m, st[l]);
      }
    }
    for (uint32_t gi = 0; gi < ng && e == hipSuccess; gi++) {
      const uint32_t l = gi % nl;
      const uint32_t mg = gi < nfull ? m : left - nfull * m;  // samples of this pass
      if (graph && mg == m) {
        e = hipGraphLaunch(ge[l], st[l]);
      } else {
        e = lam::enqueue_pass_set(ctrl[l], p->spp_offset + first + gi * m, st[l]);
        if (e == hipSuccess)
          e = lam::enqueue_pass(int(p->hero), ctx->sc, pa, nullptr, nullptr, uint32_t(off), ws[l], n * mg, ctrl[l],
                                pass_film, nullptr, wl, st[l], profile ? &prof : nullptr, gr ? &gdl[l] : nullptr,
                                false, mg);
      }
      if (nl > 1 && e == hipSuccess) {
        if (gi > 0) e = hipStreamWaitEvent(st[l], ctx->lane[(l + nl - 1) % nl].ev_film, 0);
        if (e == hipSuccess) e = lam::enqueue_film(int(p->hero), ctx->sc, pa, uint32_t(off), ws[l], n * mg, ctrl[l],
                                                   film, st[l], mg);
        if (e == hipSuccess) e = hipEventRecord(ctx->lane[l].ev_film, st[l]);
      }
      if (profile && e == hipSuccess) {
        HIPC(hipStreamSynchronize(st[0]));
        float ms = 0.f;
        for (uint32_t it = 0; it < iters; it++) {
          HIPC(hipEventElapsedTime(&ms, prof.ext0[it], prof.ext1[it]));
          ctx->timing.extend_ms += ms;
          HIPC(hipEventElapsedTime(&ms, prof.shd0[it], prof.shd1[it]));
          ctx->timing.shadow_ms += ms;
          HIPC(hipEventElapsedTime(&ms, prof.fin0[it], prof.ext0[it]));
          ctx->timing.other_ms += ms;  // iteration reset + finish kernel
          HIPC(hipEventElapsedTime(&ms, prof.ext1[it], prof.shd0[it]));
          ctx->timing.shade_ms += ms;
        }
        ctx->timing.extend_launches += iters;
        ctx->timing.shadow_launches += iters;
        ctx->timing.shade_launches += iters;
        HIPC(hipEventElapsedTime(&ms, prof.ev[0], prof.ev[1]));
        ctx->timing.raygen_ms += ms;
        HIPC(hipEventElapsedTime(&ms, prof.ev[2], prof.ev[3]));
        ctx->timing.film_ms += ms;
      }
    }
    // join: lane 0 (and whatever the caller enqueues next) waits for the others
    for (uint32_t l = 1; l < nl; l++) {
      hipError_t ej = hipEventRecord(ctx->lane[l].ev_join, st[l]);
      if (ej == hipSuccess) ej = hipStreamWaitEvent(st[0], ctx->lane[l].ev_join, 0);
      if (e == hipSuccess) e = ej;
    }
    hipError_t es = hipSuccess;
    for (uint32_t l = 0; l < nl; l++) {
      hipError_t e1 = hipStreamSynchronize(st[l]);
      if (es == hipSuccess) es = e1;
    }
    for (hipGraphExec_t x : ge)
      if (x) (void)hipGraphExecDestroy(x);
    HIPC(e);
    HIPC(es);
  }
  if (prof.counts) {
    unsigned long long c[6];
    HIPC(hipMemcpy(c, prof.counts, sizeof c, hipMemcpyDeviceToHost));
    ctx->timing.extend_nodes += c[0];
    ctx->timing.extend_tris += c[1];
    ctx->timing.extend_traced += c[2];
    ctx->timing.shadow_nodes += c[3];
    ctx->timing.shadow_tris += c[4];
    ctx->timing.shadow_traced += c[5];
  }
  return LAM_OK;
}

}  // namespace

extern "C" {

const char *lam_version(void) { return "lambda_amd 0.1 (gfx950)"; }

int lam_create(lam_ctx **out, int device) {
  if (!out) return LAM_E_INVALID;
  *out = nullptr;
  lam_ctx *ctx = new lam_ctx();
  ctx->device = device;
  int rc = LAM_OK;
  do {
    if (hipSetDevice(device) != hipSuccess) { rc = LAM_E_DEVICE; break; }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { rc = LAM_E_DEVICE; break; }
    if (hipMalloc(reinterpret_cast<void **>(&ctx->d_err), 16) != hipSuccess) { rc = LAM_E_DEVICE; break; }
    if (hipMemset(ctx->d_err, 0, 16) != hipSuccess) { rc = LAM_E_DEVICE; break; }
    if (hipEventCreate(&ctx->ev[0]) != hipSuccess || hipEventCreate(&ctx->ev[1]) != hipSuccess) {
      rc = LAM_E_DEVICE;
      break;
    }
    // the device address of the no-alloc scratch; a failure is reported by
    // lam_extend_dev_noalloc, the other entry points do not use it
    ctx->noalloc = scratch_address(device, &ctx->noalloc_rc);
  } while (false);
  if (rc != LAM_OK) {
    lam_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return LAM_OK;
}

void lam_destroy(lam_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_scene(ctx);
  for (lam_ctx::Lane &ln : ctx->lane)
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
  for (int l = 0; l < kMaxLanes; l++) free_wave(ctx, l);
  free_guide(ctx);
  if (ctx->trace_ctr) (void)hipFree(ctx->trace_ctr);
  if (ctx->tone_buf) (void)hipFree(ctx->tone_buf);
  if (ctx->fetch_counts) (void)hipFree(ctx->fetch_counts);
  for (lam_ctx::Lane &ln : ctx->lane) {
    if (ln.ctrl) (void)hipFree(ln.ctrl);
    if (ln.rays) (void)hipFree(ln.rays);
    if (ln.spill) (void)hipFree(ln.spill);
    if (ln.rep_mem) (void)hipFree(ln.rep_mem);
    for (hipEvent_t e : {ln.ev_film, ln.ev_join})
      if (e) (void)hipEventDestroy(e);
    if (ln.stream) (void)hipStreamDestroy(ln.stream);
  }
  for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
  for (int i = 0; i < 6; i++)
    if (ctx->scratch[i]) (void)hipFree(ctx->scratch[i]);
  if (ctx->d_err) (void)hipFree(ctx->d_err);
  for (int i = 0; i < 2; i++)
    if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *lam_last_error(const lam_ctx *ctx) { return ctx ? ctx->err.c_str() : "NULL context"; }

int lam_scene_upload(lam_ctx *ctx, const lam_scene_desc *d) { return lam_scene_upload_ex(ctx, d, 0); }

int lam_scene_upload_ex(lam_ctx *ctx, const lam_scene_desc *d, uint32_t flags) {
  if (!ctx) return LAM_E_INVALID;
  if (!d) return fail(ctx, LAM_E_INVALID, "scene description is NULL");
  HIPC(hipSetDevice(ctx->device));
  const uint32_t gn = d->grid.n;
  if (gn < 2 || !(d->grid.sup > d->grid.inf)) return fail(ctx, LAM_E_INVALID, "bad spectral grid");
  if (!d->cmf) return fail(ctx, LAM_E_INVALID, "cmf table missing");
  if (d->n_tris && (!d->tri_v || !d->tri_n || !d->tri_gn || !d->tri_obj || !d->tri_prim))
    return fail(ctx, LAM_E_INVALID, "triangle arrays missing");
  if (d->n_objects == 0 || !d->objects) return fail(ctx, LAM_E_INVALID, "scene has no objects");
  if (d->n_materials == 0 || !d->materials) return fail(ctx, LAM_E_INVALID, "scene has no materials");
  if (d->n_lights == 0 || !d->lights) return fail(ctx, LAM_E_INVALID, "scene has no lights");
  if (d->camera.width == 0 || d->camera.height == 0) return fail(ctx, LAM_E_INVALID, "camera has zero size");
  for (uint32_t i = 0; i < d->n_objects; i++) {
    const lam_object &o = d->objects[i];
    if (o.type < LAM_OBJ_SPHERE || o.type > LAM_OBJ_TRIMESH) return fail(ctx, LAM_E_INVALID, "bad object type");
    if (o.material >= d->n_materials) return fail(ctx, LAM_E_INVALID, "object material out of range");
  }
  auto spec_ok = [&](int32_t s) { return s >= 0 && uint32_t(s) < d->n_spectra; };
  for (uint32_t i = 0; i < d->n_materials; i++) {
    const lam_material &m = d->materials[i];
    int need = 0;
    switch (m.type) {
      case LAM_MAT_LAMBERTIAN: need = 1; break;
      case LAM_MAT_BLINN: need = 5; break;
      case LAM_MAT_THINFILM: need = 2; break;
      case LAM_MAT_EMITTER: case LAM_MAT_DIREMITTER: need = 1; break;
      case LAM_MAT_NONE: case LAM_MAT_MICROFACET: need = 0; break;
      default: return fail(ctx, LAM_E_INVALID, "bad material type");
    }
    for (int k = 0; k < need; k++)
      if (!spec_ok(m.spec[k])) return fail(ctx, LAM_E_INVALID, "material spectrum index out of range");
  }
  for (uint32_t i = 0; i < d->n_lights; i++) {
    const lam_light &l = d->lights[i];
    if (l.type > LAM_LIGHT_QUADDIR) return fail(ctx, LAM_E_INVALID, "bad light type");
    if (!spec_ok(l.spec)) return fail(ctx, LAM_E_INVALID, "light spectrum index out of range");
    if (l.type == LAM_LIGHT_SPHERE || l.type == LAM_LIGHT_QUADDIR) {
      if (l.object >= d->n_objects) return fail(ctx, LAM_E_INVALID, "light object out of range");
      const uint32_t want = l.type == LAM_LIGHT_SPHERE ? LAM_OBJ_SPHERE : LAM_OBJ_PLANE;
      if (d->objects[l.object].type != want) return fail(ctx, LAM_E_INVALID, "light object has the wrong type");
    }
  }
  for (uint32_t t = 0; t < d->n_tris; t++)
    if (d->tri_obj[t] >= d->n_objects || d->objects[d->tri_obj[t]].type != LAM_OBJ_TRIMESH)
      return fail(ctx, LAM_E_INVALID, "triangle object id invalid");

  free_scene(ctx);
  ctx->stats = lam_scene_stats{};

  const uint32_t nt = d->n_tris;
  if (nt > lamh::kLeafFirstMask) return fail(ctx, LAM_E_LIMIT, "too many triangles");
  // triangle records (v0 v1 v2 | obj prim input-index) and shading records
  // (n0 n1 n2 | geometric normal), 48 B each; order[s] = input triangle of slot s
  auto records = [&](const uint32_t *order, std::vector<float> &tri, std::vector<float> &shd) {
    tri.resize(size_t(nt) * 12);
    shd.resize(size_t(nt) * 12);
    for (uint32_t s = 0; s < nt; s++) {
      const uint32_t o = order ? order[s] : s;
      std::memcpy(&tri[size_t(s) * 12], d->tri_v + size_t(o) * 9, 36);
      uint32_t ids[3] = {d->tri_obj[o], d->tri_prim[o], o};
      std::memcpy(&tri[size_t(s) * 12 + 9], ids, 12);
      std::memcpy(&shd[size_t(s) * 12], d->tri_n + size_t(o) * 9, 36);
      std::memcpy(&shd[size_t(s) * 12 + 9], d->tri_gn + size_t(o) * 3, 12);
    }
  };
  std::vector<float> tri, shd;
  lamh::Bvh8 bvh;
  lam::GpuBvhStats gst{0, 0, 0, 0};
  const bool gpu_bvh = (flags & LAM_UPLOAD_GPU_BVH) && nt > 0;
  if (!gpu_bvh) {
    if (!lamh::build_bvh8(d->tri_v, nt, bvh)) return fail(ctx, LAM_E_LIMIT, "too many triangles");
    if (bvh.stack_bound > uint32_t(lam::kStack))
      return fail(ctx, LAM_E_BVH, "BVH needs a traversal stack of " + std::to_string(bvh.stack_bound) +
                                      " entries (limit " + std::to_string(lam::kStack) + ")");
    records(bvh.order.data(), tri, shd);
  } else {
    records(nullptr, tri, shd);
  }
  std::vector<uint32_t> analytic;
  for (uint32_t i = 0; i < d->n_objects; i++)
    if (d->objects[i].type != LAM_OBJ_TRIMESH) analytic.push_back(i);
  // scene box: triangle vertices and the transformed corners of the unit
  // sphere box / plane quad (the guiding network's position inputs)
  {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    auto grow = [&](const float *v) {
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], v[k]);
        hi[k] = std::max(hi[k], v[k]);
      }
    };
    for (size_t i = 0; i < size_t(nt) * 3; i++) grow(d->tri_v + 3 * i);
    for (uint32_t i : analytic) {
      const float *M = d->objects[i].T.M;
      const bool sphere = d->objects[i].type == LAM_OBJ_SPHERE;
      for (int c = 0; c < (sphere ? 8 : 4); c++) {
        const float x = (c & 1) ? 1.f : -1.f, y = (c & 2) ? 1.f : -1.f, z = sphere ? ((c & 4) ? 1.f : -1.f) : 0.f;
        float v[3];
        for (int r = 0; r < 3; r++) v[r] = M[4 * r] * x + M[4 * r + 1] * y + M[4 * r + 2] * z + M[4 * r + 3];
        grow(v);
      }
    }
    for (int k = 0; k < 3; k++) {
      if (!(lo[k] <= hi[k])) lo[k] = hi[k] = 0.f;
      ctx->box_lo[k] = lo[k];
      ctx->box_hi[k] = hi[k] > lo[k] ? hi[k] : lo[k] + 1.f;
    }
  }

  lam::DevScene &sc = ctx->sc;
  int rc;
  static_assert(sizeof(lamh::Node8) == sizeof(lam::Node8d), "node layout");
  if (!gpu_bvh) {
    const lamh::Node8 *nodes_dev = nullptr;
    if ((rc = upload(ctx, bvh.nodes.data(), bvh.nodes.size(), &nodes_dev))) return rc;
    sc.nodes = reinterpret_cast<const lam::Node8d *>(nodes_dev);
    const float *tri_dev = nullptr, *shd_dev = nullptr;
    if ((rc = upload(ctx, tri.data(), tri.size(), &tri_dev))) return rc;
    if ((rc = upload(ctx, shd.data(), shd.size(), &shd_dev))) return rc;
    sc.tris = reinterpret_cast<const float4 *>(tri_dev);
    sc.shade = reinterpret_cast<const float4 *>(shd_dev);
  } else if ((rc = build_bvh_on_device(ctx, tri, shd, nt, &gst))) {
    return rc;
  }
  if ((rc = upload(ctx, analytic.data(), analytic.size(), &sc.analytic))) return rc;
  if ((rc = upload(ctx, d->objects, d->n_objects, &sc.objects))) return rc;
  if ((rc = upload(ctx, d->materials, d->n_materials, &sc.materials))) return rc;
  if ((rc = upload(ctx, d->lights, d->n_lights, &sc.lights))) return rc;
  {
    // padded to whole float4s: the shade kernel stages the table in LDS 16 B at a time
    std::vector<float> spec((size_t(d->n_spectra) * gn + 3) / 4 * 4, 0.f);
    std::copy(d->spectra, d->spectra + size_t(d->n_spectra) * gn, spec.begin());
    if ((rc = upload(ctx, spec.data(), spec.size(), &sc.spectra))) return rc;
  }
  if ((rc = upload(ctx, d->cmf, size_t(3) * gn, &sc.cmf))) return rc;
  sc.n_nodes = gpu_bvh ? gst.n_nodes : uint32_t(bvh.nodes.size());
  sc.n_tris = nt;
  sc.n_analytic = uint32_t(analytic.size());
  sc.n_objects = d->n_objects;
  sc.n_materials = d->n_materials;
  sc.n_lights = d->n_lights;
  sc.n_spectra = d->n_spectra;
  sc.grid_n = gn;
  sc.grid_inf = d->grid.inf;
  sc.grid_sup = d->grid.sup;
  sc.grid_delta = (d->grid.sup - d->grid.inf) / float(gn - 1);
  sc.cam = d->camera;
  ctx->stats.n_nodes = sc.n_nodes;
  ctx->stats.n_leaves = gpu_bvh ? gst.n_leaves : bvh.n_leaves;
  ctx->stats.n_tris = nt;
  ctx->stats.depth = gpu_bvh ? gst.depth : bvh.depth;
  ctx->stats.stack_bound = gpu_bvh ? gst.stack_bound : bvh.stack_bound;
  ctx->stats.n_analytic = sc.n_analytic;
  ctx->stats.sah_cost = gpu_bvh ? 0.0 : bvh.sah_cost;
  ctx->has_scene = true;
  ctx->film_spp = 0;  // a film (and its image) belongs to a scene and camera
  ctx->img_npix = 0;
  return LAM_OK;
}

int lam_get_scene_stats(const lam_ctx *ctx, lam_scene_stats *out) {
  if (!ctx || !out) return LAM_E_INVALID;
  if (!ctx->has_scene) return LAM_E_NOSCENE;
  *out = ctx->stats;
  return LAM_OK;
}

static int trace_rays_host(lam_ctx *ctx, const lam_ray *rays, size_t n, lam_hit *hits, uint32_t *occl) {
  if (!ctx) return LAM_E_INVALID;
  if (n && (!rays || (!hits && !occl))) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (n == 0) return LAM_OK;
  if (n > (size_t(1) << 31)) return fail(ctx, LAM_E_LIMIT, "ray batch too large");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_scratch(ctx, 0, n * sizeof(lam_ray)))) return rc;
  if ((rc = ensure_scratch(ctx, 1, n * sizeof(lam_hit)))) return rc;
  if ((rc = ensure_launch(ctx)) || (rc = ensure_spill(ctx, 0))) return rc;
  HIPC(hipMemcpyAsync(ctx->scratch[0], rays, n * sizeof(lam_ray), hipMemcpyHostToDevice, ctx->stream));
  if ((rc = clear_err_flag(ctx))) return rc;
  HIPC(lam::launch_zero(ctx->trace_ctr, 16, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  HIPC(lam::launch_trace_rays(ctx->sc, static_cast<lam_ray *>(ctx->scratch[0]), n,
                              occl ? nullptr : static_cast<lam_hit *>(ctx->scratch[1]),
                              occl ? static_cast<uint32_t *>(ctx->scratch[1]) : nullptr, nullptr, ctx->d_err,
                              ctx->trace_ctr, ctx->lane[0].spill, ctx->batch_blocks, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  if (occl)
    HIPC(hipMemcpyAsync(occl, ctx->scratch[1], n * 4, hipMemcpyDeviceToHost, ctx->stream));
  else
    HIPC(hipMemcpyAsync(hits, ctx->scratch[1], n * sizeof(lam_hit), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  const double ms = event_ms(ctx);
  if (occl) { ctx->timing.shadow_ms += ms; ctx->timing.shadow_launches++; ctx->timing.shadow_rays += n; }
  else { ctx->timing.extend_ms += ms; ctx->timing.extend_launches++; ctx->timing.extend_rays += n; }
  return check_err_flag(ctx);
}

int lam_extend(lam_ctx *ctx, const lam_ray *rays, size_t n, lam_hit *hits) {
  return trace_rays_host(ctx, rays, n, hits, nullptr);
}

int lam_occluded(lam_ctx *ctx, const lam_ray *rays, size_t n, uint32_t *occluded) {
  if (ctx && n && !occluded) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  return trace_rays_host(ctx, rays, n, nullptr, occluded);
}

int lam_extend_dev(lam_ctx *ctx, const lam_ray *d_rays, size_t n, lam_hit *d_hits, uint32_t *d_counts) {
  if (!ctx) return LAM_E_INVALID;
  if (n && (!d_rays || !d_hits)) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (n == 0) return LAM_OK;
  if (n > (size_t(1) << 31)) return fail(ctx, LAM_E_LIMIT, "ray batch too large");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_launch(ctx)) || (rc = ensure_spill(ctx, 0))) return rc;
  if ((rc = clear_err_flag(ctx))) return rc;
  HIPC(lam::launch_zero(ctx->trace_ctr, 16, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  HIPC(lam::launch_trace_rays(ctx->sc, d_rays, n, d_hits, nullptr, d_counts, ctx->d_err, ctx->trace_ctr,
                              ctx->lane[0].spill, ctx->batch_blocks, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->timing.extend_ms += event_ms(ctx);
  ctx->timing.extend_launches++;
  ctx->timing.extend_rays += n;
  return check_err_flag(ctx);
}

int lam_occluded_dev(lam_ctx *ctx, const lam_ray *d_rays, size_t n, uint32_t *d_occluded, uint32_t *d_counts) {
  if (!ctx) return LAM_E_INVALID;
  if (n && (!d_rays || !d_occluded)) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (n == 0) return LAM_OK;
  if (n > (size_t(1) << 31)) return fail(ctx, LAM_E_LIMIT, "ray batch too large");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_launch(ctx)) || (rc = ensure_spill(ctx, 0))) return rc;
  if ((rc = clear_err_flag(ctx))) return rc;
  HIPC(lam::launch_zero(ctx->trace_ctr, 16, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  HIPC(lam::launch_trace_rays(ctx->sc, d_rays, n, nullptr, d_occluded, d_counts, ctx->d_err, ctx->trace_ctr,
                              ctx->lane[0].spill, ctx->batch_blocks, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->timing.shadow_ms += event_ms(ctx);
  ctx->timing.shadow_launches++;
  ctx->timing.shadow_rays += n;
  return check_err_flag(ctx);
}

int lam_occluded_dev_timed(lam_ctx *ctx, const lam_ray *d_rays, size_t n, uint32_t *d_occluded, uint32_t reps,
                           double *avg_ms) {
  if (!ctx) return LAM_E_INVALID;
  if (!d_rays || !d_occluded || !avg_ms || reps == 0 || n == 0) return fail(ctx, LAM_E_INVALID, "bad argument");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (n > (size_t(1) << 31)) return fail(ctx, LAM_E_LIMIT, "ray batch too large");
  if (reps > kTraceCounters) return fail(ctx, LAM_E_LIMIT, "at most 4096 launches per call");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_launch(ctx)) || (rc = ensure_spill(ctx, 0))) return rc;
  if ((rc = clear_err_flag(ctx))) return rc;
  HIPC(lam::launch_zero(ctx->trace_ctr, size_t(kTraceCounters) * 4, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  for (uint32_t r = 0; r < reps; r++)
    HIPC(lam::launch_trace_rays(ctx->sc, d_rays, n, nullptr, d_occluded, nullptr, ctx->d_err, ctx->trace_ctr + r,
                                ctx->lane[0].spill, ctx->batch_blocks, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  const double ms = event_ms(ctx);
  *avg_ms = ms / reps;
  ctx->timing.shadow_ms += ms;
  ctx->timing.shadow_launches += reps;
  ctx->timing.shadow_rays += n * reps;
  return check_err_flag(ctx);
}

int lam_extend_dev_noalloc(lam_ctx *ctx, const lam_ray *d_rays, size_t n, lam_hit *d_hits) {
  if (!ctx) return LAM_E_INVALID;
  if (n > (size_t(1) << 31)) return fail(ctx, LAM_E_LIMIT, "ray batch too large (at most 2^31 rays)");
  if (n && (!d_rays || !d_hits)) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (!ctx->noalloc)
    return fail(ctx, LAM_E_DEVICE, std::string("no-alloc scratch unavailable: hipGetSymbolAddress: ") +
                                       hipGetErrorString(ctx->noalloc_rc));
  if (n == 0) return LAM_OK;
  HIPC(hipSetDevice(ctx->device));
  int cus = ctx->cus;
  if (cus <= 0) HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  // 8 blocks per CU, at most the grid the spill columns are sized for
  const uint32_t blocks = std::min<uint32_t>(uint32_t(std::max(cus, 1)) * 8u, kNoAllocBlocks);
  uint32_t *fetch = ctx->noalloc, *err = ctx->noalloc + 4, *spill = ctx->noalloc + kNoAllocHeaderWords;
  // the queue counter and the overflow flag start at zero (a fill kernel,
  // not a memset node, so the call can be captured)
  HIPC(lam::launch_zero(ctx->noalloc, kNoAllocHeaderWords * 4, ctx->stream));
  HIPC(lam::launch_trace_rays(ctx->sc, d_rays, n, d_hits, nullptr, nullptr, err, fetch, spill, blocks,
                              ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  uint32_t flag = 0;
  HIPC(hipMemcpy(&flag, err, 4, hipMemcpyDeviceToHost));
  if (flag) return fail(ctx, LAM_E_BVH, "BVH traversal stack overflow");
  return LAM_OK;
}

int lam_extend_dev_timed(lam_ctx *ctx, const lam_ray *d_rays, size_t n, lam_hit *d_hits, uint32_t reps,
                         double *avg_ms) {
  if (!ctx) return LAM_E_INVALID;
  if (!d_rays || !d_hits || !avg_ms || reps == 0 || n == 0) return fail(ctx, LAM_E_INVALID, "bad argument");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (n > (size_t(1) << 31)) return fail(ctx, LAM_E_LIMIT, "ray batch too large");
  if (reps > kTraceCounters) return fail(ctx, LAM_E_LIMIT, "at most 4096 launches per call");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_launch(ctx)) || (rc = ensure_spill(ctx, 0))) return rc;
  if ((rc = clear_err_flag(ctx))) return rc;
  // one zeroed fetch counter per launch, cleared before the timed region
  HIPC(lam::launch_zero(ctx->trace_ctr, size_t(kTraceCounters) * 4, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  for (uint32_t r = 0; r < reps; r++)
    HIPC(lam::launch_trace_rays(ctx->sc, d_rays, n, d_hits, nullptr, nullptr, ctx->d_err, ctx->trace_ctr + r,
                                ctx->lane[0].spill, ctx->batch_blocks, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  const double ms = event_ms(ctx);
  *avg_ms = ms / reps;
  ctx->timing.extend_ms += ms;
  ctx->timing.extend_launches += reps;
  ctx->timing.extend_rays += n * reps;
  return check_err_flag(ctx);
}

int lam_trace_paths(lam_ctx *ctx, const lam_render_params *p, const uint32_t *pixel, const uint32_t *sample,
                    size_t n, float *out) {
  if (!ctx) return LAM_E_INVALID;
  int rc;
  if ((rc = check_params(ctx, p, false))) return rc;
  if (n == 0) return LAM_OK;
  if (!pixel || !sample || !out) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (n > (size_t(1) << 28)) return fail(ctx, LAM_E_LIMIT, "path batch too large");
  const uint64_t npix = uint64_t(ctx->sc.cam.width) * ctx->sc.cam.height;
  for (size_t i = 0; i < n; i++)
    if (pixel[i] >= npix) return fail(ctx, LAM_E_INVALID, "pixel index out of range");
  HIPC(hipSetDevice(ctx->device));
  const size_t H = p->hero;
  if ((rc = ensure_scratch(ctx, 0, n * 4))) return rc;
  if ((rc = ensure_scratch(ctx, 1, n * 4))) return rc;
  if ((rc = ensure_scratch(ctx, 2, n * H * 4))) return rc;
  HIPC(hipMemcpyAsync(ctx->scratch[0], pixel, n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->scratch[1], sample, n * 4, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = clear_err_flag(ctx))) return rc;
  const uint32_t *dpix = static_cast<uint32_t *>(ctx->scratch[0]);
  const uint32_t *dsmp = static_cast<uint32_t *>(ctx->scratch[1]);
  float *dout = static_cast<float *>(ctx->scratch[2]);
  if (p->flags & LAM_FLAG_MEGAKERNEL) {
    if ((rc = ensure_launch(ctx)) || (rc = ensure_spill(ctx, 0))) return rc;
    HIPC(lam::launch_trace_paths(ctx->sc, path_args(p), int(H), dpix, dsmp, n, dout, ctx->d_err, ctx->lane[0].spill,
                                 ctx->wl.finish_blocks, ctx->stream));
  } else {
    const size_t cap = std::min(n, kMaxWaveCap);
    if ((rc = ensure_wave(ctx, cap, int(H)))) return rc;
    HIPC(lam::launch_zero(ctx->lane[0].ctrl, lam::kCtrlWords * 4, ctx->stream));
    for (size_t off = 0; off < n; off += cap) {
      const uint32_t m = uint32_t(std::min(cap, n - off));
      HIPC(lam::enqueue_pass(int(H), ctx->sc, path_args(p), dpix + off, dsmp + off, 0, ctx->lane[0].ws, m,
                             ctx->lane[0].ctrl,
                             nullptr, dout + off * H, launch_for(ctx, p), ctx->stream, nullptr));
    }
  }
  HIPC(hipMemcpyAsync(out, dout, n * H * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if ((rc = check_ctrl_overflow(ctx))) return rc;
  return check_err_flag(ctx);
}

int lam_render(lam_ctx *ctx, const lam_render_params *p, float *linrgb, float *xyz) {
  if (!ctx) return LAM_E_INVALID;
  int rc;
  if ((rc = check_params(ctx, p))) return rc;
  if (!linrgb) return fail(ctx, LAM_E_INVALID, "NULL output buffer");
  HIPC(hipSetDevice(ctx->device));
  const uint32_t w = ctx->sc.cam.width, h = ctx->sc.cam.height;
  const size_t npix = size_t(w) * h;
  if (npix > (size_t(1) << 30)) return fail(ctx, LAM_E_LIMIT, "image too large");
  if ((rc = ensure_scratch(ctx, 4, npix * 12))) return rc;  // linRGB
  if ((rc = ensure_scratch(ctx, 5, npix * 12))) return rc;  // XYZ sums
  float *film = nullptr;
  uint64_t base_spp = 0;
  if ((rc = begin_film(ctx, p, npix, &film, &base_spp))) return rc;
  if ((rc = clear_err_flag(ctx))) return rc;
  const lam::PathArgs pa = path_args(p);
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  if (p->flags & LAM_FLAG_MEGAKERNEL) {
    if ((rc = ensure_launch(ctx)) || (rc = ensure_spill(ctx, 0))) return rc;
    for (uint32_t s = 0; s < p->spp; s++)
      HIPC(lam::launch_render_mega(ctx->sc, pa, int(p->hero), p->spp_offset + s, uint32_t(npix), film,
                                   ctx->d_err, ctx->lane[0].spill, ctx->wl.finish_blocks, ctx->stream));
  } else if ((rc = render_wavefront(ctx, p, film, npix))) {
    return rc;
  }
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  finish_film(ctx, p, npix, base_spp);
  std::vector<unsigned long long> rays;  // the megakernel does not count its rays
  if (!(p->flags & LAM_FLAG_MEGAKERNEL) && (rc = read_rays_async(ctx, rays))) return rc;
  const float inv_n = 1.f / float(ctx->film_spp * p->hero);
  HIPC(lam::launch_resolve(ctx->sc, film, w, h, inv_n, static_cast<float *>(ctx->scratch[4]),
                           xyz ? static_cast<float *>(ctx->scratch[5]) : nullptr, ctx->stream));
  HIPC(hipMemcpyAsync(linrgb, ctx->scratch[4], npix * 12, hipMemcpyDeviceToHost, ctx->stream));
  if (xyz) HIPC(hipMemcpyAsync(xyz, ctx->scratch[5], npix * 12, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->timing.render_ms = event_ms(ctx);
  add_rays(ctx, rays);
  if ((rc = check_ctrl_overflow(ctx))) return rc;
  return check_err_flag(ctx);
}

int lam_adjoint(lam_ctx *ctx, const lam_render_params *p, const lam_diff_params *dp, const float *dloss_drgb,
                float *grad) {
  if (!ctx) return LAM_E_INVALID;
  int rc;
  if ((rc = check_params(ctx, p))) return rc;
  if (!dp || !dloss_drgb || !grad) return fail(ctx, LAM_E_INVALID, "NULL argument");
  if (!dp->mat_d || !dp->mat_ior || !dp->mat_albedo) return fail(ctx, LAM_E_INVALID, "parameter maps missing");
  if (ctx->has_scene && dp->n_materials != ctx->sc.n_materials)
    return fail(ctx, LAM_E_INVALID, "parameter maps have " + std::to_string(dp->n_materials) +
                                        " entries, the uploaded scene " + std::to_string(ctx->sc.n_materials) +
                                        " materials");
  if (dp->n_params == 0) return LAM_OK;
  if (dp->n_params > lam::kMaxDiffParams)
    return fail(ctx, LAM_E_LIMIT, "at most " + std::to_string(lam::kMaxDiffParams) + " parameters");
  const uint32_t nm = ctx->sc.n_materials, gn = ctx->sc.grid_n;
  for (uint32_t m = 0; m < nm; m++) {
    const int32_t d = dp->mat_d[m], io = dp->mat_ior[m], al = dp->mat_albedo[m];
    if (d >= 0 && uint32_t(d) >= dp->n_params) return fail(ctx, LAM_E_INVALID, "mat_d index out of range");
    if (io >= 0 && uint64_t(io) + gn > dp->n_params) return fail(ctx, LAM_E_INVALID, "mat_ior range out of range");
    if (al >= 0 && uint64_t(al) + gn > dp->n_params)
      return fail(ctx, LAM_E_INVALID, "mat_albedo range out of range");
  }
  HIPC(hipSetDevice(ctx->device));
  const uint32_t w = ctx->sc.cam.width, h = ctx->sc.cam.height;
  const size_t npix = size_t(w) * h;
  if (npix > (size_t(1) << 30)) return fail(ctx, LAM_E_LIMIT, "image too large");
  if ((rc = ensure_launch(ctx))) return rc;
  // scratch: [0] parameter maps, [1] dloss, [2] gradient, then the replay
  // kernels' partial rows (shade_blocks per lane)
  const size_t maps = size_t(nm) * 3 * 4;
  const size_t gbytes = (size_t(dp->n_params) * 4 + 255) / 256 * 256;
  const uint32_t part_rows = ctx->wl.shade_blocks;
  const size_t part_bytes = size_t(part_rows) * dp->n_params * 4;
  if ((rc = ensure_scratch(ctx, 0, maps))) return rc;
  if ((rc = ensure_scratch(ctx, 1, npix * 12))) return rc;
  const size_t sbytes = (gbytes + kMaxLanes * part_bytes + 15) / 16 * 16;
  if ((rc = ensure_scratch(ctx, 2, sbytes))) return rc;
  int32_t *dmaps = static_cast<int32_t *>(ctx->scratch[0]);
  HIPC(hipMemcpyAsync(dmaps, dp->mat_d, nm * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dmaps + nm, dp->mat_ior, nm * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(dmaps + 2 * nm, dp->mat_albedo, nm * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->scratch[1], dloss_drgb, npix * 12, hipMemcpyHostToDevice, ctx->stream));
  float *dgrad = static_cast<float *>(ctx->scratch[2]);
  float *dpart = dgrad + gbytes / 4;
  HIPC(lam::launch_zero(dgrad, sbytes, ctx->stream));
  if ((rc = clear_err_flag(ctx))) return rc;
  // Passes of m samples of every pixel of a chunk, as lam_render's: the
  // forward pass (the render's wavefront pass without a film, which leaves
  // each path's total radiance in the lane's path state), then the replay
  // pass of the same slots on the same lane.  Consecutive passes go
  // round-robin to the pass lanes, so one lane's forward pass runs beside
  // another's replay; every replay adds into the one gradient with atomics.
  const size_t cpix = pass_pixels(npix);
  const uint32_t m = pass_samples(ctx, npix, p->spp, (p->flags & LAM_FLAG_NO_PIPE) ? 1u : ctx->lanes);
  const size_t cap = cpix * m;
  const uint32_t npass = (p->spp + m - 1) / m;
  const uint32_t lanes = (p->flags & LAM_FLAG_NO_PIPE) ? 1u : std::min(ctx->lanes, npass);
  hipStream_t st[kMaxLanes];
  for (uint32_t l = 0; l < lanes; l++) {
    if ((rc = ensure_wave(ctx, cap, int(p->hero), int(l)))) return rc;
    if ((rc = ensure_replay(ctx, int(l), int(p->hero)))) return rc;
    st[l] = l == 0 ? ctx->stream : ctx->lane[l].stream;
    HIPC(lam::launch_zero(ctx->lane[l].ctrl, lam::kCtrlWords * 4, ctx->stream));
  }
  const lam::PathArgs pa = path_args(p);
  lam::WaveLaunch wl = launch_for(ctx, p);
  const float inv_n = 1.f / float(uint64_t(p->spp) * p->hero);
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  HIPC(hipEventRecord(ctx->lane[0].ev_join, ctx->stream));
  for (uint32_t l = 1; l < lanes; l++) HIPC(hipStreamWaitEvent(st[l], ctx->lane[0].ev_join, 0));
  hipError_t e = hipSuccess;
  uint32_t gi = 0;
  for (size_t off = 0; off < npix && e == hipSuccess; off += cpix) {
    const uint32_t n = uint32_t(std::min(cpix, npix - off));
    for (uint32_t first = 0; first < p->spp && e == hipSuccess; first += m, gi++) {
      const uint32_t mg = std::min(m, p->spp - first);
      const uint32_t l = gi % lanes;
      lam_ctx::Lane &ln = ctx->lane[l];
      wl.tail_threshold = std::min(launch_for(ctx, p).tail_threshold, n * mg / ctx->tail_div);
      e = lam::enqueue_pass_set(ln.ctrl, p->spp_offset + first, st[l]);
      if (e == hipSuccess)
        e = lam::enqueue_pass(int(p->hero), ctx->sc, pa, nullptr, nullptr, uint32_t(off), ln.ws, n * mg, ln.ctrl,
                              nullptr, nullptr, wl, st[l], nullptr, nullptr, false, mg);
      if (e == hipSuccess)
        e = lam::enqueue_replay(int(p->hero), ctx->sc, pa, uint32_t(off), mg, ln.ws, ln.rep, dmaps, dmaps + nm,
                                dmaps + 2 * nm, dp->n_params, dpart + size_t(l) * part_rows * dp->n_params, n * mg,
                                ln.ctrl, static_cast<float *>(ctx->scratch[1]), inv_n, wl, st[l]);
    }
  }
  // join: stream 0 waits for the other lanes
  for (uint32_t l = 1; l < lanes; l++) {
    hipError_t ej = hipEventRecord(ctx->lane[l].ev_join, st[l]);
    if (ej == hipSuccess) ej = hipStreamWaitEvent(ctx->stream, ctx->lane[l].ev_join, 0);
    if (e == hipSuccess) e = ej;
  }
  HIPC(e);
  HIPC(lam::enqueue_sum_rows(dpart, lanes * part_rows, dp->n_params, dgrad, ctx->stream));
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  HIPC(hipMemcpyAsync(grad, dgrad, size_t(dp->n_params) * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->timing.render_ms = event_ms(ctx);
  if ((rc = check_ctrl_overflow(ctx))) return rc;
  return check_err_flag(ctx);
}

// ---- guiding --------------------------------------------------------------------------
size_t lam_guide_param_count(void) { return size_t(lam::kGParams); }

int lam_guide_reset(lam_ctx *ctx, uint64_t seed) {
  if (!ctx) return LAM_E_INVALID;
  HIPC(hipSetDevice(ctx->device));
  std::vector<float> th;
  guide_init_params(seed, th);
  return guide_upload_params(ctx, th.data());
}

int lam_guide_set_params(lam_ctx *ctx, const float *theta) {
  if (!ctx) return LAM_E_INVALID;
  if (!theta) return fail(ctx, LAM_E_INVALID, "NULL parameters");
  for (int i = 0; i < lam::kGParams; i++)
    if (!std::isfinite(theta[i])) return fail(ctx, LAM_E_INVALID, "parameters must be finite");
  HIPC(hipSetDevice(ctx->device));
  return guide_upload_params(ctx, theta);
}

int lam_guide_get_params(lam_ctx *ctx, float *theta) {
  if (!ctx) return LAM_E_INVALID;
  if (!theta) return fail(ctx, LAM_E_INVALID, "NULL output");
  if (!ctx->guide_ready) return fail(ctx, LAM_E_INVALID, "guiding network not initialised");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpyAsync(theta, ctx->g_th, size_t(lam::kGParams) * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return LAM_OK;
}

int lam_guide_infer(lam_ctx *ctx, const float *in, size_t n, float *raw) {
  if (!ctx) return LAM_E_INVALID;
  if (n == 0) return LAM_OK;
  if (!in || !raw) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (!ctx->guide_ready) return fail(ctx, LAM_E_INVALID, "guiding network not initialised");
  if (n > (size_t(1) << 24)) return fail(ctx, LAM_E_LIMIT, "batch too large");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_scratch(ctx, 0, n * lam::kGIn * 4))) return rc;
  if ((rc = ensure_scratch(ctx, 1, n * lam::kGOut * 4))) return rc;
  HIPC(hipMemcpyAsync(ctx->scratch[0], in, n * lam::kGIn * 4, hipMemcpyHostToDevice, ctx->stream));
  int cus = 0;
  HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const uint32_t blocks = uint32_t(std::min<size_t>((n + 255) / 256, size_t(cus > 0 ? cus : 256) * 3));
  HIPC(lam::guide_infer(ctx->g_pack, static_cast<const float *>(ctx->scratch[0]), uint32_t(n),
                        static_cast<float *>(ctx->scratch[1]), blocks, ctx->stream));
  HIPC(hipMemcpyAsync(raw, ctx->scratch[1], n * lam::kGOut * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return LAM_OK;
}

static int guide_rows_in(lam_ctx *ctx, const float *rows, size_t n) {
  if (!rows) return fail(ctx, LAM_E_INVALID, "NULL rows");
  if (n == 0 || n > 65536) return fail(ctx, LAM_E_LIMIT, "batch must hold 1 to 65536 rows");
  if (!ctx->guide_ready) return fail(ctx, LAM_E_INVALID, "guiding network not initialised");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_guide_ws(ctx, uint32_t(n)))) return rc;
  if ((rc = ensure_scratch(ctx, 0, n * 40))) return rc;
  HIPC(hipMemcpyAsync(ctx->scratch[0], rows, n * 40, hipMemcpyHostToDevice, ctx->stream));
  return LAM_OK;
}

int lam_guide_loss_grad(lam_ctx *ctx, const float *rows, size_t n, float *loss, float *grad, float *raw) {
  if (!ctx) return LAM_E_INVALID;
  int rc;
  if ((rc = guide_rows_in(ctx, rows, n))) return rc;
  if (raw && (rc = ensure_scratch(ctx, 1, n * lam::kGOut * 4))) return rc;
  lam::GuideTrain gt = guide_train_state(ctx, uint32_t(n), 0.f, 0);
  float *d_raw = raw ? static_cast<float *>(ctx->scratch[1]) : nullptr;
  HIPC(lam::guide_loss_grad_rows(gt, static_cast<const float *>(ctx->scratch[0]), uint32_t(n), ctx->g_grad, d_raw,
                                 ctx->stream));
  float sc[2] = {0.f, 0.f};
  HIPC(hipMemcpyAsync(sc, ctx->g_scal, 8, hipMemcpyDeviceToHost, ctx->stream));
  if (grad) HIPC(hipMemcpyAsync(grad, ctx->g_grad, size_t(lam::kGParams) * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (raw) HIPC(hipMemcpyAsync(raw, d_raw, n * lam::kGOut * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if (loss) *loss = sc[1];
  return LAM_OK;
}

int lam_guide_train(lam_ctx *ctx, const float *rows, size_t n, uint32_t steps, float lr, float *loss) {
  if (!ctx) return LAM_E_INVALID;
  if (!(lr > 0.f)) return fail(ctx, LAM_E_INVALID, "lr must be positive");
  int rc;
  if ((rc = guide_rows_in(ctx, rows, n))) return rc;
  lam::GuideTrain gt = guide_train_state(ctx, uint32_t(n), lr, 0);
  HIPC(lam::guide_train_rows(gt, static_cast<const float *>(ctx->scratch[0]), uint32_t(n), steps, ctx->stream));
  ctx->g_step = gt.step;
  float sc[2] = {0.f, 0.f};
  HIPC(hipMemcpyAsync(sc, ctx->g_scal, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if (loss) *loss = sc[1];
  return LAM_OK;
}

int lam_render_guided(lam_ctx *ctx, const lam_render_params *p, const lam_guide_params *g, float *linrgb,
                      float *xyz) {
  if (!ctx) return LAM_E_INVALID;
  int rc;
  if ((rc = check_params(ctx, p))) return rc;
  if (!g) return fail(ctx, LAM_E_INVALID, "guide params are NULL");
  if (!linrgb) return fail(ctx, LAM_E_INVALID, "NULL output buffer");
  if (!(g->alpha >= 0.f && g->alpha < 1.f)) return fail(ctx, LAM_E_INVALID, "alpha must be in [0, 1)");
  if (g->train_passes && (g->batch < 256 || g->batch > 65536 || g->batch % 256))
    return fail(ctx, LAM_E_INVALID, "batch must be a multiple of 256 in [256, 65536]");
  if (g->train_passes && !(g->lr > 0.f)) return fail(ctx, LAM_E_INVALID, "lr must be positive");
  if (p->flags & LAM_FLAG_MEGAKERNEL) return fail(ctx, LAM_E_INVALID, "guiding needs the wavefront path");
  HIPC(hipSetDevice(ctx->device));
  if (!ctx->guide_ready && (rc = lam_guide_reset(ctx, g->seed))) return rc;
  const uint32_t w = ctx->sc.cam.width, h = ctx->sc.cam.height;
  const size_t npix = size_t(w) * h;
  if (npix > (size_t(1) << 30)) return fail(ctx, LAM_E_LIMIT, "image too large");
  // the training passes trace train_pass_samples samples of every pixel
  const size_t train_slots = pass_pixels(npix) * train_pass_samples(ctx, npix, p->spp, g->train_passes);
  if ((rc = ensure_wave(ctx, train_slots, int(p->hero)))) return rc;
  const size_t wcap = ctx->lane[0].ws.cap;
  if (g->train_passes) {
    if (ctx->g_rec_slot_cap < wcap) {
      if (ctx->g_rec_of_slot) HIPC(hipFree(ctx->g_rec_of_slot));
      ctx->g_rec_of_slot = nullptr;
      ctx->g_rec_slot_cap = 0;
      HIPC(hipMalloc(reinterpret_cast<void **>(&ctx->g_rec_of_slot), wcap * 4));
      ctx->g_rec_slot_cap = wcap;
    }
    const uint32_t rcap = uint32_t(std::min<size_t>(wcap * 8, size_t(1) << 25));
    if (ctx->g_rec_cap < rcap) {
      if (ctx->g_rec) HIPC(hipFree(ctx->g_rec));
      ctx->g_rec = nullptr;
      ctx->g_rec_cap = 0;
      HIPC(hipMalloc(reinterpret_cast<void **>(&ctx->g_rec), size_t(rcap) * 64));
      ctx->g_rec_cap = rcap;
    }
    if (!ctx->g_rec_count) HIPC(hipMalloc(reinterpret_cast<void **>(&ctx->g_rec_count), 64));
    if ((rc = ensure_guide_ws(ctx, g->batch))) return rc;
  }
  GuideRun gr{};
  gr.gp = g;
  lam::GuideDev &gd = gr.gd_render;
  gd.frag = static_cast<const uint4 *>(ctx->g_pack);
  gd.mix = nullptr;  // per lane, set by render_wavefront
  gd.alpha = g->alpha;
  gd.mat_mask = g->materials ? g->materials
                             : (1u << LAM_MAT_LAMBERTIAN) | (1u << LAM_MAT_BLINN) | (1u << LAM_MAT_THINFILM);
  gd.max_vertex = g->max_vertex ? g->max_vertex : 0xffffffffu;
  gd.lobes = g->lobes ? g->lobes : (LAM_GUIDE_LAMBERTIAN | LAM_GUIDE_BLINN_KD);
  gd.lo = lam::f3{ctx->box_lo[0], ctx->box_lo[1], ctx->box_lo[2]};
  gd.inv_ext = lam::f3{1.f / (ctx->box_hi[0] - ctx->box_lo[0]), 1.f / (ctx->box_hi[1] - ctx->box_lo[1]),
                       1.f / (ctx->box_hi[2] - ctx->box_lo[2])};
  gr.gd_train = gd;
  gr.gd_train.rec = ctx->g_rec;
  gr.gd_train.rec_count = ctx->g_rec_count;
  gr.gd_train.rec_cap = ctx->g_rec_cap;
  gr.gd_train.rec_of_slot = ctx->g_rec_of_slot;
  if ((rc = ensure_scratch(ctx, 4, npix * 12))) return rc;
  if ((rc = ensure_scratch(ctx, 5, npix * 12))) return rc;
  float *film = nullptr;
  uint64_t base_spp = 0;
  if ((rc = begin_film(ctx, p, npix, &film, &base_spp))) return rc;
  if ((rc = clear_err_flag(ctx))) return rc;
  HIPC(hipEventRecord(ctx->ev[0], ctx->stream));
  if ((rc = render_wavefront(ctx, p, film, npix, &gr))) return rc;
  HIPC(hipEventRecord(ctx->ev[1], ctx->stream));
  finish_film(ctx, p, npix, base_spp);
  uint32_t gc[kMaxLanes][2] = {};
  for (int l = 0; l < kMaxLanes; l++)
    if (ctx->lane[l].ctrl) HIPC(hipMemcpyAsync(gc[l], ctx->lane[l].ctrl + 12, 8, hipMemcpyDeviceToHost, ctx->stream));
  std::vector<unsigned long long> rays;
  if ((rc = read_rays_async(ctx, rays))) return rc;
  const float inv_n = 1.f / float(ctx->film_spp * p->hero);
  HIPC(lam::launch_resolve(ctx->sc, film, w, h, inv_n, static_cast<float *>(ctx->scratch[4]),
                           xyz ? static_cast<float *>(ctx->scratch[5]) : nullptr, ctx->stream));
  HIPC(hipMemcpyAsync(linrgb, ctx->scratch[4], npix * 12, hipMemcpyDeviceToHost, ctx->stream));
  if (xyz) HIPC(hipMemcpyAsync(xyz, ctx->scratch[5], npix * 12, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->timing.render_ms = event_ms(ctx);
  add_rays(ctx, rays);
  ctx->timing.guide_dropped = ctx->timing.guide_zero_pdf = 0;
  for (int l = 0; l < kMaxLanes; l++) {
    ctx->timing.guide_dropped += gc[l][0];
    ctx->timing.guide_zero_pdf += gc[l][1];
  }
  if ((rc = check_ctrl_overflow(ctx))) return rc;
  return check_err_flag(ctx);
}

int lam_debug_math(lam_ctx *ctx, int fn, const float *a, const float *b, size_t n, float *out) {
  if (!ctx) return LAM_E_INVALID;
  if (n == 0) return LAM_OK;
  if (!a || !b || !out) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (n > (size_t(1) << 28)) return fail(ctx, LAM_E_LIMIT, "batch too large");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_scratch(ctx, 0, n * 4))) return rc;
  if ((rc = ensure_scratch(ctx, 1, n * 4))) return rc;
  if ((rc = ensure_scratch(ctx, 2, n * 4))) return rc;
  HIPC(hipMemcpyAsync(ctx->scratch[0], a, n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->scratch[1], b, n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(lam::launch_math(fn, static_cast<float *>(ctx->scratch[0]), static_cast<float *>(ctx->scratch[1]), n,
                        static_cast<float *>(ctx->scratch[2]), ctx->stream));
  HIPC(hipMemcpyAsync(out, ctx->scratch[2], n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return LAM_OK;
}

int lam_debug_radix_sort(lam_ctx *ctx, const uint32_t *keys, const uint32_t *vals, size_t n, uint32_t bits,
                         uint32_t *keys_out, uint32_t *vals_out) {
  if (!ctx) return LAM_E_INVALID;
  if (n == 0) return LAM_OK;
  if (!keys || !vals || !keys_out || !vals_out) return fail(ctx, LAM_E_INVALID, "NULL buffer");
  if (bits < 1 || bits > 32) return fail(ctx, LAM_E_INVALID, "bits must be 1..32");
  if (n > (size_t(1) << 28)) return fail(ctx, LAM_E_LIMIT, "batch too large");
  HIPC(hipSetDevice(ctx->device));
  int rc;
  // (slot 3 holds the film of the last render: not used here)
  const int slots[4] = {0, 1, 2, 4};
  for (int i : slots)
    if ((rc = ensure_scratch(ctx, i, n * 4))) return rc;
  if ((rc = ensure_scratch(ctx, 5, lam::radix_sort_scratch(uint32_t(n))))) return rc;
  uint32_t *kin = static_cast<uint32_t *>(ctx->scratch[0]), *vin = static_cast<uint32_t *>(ctx->scratch[1]);
  uint32_t *kout = static_cast<uint32_t *>(ctx->scratch[2]), *vout = static_cast<uint32_t *>(ctx->scratch[4]);
  HIPC(hipMemcpyAsync(kin, keys, n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(vin, vals, n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(lam::radix_sort_pairs(kin, vin, kout, vout, uint32_t(n), bits, ctx->scratch[5], ctx->stream));
  HIPC(hipMemcpyAsync(keys_out, kout, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipMemcpyAsync(vals_out, vout, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return LAM_OK;
}

int lam_get_timing(const lam_ctx *ctx, lam_timing *out) {
  if (!ctx || !out) return LAM_E_INVALID;
  *out = ctx->timing;
  return LAM_OK;
}

int lam_reset_timing(lam_ctx *ctx) {
  if (!ctx) return LAM_E_INVALID;
  ctx->timing = lam_timing{};
  return LAM_OK;
}

int lam_tone_map(lam_ctx *ctx, float y_mid, uint8_t *rgb8) {
  if (!ctx) return LAM_E_INVALID;
  if (!rgb8) return fail(ctx, LAM_E_INVALID, "NULL output buffer");
  if (!(y_mid > 0.f)) return fail(ctx, LAM_E_INVALID, "y_mid must be positive");
  if (ctx->img_npix == 0) return fail(ctx, LAM_E_INVALID, "no rendered image");
  HIPC(hipSetDevice(ctx->device));
  if (!ctx->tone_buf) HIPC(hipMalloc(&ctx->tone_buf, 1024 * 8 + 16));
  int rc;
  if ((rc = ensure_scratch(ctx, 5, ctx->img_npix * 12))) return rc;
  double *partial = static_cast<double *>(ctx->tone_buf);
  float *exposure = reinterpret_cast<float *>(partial + 1024);
  uint8_t *d_out = static_cast<uint8_t *>(ctx->scratch[5]);
  HIPC(lam::launch_tonemap8(static_cast<const float *>(ctx->scratch[4]), uint32_t(ctx->img_npix), y_mid, partial,
                            exposure, d_out, ctx->stream));
  HIPC(hipMemcpyAsync(rgb8, d_out, ctx->img_npix * 3, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return LAM_OK;
}

int lam_film_get(lam_ctx *ctx, float *film, uint32_t *spp, uint32_t *hero) {
  if (!ctx) return LAM_E_INVALID;
  if (!film || !spp || !hero) return fail(ctx, LAM_E_INVALID, "NULL argument");
  if (ctx->film_spp == 0) return fail(ctx, LAM_E_INVALID, "no film");
  if (ctx->film_spp > UINT32_MAX) return fail(ctx, LAM_E_LIMIT, "film holds more than 2^32-1 samples per pixel");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpyAsync(film, ctx->scratch[3], ctx->film_npix * 32, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  *spp = uint32_t(ctx->film_spp);
  *hero = ctx->film_hero;
  return LAM_OK;
}

int lam_film_set(lam_ctx *ctx, const float *film, uint32_t spp, uint32_t hero) {
  if (!ctx) return LAM_E_INVALID;
  if (!film) return fail(ctx, LAM_E_INVALID, "NULL film");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (spp == 0) return fail(ctx, LAM_E_INVALID, "a film holds at least one sample per pixel");
  if (hero != 4 && hero != 16) return fail(ctx, LAM_E_INVALID, "hero must be 4 or 16");
  const size_t npix = size_t(ctx->sc.cam.width) * ctx->sc.cam.height;
  HIPC(hipSetDevice(ctx->device));
  int rc;
  if ((rc = ensure_scratch(ctx, 3, npix * 32))) return rc;
  HIPC(hipMemcpyAsync(ctx->scratch[3], film, npix * 32, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->film_npix = npix;
  ctx->film_hero = hero;
  ctx->film_spp = spp;
  return LAM_OK;
}

int lam_set_camera(lam_ctx *ctx, const lam_camera *cam) {
  if (!ctx) return LAM_E_INVALID;
  if (!cam) return fail(ctx, LAM_E_INVALID, "NULL camera");
  if (!ctx->has_scene) return fail(ctx, LAM_E_NOSCENE, "no scene uploaded");
  if (cam->width == 0 || cam->height == 0) return fail(ctx, LAM_E_INVALID, "empty image");
  if (size_t(cam->width) * cam->height > (size_t(1) << 30)) return fail(ctx, LAM_E_LIMIT, "image too large");
  ctx->sc.cam = *cam;
  ctx->film_spp = 0;
  ctx->img_npix = 0;
  return LAM_OK;
}

int lam_get_wave_config(const lam_ctx *ctx, lam_wave_config *out) {
  if (!ctx || !out) return LAM_E_INVALID;
  out->lanes = ctx->lanes;
  out->tail_threshold = ctx->wl.tail_threshold;
  out->tail_div = ctx->tail_div;
  out->max_iters = ctx->wl.max_iters;
  out->trace_per_cu = ctx->trace_per_cu;
  out->batch_per_cu = ctx->batch_per_cu;
  out->finish_per_cu = ctx->finish_per_cu;
  out->shade_per_cu = ctx->shade_per_cu;
  out->samples_per_pass = ctx->samples_per_pass;
  return LAM_OK;
}

int lam_set_wave_config(lam_ctx *ctx, const lam_wave_config *cfg) {
  if (!ctx) return LAM_E_INVALID;
  if (!cfg) return fail(ctx, LAM_E_INVALID, "NULL config");
  if (cfg->lanes < 1 || cfg->lanes > uint32_t(kMaxLanes)) return fail(ctx, LAM_E_INVALID, "lanes must be 1..4");
  if (cfg->tail_div < 1) return fail(ctx, LAM_E_INVALID, "tail_div must be >= 1");
  // the finish kernel runs only where max_iters is set (LAM_FLAG_NO_TAIL
  // gives the pure wavefront)
  if (cfg->max_iters < 1) return fail(ctx, LAM_E_INVALID, "max_iters must be >= 1");
  if (cfg->trace_per_cu < 1 || cfg->trace_per_cu > 32) return fail(ctx, LAM_E_INVALID, "trace_per_cu must be 1..32");
  if (cfg->batch_per_cu < 1 || cfg->batch_per_cu > 32) return fail(ctx, LAM_E_INVALID, "batch_per_cu must be 1..32");
  if (cfg->finish_per_cu < 1 || cfg->finish_per_cu > 32)
    return fail(ctx, LAM_E_INVALID, "finish_per_cu must be 1..32");
  if (cfg->shade_per_cu < 1 || cfg->shade_per_cu > 8) return fail(ctx, LAM_E_INVALID, "shade_per_cu must be 1..8");
  if (cfg->samples_per_pass < 1 || cfg->samples_per_pass > 32)
    return fail(ctx, LAM_E_INVALID, "samples_per_pass must be 1..32");
  ctx->lanes = cfg->lanes;
  ctx->wl.tail_threshold = cfg->tail_threshold;
  ctx->tail_div = cfg->tail_div;
  ctx->wl.max_iters = cfg->max_iters;
  ctx->trace_per_cu = cfg->trace_per_cu;
  ctx->batch_per_cu = cfg->batch_per_cu;
  ctx->finish_per_cu = cfg->finish_per_cu;
  ctx->shade_per_cu = cfg->shade_per_cu;
  ctx->samples_per_pass = cfg->samples_per_pass;
  // before the first launch the CU count is not known yet: ensure_launch
  // sets the grids from the per-CU counts then
  if (ctx->cus > 0) {
    ctx->wl.trace_blocks = uint32_t(ctx->cus) * cfg->trace_per_cu;
    ctx->batch_blocks = uint32_t(ctx->cus) * cfg->batch_per_cu;
    ctx->wl.finish_blocks = uint32_t(ctx->cus) * cfg->finish_per_cu;
    ctx->wl.shade_blocks = uint32_t(ctx->cus) * cfg->shade_per_cu;
  }
  return LAM_OK;
}

}  // extern "C"

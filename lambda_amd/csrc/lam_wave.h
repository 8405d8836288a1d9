// Pieces of the wavefront integrator shared by its kernel files
// (lam_wavefront.hip: the render passes; lam_adjoint.hip: the adjoint's
// replay passes): the control-block words, path identities and the
// block-aggregated queue appends.
#pragma once
#include <algorithm>
#include "lam_launch.h"
#include "lam_shade.h"

namespace lam {

// control block (uint32 words)
enum : uint32_t {
  CTRL_COUNT0 = 0,  // extend queue ping
  CTRL_COUNT1 = 1,  // extend queue pong
  CTRL_SHADOW = 2,  // shadow queue count
  CTRL_FETCH_EXT = 3,
  CTRL_FETCH_SHD = 4,
  CTRL_TAIL = 5,    // 1: the current queue is finished by k_finish
  CTRL_FETCH_FIN = 6,
  CTRL_PASS = 8,    // current pass (sample) index
  CTRL_GUIDE_DROP = 12,  // guided continuations dropped (weight not finite)
  CTRL_GUIDE_ZERO = 13,  // guide-branch samples with a zero mixture density
  CTRL_WORDS = 16
};

// threads per shade block (the block appends aggregate this many entries)
constexpr uint32_t kShadeBlock = 256;
// largest spectra table the shade kernel stages in LDS (3 blocks per CU)
constexpr uint32_t kSpecLdsBytes = 48 * 1024;
// largest spectra + materials + lights + objects the shade kernel stages
// in LDS together (3 blocks per CU within the CU's 160 KB)
constexpr uint32_t kTabLdsBytes = 52 * 1024;

// LDS footprint of the staged tables in float4 units: spectra; materials
// (56 B each, rounded up as a whole); lights (48 B); objects (112 B)
struct TabLds {
  uint32_t spec4, mat4, light4, obj4;
  __host__ __device__ uint32_t total4() const { return spec4 + mat4 + light4 + obj4; }
};
__host__ __device__ inline TabLds tab_lds(const DevScene &sc) {
  static_assert(sizeof(lam_material) == 56 && sizeof(lam_light) == 48 && sizeof(lam_object) == 112,
                "table layouts");
  TabLds t;
  t.spec4 = (sc.n_spectra * sc.grid_n + 3u) / 4u;
  t.mat4 = (sc.n_materials * 56u + 15u) / 16u;
  t.light4 = sc.n_lights * 3u;
  t.obj4 = sc.n_objects * 7u;
  return t;
}

// Copies the scene's read-only tables into the block's dynamic LDS and
// points sc at the copies: STAGE bit 1 the coefficient spectra, bit 2 the
// materials, lights and objects (after the spectra when both), so that the
// shading of a hit waits on one dependent global fetch (the triangle
// records) instead of a chain of object -> material -> light -> light
// object fetches.  Every thread of the block calls it (one barrier); NT
// threads per block.
constexpr int kStageSpectra = 1, kStageTables = 2;
template <int STAGE, uint32_t NT>
__device__ __forceinline__ void stage_tables(DevScene &sc, uint32_t off4 = 0) {
  extern __shared__ float4 lds_dyn[];
  float4 *lds_tab = lds_dyn + off4;  // after off4 float4s the caller keeps
  const TabLds t = tab_lds(sc);
  float4 *mat = lds_tab;
  if constexpr (STAGE & kStageSpectra) {
    const float4 *sp = reinterpret_cast<const float4 *>(sc.spectra);
    for (uint32_t i = threadIdx.x; i < t.spec4; i += NT) lds_tab[i] = sp[i];
    mat += t.spec4;
  }
  if constexpr (STAGE & kStageTables) {
    float4 *lt = mat + t.mat4;
    float4 *ob = lt + t.light4;
    // materials are 8-byte aligned in HBM: copied as uint2
    const uint2 *ms = reinterpret_cast<const uint2 *>(sc.materials);
    uint2 *md = reinterpret_cast<uint2 *>(mat);
    for (uint32_t i = threadIdx.x; i < sc.n_materials * 7u; i += NT) md[i] = ms[i];
    const float4 *ls = reinterpret_cast<const float4 *>(sc.lights);
    for (uint32_t i = threadIdx.x; i < t.light4; i += NT) lt[i] = ls[i];
    const float4 *os = reinterpret_cast<const float4 *>(sc.objects);
    for (uint32_t i = threadIdx.x; i < t.obj4; i += NT) ob[i] = os[i];
  }
  __syncthreads();
  if constexpr (STAGE & kStageSpectra) sc.spectra = reinterpret_cast<const float *>(lds_tab);
  if constexpr (STAGE & kStageTables) {
    sc.materials = reinterpret_cast<const lam_material *>(mat);
    sc.lights = reinterpret_cast<const lam_light *>(mat + t.mat4);
    sc.objects = reinterpret_cast<const lam_object *>(mat + t.mat4 + t.light4);
  }
}

// Staging choice for a shade launch and its dynamic LDS bytes: spectra and
// tables together when they fit kTabLdsBytes, else the tables alone, else
// (spectra_alone) the spectra when they fit kSpecLdsBytes, else nothing.
struct StagePlan {
  int stage;
  uint32_t bytes;
};
inline StagePlan stage_plan(const DevScene &sc, bool spectra_alone, uint32_t budget = kTabLdsBytes) {
  const TabLds t = tab_lds(sc);
  const uint32_t spec = t.spec4 * 16u, tab = (t.mat4 + t.light4 + t.obj4) * 16u;
  if (spec + tab <= budget) return {kStageSpectra | kStageTables, spec + tab};
  if (tab <= budget) return {kStageTables, tab};
  if (spectra_alone && spec <= std::min(budget, kSpecLdsBytes)) return {kStageSpectra, spec};
  return {0, 0u};
}

// L[k] += v[k] over the H wavelengths of slot p, all loads first: the
// path-state arrays may alias as far as the compiler knows, so a
// wavelength-by-wavelength += would wait for each load in turn
template <int H>
__device__ __forceinline__ void add_to(float *__restrict__ L, size_t cap, uint32_t p, const float (&v)[H]) {
  float cur[H];
#pragma unroll
  for (int k = 0; k < H; k++) cur[k] = L[size_t(k) * cap + p];
#pragma unroll
  for (int k = 0; k < H; k++) L[size_t(k) * cap + p] = cur[k] + v[k];
}
// L[k] -= v[k], the same way
template <int H>
__device__ __forceinline__ void sub_from(float *__restrict__ L, size_t cap, uint32_t p, const float (&v)[H]) {
  float cur[H];
#pragma unroll
  for (int k = 0; k < H; k++) cur[k] = L[size_t(k) * cap + p];
#pragma unroll
  for (int k = 0; k < H; k++) L[size_t(k) * cap + p] = cur[k] - v[k];
}

// Identity of a path slot.  From lists (test hooks), or a pass of m samples
// of npix pixels: slot = q npix + i is pixel pixel_offset + i, sample
// ctrl[CTRL_PASS] + q (npix = 0: one sample, slot = i).
struct PathIds {
  const uint32_t *pixel;   // nullable
  const uint32_t *sample;  // nullable
  uint32_t pixel_offset;
  uint32_t npix;
};

// (pass = ctrl[CTRL_PASS], read once per kernel by the callers that loop)
__device__ __forceinline__ void path_id_of(const PathIds &ids, uint32_t pass, uint32_t slot, uint32_t &pixel,
                                           uint32_t &sample) {
  if (ids.pixel) {
    pixel = ids.pixel[slot];
    sample = ids.sample[slot];
    return;
  }
  uint32_t q = 0;
  if (ids.npix) {  // (a uniform divisor: one division sequence per lane)
    q = slot / ids.npix;
    slot -= q * ids.npix;
  }
  pixel = slot + ids.pixel_offset;
  sample = pass + q;
}
__device__ __forceinline__ void path_id(const PathIds &ids, const uint32_t *ctrl, uint32_t slot, uint32_t &pixel,
                                        uint32_t &sample) {
  path_id_of(ids, ids.pixel ? 0u : ctrl[CTRL_PASS], slot, pixel, sample);
}

template <int H>
__device__ __forceinline__ float lam_for(const DevScene &sc, uint64_t key, Lam<H> &L) {
  return hero_lambdas<H>(sc, lam_draw_f(key, LAM_DIM_HERO), L);
}

// Block-aggregated append to NQ queues: one atomic per queue per block
// iteration instead of one per wave.  Same-address atomics on the queue
// counts are device-scope and serialise at the memory side across the 8
// XCDs; per-wave appends cost the shade kernel more than half its time.
// Every thread of the block calls this (two barriers).  The per-wave counts
// are double-buffered by iteration parity `par`, so the next iteration may
// start writing before every wave has read this one's offsets.  Within a
// block, queue order is wave order then lane order.
template <int WAVES, int NQ>
struct BlockQueues {
  uint32_t cnt[2][NQ][WAVES];
  uint32_t base[2][NQ];
};
template <int WAVES, int NQ>
__device__ __forceinline__ void block_append(BlockQueues<WAVES, NQ> &sm, uint32_t par, const bool (&want)[NQ],
                                             uint32_t *const (&count)[NQ], uint32_t (&slot)[NQ]) {
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  uint64_t m[NQ];
#pragma unroll
  for (int k = 0; k < NQ; k++) {
    m[k] = __ballot(want[k]);
    if (lane == 0) sm.cnt[par][k][wave] = __popcll(m[k]);
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    const uint32_t k = threadIdx.x;
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
      const uint32_t c = sm.cnt[par][k][w];
      sm.cnt[par][k][w] = tot;
      tot += c;
    }
    uint32_t *ck = count[0];
#pragma unroll
    for (int j = 1; j < NQ; j++)
      if (k == uint32_t(j)) ck = count[j];
    sm.base[par][k] = tot ? atomicAdd(ck, tot) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NQ; k++) slot[k] = sm.base[par][k] + sm.cnt[par][k][wave] + __popcll(m[k] & below);
}

}  // namespace lam

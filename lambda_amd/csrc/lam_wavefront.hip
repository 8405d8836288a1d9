// Wavefront integrator for gfx950: the recursive Renderer::tracePath
// (src/renderer.cpp:64-146) split into kernels linked by queues.
//
// One pass traces one path for each of N path slots.  Slot p's identity
// (pixel, sample) comes either from lists (test hook) or from (p + offset,
// pass index).  All paths of a pass start together, so every entry of the
// extend queue at iteration s is at path vertex s.
//
// Per iteration s:
//   k_iter_begin  reset the next-queue / shadow-queue counts and the work
//                 fetch counters (a kernel, so the pass can be captured in a
//                 hipGraph without memset nodes)
//   k_extend      persistent closest-hit over the current queue: each wave
//                 fetches 64 rays with one atomic
//   k_shade       hit -> surface, emitter, NEE light sample, BSDF sample,
//                 MIS weights; appends the continuation ray and the shadow ray
//                 to compacted queues (wave ballots + a block prefix count in
//                 LDS + one atomic per block iteration per queue)
//   k_shadow      persistent any-hit; adds the pending NEE term when visible
// then k_film splats each path's spectral radiance into the Kahan XYZ film.
//
// Path state is SoA ([k][slot]) so every access is coalesced; it holds only
// the throughput β, the radiance L and the pending NEE term, H floats each.
// Random numbers, hero wavelengths and the vertex index are recomputed from
// (pixel, sample) and s, so replaying a path reproduces it exactly.
#include "lam_guide.h"
#include "lam_launch.h"
#include "lam_shade.h"
#include "lam_wave.h"

namespace lam {

// ---- ray generation: render (src/renderer.cpp:35-40) ------------------------------
template <int H>
__global__ __launch_bounds__(256) void k_raygen(DevScene sc, PathArgs pa, PathIds ids, WaveState ws, uint32_t n,
                                                uint32_t *__restrict__ ctrl, uint32_t qsel) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    ctrl[CTRL_COUNT0 + qsel] = n;
    ctrl[CTRL_FETCH_EXT] = 0;
    ctrl[CTRL_TAIL] = 0;
  }
  if (p >= n) return;
  uint32_t pixel, sample;
  path_id(ids, ctrl, p, pixel, sample);
  const uint64_t key = lam_path_key(pa.seed, pixel, sample);
  const uint32_t w = sc.cam.width;
  const uint32_t row = pixel / w, colj = pixel - row * w;
  const float u = float(colj) + lam_draw_f(key, LAM_DIM_U);
  const float v = float(row) + lam_draw_f(key, LAM_DIM_V);
  f3 o, d;
  camera_ray(sc.cam, u, v, lam_draw_f(key, LAM_DIM_LENS_X), lam_draw_f(key, LAM_DIM_LENS_Y), o, d);
  float4 *q = reinterpret_cast<float4 *>(ws.ray[qsel]) + 2 * size_t(p);
  q[0] = make_float4(o.x, o.y, o.z, kFltMax);
  q[1] = make_float4(d.x, d.y, d.z, __uint_as_float(p));
#pragma unroll
  for (int k = 0; k < H; k++) {
    ws.beta[size_t(k) * ws.cap + p] = 1.f;
    ws.L[size_t(k) * ws.cap + p] = 0.f;
  }
}

// Resets the counters of iteration s and decides whether the paths left in
// the current queue are finished by one thread each (k_finish): when the
// queue is short enough that a synchronised bounce costs about as much as
// its slowest ray, or at the last scheduled iteration (force).
__global__ void k_iter_begin(uint32_t *__restrict__ ctrl, uint32_t next_q, uint32_t tail_threshold,
                             uint32_t force, unsigned long long *__restrict__ rays) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    // the shadow rays of the previous iteration were traced; the queue of
    // this one is traced by k_extend unless the finish kernel takes it (it
    // counts its own rays)
    if (rays) rays[1] += ctrl[CTRL_SHADOW];
    ctrl[CTRL_COUNT0 + next_q] = 0;
    ctrl[CTRL_SHADOW] = 0;
    ctrl[CTRL_FETCH_EXT] = 0;
    ctrl[CTRL_FETCH_SHD] = 0;
    ctrl[CTRL_FETCH_FIN] = 0;
    const uint32_t c = ctrl[CTRL_COUNT0 + (next_q ^ 1u)];
    if (c > 0 && (c <= tail_threshold || force)) ctrl[CTRL_TAIL] = 1;
    if (rays && !ctrl[CTRL_TAIL]) rays[0] += c;
  }
}

__global__ void k_pass_set(uint32_t *__restrict__ ctrl, uint32_t pass) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctrl[CTRL_PASS] = pass;
}
__global__ void k_pass_next(uint32_t *__restrict__ ctrl, uint32_t stride) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctrl[CTRL_PASS] = ctrl[CTRL_PASS] + stride;
}

// ---- extend: persistent closest hit ---------------------------------------------------
template <bool COUNT>
__global__ __launch_bounds__(64, 6) void k_extend(DevScene sc, const float4 *__restrict__ rays, Hit *__restrict__ hits,
                                               uint32_t *__restrict__ ctrl, uint32_t qsel, uint32_t *__restrict__ spill,
                                               unsigned long long *__restrict__ counts) {
  __shared__ uint32_t stack[kLdsStack * 64];
  if (ctrl[CTRL_TAIL]) return;
  const uint32_t n = ctrl[CTRL_COUNT0 + qsel];
  TravCounts c{0, 0};
  uint32_t nrays = 0;
  auto load = [&](uint32_t i) {
    const float4 a = rays[2 * size_t(i)], b = rays[2 * size_t(i) + 1];
    RayItem it;
    it.r.o = mk(a.x, a.y, a.z);
    it.r.tmax = a.w;
    it.r.d = mk(b.x, b.y, b.z);
    it.r.skip = LAM_NONE;
    return it;
  };
  auto done = [&](uint32_t i, const RayItem &, const Trav &tv, bool ovf, const TravCounts &rc) {
    Hit h = tv.best;
    if (ovf) {  // cannot happen under the host-checked stack bound
      atomicOr(ctrl + 15, 1u);
      h.obj = LAM_NONE;
    }
    hits[i] = h;
    c.nodes += rc.nodes;
    c.tris += rc.tris;
    nrays++;
  };
  const SplitStack stk = SplitStack::of(stack, spill, blockIdx.x, gridDim.x);
  trace_queue<false, COUNT>(sc, n, ctrl + CTRL_FETCH_EXT, stk, load, done);
  if (COUNT) {
    atomicAdd(counts + 0, (unsigned long long)c.nodes);
    atomicAdd(counts + 1, (unsigned long long)c.tris);
    atomicAdd(counts + 2, (unsigned long long)nrays);
  }
}

// ---- shade ------------------------------------------------------------------------------
// H = 4: at least 3 waves per SIMD; the register budget then stays under
// 168 VGPRs without spills (2 waves at the default allocation), and the
// latency-bound kernel gains ~10% from the third wave.  H = 16 needs the
// whole register file (1 wave).
// STAGE (stage_tables): bit 1 the coefficient spectra, bit 2 the materials,
// lights and objects are staged in LDS at the start of the block and every
// lookup of the block reads them there (stage_plan picks what fits).
// The ray and hit of the thread's next queue entry are loaded at the end of
// the current one, so that their round trip overlaps its stores and the
// block's barrier (loads and stores share one in-order wait counter, so a
// load issued earlier would be waited for by the entry's own loads).
template <int H, int STAGE>
__global__ __launch_bounds__(kShadeBlock, H == 4 ? 3 : 1) void k_shade(DevScene sc, PathArgs pa, PathIds ids, WaveState ws, uint32_t s,
                                               uint32_t qsel, uint32_t *__restrict__ ctrl) {
  if (ctrl[CTRL_TAIL]) return;
  if constexpr (STAGE > 0) stage_tables<STAGE, kShadeBlock>(sc);
  const uint32_t n = ctrl[CTRL_COUNT0 + qsel];
  const uint32_t pass = ctrl[CTRL_PASS];
  const float4 *__restrict__ rin = reinterpret_cast<const float4 *>(ws.ray[qsel]);
  float4 *__restrict__ rout = reinterpret_cast<float4 *>(ws.ray[qsel ^ 1u]);
  const size_t cap = ws.cap;
  const uint32_t stride = gridDim.x * kShadeBlock;
  __shared__ BlockQueues<kShadeBlock / 64, 2> sq;
  uint32_t par = 0;
  float4 next_a = make_float4(0.f, 0.f, 0.f, 0.f), next_b = next_a;
  Hit next_hit{0.f, LAM_NONE, LAM_NONE, LAM_NONE};
  auto fetch = [&](uint32_t q) {
    if (q < n) {
      next_a = rin[2 * size_t(q)];
      next_b = rin[2 * size_t(q) + 1];
      next_hit = ws.hits[q];
    }
  };
  fetch(blockIdx.x * kShadeBlock + threadIdx.x);
  // uniform trip count across the block: every thread takes part in the
  // block appends
  for (uint32_t bbase = blockIdx.x * kShadeBlock; bbase < n; bbase += stride, par ^= 1u) {
    const uint32_t q = bbase + threadIdx.x;
    bool push_next = false, push_shadow = false;
    uint32_t p = 0;
    Ray nxt, shd;
    float spend[H];  // the shadow ray's NEE term
    const float4 a = next_a, b = next_b;
    const Hit hit = next_hit;
    if (q < n) {
      const f3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
      p = __float_as_uint(b.w);
      if (hit.obj != LAM_NONE) {
        uint32_t pixel, sample;
        path_id_of(ids, pass, p, pixel, sample);
        const uint64_t key = lam_path_key(pa.seed, pixel, sample);
        Lam<H> lam;
        lam_for<H>(sc, key, lam);
        float beta[H];
#pragma unroll
        for (int k = 0; k < H; k++) beta[k] = ws.beta[size_t(k) * cap + p];
        const Surf sf = surface(sc, hit, ro, rd);
        const f3 o = neg(normalized(rd));
        const f3 nn = sf.n();
        const lam_material &mat = sc.materials[sf.mat];
        if (mat.type == LAM_MAT_EMITTER || mat.type == LAM_MAT_DIREMITTER) {
          float le[H];
          spec_eval<H>(sc, mat.spec[0], lam, le);
          float f = 1.f;
          if (mat.type == LAM_MAT_DIREMITTER) f = lam_powf(fmax_ref(dot(o, sf.F.z), 0.f), mat.specular);
          float add[H];
#pragma unroll
          for (int k = 0; k < H; k++) add[k] = beta[k] * (mat.type == LAM_MAT_DIREMITTER ? le[k] * f : le[k]);
          add_to<H>(ws.L, cap, p, add);
        } else if (s <= pa.max_scat) {
          const uint32_t nl = sc.n_lights;
          const uint32_t li = lam_draw_u32(key, lam_vdim(s, LAM_VDIM_LIGHT_PICK)) % nl;
          const float pl = 1.f / float(nl);
          const lam_light &lt = sc.lights[li];
          LightSample<H> ls;
          sample_light<H>(sc, lt, lam, sf, lam_draw_f(key, lam_vdim(s, LAM_VDIM_LIGHT_A)),
                          lam_draw_f(key, lam_vdim(s, LAM_VDIM_LIGHT_B)), ls);
          float coef[H];
          bxdf<H>(sc, mat, lam, ls.dir, o, nn, sf.front, coef);
          float lis[H];
#pragma unroll
          for (int k = 0; k < H; k++) lis[k] = ls.mult[k] * coef[k] / (ls.prob * pl);
          MatSample<H> ms;
          sample_material<H>(sc, mat, lam, sf, o, lam_draw_f(key, lam_vdim(s, LAM_VDIM_LOBE)),
                             lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_A)),
                             lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_B)), pa.sample_p, ms);
          float pend[H];
          if (!ms.ok) {
#pragma unroll
            for (int k = 0; k < H; k++) pend[k] = beta[k] * lis[k];
          } else {
            const float p_L_mati = prob_light(sc, lt, sf.p, ms.dir) * pl;
            const float p_mat_Li = prob_material(sc, mat, sf, ls.dir, o, lam.l[0], pa.sample_p);
            const float p_L_i = ls.prob * pl;
            const float w_L = p_L_i * p_L_i / (p_L_i * p_L_i + p_mat_Li * p_mat_Li);
            const float w_mat = ms.prob * ms.prob / (ms.prob * ms.prob + p_L_mati * p_L_mati);
#pragma unroll
            for (int k = 0; k < H; k++) {
              pend[k] = beta[k] * (lis[k] * w_L);
              ws.beta[size_t(k) * cap + p] = beta[k] * (ms.mult[k] / ms.prob * w_mat);
            }
            push_next = true;
            nxt.o = sf.p;
            nxt.d = ms.dir;
          }
          if (ls.shadow) {
#pragma unroll
            for (int k = 0; k < H; k++) spend[k] = pend[k];
            push_shadow = true;
            shd = ls.sray;
          } else {
            add_to<H>(ws.L, cap, p, pend);
          }
        }
      }
    }
    uint32_t slot[2];
    block_append(sq, par, {push_next, push_shadow}, {ctrl + CTRL_COUNT0 + (qsel ^ 1u), ctrl + CTRL_SHADOW}, slot);
    // the next entry's ray and hit: issued after this entry's last wait on
    // memory (the block append's atomic), in flight across its stores
    fetch(q + stride);
    if (push_next) {
      rout[2 * size_t(slot[0])] = make_float4(nxt.o.x, nxt.o.y, nxt.o.z, kFltMax);
      rout[2 * size_t(slot[0]) + 1] = make_float4(nxt.d.x, nxt.d.y, nxt.d.z, __uint_as_float(p));
    }
    if (push_shadow) {
      float4 *sr = reinterpret_cast<float4 *>(ws.sray);
      sr[2 * size_t(slot[1])] = make_float4(shd.o.x, shd.o.y, shd.o.z, shd.tmax);
      sr[2 * size_t(slot[1]) + 1] = make_float4(shd.d.x, shd.d.y, shd.d.z, __uint_as_float(shd.skip));
      ws.spath[slot[1]] = p;
#pragma unroll
      for (int k = 0; k < H; k++) ws.pend[size_t(k) * cap + slot[1]] = spend[k];
    }
  }
}

// ---- guided shade ----------------------------------------------------------------------
// k_shade with the learned proposal.  Where the reference's sampler picks a
// lobe the mixture stands in for (guide_draw: a Lambertian cosine lobe, the
// Kd lobe or the specular reflection lobe of a Blinn material, with the same
// lobe choice from dimension LOBE), the direction is drawn from the vMF
// mixture with probability α (dimension GUIDE) and from the lobe's own
// distribution otherwise, and both are weighted by one-sample MIS over the
// mixture of the two:
//   β ← β ⊙ f(ω) w(ω) r(ω) / (pp M(ω)),  M = (1-α) p_lobe + α p_g,
// pp = the lobe's selection probability, f its BSDF·cos, w the reference's
// MIS weight against the light (with the reference's lobe pdf q) and
// r = p_lobe / (q / pp) the ratio of the lobe's true sampling density to the
// reference's pdf (1 for cosine lobes, 1/(cos θh |o·h|) for the Blinn
// reflection).  The lobe's term keeps the plain expectation ∫ f w r L.
// Other lobes, the NEE term and its weight, and roulette are unchanged;
// α = 0 reproduces k_shade.
// TRAIN: each guided continuation appends a training record (network inputs,
// ω, Σβ, ΣL before the vertex, ΣNEE term, slot; k_shadow marks a visible
// NEE term).
// k-th set bit (from 0) of m, for k < popcount(m): binary search over halves
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t k) {
  uint32_t base = 0;
#pragma unroll
  for (uint32_t w = 32; w >= 1; w >>= 1) {
    const uint64_t lo = m & ((uint64_t(1) << w) - 1);
    const uint32_t c = __popcll(lo);
    const bool hi = k >= c;
    k = hi ? k - c : k;
    m = hi ? m >> w : lo;
    base += hi ? w : 0u;
  }
  return base;
}

// The network for the entries of the extend queue (after k_extend) whose
// continuation k_shade_guided will draw from the mixture (guide_draw of
// bounce s): inputs from the hit (position, outgoing direction) and the
// path's hero wavelength; 40 outputs stored as fp16 at the entry.  A
// wave scans 64 entries at a time and packs the ones that need the network
// into full waves of 64 queries (ballot, k-th set bit, lane shuffles; one
// wave of overflow), so that each MFMA evaluation serves 64 queries; the
// packed parameters are staged in LDS once per block.
template <int H>
__global__ __launch_bounds__(256) void k_guide_queue(DevScene sc, PathArgs pa, PathIds ids, WaveState ws,
                                                     GuideDev gd, uint32_t s, uint32_t qsel,
                                                     const uint32_t *__restrict__ ctrl) {
  __shared__ uint4 sfrag[kGFrags * 64];
  __shared__ float sbias[kGBiasWords];
  if (ctrl[CTRL_TAIL]) return;
  {
    const float *gb = reinterpret_cast<const float *>(gd.frag + kGFrags * 64);
    for (uint32_t i = threadIdx.x; i < uint32_t(kGFrags) * 64; i += blockDim.x) sfrag[i] = gd.frag[i];
    for (uint32_t i = threadIdx.x; i < uint32_t(kGBiasWords); i += blockDim.x) sbias[i] = gb[i];
  }
  __syncthreads();
  const uint32_t n = ctrl[CTRL_COUNT0 + qsel];
  const float4 *__restrict__ rin = reinterpret_cast<const float4 *>(ws.ray[qsel]);
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t lane = threadIdx.x % 64, ln = lane % 16, lg = lane / 16;
  uint2 *mix = reinterpret_cast<uint2 *>(gd.mix);
  // queries waiting for the network: lanes [0, P) of slot A; slot B holds
  // what a scan adds beyond 64
  float ain[kGIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, bin[kGIn];
  uint32_t aq = 0, bq = 0, P = 0;
  auto evaluate = [&](uint32_t nvalid) {
    f32x4 o[4][3];
    mlp_wave(sfrag, sbias, ain, o);
    // lane (n, g) holds rows 16t + 4g .. +3 of the query in lane 16G + n
#pragma unroll
    for (int G = 0; G < 4; G++) {
      const uint32_t src = 16u * G + ln;
      const uint32_t q = __shfl(aq, int(src), 64);
      if (src >= nvalid) continue;
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const uint32_t row = 16 * t + 4 * lg;
        if (row < uint32_t(kGOut)) {
          const f32x4 v = o[G][t];
          const uint32_t w0 = uint32_t(__builtin_bit_cast(uint16_t, _Float16(v[0]))) |
                              (uint32_t(__builtin_bit_cast(uint16_t, _Float16(v[1]))) << 16);
          const uint32_t w1 = uint32_t(__builtin_bit_cast(uint16_t, _Float16(v[2]))) |
                              (uint32_t(__builtin_bit_cast(uint16_t, _Float16(v[3]))) << 16);
          mix[size_t(q) * 10 + row / 4] = make_uint2(w0, w1);
        }
      }
    }
  };
  const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t wbase = first - lane; wbase < n; wbase += stride) {
    const uint32_t q = wbase + lane;
    float gin[kGIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool need = false;
    if (q < n && s <= pa.max_scat) {
      const Hit hit = ws.hits[q];
      if (hit.obj != LAM_NONE) {
        const lam_material &mat = sc.materials[sc.objects[hit.obj].material];
        if (mat.type != LAM_MAT_EMITTER && mat.type != LAM_MAT_DIREMITTER) {
          const float4 b = rin[2 * size_t(q) + 1];
          uint32_t pixel, sample;
          path_id(ids, ctrl, __float_as_uint(b.w), pixel, sample);
          const uint64_t key = lam_path_key(pa.seed, pixel, sample);
          float pp;
          const float4 a = rin[2 * size_t(q)];
          const f3 rd = mk(b.x, b.y, b.z);
          const Surf sf = surface(sc, hit, mk(a.x, a.y, a.z), rd);
          if (guide_draw(gd, mat, s, lam_draw_f(key, lam_vdim(s, LAM_VDIM_LOBE)), pa.sample_p, sf.front, pp) !=
              kGuideNone) {
            Lam<H> lam;
            lam_for<H>(sc, key, lam);
            guide_inputs<H>(gd, sc, sf.p, neg(normalized(rd)), lam, gin);
            need = true;
          }
        }
      }
    }
    const uint64_t m = __ballot(need);
    const uint32_t c = __popcll(m);
    if (c == 0) continue;
    // lane d of slot A takes new query d - P, lane d of slot B new query d + 64 - P
    const bool take_a = lane >= P && lane - P < c;
    const bool take_b = lane + 64u - P < c;
    const int sa = int(take_a ? nth_set_bit(m, lane - P) : lane);
    const int sb = int(take_b ? nth_set_bit(m, lane + 64u - P) : lane);
#pragma unroll
    for (int i = 0; i < kGIn; i++) {
      const float va = __shfl(gin[i], sa, 64), vb = __shfl(gin[i], sb, 64);
      if (take_a) ain[i] = va;
      bin[i] = vb;
    }
    const uint32_t qa = __shfl(q, sa, 64), qb = __shfl(q, sb, 64);
    if (take_a) aq = qa;
    bq = qb;
    P += c;
    if (P >= 64) {
      evaluate(64);
#pragma unroll
      for (int i = 0; i < kGIn; i++) ain[i] = bin[i];
      aq = bq;
      P -= 64;
    }
  }
  if (P > 0) evaluate(P);
}

__device__ __forceinline__ void load_mix(const uint4 *src, float *raw) {
#pragma unroll
  for (int c = 0; c < 5; c++) {
    const uint4 v = src[c];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      raw[8 * c + 2 * j] = float(__builtin_bit_cast(_Float16, uint16_t(w[j] & 0xFFFFu)));
      raw[8 * c + 2 * j + 1] = float(__builtin_bit_cast(_Float16, uint16_t(w[j] >> 16)));
    }
  }
}

template <int H, bool TRAIN, int STAGE>
__global__ __launch_bounds__(kShadeBlock) void k_shade_guided(DevScene sc, PathArgs pa, PathIds ids, WaveState ws,
                                                      GuideDev gd, uint32_t s, uint32_t qsel,
                                                      uint32_t *__restrict__ ctrl) {
  if (ctrl[CTRL_TAIL]) return;
  if constexpr (STAGE > 0) stage_tables<STAGE, kShadeBlock>(sc);
  const uint32_t n = ctrl[CTRL_COUNT0 + qsel];
  const uint32_t pass = ctrl[CTRL_PASS];
  const float4 *__restrict__ rin = reinterpret_cast<const float4 *>(ws.ray[qsel]);
  float4 *__restrict__ rout = reinterpret_cast<float4 *>(ws.ray[qsel ^ 1u]);
  const size_t cap = ws.cap;
  const uint32_t stride = gridDim.x * kShadeBlock;
  __shared__ BlockQueues<kShadeBlock / 64, TRAIN ? 3 : 2> sq;
  uint32_t par = 0;
  // the next entry's ray and hit, loaded at the end of the current one (as
  // k_shade)
  float4 next_a = make_float4(0.f, 0.f, 0.f, 0.f), next_b = next_a;
  Hit next_hit{0.f, LAM_NONE, LAM_NONE, LAM_NONE};
  auto fetch = [&](uint32_t q) {
    if (q < n) {
      next_a = rin[2 * size_t(q)];
      next_b = rin[2 * size_t(q) + 1];
      next_hit = ws.hits[q];
    }
  };
  fetch(blockIdx.x * kShadeBlock + threadIdx.x);
  for (uint32_t bbase = blockIdx.x * kShadeBlock; bbase < n; bbase += stride, par ^= 1u) {
    const uint32_t q = bbase + threadIdx.x;
    bool push_next = false, push_shadow = false, live = false, guided = false, want_rec = false;
    uint32_t p = 0, mat_id = 0;
    Ray nxt, shd;
    float spend[H];  // the shadow ray's NEE term
    uint64_t key = 0;
    Lam<H> lam;
    float beta[H];
    Surf sf;
    f3 o = mk(0.f, 0.f, 1.f);
    float gin[kGIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float4 a = next_a, b = next_b;
    const Hit hit = next_hit;
    if (q < n) {
      const f3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
      p = __float_as_uint(b.w);
      if (hit.obj != LAM_NONE) {
        uint32_t pixel, sample;
        path_id_of(ids, pass, p, pixel, sample);
        key = lam_path_key(pa.seed, pixel, sample);
        lam_for<H>(sc, key, lam);
#pragma unroll
        for (int k = 0; k < H; k++) beta[k] = ws.beta[size_t(k) * cap + p];
        sf = surface(sc, hit, ro, rd);
        o = neg(normalized(rd));
        mat_id = sf.mat;
        const lam_material &mat = sc.materials[mat_id];
        if (mat.type == LAM_MAT_EMITTER || mat.type == LAM_MAT_DIREMITTER) {
          float le[H];
          spec_eval<H>(sc, mat.spec[0], lam, le);
          float f = 1.f;
          if (mat.type == LAM_MAT_DIREMITTER) f = lam_powf(fmax_ref(dot(o, sf.F.z), 0.f), mat.specular);
          float add[H];
#pragma unroll
          for (int k = 0; k < H; k++) add[k] = beta[k] * (mat.type == LAM_MAT_DIREMITTER ? le[k] * f : le[k]);
          add_to<H>(ws.L, cap, p, add);
        } else if (s <= pa.max_scat) {
          live = true;
          guided = guide_vertex(gd, mat, s);
          if (TRAIN && guided) guide_inputs<H>(gd, sc, sf.p, o, lam, gin);
        }
      }
    }
    // phase 2: NEE and the continuation
    float rec_b = 0.f, rec_l = 0.f, rec_p = 0.f, rec_v = 0.f;  // training record sums
    if (live) {
      const lam_material &mat = sc.materials[mat_id];
      const f3 nn = sf.n();
      const uint32_t nl = sc.n_lights;
      const uint32_t li = lam_draw_u32(key, lam_vdim(s, LAM_VDIM_LIGHT_PICK)) % nl;
      const float pl = 1.f / float(nl);
      const lam_light &lt = sc.lights[li];
      LightSample<H> ls;
      sample_light<H>(sc, lt, lam, sf, lam_draw_f(key, lam_vdim(s, LAM_VDIM_LIGHT_A)),
                      lam_draw_f(key, lam_vdim(s, LAM_VDIM_LIGHT_B)), ls);
      float coef[H];
      bxdf<H>(sc, mat, lam, ls.dir, o, nn, sf.front, coef);
      float lis[H];
#pragma unroll
      for (int k = 0; k < H; k++) lis[k] = ls.mult[k] * coef[k] / (ls.prob * pl);
      const float alpha = guided ? gd.alpha : 0.f;
      const float P = pa.sample_p;
      const float xl = lam_draw_f(key, lam_vdim(s, LAM_VDIM_LOBE));
      // does the reference's sampler take a lobe the mixture stands in for
      // (cg.hpp:1768, 1809-1853)?
      float pp = 0.f;
      const int kind = guide_draw(gd, mat, s, xl, P, sf.front, pp);
      bool ok = false;
      f3 wi = mk(0.f, 0.f, 0.f);
      float fac[H];
      if (kind != kGuideNone) {
        ok = true;
        const float xa = lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_A)), xb = lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_B));
        GuideMix gm;
        {
          float raw[kGOut];
          load_mix(gd.mix + 5 * size_t(q), raw);
          mix_from_raw(raw, gm);
        }
        const bool from_guide = lam_draw_f(key, lam_vdim(s, LAM_VDIM_GUIDE)) < alpha;
        if (from_guide) {
          wi = mix_sample(gm, lam_draw_f(key, lam_vdim(s, LAM_VDIM_SPARE)), xa, xb);
        } else if (kind == kGuideCosine) {  // the lobe's own cosine sample (cg.hpp:1762-1770)
          f3 dir = cos_hemi(xa, xb);
          if (!sf.front) dir.z *= -1;
          wi = to_basis(sf.F, dir);
        } else {  // the Blinn half vector, reflected (cg.hpp:1815-1820)
          const f3 h = to_basis(sf.F, blinn_h(mat.alpha, xa, xb));
          wi = reflect(o, h);
        }
        const float pg = mix_pdf(gm, wi);
        const float pLw = prob_light(sc, lt, sf.p, wi) * pl;
        if (kind == kGuideCosine) {
          // β ← β f_d w_d / (pp M), M = (1-α) cos/π + α p_g, w_d against the
          // light with the reference's lobe pdf pp cos/π
          const float c = dot(wi, nn);
          const float M = (1.f - alpha) * fmax_ref(c, 0.f) * kInvPi + alpha * pg;
          if (c > 0.f && M > 0.f) {
            float kd[H];
            spec_eval<H>(sc, mat.spec[0], lam, kd);
            const float pi = pp * (c * kInvPi);
            const float w_d = mis_w(pi, pLw * pLw);
            const float g = w_d / (pp * M);
#pragma unroll
            for (int k = 0; k < H; k++) fac[k] = kd[k] * kInvPi * c * g;
          } else {
#pragma unroll
            for (int k = 0; k < H; k++) fac[k] = 0.f;
            if (!(M > 0.f)) atomicAdd(ctrl + CTRL_GUIDE_ZERO, 1u);
          }
        } else {
          // Blinn reflection (front face, Kt = 0): the half vector h of wi on
          // the normal's side, the lobe's true density p_h(h)/(4|o·h|) with
          // p_h = (α+1)/(2π) cos^α θh, and the reference's pdf
          // q = pp (α+1)/(2π) cos^(α+1) θh / 4 (no 1/(o·h)).  The plain lobe
          // term has the expectation ∫ f w_q L / (cos θh |o·h|) dω (its
          // samples have the true density but are divided by q); the mixture
          // estimate f w_q L / (cos θh |o·h| pp M) keeps it.
          f3 h = normalized(add(wi, o));
          if (dot(h, nn) < 0.f) h = neg(h);
          const float ch = dot(h, nn), oh = fabsf(dot(o, h));
          const float ap1 = mat.alpha + 1.f;
          const float pw = lam_powf(ch, mat.alpha);
          const float plobe = oh > 0.f ? ap1 * .5f * kInvPi * pw / (4.f * oh) : 0.f;
          const float M = (1.f - alpha) * plobe + alpha * pg;
          if (ch > 0.f && oh > 0.f && M > 0.f) {
            float ks[H];
            spec_eval<H>(sc, mat.spec[1], lam, ks);
            const float qref = pp * (ap1 * .5f * kInvPi * (pw * ch)) * .25f;
            const float w_q = mis_w(qref, pLw * pLw);
            const float g = w_q / (ch * oh * pp * M);
#pragma unroll
            for (int k = 0; k < H; k++) fac[k] = ks[k] * (mat.alpha + 2) * .125f * kInvPi * pw * g;
          } else {
#pragma unroll
            for (int k = 0; k < H; k++) fac[k] = 0.f;
            if (!(M > 0.f)) atomicAdd(ctrl + CTRL_GUIDE_ZERO, 1u);
          }
        }
      } else {
        MatSample<H> ms;
        sample_material<H>(sc, mat, lam, sf, o, xl, lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_A)),
                           lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_B)), P, ms);
        if (ms.ok) {
          ok = true;
          wi = ms.dir;
          const float p_L_mati = prob_light(sc, lt, sf.p, ms.dir) * pl;
          const float w_mat = ms.prob * ms.prob / (ms.prob * ms.prob + p_L_mati * p_L_mati);
#pragma unroll
          for (int k = 0; k < H; k++) fac[k] = ms.mult[k] / ms.prob * w_mat;
        }
      }
      float pend[H];
      if (!ok) {
#pragma unroll
        for (int k = 0; k < H; k++) pend[k] = beta[k] * lis[k];
      } else {
        const float p_mat_Li = prob_material(sc, mat, sf, ls.dir, o, lam.l[0], P);
        const float p_L_i = ls.prob * pl;
        const float w_L = p_L_i * p_L_i / (p_L_i * p_L_i + p_mat_Li * p_mat_Li);
        bool any = false, finite = true;
#pragma unroll
        for (int k = 0; k < H; k++) {
          pend[k] = beta[k] * (lis[k] * w_L);
          const float nb = beta[k] * fac[k];
          any = any || nb != 0.f;
          finite = finite && isfinite(fac[k]);
          ws.beta[size_t(k) * cap + p] = nb;
        }
        // a zero-weight continuation carries nothing: drop it (the plain
        // kernel traces it with β = 0); so is one whose weight overflowed in a
        // degenerate configuration (never with α = 0; counted)
        push_next = (any && finite) || alpha == 0.f;
        if (!finite && alpha > 0.f) atomicAdd(ctrl + CTRL_GUIDE_DROP, 1u);
        nxt.o = sf.p;
        nxt.d = wi;
      }
      if (TRAIN) want_rec = kind != kGuideNone && push_next;
      if (TRAIN && want_rec) {
#pragma unroll
        for (int k = 0; k < H; k++) {
          rec_b += beta[k];
          rec_l += ws.L[size_t(k) * cap + p];
          rec_p += pend[k];
        }
        rec_v = ls.shadow ? 0.f : 1.f;
      }
      if (ls.shadow) {
#pragma unroll
        for (int k = 0; k < H; k++) spend[k] = pend[k];
        push_shadow = true;
        shd = ls.sray;
      } else {
        add_to<H>(ws.L, cap, p, pend);
      }
    }
    uint32_t slot[TRAIN ? 3 : 2];
    uint32_t *const next_count = ctrl + CTRL_COUNT0 + (qsel ^ 1u);
    if constexpr (TRAIN)
      block_append(sq, par, {push_next, push_shadow, want_rec}, {next_count, ctrl + CTRL_SHADOW, gd.rec_count}, slot);
    else
      block_append(sq, par, {push_next, push_shadow}, {next_count, ctrl + CTRL_SHADOW}, slot);
    fetch(q + stride);
    if constexpr (TRAIN) {
      const uint32_t r = slot[TRAIN ? 2 : 0];
      if (live) gd.rec_of_slot[p] = (want_rec && r < gd.rec_cap) ? r : LAM_NONE;
      if (want_rec && r < gd.rec_cap) {
        float4 *R = gd.rec + 4 * size_t(r);
        R[0] = make_float4(gin[0], gin[1], gin[2], gin[3]);
        R[1] = make_float4(gin[4], gin[5], rec_b, rec_l);
        R[2] = make_float4(nxt.d.x, nxt.d.y, nxt.d.z, rec_p);
        R[3] = make_float4(__uint_as_float(p), rec_v, 0.f, 0.f);
      }
    }
    if (push_next) {
      rout[2 * size_t(slot[0])] = make_float4(nxt.o.x, nxt.o.y, nxt.o.z, kFltMax);
      rout[2 * size_t(slot[0]) + 1] = make_float4(nxt.d.x, nxt.d.y, nxt.d.z, __uint_as_float(p));
    }
    if (push_shadow) {
      float4 *sr = reinterpret_cast<float4 *>(ws.sray);
      sr[2 * size_t(slot[1])] = make_float4(shd.o.x, shd.o.y, shd.o.z, shd.tmax);
      sr[2 * size_t(slot[1]) + 1] = make_float4(shd.d.x, shd.d.y, shd.d.z, __uint_as_float(shd.skip));
      ws.spath[slot[1]] = p;
#pragma unroll
      for (int k = 0; k < H; k++) ws.pend[size_t(k) * cap + slot[1]] = spend[k];
    }
  }
}

// training weight of each record of the pass just traced:
//   w = max(0, ΣL_final - ΣL_before - [NEE visible] ΣNEE) / Σβ
template <int H>
__global__ void k_guide_weights(WaveState ws, float4 *__restrict__ rec, const uint32_t *__restrict__ rec_count,
                                uint32_t rec_cap) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nr = min(*rec_count, rec_cap);
  if (r >= nr) return;
  const float4 b = rec[4 * size_t(r) + 1], c = rec[4 * size_t(r) + 2], d = rec[4 * size_t(r) + 3];
  const uint32_t p = __float_as_uint(d.x);
  float Lf = 0.f;
#pragma unroll
  for (int k = 0; k < H; k++) Lf += ws.L[size_t(k) * ws.cap + p];
  const float dl = Lf - b.w - d.y * c.w;
  const float w = b.z > 0.f ? fmaxf(dl, 0.f) / b.z : 0.f;
  rec[4 * size_t(r) + 3].z = isfinite(w) ? w : 0.f;
}

hipError_t guide_weights(int hero, const WaveState &ws, const GuideDev &gd, hipStream_t st) {
  const uint32_t blocks = (gd.rec_cap + 255) / 256;
  if (hero == 16)
    hipLaunchKernelGGL((k_guide_weights<16>), dim3(blocks), dim3(256), 0, st, ws, gd.rec, gd.rec_count, gd.rec_cap);
  else
    hipLaunchKernelGGL((k_guide_weights<4>), dim3(blocks), dim3(256), 0, st, ws, gd.rec, gd.rec_count, gd.rec_cap);
  return hipGetLastError();
}

// ---- shadow: persistent any-hit, then the pending NEE term -----------------------------
template <int H, bool COUNT, bool REC = false>
__global__ __launch_bounds__(64, H == 4 ? 6 : 1) void k_shadow(DevScene sc, WaveState ws, uint32_t *__restrict__ ctrl,
                                               unsigned long long *__restrict__ counts,
                                               float4 *__restrict__ rec = nullptr,
                                               const uint32_t *__restrict__ rec_of_slot = nullptr,
                                               uint32_t rec_cap = 0) {
  __shared__ uint32_t stack[kLdsStack * 64];
  if (ctrl[CTRL_TAIL]) return;
  const uint32_t n = ctrl[CTRL_SHADOW];
  const float4 *__restrict__ sr = reinterpret_cast<const float4 *>(ws.sray);
  const size_t cap = ws.cap;
  TravCounts c{0, 0};
  uint32_t nrays = 0;
  // the entry's ray, its path slot and NEE term (queue-indexed, so all are
  // loaded ahead with the ray)
  struct Item {
    Ray r;
    uint32_t p;
    float pe[H];
  };
  auto load = [&](uint32_t i) {
    const float4 a = sr[2 * size_t(i)], b = sr[2 * size_t(i) + 1];
    Item it;
    it.r.o = mk(a.x, a.y, a.z);
    it.r.tmax = a.w;
    it.r.d = mk(b.x, b.y, b.z);
    it.r.skip = __float_as_uint(b.w);
    it.p = ws.spath[i];
#pragma unroll
    for (int k = 0; k < H; k++) it.pe[k] = ws.pend[size_t(k) * cap + i];
    return it;
  };
  auto done = [&](uint32_t i, const Item &it, const Trav &tv, bool ovf, const TravCounts &rc) {
    if (ovf) atomicOr(ctrl + 15, 1u);
    if (tv.best.obj == LAM_NONE && !ovf) {
      // L += NEE term: one shadow ray per path and iteration, so the add
      // is the sum L + pe exactly; the atomic form needs no wait
#pragma unroll
      for (int k = 0; k < H; k++) unsafeAtomicAdd(ws.L + size_t(k) * cap + it.p, it.pe[k]);
      if (REC) {  // the vertex's training record: its NEE term reached L
        const uint32_t r = rec_of_slot[it.p];
        if (r < rec_cap) rec[4 * size_t(r) + 3].y = 1.f;
      }
    }
    c.nodes += rc.nodes;
    c.tris += rc.tris;
    nrays++;
  };
  const SplitStack stk = SplitStack::of(stack, ws.spill, blockIdx.x, gridDim.x);
  trace_queue<true, COUNT>(sc, n, ctrl + CTRL_FETCH_SHD, stk, load, done);
  if (COUNT) {
    atomicAdd(counts + 3, (unsigned long long)c.nodes);
    atomicAdd(counts + 4, (unsigned long long)c.tris);
    atomicAdd(counts + 5, (unsigned long long)nrays);
  }
}

// ---- finish: the tail of the queue, one path per lane to its end ------------------------
// A lane traces its path one vertex per step and takes the next path of the
// queue when its own ends (per-lane refill from a wave-local pool, as the
// trace kernels take rays), so a wave stays busy until the queue is empty
// instead of idling behind its longest path.
template <int H>
__global__ __launch_bounds__(64) void k_finish(DevScene sc, PathArgs pa, PathIds ids, WaveState ws, uint32_t s0,
                                               uint32_t qsel, uint32_t *__restrict__ ctrl) {
  __shared__ uint32_t stack[kLdsStack * 64];
  if (!ctrl[CTRL_TAIL]) return;
  const SplitStack stk = SplitStack::of(stack, ws.spill, blockIdx.x, gridDim.x);
  const uint32_t n = ctrl[CTRL_COUNT0 + qsel];
  const float4 *__restrict__ rin = reinterpret_cast<const float4 *>(ws.ray[qsel]);
  const size_t cap = ws.cap;
  QueuePool qp;
  qp.init(n);
  bool active = false, ovf = false;
  uint32_t p = 0, s = 0, n_ext = 0, n_shd = 0;
  uint64_t key = 0;
  Lam<H> lam;
  Ray r;
  float beta[H], L[H];
  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (qp.done(nidle)) break;
    if (qp.can_refill(nidle)) {
      const uint32_t q = qp.take(idle, nidle, !active, ctrl + CTRL_FETCH_FIN);
      if (!active && q < n) {
        const float4 a = rin[2 * size_t(q)], b = rin[2 * size_t(q) + 1];
        r.o = mk(a.x, a.y, a.z);
        r.tmax = a.w;
        r.d = mk(b.x, b.y, b.z);
        r.skip = LAM_NONE;
        p = __float_as_uint(b.w);
        uint32_t pixel, sample;
        path_id(ids, ctrl, p, pixel, sample);
        key = lam_path_key(pa.seed, pixel, sample);
        lam_for<H>(sc, key, lam);
#pragma unroll
        for (int k = 0; k < H; k++) {
          beta[k] = ws.beta[size_t(k) * cap + p];
          L[k] = ws.L[size_t(k) * cap + p];
        }
        s = s0;
        ovf = false;
        active = true;
      }
    }
    if (active) {
      n_ext++;
      if (path_vertex<H>(sc, pa.max_scat, pa.sample_p, key, lam, s, r, beta, L, stk, ovf, &n_shd)) {
        s++;
      } else {
        if (ovf) atomicOr(ctrl + 15, 1u);
#pragma unroll
        for (int k = 0; k < H; k++) ws.L[size_t(k) * cap + p] = L[k];
        active = false;
      }
    }
  }
  if (ws.rays) {  // one pair of atomics per wave, on the block's slot
    for (int o = 32; o > 0; o >>= 1) {
      n_ext += __shfl_xor(n_ext, o, 64);
      n_shd += __shfl_xor(n_shd, o, 64);
    }
    if (threadIdx.x % 64 == 0) {
      unsigned long long *slot = ws.rays + size_t(blockIdx.x % kRaySlots) * kRaySlotWords;
      atomicAdd(slot, (unsigned long long)n_ext);
      atomicAdd(slot + 1, (unsigned long long)n_shd);
    }
  }
}

// ---- film / output ----------------------------------------------------------------------
// One thread per pixel of the pass: its m samples (slots i, i + npix, ...)
// are added to the film in sample order, so the film receives the sums of a
// one-sample-per-pass render in the same order.
// STAGE: the colour-matching table is copied to LDS first (every splat
// reads six of its entries).  A pixel's next sample's radiance is loaded
// before the current sample is splatted, so that its round trip overlaps
// the splat's arithmetic.
constexpr uint32_t kFilmLdsBytes = 32 * 1024;
template <int H, int STAGE>
__global__ __launch_bounds__(256) void k_film(DevScene sc, PathArgs pa, PathIds ids, WaveState ws, uint32_t n,
                                              const uint32_t *__restrict__ ctrl, float *__restrict__ film,
                                              float *__restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (out) {  // test hook: raw spectral radiance per path
    if (t >= n) return;
#pragma unroll
    for (int k = 0; k < H; k++) out[size_t(t) * H + k] = ws.L[size_t(k) * ws.cap + t];
    return;
  }
  if constexpr (STAGE) {
    extern __shared__ float cmf_lds[];
    for (uint32_t i = threadIdx.x; i < 3 * sc.grid_n; i += blockDim.x) cmf_lds[i] = sc.cmf[i];
    __syncthreads();
    sc.cmf = cmf_lds;
  }
  const uint32_t npx = ids.npix ? ids.npix : n;
  if (t >= npx) return;
  const float inv_d = 1.f / sc.grid_delta;
  const size_t cap = ws.cap;
  auto sample_xyz = [&](const float (&L)[H], uint32_t pixel, uint32_t sample, float &X, float &Y, float &Z) {
    const uint64_t key = lam_path_key(pa.seed, pixel, sample);
    Lam<H> lam;
    const float inv_om = lam_for<H>(sc, key, lam);
    X = Y = Z = 0.f;
#pragma unroll
    for (int k = 0; k < H; k++) splat_xyz(sc, lam.l[k], L[k] / inv_om, X, Y, Z, inv_d);
  };
  if (ids.pixel) {  // slots from lists
    for (uint32_t p = t; p < n; p += npx) {
      const uint32_t pixel = ids.pixel[p];
      float L[H];
#pragma unroll
      for (int k = 0; k < H; k++) L[k] = ws.L[size_t(k) * cap + p];
      float X, Y, Z;
      sample_xyz(L, pixel, ids.sample[p], X, Y, Z);
      film_add(film, pixel, X, Y, Z);
    }
    return;
  }
  // slot p = q npx + t: pixel t (+ offset), sample pass + q.  The pixel's
  // sums and compensations stay in registers over its m samples (the same
  // operations in the same order as m separate film_add calls).
  const uint32_t pass = ctrl[CTRL_PASS];
  const uint32_t pixel = t + ids.pixel_offset;
  float4 *f = reinterpret_cast<float4 *>(film) + size_t(pixel) * 2;
  float4 s = f[0], c = f[1];
  float L[H];
#pragma unroll
  for (int k = 0; k < H; k++) L[k] = ws.L[size_t(k) * cap + t];
  for (uint32_t p = t, q = 0; p < n; p += npx, q++) {
    float Ln[H];
    const uint32_t pn = p + npx;
#pragma unroll
    for (int k = 0; k < H; k++) Ln[k] = pn < n ? ws.L[size_t(k) * cap + pn] : 0.f;
    float X, Y, Z;
    sample_xyz(L, pixel, pass + q, X, Y, Z);
    kahan_add(s, c, X, Y, Z);
#pragma unroll
    for (int k = 0; k < H; k++) L[k] = Ln[k];
  }
  f[0] = s;
  f[1] = c;
}

template <int H>
static void launch_film(const DevScene &sc, const PathArgs &pa, const PathIds &ids, const WaveState &ws, uint32_t n,
                        const uint32_t *ctrl, float *film, float *out, uint32_t blocks, hipStream_t st) {
  const uint32_t bytes = 3u * sc.grid_n * 4u;
  if (!out && bytes <= kFilmLdsBytes)
    hipLaunchKernelGGL((k_film<H, 1>), dim3(blocks), dim3(256), bytes, st, sc, pa, ids, ws, n, ctrl, film, out);
  else
    hipLaunchKernelGGL((k_film<H, 0>), dim3(blocks), dim3(256), 0, st, sc, pa, ids, ws, n, ctrl, film, out);
}

// ---------------------------------------------------------------------------
// host side: one pass = raygen, (max_scat + 2) iterations, film
// ---------------------------------------------------------------------------
template <int H>
static hipError_t enqueue_pass_t(const DevScene &sc, const PathArgs &pa, const PathIds &ids, const WaveState &ws,
                                 uint32_t n, uint32_t *ctrl, float *film, float *out, const WaveLaunch &wl,
                                 hipStream_t st, WaveProfile *prof, const GuideDev *gd, bool train) {
  const uint32_t b256 = (n + 255) / 256;
  // (rounded up to whole float4s, as the staging copies them)
  const StagePlan plain = stage_plan(sc, true), guided = stage_plan(sc, false);
  auto ev = [&](int slot) {
    if (prof) (void)hipEventRecord(prof->ev[slot], st);
  };
  ev(0);
  hipLaunchKernelGGL((k_raygen<H>), dim3(b256), dim3(256), 0, st, sc, pa, ids, ws, n, ctrl, 0u);
  ev(1);
  uint32_t q = 0;
  const uint32_t iters = wave_iterations(pa.max_scat, wl);
  for (uint32_t s = 0; s < iters; s++) {
    const bool last = wl.max_iters && s + 1 == iters;
    if (prof) (void)hipEventRecord(prof->fin0[s], st);
    hipLaunchKernelGGL(k_iter_begin, dim3(1), dim3(64), 0, st, ctrl, q ^ 1u, wl.tail_threshold, last ? 1u : 0u,
                       ws.rays);
    if (wl.max_iters)
      hipLaunchKernelGGL((k_finish<H>), dim3(wl.finish_blocks), dim3(64), 0, st, sc, pa, ids, ws, s, q, ctrl);
    if (prof) {
      (void)hipEventRecord(prof->ext0[s], st);
      if (prof->counts)
        hipLaunchKernelGGL((k_extend<true>), dim3(wl.trace_blocks), dim3(64), 0, st, sc,
                           reinterpret_cast<const float4 *>(ws.ray[q]), ws.hits, ctrl, q, ws.spill, prof->counts);
      else
        hipLaunchKernelGGL((k_extend<false>), dim3(wl.trace_blocks), dim3(64), 0, st, sc,
                           reinterpret_cast<const float4 *>(ws.ray[q]), ws.hits, ctrl, q, ws.spill, nullptr);
      (void)hipEventRecord(prof->ext1[s], st);
    } else {
      hipLaunchKernelGGL((k_extend<false>), dim3(wl.trace_blocks), dim3(64), 0, st, sc,
                         reinterpret_cast<const float4 *>(ws.ray[q]), ws.hits, ctrl, q, ws.spill, nullptr);
    }
    // vertices past the guided ones are shaded by the plain kernels (the guided
    // kernels with no guided vertex compute the same values)
    const GuideDev *gv = (gd && s < gd->max_vertex) ? gd : nullptr;
    if (gv)
      hipLaunchKernelGGL((k_guide_queue<H>), dim3(wl.guide_blocks), dim3(256), 0, st, sc, pa, ids, ws, *gv, s, q,
                         ctrl);
    if (!gv) {
      const dim3 g(wl.shade_blocks), b(kShadeBlock);
      switch (plain.stage) {
        case 3: hipLaunchKernelGGL((k_shade<H, 3>), g, b, plain.bytes, st, sc, pa, ids, ws, s, q, ctrl); break;
        case 2: hipLaunchKernelGGL((k_shade<H, 2>), g, b, plain.bytes, st, sc, pa, ids, ws, s, q, ctrl); break;
        case 1: hipLaunchKernelGGL((k_shade<H, 1>), g, b, plain.bytes, st, sc, pa, ids, ws, s, q, ctrl); break;
        default: hipLaunchKernelGGL((k_shade<H, 0>), g, b, 0, st, sc, pa, ids, ws, s, q, ctrl); break;
      }
    } else {
      const dim3 g(wl.shade_blocks), b(kShadeBlock);
      const uint32_t lb = guided.bytes;
      if (train) {
        switch (guided.stage) {
          case 3: hipLaunchKernelGGL((k_shade_guided<H, true, 3>), g, b, lb, st, sc, pa, ids, ws, *gv, s, q, ctrl); break;
          case 2: hipLaunchKernelGGL((k_shade_guided<H, true, 2>), g, b, lb, st, sc, pa, ids, ws, *gv, s, q, ctrl); break;
          default: hipLaunchKernelGGL((k_shade_guided<H, true, 0>), g, b, 0, st, sc, pa, ids, ws, *gv, s, q, ctrl); break;
        }
      } else {
        switch (guided.stage) {
          case 3: hipLaunchKernelGGL((k_shade_guided<H, false, 3>), g, b, lb, st, sc, pa, ids, ws, *gv, s, q, ctrl); break;
          case 2: hipLaunchKernelGGL((k_shade_guided<H, false, 2>), g, b, lb, st, sc, pa, ids, ws, *gv, s, q, ctrl); break;
          default: hipLaunchKernelGGL((k_shade_guided<H, false, 0>), g, b, 0, st, sc, pa, ids, ws, *gv, s, q, ctrl); break;
        }
      }
    }
    if (prof) (void)hipEventRecord(prof->shd0[s], st);
    if (gv && train)
      hipLaunchKernelGGL((k_shadow<H, false, true>), dim3(wl.trace_blocks), dim3(64), 0, st, sc, ws, ctrl, nullptr,
                         gv->rec, gv->rec_of_slot, gv->rec_cap);
    else if (prof && prof->counts)
      hipLaunchKernelGGL((k_shadow<H, true>), dim3(wl.trace_blocks), dim3(64), 0, st, sc, ws, ctrl, prof->counts,
                         nullptr, nullptr, 0u);
    else
      hipLaunchKernelGGL((k_shadow<H, false>), dim3(wl.trace_blocks), dim3(64), 0, st, sc, ws, ctrl, nullptr,
                         nullptr, nullptr, 0u);
    if (prof) (void)hipEventRecord(prof->shd1[s], st);
    q ^= 1u;
  }
  ev(2);
  if (film || out) launch_film<H>(sc, pa, ids, ws, n, ctrl, film, out, b256, st);
  ev(3);
  return hipGetLastError();
}

hipError_t enqueue_pass(int hero, const DevScene &sc, const PathArgs &pa, const uint32_t *pixel,
                        const uint32_t *sample, uint32_t pixel_offset, const WaveState &ws, uint32_t n,
                        uint32_t *ctrl, float *film, float *out, const WaveLaunch &wl, hipStream_t st,
                        WaveProfile *prof, const GuideDev *gd, bool train, uint32_t m) {
  PathIds ids{pixel, sample, pixel_offset, m > 1 ? n / m : 0u};
  if (hero == 16) return enqueue_pass_t<16>(sc, pa, ids, ws, n, ctrl, film, out, wl, st, prof, gd, train);
  return enqueue_pass_t<4>(sc, pa, ids, ws, n, ctrl, film, out, wl, st, prof, gd, train);
}

hipError_t enqueue_film(int hero, const DevScene &sc, const PathArgs &pa, uint32_t pixel_offset,
                        const WaveState &ws, uint32_t n, const uint32_t *ctrl, float *film, hipStream_t st,
                        uint32_t m) {
  const PathIds ids{nullptr, nullptr, pixel_offset, m > 1 ? n / m : 0u};
  const uint32_t b256 = (n / m + 255) / 256;
  if (hero == 16)
    launch_film<16>(sc, pa, ids, ws, n, ctrl, film, nullptr, b256, st);
  else
    launch_film<4>(sc, pa, ids, ws, n, ctrl, film, nullptr, b256, st);
  return hipGetLastError();
}

hipError_t enqueue_iter_extend(const DevScene &sc, const WaveState &ws, uint32_t *ctrl, uint32_t q,
                               const WaveLaunch &wl, hipStream_t st) {
  hipLaunchKernelGGL(k_iter_begin, dim3(1), dim3(64), 0, st, ctrl, q ^ 1u, 0u, 0u, nullptr);
  hipLaunchKernelGGL((k_extend<false>), dim3(wl.trace_blocks), dim3(64), 0, st, sc,
                     reinterpret_cast<const float4 *>(ws.ray[q]), ws.hits, ctrl, q, ws.spill, nullptr);
  return hipGetLastError();
}

hipError_t enqueue_pass_set(uint32_t *ctrl, uint32_t pass, hipStream_t st) {
  hipLaunchKernelGGL(k_pass_set, dim3(1), dim3(64), 0, st, ctrl, pass);
  return hipGetLastError();
}
hipError_t enqueue_pass_next(uint32_t *ctrl, uint32_t stride, hipStream_t st) {
  hipLaunchKernelGGL(k_pass_next, dim3(1), dim3(64), 0, st, ctrl, stride);
  return hipGetLastError();
}

}  // namespace lam

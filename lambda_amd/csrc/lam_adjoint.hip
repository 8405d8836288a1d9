// Path-replay adjoint of the render loop (gradients of an image loss with
// respect to thin-film thickness, thin-film IOR and Lambertian albedo).
//
// For one path the forward estimator is a sum of terms, each a product of
// per-vertex factors (src/renderer.cpp:64-146 in forward form):
//   L = Σ_s β_s ⊙ w_L,s L_IS,s + β_{s+1} ⊙ Le,   β_{s+1} = β_s ⊙ f_s / p_s w_mat,s
// A parameter θ of the material at vertex j enters through the sampled
// BSDF value f_j (every later term) and the NEE BSDF value coef_j (the NEE
// term of vertex j).  With the random numbers of the path fixed:
//   ∂L/∂θ = Σ_j (∂f_j/∂θ / f_j) ⊙ L_after(j) + Σ_j β_j ⊙ w_L,j mult_j ∂coef_j/∂θ / p_L,j
// where L_after(j) is the radiance collected after vertex j's NEE term.
// None of the differentiated parameters changes a sampled direction, pdf or
// Russian-roulette decision (Lambertian sampling is cosine, thin-film
// sampling is the Blinn half-vector of α, cg.hpp:1755-1931).
//
// The replay is a wavefront pass over the same path slots as the forward
// pass before it (the render's pass without a film, which leaves each path's
// total radiance L in the path state), and shares its extend kernel and
// queues.  L then holds the radiance a path still collects: every collected
// term is subtracted as the replay meets it, so nothing is stored per vertex.
// Per iteration:
//   k_extend       closest hits of the queue (the render's kernel)
//   k_adj_pending  at the path's next vertex, the pending terms of the
//                  previous one: its sampled factor, times the L left now
//                  (= L_after), and its NEE term if the light was visible
//   k_adj_shade    the vertex itself: light sample, BSDF sample, MIS
//                  weights as in k_shade, the NEE gradient held pending
//                  with its shadow ray, the sampled factor's ∂f/f held
//                  pending for the next vertex, β and the continuation
//   k_adj_shadow   any hit; a visible light subtracts the NEE term from L
//                  and marks the pending NEE gradient as due
// After the last iteration k_adj_final adds the NEE terms of paths that
// ended at a visible light sample.  The gradient terms gather in LDS per
// block (float atomics) and each block adds its sums into a partial row of
// its own; k_adj_sum_rows adds the rows into the gradient at the end.
#include "lam_launch.h"
#include "lam_shade.h"
#include "lam_wave.h"

namespace lam {

struct DiffDev {
  const int32_t *mat_d, *mat_ior, *mat_albedo;
  uint32_t n_params;
  // per-block partial gradients of the replay kernels: block b of any of
  // them adds its LDS sums into row b (n_params floats; rows of one lane,
  // whose kernels run in stream order)
  float *part;
};

// the block's gradient sums in LDS (n_params floats), zeroed at the start
__device__ __forceinline__ void grad_lds_init(float *gl, const DiffDev &dv) {
  for (uint32_t j = threadIdx.x; j < dv.n_params; j += blockDim.x) gl[j] = 0.f;
  __syncthreads();
}
// ... and added into the block's partial row at the end
__device__ __forceinline__ void grad_lds_flush(const float *gl, const DiffDev &dv) {
  __syncthreads();
  float *row = dv.part + size_t(blockIdx.x) * dv.n_params;
  for (uint32_t j = threadIdx.x; j < dv.n_params; j += blockDim.x) {
    const float g = gl[j];
    if (g != 0.f) row[j] += g;
  }
}

// thinfilm::BRDFcosθ (cg.hpp:1031-1051) with its derivatives with respect to
// the film thickness d and the film IOR n1 at each hero wavelength.  The
// value is evaluated with the same operations as thinfilm_f.
template <int H>
__device__ void thinfilm_f_grad(const DevScene &sc, const lam_material &m, const Lam<H> &L, f3 i, f3 o, f3 n,
                                bool front, float *val, float *dd, float *dn) {
  if (!front || dot(i, n) < 0.f || dot(o, n) < 0.f) {
#pragma unroll
    for (int k = 0; k < H; k++) { val[k] = 0.f; dd[k] = 0.f; dn[k] = 0.f; }
    return;
  }
  const f3 h = half_vec(i, o, n);
  const float ci = dot(i, h);
  float n1[H], n2[H];
  spec_eval<H>(sc, m.spec[0], L, n1);
  spec_eval<H>(sc, m.spec[1], L, n2);
  const float pw5 = lam_powf(1.f - ci, 5.f);
  const float c = (m.alpha + 2) * .125f * kInvPi;
  const float pw = lam_powf(dot(h, n), m.alpha);
  const float Q = 1.f - ci * ci;
#pragma unroll
  for (int k = 0; k < H; k++) {
    const float n01 = 1.f / n1[k];
    const float S = sqrtf(1.f - n01 * n01);
    const float cos_l = S * (1.f - ci * ci);
    const float arg = 4 * kPi * n1[k] * m.d * cos_l / L.l[k];
    float sin_d, cos_d;
    lam_sincosf(arg, &sin_d, &cos_d);
    const float a01 = 1.f - n1[k], b01 = 1.f + n1[k];
    const float r0_01 = a01 * a01 / (b01 * b01);
    const float a12 = n1[k] - n2[k], b12 = n1[k] + n2[k];
    const float r0_12 = a12 * a12 / (b12 * b12);
    const float R01 = r0_01 + (1.f - r0_01) * pw5;
    const float R12 = r0_12 + (1.f - r0_12) * pw5;
    const float sq = sqrtf(R01 * R12);
    const float ravg = 2 * sq;
    const float num = R01 + R12 + ravg * cos_d;
    const float den = R01 * R12 + 1.f + ravg * cos_d;
    val[k] = num / den * c * pw;
    // tangents
    const float dcos_dd = -sin_d * (4 * kPi * n1[k] * cos_l / L.l[k]);
    const float dS = (S > 0.f) ? n01 * n01 * n01 / S : 0.f;  // d sqrt(1-1/n1^2) / d n1
    const float dcos_l = dS * Q;
    const float darg_dn = 4 * kPi * m.d * (cos_l + n1[k] * dcos_l) / L.l[k];
    const float dcos_dn = -sin_d * darg_dn;
    const float dr001 = -4.f * a01 / (b01 * b01 * b01);
    const float dr012 = 4.f * a12 * n2[k] / (b12 * b12 * b12);
    const float dR01 = dr001 * (1.f - pw5);
    const float dR12 = dr012 * (1.f - pw5);
    const float dravg = sq > 0.f ? (dR01 * R12 + R01 * dR12) / sq : 0.f;
    const float inv_den2 = 1.f / (den * den);
    // ∂/∂d: only cos δ moves
    dd[k] = ravg * dcos_dd * (den - num) * inv_den2 * c * pw;
    const float dnum = dR01 + dR12 + dravg * cos_d + ravg * dcos_dn;
    const float dden = dR01 * R12 + R01 * dR12 + dravg * cos_d + ravg * dcos_dn;
    dn[k] = (dnum * den - num * dden) * inv_den2 * c * pw;
  }
}

template <int H>
__device__ bool differentiable(const DiffDev &dv, const lam_material &m, uint32_t mat) {
  if (m.type == LAM_MAT_LAMBERTIAN) return dv.mat_albedo[mat] >= 0;
  if (m.type == LAM_MAT_THINFILM) return dv.mat_d[mat] >= 0 || dv.mat_ior[mat] >= 0;
  return false;
}

// ---- pending gradient terms and their records ------------------------------------------
// A term of one vertex is (material, u[k], v[k]) per hero wavelength k, for
// a Lambertian material u = the weight of the albedo at λ_k (shared by the
// two grid coefficients the lookup interpolates, (1-t) and t); for thin film
// u = the weight of the thickness d (summed over k) and v = the weight of
// the film IOR at λ_k.  Its records multiply by m[k] (L_after for a sampled
// factor, 1 for an NEE term).
constexpr uint32_t kPendMark = 0x80000000u;  // pn_mat: NEE light not (yet) seen visible

// a term's contributions, times m[k], added into the block's LDS sums
template <int H>
__device__ __forceinline__ void accumulate_term(float *gl, const DiffDev &dv, const lam_material &m, uint32_t mat,
                                                const Lam<H> &L, const float *u, const float *v, const float *mult) {
  if (m.type == LAM_MAT_LAMBERTIAN) {
    const int32_t b = dv.mat_albedo[mat];
    if (b < 0) return;
#pragma unroll
    for (int k = 0; k < H; k++) {
      const float g = u[k] * mult[k];
      if (g == 0.f) continue;
      atomicAdd(gl + b + L.i[k], g * (1.f - L.t[k]));
      atomicAdd(gl + b + L.i[k] + 1, g * L.t[k]);
    }
  } else if (m.type == LAM_MAT_THINFILM) {
    const int32_t bd = dv.mat_d[mat], bn = dv.mat_ior[mat];
    if (bd >= 0) {
      float gd = 0.f;
#pragma unroll
      for (int k = 0; k < H; k++) gd += u[k] * mult[k];
      if (gd != 0.f) atomicAdd(gl + bd, gd);
    }
    if (bn >= 0)
#pragma unroll
      for (int k = 0; k < H; k++) {
        const float g = v[k] * mult[k];
        if (g == 0.f) continue;
        atomicAdd(gl + bn + L.i[k], g * (1.f - L.t[k]));
        atomicAdd(gl + bn + L.i[k] + 1, g * L.t[k]);
      }
  }
}

// The term of the BSDF value at (i, o): ratio = false: ∂f/∂θ scaled by w[k]
// (an NEE term); ratio = true: (∂f/∂θ)/f scaled by w[k] (a sampled factor).
template <int H>
__device__ __forceinline__ void bsdf_term(const DevScene &sc, const lam_material &m, const Lam<H> &L, f3 i, f3 o,
                                          f3 n, bool front, const float *w, bool ratio, float *u, float *v) {
  if (m.type == LAM_MAT_LAMBERTIAN) {
    // albedo(λ) = (1-t) v_i + t v_{i+1}; f = albedo · cos/π
    float rho[H];
    spec_eval<H>(sc, m.spec[0], L, rho);
    const float cs = fmax_ref(0.f, dot(i, n)) * kInvPi;
#pragma unroll
    for (int k = 0; k < H; k++) {
      u[k] = ratio ? (rho[k] != 0.f ? w[k] / rho[k] : 0.f) : w[k] * cs;
      v[k] = 0.f;
    }
  } else {
    float val[H], dd[H], dn[H];
    thinfilm_f_grad<H>(sc, m, L, i, o, n, front, val, dd, dn);
#pragma unroll
    for (int k = 0; k < H; k++) {
      const float s = ratio ? (val[k] != 0.f ? w[k] / val[k] : 0.f) : w[k];
      u[k] = s * dd[k];
      v[k] = s * dn[k];
    }
  }
}

// Replay state beyond the path state (per slot, [k][slot] for per-wavelength
// values): the adjoint weights and the pending sampled factor and NEE term.
struct ReplayDev {
  float *a;
  uint32_t *pf_mat;  // LAM_NONE: no pending sampled factor
  float *pf_u, *pf_v;
  uint32_t *pn_mat;  // LAM_NONE: none; with kPendMark: the light not seen visible
  float *pn_u, *pn_v;
};

// adjoint weights of a path's hero wavelengths: ∂loss/∂L_k through the
// film's linear map (cg.hpp:144-148, 306-341)
template <int H>
__device__ __forceinline__ void adjoint_weights(const DevScene &sc, const float *__restrict__ dloss, uint32_t pixel,
                                                const Lam<H> &lam, float prob_l, float inv_n, float *a) {
  const uint32_t w = sc.cam.width, h = sc.cam.height;
  const uint32_t row = pixel / w, colj = pixel - row * w;
  const size_t img = (size_t(h - row - 1) * w + colj) * 3;
  const float g0 = dloss[img], g1 = dloss[img + 1], g2 = dloss[img + 2];
  const float gx = 3.2404542f * g0 - 0.9692660f * g1 + 0.0556434f * g2;
  const float gy = -1.5371385f * g0 + 1.8760108f * g1 - 0.2040259f * g2;
  const float gz = -0.4985314f * g0 + 0.0415560f * g1 + 1.0572252f * g2;
#pragma unroll
  for (int k = 0; k < H; k++) {
    float X = 0.f, Y = 0.f, Z = 0.f;
    splat_xyz(sc, lam.l[k], 1.f, X, Y, Z, 1.f / sc.grid_delta);
    a[k] = (gx * X + gy * Y + gz * Z) * inv_n * sc.grid_delta / prob_l;
  }
}

// ---- replay kernels -------------------------------------------------------------------
// the camera rays of the pass's slots (as k_raygen); L keeps the forward total
template <int H>
__global__ __launch_bounds__(256) void k_adj_raygen(DevScene sc, PathArgs pa, PathIds ids, WaveState ws,
                                                    ReplayDev rd, uint32_t n, uint32_t *__restrict__ ctrl,
                                                    const float *__restrict__ dloss, float inv_n) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    ctrl[CTRL_COUNT0] = n;
    ctrl[CTRL_FETCH_EXT] = 0;
    ctrl[CTRL_TAIL] = 0;
  }
  if (p >= n) return;
  uint32_t pixel, sample;
  path_id(ids, ctrl, p, pixel, sample);
  const uint64_t key = lam_path_key(pa.seed, pixel, sample);
  const uint32_t w = sc.cam.width;
  const uint32_t row = pixel / w, colj = pixel - row * w;
  const float uu = float(colj) + lam_draw_f(key, LAM_DIM_U);
  const float vv = float(row) + lam_draw_f(key, LAM_DIM_V);
  f3 o, d;
  camera_ray(sc.cam, uu, vv, lam_draw_f(key, LAM_DIM_LENS_X), lam_draw_f(key, LAM_DIM_LENS_Y), o, d);
  float4 *q = reinterpret_cast<float4 *>(ws.ray[0]) + 2 * size_t(p);
  q[0] = make_float4(o.x, o.y, o.z, kFltMax);
  q[1] = make_float4(d.x, d.y, d.z, __uint_as_float(p));
  Lam<H> lam;
  const float prob_l = lam_for<H>(sc, key, lam);
  float a[H];
  adjoint_weights<H>(sc, dloss, pixel, lam, prob_l, inv_n, a);
#pragma unroll
  for (int k = 0; k < H; k++) {
    ws.beta[size_t(k) * ws.cap + p] = 1.f;
    rd.a[size_t(k) * ws.cap + p] = a[k];
  }
  rd.pf_mat[p] = LAM_NONE;
  rd.pn_mat[p] = LAM_NONE;
}

// At the path's next vertex (every entry of the extend queue, hit or miss),
// before k_adj_shade: the previous vertex's pending terms, its sampled
// factor times the radiance the path still collects (L_after) and its NEE
// term if the light was visible
template <int H>
__global__ __launch_bounds__(kShadeBlock) void k_adj_pending(DevScene sc, PathArgs pa, PathIds ids, WaveState ws,
                                                             ReplayDev rd, DiffDev dv, uint32_t qsel,
                                                             const uint32_t *__restrict__ ctrl) {
  extern __shared__ float gl[];
  const uint32_t n = ctrl[CTRL_COUNT0 + qsel];
  if (blockIdx.x * kShadeBlock >= n) return;
  grad_lds_init(gl, dv);
  const float4 *__restrict__ rin = reinterpret_cast<const float4 *>(ws.ray[qsel]);
  const size_t cap = ws.cap;
  for (uint32_t q = blockIdx.x * kShadeBlock + threadIdx.x; q < n; q += gridDim.x * kShadeBlock) {
    const uint32_t p = __float_as_uint(rin[2 * size_t(q) + 1].w);
    const uint32_t mf = rd.pf_mat[p], mn = rd.pn_mat[p];
    if (mf == LAM_NONE && mn == LAM_NONE) continue;
    uint32_t pixel, sample;
    path_id(ids, ctrl, p, pixel, sample);
    Lam<H> lam;
    lam_for<H>(sc, lam_path_key(pa.seed, pixel, sample), lam);
    if (mf != LAM_NONE) {
      float uf[H], vf[H], Lr[H];
#pragma unroll
      for (int k = 0; k < H; k++) {
        uf[k] = rd.pf_u[size_t(k) * cap + p];
        vf[k] = rd.pf_v[size_t(k) * cap + p];
        Lr[k] = ws.L[size_t(k) * cap + p];
      }
      accumulate_term<H>(gl, dv, sc.materials[mf], mf, lam, uf, vf, Lr);
      rd.pf_mat[p] = LAM_NONE;
    }
    if (mn != LAM_NONE) {
      if (!(mn & kPendMark)) {
        float un[H], vn[H], ones[H];
#pragma unroll
        for (int k = 0; k < H; k++) {
          un[k] = rd.pn_u[size_t(k) * cap + p];
          vn[k] = rd.pn_v[size_t(k) * cap + p];
          ones[k] = 1.f;
        }
        accumulate_term<H>(gl, dv, sc.materials[mn], mn, lam, un, vn, ones);
      }
      rd.pn_mat[p] = LAM_NONE;
    }
  }
  grad_lds_flush(gl, dv);
}

// STAGE: scene tables staged in LDS after the gradient sums (stage_tables);
// the next entry's ray and hit are loaded after the block append (as
// k_shade)
template <int H, int STAGE>
__global__ __launch_bounds__(kShadeBlock) void k_adj_shade(DevScene sc, PathArgs pa, PathIds ids, WaveState ws,
                                                           ReplayDev rd, DiffDev dv, uint32_t s, uint32_t qsel,
                                                           uint32_t *__restrict__ ctrl) {
  const uint32_t n = ctrl[CTRL_COUNT0 + qsel];
  const uint32_t pass = ctrl[CTRL_PASS];
  const float4 *__restrict__ rin = reinterpret_cast<const float4 *>(ws.ray[qsel]);
  float4 *__restrict__ rout = reinterpret_cast<float4 *>(ws.ray[qsel ^ 1u]);
  const size_t cap = ws.cap;
  const uint32_t stride = gridDim.x * kShadeBlock;
  __shared__ BlockQueues<kShadeBlock / 64, 2> sq;
  extern __shared__ float gl[];
  if (blockIdx.x * kShadeBlock >= n) return;
  if constexpr (STAGE > 0) stage_tables<STAGE, kShadeBlock>(sc, (dv.n_params + 3u) / 4u);
  grad_lds_init(gl, dv);
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
  for (uint32_t bbase = blockIdx.x * kShadeBlock; bbase < n; bbase += stride, par ^= 1u) {
    const uint32_t q = bbase + threadIdx.x;
    bool push_next = false, push_shadow = false;
    uint32_t p = 0;
    Ray nxt, shd;
    Lam<H> lam;
    const float4 ra = next_a, rb = next_b;
    const Hit hit = next_hit;
    if (q < n) {
      const f3 ro = mk(ra.x, ra.y, ra.z), rdir = mk(rb.x, rb.y, rb.z);
      p = __float_as_uint(rb.w);
      uint32_t pixel, sample;
      path_id_of(ids, pass, p, pixel, sample);
      const uint64_t key = lam_path_key(pa.seed, pixel, sample);
      lam_for<H>(sc, key, lam);
      if (hit.obj != LAM_NONE && s <= pa.max_scat) {
        const Surf sf = surface(sc, hit, ro, rdir);
        const f3 o = neg(normalized(rdir));
        const f3 nn = sf.n();
        const lam_material &mat = sc.materials[sf.mat];
        if (mat.type != LAM_MAT_EMITTER && mat.type != LAM_MAT_DIREMITTER) {
          float beta[H], a[H];
#pragma unroll
          for (int k = 0; k < H; k++) {
            beta[k] = ws.beta[size_t(k) * cap + p];
            a[k] = rd.a[size_t(k) * cap + p];
          }
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
          const bool diff = differentiable<H>(dv, mat, sf.mat);
          float w_L = 1.f, w_mat = 0.f;
          if (ms.ok) {
            const float p_L_mati = prob_light(sc, lt, sf.p, ms.dir) * pl;
            const float p_mat_Li = prob_material(sc, mat, sf, ls.dir, o, lam.l[0], pa.sample_p);
            const float p_L_i = ls.prob * pl;
            w_L = p_L_i * p_L_i / (p_L_i * p_L_i + p_mat_Li * p_mat_Li);
            w_mat = ms.prob * ms.prob / (ms.prob * ms.prob + p_L_mati * p_L_mati);
          }
          // NEE: the term, and ∂term/∂coef = β w_L mult / (p_L p_l)
          float pend[H], wn[H];
#pragma unroll
          for (int k = 0; k < H; k++) {
            pend[k] = ms.ok ? beta[k] * (lis[k] * w_L) : beta[k] * lis[k];
            wn[k] = a[k] * (beta[k] * w_L * ls.mult[k] / (ls.prob * pl));
          }
          float un[H], vn[H];
          if (diff) bsdf_term<H>(sc, mat, lam, ls.dir, o, nn, sf.front, wn, false, un, vn);
          if (ls.shadow) {
#pragma unroll
            for (int k = 0; k < H; k++) ws.pend[size_t(k) * cap + p] = pend[k];
            if (diff) {
              rd.pn_mat[p] = sf.mat | kPendMark;
#pragma unroll
              for (int k = 0; k < H; k++) {
                rd.pn_u[size_t(k) * cap + p] = un[k];
                rd.pn_v[size_t(k) * cap + p] = vn[k];
              }
            }
            push_shadow = true;
            shd = ls.sray;
          } else {
            sub_from<H>(ws.L, cap, p, pend);
            if (diff) {  // no shadow ray decides this NEE term: it is due now
              float ones[H];
#pragma unroll
              for (int k = 0; k < H; k++) ones[k] = 1.f;
              accumulate_term<H>(gl, dv, mat, sf.mat, lam, un, vn, ones);
            }
          }
          if (ms.ok) {
            if (diff) {  // the sampled factor, (∂f/∂θ)/f a, waits for L_after
              float uf[H], vf[H];
              bsdf_term<H>(sc, mat, lam, ms.dir, o, nn, sf.front, a, true, uf, vf);
              rd.pf_mat[p] = sf.mat;
#pragma unroll
              for (int k = 0; k < H; k++) {
                rd.pf_u[size_t(k) * cap + p] = uf[k];
                rd.pf_v[size_t(k) * cap + p] = vf[k];
              }
            }
#pragma unroll
            for (int k = 0; k < H; k++) ws.beta[size_t(k) * cap + p] = beta[k] * (ms.mult[k] / ms.prob * w_mat);
            push_next = true;
            nxt.o = sf.p;
            nxt.d = ms.dir;
          }
        }
      }
    }
    uint32_t slot[2];
    block_append(sq, par, {push_next, push_shadow}, {ctrl + CTRL_COUNT0 + (qsel ^ 1u), ctrl + CTRL_SHADOW}, slot);
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
    }
  }
  grad_lds_flush(gl, dv);
}

// any hit; a visible light subtracts the NEE term from L and makes the
// pending NEE gradient due
template <int H>
__global__ __launch_bounds__(64) void k_adj_shadow(DevScene sc, WaveState ws, ReplayDev rd,
                                                   uint32_t *__restrict__ ctrl) {
  __shared__ uint32_t stack[kLdsStack * 64];
  const uint32_t n = ctrl[CTRL_SHADOW];
  const float4 *__restrict__ sr = reinterpret_cast<const float4 *>(ws.sray);
  const size_t cap = ws.cap;
  auto load = [&](uint32_t i) {
    const float4 a = sr[2 * size_t(i)], b = sr[2 * size_t(i) + 1];
    RayItem it;
    it.r.o = mk(a.x, a.y, a.z);
    it.r.tmax = a.w;
    it.r.d = mk(b.x, b.y, b.z);
    it.r.skip = __float_as_uint(b.w);
    return it;
  };
  auto done = [&](uint32_t i, const RayItem &, const Trav &tv, bool ovf, const TravCounts &) {
    if (ovf) atomicOr(ctrl + 15, 1u);
    if (tv.best.obj == LAM_NONE && !ovf) {
      const uint32_t p = ws.spath[i];
      float pe[H];
#pragma unroll
      for (int k = 0; k < H; k++) pe[k] = ws.pend[size_t(k) * cap + p];
      sub_from<H>(ws.L, cap, p, pe);
      const uint32_t m = rd.pn_mat[p];
      if (m != LAM_NONE) rd.pn_mat[p] = m & ~kPendMark;
    }
  };
  const SplitStack stk = SplitStack::of(stack, ws.spill, blockIdx.x, gridDim.x);
  trace_queue<true, false>(sc, n, ctrl + CTRL_FETCH_SHD, stk, load, done);
}

// after the last iteration: the NEE terms of paths that ended with a visible
// light sample (the slots of the pass, grid-stride)
template <int H>
__global__ __launch_bounds__(kShadeBlock) void k_adj_final(DevScene sc, PathArgs pa, PathIds ids, WaveState ws,
                                                           ReplayDev rd, DiffDev dv, uint32_t n,
                                                           const uint32_t *__restrict__ ctrl) {
  extern __shared__ float gl[];
  if (blockIdx.x * kShadeBlock >= n) return;
  grad_lds_init(gl, dv);
  for (uint32_t p = blockIdx.x * kShadeBlock + threadIdx.x; p < n; p += gridDim.x * kShadeBlock) {
    const uint32_t m = rd.pn_mat[p];
    if (m == LAM_NONE) continue;
    if (!(m & kPendMark)) {
      uint32_t pixel, sample;
      path_id(ids, ctrl, p, pixel, sample);
      Lam<H> lam;
      lam_for<H>(sc, lam_path_key(pa.seed, pixel, sample), lam);
      float u[H], v[H], ones[H];
#pragma unroll
      for (int k = 0; k < H; k++) {
        u[k] = rd.pn_u[size_t(k) * ws.cap + p];
        v[k] = rd.pn_v[size_t(k) * ws.cap + p];
        ones[k] = 1.f;
      }
      accumulate_term<H>(gl, dv, sc.materials[m], m, lam, u, v, ones);
    }
    rd.pn_mat[p] = LAM_NONE;
  }
  grad_lds_flush(gl, dv);
}

// grad[j] += Σ_r part[r][j] over `rows` partial rows (one thread per parameter)
__global__ void k_adj_sum_rows(const float *__restrict__ part, uint32_t rows, uint32_t n_params,
                               float *__restrict__ grad) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_params) return;
  float s = 0.f;
  for (uint32_t r0 = 0; r0 < rows; r0 += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = r0 + u < rows ? part[size_t(r0 + u) * n_params + j] : 0.f;
#pragma unroll
    for (int u = 0; u < 16; u++) s += v[u];
  }
  grad[j] += s;
}

// ---- host side ---------------------------------------------------------------------------
template <int H>
static hipError_t enqueue_replay_t(const DevScene &sc, const PathArgs &pa, const PathIds &ids, const WaveState &ws,
                                   const ReplayDev &rd, const DiffDev &dv, uint32_t n, uint32_t *ctrl,
                                   const float *dloss, float inv_n, const WaveLaunch &wl, hipStream_t st) {
  const uint32_t b256 = (n + 255) / 256;
  const size_t gl_bytes = size_t(dv.n_params) * 4;
  // the replay shade stages the scene tables after the gradient sums when
  // both fit 64 KB
  const size_t gl4 = (size_t(dv.n_params) + 3) / 4 * 16;
  const StagePlan plan = gl4 < 64 * 1024 ? stage_plan(sc, false, uint32_t(64 * 1024 - gl4)) : StagePlan{0, 0u};
  hipLaunchKernelGGL((k_adj_raygen<H>), dim3(b256), dim3(256), 0, st, sc, pa, ids, ws, rd, n, ctrl, dloss, inv_n);
  uint32_t q = 0;
  const uint32_t iters = pa.max_scat + 2;
  hipError_t e = hipSuccess;
  for (uint32_t s = 0; s < iters && e == hipSuccess; s++) {
    e = enqueue_iter_extend(sc, ws, ctrl, q, wl, st);
    hipLaunchKernelGGL((k_adj_pending<H>), dim3(wl.shade_blocks), dim3(kShadeBlock), gl_bytes, st, sc, pa, ids, ws,
                       rd, dv, q, ctrl);
    const dim3 g(wl.shade_blocks), b(kShadeBlock);
    const size_t sb = gl4 + plan.bytes;
    switch (plan.stage) {
      case 3: hipLaunchKernelGGL((k_adj_shade<H, 3>), g, b, sb, st, sc, pa, ids, ws, rd, dv, s, q, ctrl); break;
      case 2: hipLaunchKernelGGL((k_adj_shade<H, 2>), g, b, sb, st, sc, pa, ids, ws, rd, dv, s, q, ctrl); break;
      default: hipLaunchKernelGGL((k_adj_shade<H, 0>), g, b, gl_bytes, st, sc, pa, ids, ws, rd, dv, s, q, ctrl); break;
    }
    hipLaunchKernelGGL((k_adj_shadow<H>), dim3(wl.trace_blocks), dim3(64), 0, st, sc, ws, rd, ctrl);
    q ^= 1u;
  }
  hipLaunchKernelGGL((k_adj_final<H>), dim3(std::min(b256, wl.shade_blocks)), dim3(kShadeBlock), gl_bytes, st, sc, pa,
                     ids, ws, rd, dv, n, ctrl);
  return e != hipSuccess ? e : hipGetLastError();
}

hipError_t enqueue_replay(int hero, const DevScene &sc, const PathArgs &pa, uint32_t pixel_offset, uint32_t m,
                          const WaveState &ws, const ReplayBufs &rb, const int32_t *mat_d, const int32_t *mat_ior,
                          const int32_t *mat_albedo, uint32_t n_params, float *part, uint32_t n, uint32_t *ctrl,
                          const float *dloss, float inv_n, const WaveLaunch &wl, hipStream_t st) {
  const PathIds ids{nullptr, nullptr, pixel_offset, m > 1 ? n / m : 0u};
  const DiffDev dv{mat_d, mat_ior, mat_albedo, n_params, part};
  const size_t cap = ws.cap;
  const size_t hf = size_t(hero) * cap;
  ReplayDev rd;
  rd.a = rb.f;
  rd.pf_u = rb.f + hf;
  rd.pf_v = rb.f + 2 * hf;
  rd.pn_u = rb.f + 3 * hf;
  rd.pn_v = rb.f + 4 * hf;
  rd.pf_mat = rb.u;
  rd.pn_mat = rb.u + cap;
  if (hero == 16) return enqueue_replay_t<16>(sc, pa, ids, ws, rd, dv, n, ctrl, dloss, inv_n, wl, st);
  return enqueue_replay_t<4>(sc, pa, ids, ws, rd, dv, n, ctrl, dloss, inv_n, wl, st);
}

hipError_t enqueue_sum_rows(const float *part, uint32_t rows, uint32_t n_params, float *grad, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_sum_rows, dim3((n_params + 255) / 256), dim3(256), 0, st, part, rows, n_params, grad);
  return hipGetLastError();
}

}  // namespace lam

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
// The forward pass (the render's wavefront pass) leaves each path's total
// L; the replay walks the path again subtracting each collected term, so
// nothing is stored per vertex.
// None of the differentiated parameters changes a sampled direction, pdf or
// Russian-roulette decision (Lambertian sampling is cosine, thin-film
// sampling is the Blinn half-vector of α, cg.hpp:1755-1931).
#include "lam_launch.h"
#include "lam_shade.h"

namespace lam {

struct DiffDev {
  const int32_t *mat_d, *mat_ior, *mat_albedo;
  uint32_t n_params;
};

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
    const float cos_d = lam_cosf(arg);
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
    const float sin_d = lam_sinf(arg);
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

// gradient of one BSDF factor of the material at a vertex, scaled per
// wavelength by w[k], accumulated into the block's gradient in LDS.
// ratio: the factor value divides w (sampled-factor form ∂f/f · L_after)
template <int H>
__device__ void accumulate(const DevScene &sc, const DiffDev &dv, const lam_material &m, uint32_t mat,
                           const Lam<H> &L, f3 i, f3 o, f3 n, bool front, const float *w, bool ratio,
                           float *gl) {
  if (m.type == LAM_MAT_LAMBERTIAN) {
    const int32_t base = dv.mat_albedo[mat];
    if (base < 0) return;
    // albedo(λ) = (1-t) v_i + t v_{i+1}; f = albedo · cos/π
    float rho[H];
    spec_eval<H>(sc, m.spec[0], L, rho);
    const float cs = fmax_ref(0.f, dot(i, n)) * kInvPi;
#pragma unroll
    for (int k = 0; k < H; k++) {
      float g;
      if (ratio) g = rho[k] != 0.f ? w[k] / rho[k] : 0.f;
      else g = w[k] * cs;
      if (g == 0.f) continue;
      atomicAdd(gl + base + L.i[k], g * (1.f - L.t[k]));
      atomicAdd(gl + base + L.i[k] + 1, g * L.t[k]);
    }
  } else if (m.type == LAM_MAT_THINFILM) {
    const int32_t bd = dv.mat_d[mat], bn = dv.mat_ior[mat];
    if (bd < 0 && bn < 0) return;
    float val[H], dd[H], dn[H];
    thinfilm_f_grad<H>(sc, m, L, i, o, n, front, val, dd, dn);
    float gd = 0.f;
#pragma unroll
    for (int k = 0; k < H; k++) {
      float s = w[k];
      if (ratio) s = val[k] != 0.f ? w[k] / val[k] : 0.f;
      if (s == 0.f) continue;
      gd += s * dd[k];
      if (bn >= 0) {
        const float gn = s * dn[k];
        atomicAdd(gl + bn + L.i[k], gn * (1.f - L.t[k]));
        atomicAdd(gl + bn + L.i[k] + 1, gn * L.t[k]);
      }
    }
    if (bd >= 0 && gd != 0.f) atomicAdd(gl + bd, gd);
  }
}

template <int H>
__device__ bool differentiable(const DiffDev &dv, const lam_material &m, uint32_t mat) {
  if (m.type == LAM_MAT_LAMBERTIAN) return dv.mat_albedo[mat] >= 0;
  if (m.type == LAM_MAT_THINFILM) return dv.mat_d[mat] >= 0 || dv.mat_ior[mat] >= 0;
  return false;
}

// One vertex of the replay, with the operations of path_vertex (lam_shade.h):
// L holds the radiance the path still collects; each collected term is
// subtracted, and the parameter gradients of the vertex are accumulated
// into gl, weighted by a[k] = ∂loss/∂L_k.  Returns false at the path's end.
template <int H, class Stack>
__device__ __forceinline__ bool replay_vertex(const DevScene &sc, const DiffDev &dv, uint32_t max_scat,
                                              float sample_p, uint64_t key, const Lam<H> &lam, uint32_t s, Ray &ray,
                                              float *beta, float *L, const float *a, float *gl, const Stack &stk,
                                              bool &ovf) {
  TravCounts tc{0, 0};
  const Hit hit = trace<false, false>(sc, ray, stk, tc, ovf);
  if (hit.obj == LAM_NONE || ovf) return false;
  const Surf sf = surface(sc, hit, ray.o, ray.d);
  const f3 o = neg(normalized(ray.d));
  const f3 n = sf.n();
  const lam_material &mat = sc.materials[sf.mat];
  // an emitter's term is the path's last: it is what L still holds
  if (mat.type == LAM_MAT_EMITTER || mat.type == LAM_MAT_DIREMITTER) return false;
  if (s > max_scat) return false;
  const uint32_t nl = sc.n_lights;
  const uint32_t li = lam_draw_u32(key, lam_vdim(s, LAM_VDIM_LIGHT_PICK)) % nl;
  const float pl = 1.f / float(nl);
  const lam_light &lt = sc.lights[li];
  LightSample<H> ls;
  sample_light<H>(sc, lt, lam, sf, lam_draw_f(key, lam_vdim(s, LAM_VDIM_LIGHT_A)),
                  lam_draw_f(key, lam_vdim(s, LAM_VDIM_LIGHT_B)), ls);
  float coef[H];
  bxdf<H>(sc, mat, lam, ls.dir, o, n, sf.front, coef);
  bool vis = true;
  if (ls.shadow) {
    const Hit sh = trace<true, false>(sc, ls.sray, stk, tc, ovf);
    vis = sh.obj == LAM_NONE;
  }
  float lis[H];
#pragma unroll
  for (int k = 0; k < H; k++) lis[k] = ls.mult[k] * coef[k] / (ls.prob * pl);
  MatSample<H> ms;
  sample_material<H>(sc, mat, lam, sf, o, lam_draw_f(key, lam_vdim(s, LAM_VDIM_LOBE)),
                     lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_A)), lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_B)),
                     sample_p, ms);
  const bool diff = differentiable<H>(dv, mat, sf.mat);
  float w_L = 1.f, w_mat = 0.f;
  if (ms.ok) {
    const float p_L_mati = prob_light(sc, lt, sf.p, ms.dir) * pl;
    const float p_mat_Li = prob_material(sc, mat, sf, ls.dir, o, lam.l[0], sample_p);
    const float p_L_i = ls.prob * pl;
    w_L = p_L_i * p_L_i / (p_L_i * p_L_i + p_mat_Li * p_mat_Li);
    w_mat = ms.prob * ms.prob / (ms.prob * ms.prob + p_L_mati * p_L_mati);
  }
  if (vis) {
    float w[H];
#pragma unroll
    for (int k = 0; k < H; k++) {
      L[k] -= ms.ok ? beta[k] * (lis[k] * w_L) : beta[k] * lis[k];
      // ∂term/∂coef = β w_L mult / (p_L p_l)
      w[k] = a[k] * (beta[k] * w_L * ls.mult[k] / (ls.prob * pl));
    }
    if (diff) accumulate<H>(sc, dv, mat, sf.mat, lam, ls.dir, o, n, sf.front, w, false, gl);
  }
  if (!ms.ok) return false;
  if (diff) {
    float w[H];
#pragma unroll
    for (int k = 0; k < H; k++) w[k] = a[k] * L[k];
    accumulate<H>(sc, dv, mat, sf.mat, lam, ms.dir, o, n, sf.front, w, true, gl);
  }
#pragma unroll
  for (int k = 0; k < H; k++) beta[k] = beta[k] * (ms.mult[k] / ms.prob * w_mat);
  ray.o = sf.p;
  ray.d = ms.dir;
  ray.tmax = kFltMax;
  ray.skip = LAM_NONE;
  return true;
}

// camera ray of a path (renderer.cpp:33-37 with the counter RNG)
__device__ __forceinline__ Ray adjoint_camera_ray(const DevScene &sc, uint64_t key, uint32_t pixel) {
  const uint32_t w = sc.cam.width;
  const uint32_t row = pixel / w, colj = pixel - row * w;
  const float u = float(colj) + lam_draw_f(key, LAM_DIM_U);
  const float v = float(row) + lam_draw_f(key, LAM_DIM_V);
  Ray ray;
  camera_ray(sc.cam, u, v, lam_draw_f(key, LAM_DIM_LENS_X), lam_draw_f(key, LAM_DIM_LENS_Y), ray.o, ray.d);
  ray.tmax = kFltMax;
  ray.skip = LAM_NONE;
  return ray;
}

// The replay of one sample per pixel for the n paths of a chunk (pixels
// pixel_offset .. pixel_offset + n - 1).  The forward pass is the render's
// wavefront pass, run just before on the same paths: Lfwd[k·lcap + i] is the
// total radiance of path i at wavelength k (the same value the vertex
// operations of the replay add up to).  Adjoint weights come from
// ∂loss/∂rgb of the path's pixel (the film's linear map, cg.hpp:144-148 and
// 306-341).  Each lane replays its own path one vertex per step and takes
// the next path when its replay ends (the queue pool of the traversal
// kernels: a block starts with paths 64b..64b+63, then reserves runs from
// *ctr, which starts at 0), so a wave never waits for its longest path.
// Gradients gather in LDS and leave with one atomic per parameter per block.
template <int H>
__global__ __launch_bounds__(64) void k_adjoint(DevScene sc, PathArgs pa, DiffDev dv, uint32_t sample, uint32_t n,
                                                uint32_t pixel_offset, const float *__restrict__ Lfwd, size_t lcap,
                                                float inv_n, const float *__restrict__ dloss, float *__restrict__ grad,
                                                uint32_t *__restrict__ ctr, uint32_t *__restrict__ err,
                                                uint32_t *__restrict__ spill) {
  __shared__ uint32_t stack[kLdsStack * 64];
  extern __shared__ float gl[];
  const uint32_t lane = threadIdx.x;
  const SplitStack stk = SplitStack::of(stack, spill, blockIdx.x, gridDim.x);
  for (uint32_t j = lane; j < dv.n_params; j += 64) gl[j] = 0.f;
  __syncthreads();
  const uint32_t w = sc.cam.width, h = sc.cam.height;
  QueuePool qp;
  qp.init(n);
  bool active = false, ovf = false;
  uint32_t s = 0;
  uint64_t key = 0;
  Lam<H> lam;
  float a[H], beta[H], L[H];
  Ray ray;
  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (qp.done(nidle)) break;
    if (qp.can_refill(nidle)) {
      const uint32_t i = qp.take(idle, nidle, !active, ctr);
      if (!active && i < n) {
        const uint32_t pixel = pixel_offset + i;
        key = lam_path_key(pa.seed, pixel, sample);
        ray = adjoint_camera_ray(sc, key, pixel);
        const float prob_l = hero_lambdas<H>(sc, lam_draw_f(key, LAM_DIM_HERO), lam);
        // adjoint weight of each wavelength: ∂loss/∂L_k
        const uint32_t row = pixel / w, colj = pixel - row * w;
        const size_t img = (size_t(h - row - 1) * w + colj) * 3;
        const float g0 = dloss[img], g1 = dloss[img + 1], g2 = dloss[img + 2];
        const float gx = 3.2404542f * g0 - 0.9692660f * g1 + 0.0556434f * g2;
        const float gy = -1.5371385f * g0 + 1.8760108f * g1 - 0.2040259f * g2;
        const float gz = -0.4985314f * g0 + 0.0415560f * g1 + 1.0572252f * g2;
        bool any = false;
#pragma unroll
        for (int k = 0; k < H; k++) {
          float X = 0.f, Y = 0.f, Z = 0.f;
          splat_xyz(sc, lam.l[k], 1.f, X, Y, Z);
          a[k] = (gx * X + gy * Y + gz * Z) * inv_n * sc.grid_delta / prob_l;
          any = any || a[k] != 0.f;
          beta[k] = 1.f;
          L[k] = Lfwd[size_t(k) * lcap + i];
        }
        active = any;
        s = 0;
      }
    }
    if (active) {
      active = replay_vertex<H>(sc, dv, pa.max_scat, pa.sample_p, key, lam, s, ray, beta, L, a, gl, stk, ovf);
      s++;
      if (ovf) {
        atomicOr(err, 1u);
        active = false;
      }
    }
  }
  __syncthreads();
  for (uint32_t j = lane; j < dv.n_params; j += 64)
    if (gl[j] != 0.f) atomicAdd(grad + j, gl[j]);
}

hipError_t launch_adjoint(int hero, const DevScene &sc, const PathArgs &pa, const int32_t *mat_d,
                          const int32_t *mat_ior, const int32_t *mat_albedo, uint32_t n_params, uint32_t sample,
                          uint32_t n, uint32_t pixel_offset, const float *Lfwd, size_t lcap, float inv_n,
                          const float *dloss, float *grad, uint32_t *ctr, uint32_t *err, uint32_t *spill,
                          uint32_t blocks, hipStream_t st) {
  DiffDev dv{mat_d, mat_ior, mat_albedo, n_params};
  const size_t shmem = size_t(n_params) * 4;
  if (hero == 16)
    hipLaunchKernelGGL((k_adjoint<16>), dim3(blocks), dim3(64), shmem, st, sc, pa, dv, sample, n, pixel_offset, Lfwd,
                       lcap, inv_n, dloss, grad, ctr, err, spill);
  else
    hipLaunchKernelGGL((k_adjoint<4>), dim3(blocks), dim3(64), shmem, st, sc, pa, dv, sample, n, pixel_offset, Lfwd,
                       lcap, inv_n, dloss, grad, ctr, err, spill);
  return hipGetLastError();
}

}  // namespace lam

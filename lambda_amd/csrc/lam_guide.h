// Learned path guiding: a small MLP maps a path vertex (position, outgoing
// direction, hero wavelength) to a mixture of 8 von Mises-Fisher lobes over
// the continuation direction.  It is evaluated with bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, fp32 accumulate) one wave at a time, and mixed
// with the sampling of a diffuse or glossy lobe under one-sample MIS, so
// the estimator keeps the expectation of the plain render loop.
//
// Network: 6 inputs in [0,1] -> frequency encoding (sin/cos(π 2^j x) for
// j = 0..11 and the raw inputs: 150 features, padded to 160) -> 64 -> 64 ->
// 64 -> 64 (ReLU) -> 40 = 8 lobes x (3-component mean direction, log κ,
// weight logit).  The mean direction is a 3-vector normalised after the
// network (no angle wrap).
//
// MFMA operand layouts (16x16x32 bf16; lane l, n = l % 16, g = l / 16):
//   A[m][k]: lane holds A[n][8g + j], j = 0..7
//   B[k][n]: lane holds B[8g + j][n]
//   D[m][n]: lane holds D[4g + i][n], i = 0..3
// Layers are computed transposed, D = W · Xᵀ (rows = neurons, columns = 16
// queries), so a layer's output registers already hold, per lane, the 16
// neurons {16t + 4g + i} of one query.  The next layer takes them as its B
// operand unchanged by storing its weight columns in the matching order:
// K index 32c + 8g + j  <->  neuron 32c + 16(j / 4) + 4g + (j % 4).
#pragma once
#include "lam_launch.h"
#include "lam_shade.h"

namespace lam {

constexpr int kGIn = 6, kGFreq = 12, kGFeat = 150, kGFeatPad = 160, kGHid = 64, kGLobes = 8, kGOut = 40,
              kGOutPad = 48;
// fp32 parameter vector: W1[64][150] b1 W2[64][64] b2 W3 b3 W4 b4 W5[40][64] b5
constexpr int kGW1 = 0, kGB1 = kGW1 + kGHid * kGFeat, kGW2 = kGB1 + kGHid, kGB2 = kGW2 + kGHid * kGHid,
              kGW3 = kGB2 + kGHid, kGB3 = kGW3 + kGHid * kGHid, kGW4 = kGB3 + kGHid, kGB4 = kGW4 + kGHid * kGHid,
              kGW5 = kGB4 + kGHid, kGB5 = kGW5 + kGOut * kGHid, kGParams = kGB5 + kGOut;
// packed weight fragments: layer 1 (4 row tiles x 5 K chunks), layers 2-4
// (4 x 2 each), layer 5 (3 x 2); 64 lanes x 16 bytes each; then the biases
// as fp32 (64 x 4 + 48); then the training backward's fragments of the
// transposed weights of layers 2-5 (4 input tiles x 2 K chunks each; K runs
// over the layer's outputs, permuted as the forward's B operand for layers
// 2-4, in natural order for layer 5)
constexpr int kGFragL1 = 0, kGFragL2 = 20, kGFragL5 = 44, kGFrags = 50;
constexpr int kGBiasWords = 4 * kGHid + kGOutPad;
constexpr int kGBFragL2 = 0, kGBFragL5 = 24, kGBFrags = 32;
constexpr size_t kGPackFwdBytes = size_t(kGFrags) * 64 * 16 + size_t(kGBiasWords) * 4;
constexpr size_t kGPackBytes = kGPackFwdBytes + size_t(kGBFrags) * 64 * 16;
constexpr float kGLogKMin = -4.f, kGLogKMax = 9.f;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = uint32_t(b) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ inline float bfround(float f) { return bf2f(f2bf(f)); }

// K position (within a 64-wide hidden layer) of neuron k's column
__host__ __device__ inline int g_perm(int k) {
  const int c = k >> 5, g = (k >> 3) & 3, j = k & 7;
  return 32 * c + 16 * (j >> 2) + 4 * g + (j & 3);
}
// its inverse: the K position that holds neuron k
__host__ __device__ inline int g_perm_inv(int k) {
  return 32 * (k >> 5) + 8 * ((k >> 2) & 3) + 4 * ((k >> 4) & 1) + (k & 3);
}

// sin and cos of 2π fr for fr in [0, 1): reduction to the nearest quarter
// turn (exact), then Taylor polynomials of degree 9 (sin) and 10 (cos) on
// [-π/4, π/4] (truncation < 2e-9), float arithmetic in this fixed order, so
// the host restatement gives the same bits
__host__ __device__ inline void g_sincos(float fr, float &s, float &c) {
  const float qf = floorf(fr * 4.f + .5f);
  const float t = fr - qf * .25f;
  const float x = t * 6.28318530717958647692f;
  const float x2 = x * x;
  const float sp = x + x * x2 * (-1.f / 6.f + x2 * (1.f / 120.f + x2 * (-1.f / 5040.f + x2 * (1.f / 362880.f))));
  const float cp =
      1.f + x2 * (-.5f + x2 * (1.f / 24.f + x2 * (-1.f / 720.f + x2 * (1.f / 40320.f + x2 * (-1.f / 3628800.f)))));
  const int q = int(qf) & 3;
  s = q == 0 ? sp : (q == 1 ? cp : (q == 2 ? -sp : -cp));
  c = q == 0 ? cp : (q == 1 ? -sp : (q == 2 ? -cp : sp));
}

// sin and cos of π 2^j x for j = j0 .. j0 + 3, j0 a multiple of 4: the
// quarter-turn evaluation above at j0, then double-angle steps
// (sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a)), which
// cost 5 operations instead of a polynomial and double an angle error of a
// few ulp at most three times
__host__ __device__ inline void g_sincos_j(float x, int j, float &s, float &c) {
  const int j0 = j & ~3;
  const float r = x * float(1 << j0) * .5f;
  g_sincos(r - floorf(r), s, c);
  for (int k = j0; k < j; k++) {
    const float s2 = (s * c) * 2.f, c2 = (c - s) * (c + s);
    s = s2;
    c = c2;
  }
}

// Feature f of the encoding: f = 24 d + 2 j + c (f < 144) is sin (c = 0) or
// cos (c = 1) of π 2^j x_d (g_sincos_j); f = 144 + d is x_d itself;
// 150..159 are zero.  Any 8-aligned block of features holds one input and
// four frequencies j0 .. j0 + 3 with j0 in {0, 4, 8}.
__device__ inline float g_feature(int f, const float *in) {
  if (f >= kGFeat) return 0.f;
  if (f >= 144) return in[f - 144];
  const int d = f / 24, j = (f % 24) >> 1;
  float sv, cv;
  g_sincos_j(in[d], j, sv, cv);
  return (f & 1) ? cv : sv;
}

// two floats as bf16 (round to nearest even, as f2bf) in one word: gfx950's
// packed conversion instruction
__device__ inline uint32_t pack_bf2(float a, float b) {
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  typedef float f32x2_t __attribute__((ext_vector_type(2)));
  const f32x2_t v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ inline bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ inline f32x4 mfma16(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(a), as_bf8(b), c, 0, 0, 0);
}
// ReLU of a layer output (4 row tiles) as the next layer's two bf16 B chunks
__device__ inline void relu_pack(const f32x4 *acc, uint4 *xb) {
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const f32x4 a = acc[2 * c], b = acc[2 * c + 1];
    xb[c].x = pack_bf2(fmaxf(a[0], 0.f), fmaxf(a[1], 0.f));
    xb[c].y = pack_bf2(fmaxf(a[2], 0.f), fmaxf(a[3], 0.f));
    xb[c].z = pack_bf2(fmaxf(b[0], 0.f), fmaxf(b[1], 0.f));
    xb[c].w = pack_bf2(fmaxf(b[2], 0.f), fmaxf(b[3], 0.f));
  }
}

// the six inputs of one query as named scalars (an array indexed by a
// lane-dependent input number would be placed in scratch memory)
struct GIn {
  float x0, x1, x2, x3, x4, x5;
};
__device__ inline float gin_at(const GIn &q, int d) {
  return d == 0 ? q.x0 : d == 1 ? q.x1 : d == 2 ? q.x2 : d == 3 ? q.x3 : d == 4 ? q.x4 : q.x5;
}

// the 8 bf16 features 32c + 8g .. +7 of one query (one input, four frequencies)
__device__ inline uint4 feature_chunk(int c, uint32_t g, const GIn &q) {
  const int f0 = 32 * c + 8 * int(g);
  float v[8];
  if (f0 < 144) {
    const int d = f0 / 24, j0 = (f0 % 24) >> 1;
    const float r = gin_at(q, d) * float(1 << j0) * .5f;
    float sv, cv;
    g_sincos(r - floorf(r), sv, cv);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v[2 * k] = sv;
      v[2 * k + 1] = cv;
      const float s2 = (sv * cv) * 2.f, c2 = (cv - sv) * (cv + sv);
      sv = s2;
      cv = c2;
    }
  } else if (f0 == 144) {
    v[0] = q.x0; v[1] = q.x1; v[2] = q.x2; v[3] = q.x3; v[4] = q.x4; v[5] = q.x5;
    v[6] = v[7] = 0.f;
  } else {
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = 0.f;
  }
  return make_uint4(pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]), pack_bf2(v[4], v[5]), pack_bf2(v[6], v[7]));
}

// The network over the 64 queries of a wave (one per lane) as 4 groups of
// 16 MFMA columns; each weight fragment is read once and applied to all four
// groups.  Every lane of the wave must call it (lanes without a query pass
// any input).  frag/bias: packed parameters (LDS or global).  The outputs
// stay in the MFMA result layout: o[G][t][i] of lane (n, g) is output row
// 16t + 4g + i of the query in lane 16G + n (rows >= 40 are padding).
__device__ inline void mlp_wave(const uint4 *__restrict__ frag, const float *__restrict__ bias, const float *in,
                                f32x4 (&o)[4][3]) {
  const uint32_t lane = threadIdx.x & 63u, n = lane & 15u, g = lane >> 4;
  GIn q[4];
#pragma unroll
  for (int G = 0; G < 4; G++) {
    const int src = int(16 * G + n);
    q[G] = GIn{__shfl(in[0], src, 64), __shfl(in[1], src, 64), __shfl(in[2], src, 64),
               __shfl(in[3], src, 64), __shfl(in[4], src, 64), __shfl(in[5], src, 64)};
  }
  f32x4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    f32x4 b;
#pragma unroll
    for (int i = 0; i < 4; i++) b[i] = bias[16 * t + 4 * g + i];
#pragma unroll
    for (int G = 0; G < 4; G++) acc[G][t] = b;
  }
#pragma unroll 1
  for (int c = 0; c < 5; c++) {
    uint4 xb[4];
#pragma unroll
    for (int G = 0; G < 4; G++) xb[G] = feature_chunk(c, g, q[G]);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const uint4 a = frag[(kGFragL1 + t * 5 + c) * 64 + lane];
#pragma unroll
      for (int G = 0; G < 4; G++) acc[G][t] = mfma16(a, xb[G], acc[G][t]);
    }
  }
#pragma unroll 1
  for (int l = 0; l < 3; l++) {
    uint4 xb[4][2];
#pragma unroll
    for (int G = 0; G < 4; G++) relu_pack(acc[G], xb[G]);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      f32x4 b;
#pragma unroll
      for (int i = 0; i < 4; i++) b[i] = bias[kGHid * (l + 1) + 16 * t + 4 * g + i];
#pragma unroll
      for (int G = 0; G < 4; G++) acc[G][t] = b;
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const uint4 a = frag[(kGFragL2 + 8 * l + t * 2 + c) * 64 + lane];
#pragma unroll
        for (int G = 0; G < 4; G++) acc[G][t] = mfma16(a, xb[G][c], acc[G][t]);
      }
    }
  }
  uint4 xb[4][2];
#pragma unroll
  for (int G = 0; G < 4; G++) relu_pack(acc[G], xb[G]);
#pragma unroll
  for (int t = 0; t < 3; t++) {
    f32x4 b;
#pragma unroll
    for (int i = 0; i < 4; i++) b[i] = bias[4 * kGHid + 16 * t + 4 * g + i];
#pragma unroll
    for (int G = 0; G < 4; G++) o[G][t] = b;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const uint4 a = frag[(kGFragL5 + t * 2 + c) * 64 + lane];
#pragma unroll
      for (int G = 0; G < 4; G++) o[G][t] = mfma16(a, xb[G][c], o[G][t]);
    }
  }
}

// ---- the mixture ----------------------------------------------------------------------
struct GuideMix {
  f3 mu[kGLobes];
  float kappa[kGLobes], w[kGLobes];
};

__device__ __forceinline__ void mix_from_raw(const float *raw, GuideMix &m) {
  float mx = -3.4e38f;
#pragma unroll
  for (int L = 0; L < kGLobes; L++) mx = fmaxf(mx, raw[5 * L + 4]);
  float s = 0.f;
#pragma unroll
  for (int L = 0; L < kGLobes; L++) {
    const f3 v = mk(raw[5 * L], raw[5 * L + 1], raw[5 * L + 2]);
    const float l2 = len2(v);
    m.mu[L] = l2 > 1e-24f ? mul(v, 1.f / sqrtf(l2)) : mk(0.f, 0.f, 1.f);
    m.kappa[L] = expf(fminf(fmaxf(raw[5 * L + 3], kGLogKMin), kGLogKMax));
    m.w[L] = expf(raw[5 * L + 4] - mx);
    s += m.w[L];
  }
  const float inv = 1.f / s;
#pragma unroll
  for (int L = 0; L < kGLobes; L++) m.w[L] *= inv;
}

// vMF density κ / (2π (1 - e^{-2κ})) · e^{κ(μ·ω - 1)}
__device__ __forceinline__ float vmf_pdf(f3 mu, float k, f3 w) {
  return k / (2.f * kPi * -expm1f(-2.f * k)) * expf(k * (dot(mu, w) - 1.f));
}
__device__ __forceinline__ float mix_pdf(const GuideMix &m, f3 w) {
  float p = 0.f;
#pragma unroll
  for (int L = 0; L < kGLobes; L++) p += m.w[L] * vmf_pdf(m.mu[L], m.kappa[L], w);
  return p;
}
// orthonormal basis around a unit vector (Duff et al. 2017, branchless)
__device__ __forceinline__ void basis_of(f3 n, f3 &t, f3 &b) {
  const float sg = copysignf(1.f, n.z);
  const float a = -1.f / (sg + n.z);
  const float c = n.x * n.y * a;
  t = mk(1.f + sg * n.x * n.x * a, sg * c, -sg * n.x);
  b = mk(c, sg + n.y * n.y * a, -n.y);
}
// vMF sampling, w = 1 + log(ξ + (1-ξ) e^{-2κ}) / κ (Jakob 2012)
__device__ __forceinline__ f3 vmf_sample(f3 mu, float k, float u1, float u2) {
  const float w = fminf(fmaxf(1.f + logf(u1 + (1.f - u1) * expf(-2.f * k)) / k, -1.f), 1.f);
  const float r = sqrtf(fmaxf(0.f, 1.f - w * w));
  const float ph = 2.f * kPi * u2;
  f3 t, b;
  basis_of(mu, t, b);
  return normalized(add(add(mul(mu, w), mul(t, r * cosf(ph))), mul(b, r * sinf(ph))));
}
__device__ __forceinline__ f3 mix_sample(const GuideMix &m, float ul, float u1, float u2) {
  // lobe L with cumulative weight above ul; its parameters are gathered with
  // 0/1 weights (selects of array elements get folded into an indexed load,
  // which would move the whole mixture to scratch memory)
  int sel = kGLobes - 1;
  float c = 0.f;
  bool found = false;
#pragma unroll
  for (int L = 0; L < kGLobes - 1; L++) {
    c += m.w[L];
    const bool take = !found && ul < c;
    sel = take ? L : sel;
    found = found || take;
  }
  f3 mu = mk(0.f, 0.f, 0.f);
  float k = 0.f;
#pragma unroll
  for (int L = 0; L < kGLobes; L++) {
    const float f = sel == L ? 1.f : 0.f;
    mu = add(mu, mul(m.mu[L], f));
    k += m.kappa[L] * f;
  }
  return vmf_sample(mu, k, u1, u2);
}

// ---- where the proposal applies ----------------------------------------------------------
// The learned proposal stands in for one lobe of the reference's sampler:
// the cosine lobe of a Lambertian material (cg.hpp:1755-1778), the Kd lobe
// of a Blinn material (cg.hpp:1853-1861) or the specular reflection lobe of
// a Blinn material without transmission (Kt = 0, so fresnelSplit always
// reflects; cg.hpp:1815-1847), as selected by GuideDev::lobes.  Thin film
// and transmitting Blinn lobes keep their own samplers.
constexpr int kGuideNone = 0, kGuideCosine = 1, kGuideBlinnSpec = 2;
__device__ __forceinline__ bool guide_supported(const GuideDev &gd, const lam_material &m) {
  if (m.type == LAM_MAT_LAMBERTIAN) return (gd.lobes & LAM_GUIDE_LAMBERTIAN) != 0;
  if (m.type != LAM_MAT_BLINN) return false;
  return ((gd.lobes & LAM_GUIDE_BLINN_KD) && m.p_d > 0.f) ||
         ((gd.lobes & LAM_GUIDE_BLINN_SPEC) && m.w_t == 0.f && m.p_spec > 0.f);
}

// Does the guided shade draw this vertex's continuation from a lobe the
// proposal applies to, i.e. read the network's mixture?  The reference's
// lobe choice from the LOBE dimension xl (cg.hpp:1768, 1809-1853) with
// roulette probability P; pp = the chosen lobe's selection probability.
// The network is evaluated for exactly these vertices.
__device__ __forceinline__ bool guide_vertex(const GuideDev &gd, const lam_material &m, uint32_t s) {
  return gd.alpha > 0.f && ((gd.mat_mask >> m.type) & 1u) && guide_supported(gd, m) && s < gd.max_vertex;
}
__device__ __forceinline__ int guide_draw(const GuideDev &gd, const lam_material &m, uint32_t s, float xl, float P,
                                          bool front, float &pp) {
  pp = 0.f;
  if (!guide_vertex(gd, m, s)) return kGuideNone;
  if (m.type == LAM_MAT_LAMBERTIAN) {
    pp = P;
    return !(xl > P) ? kGuideCosine : kGuideNone;
  }
  if (!(xl < P)) return kGuideNone;
  const float pp_spec = P * m.p_spec, pp_d = P * m.p_d;
  if (xl < pp_spec) {
    // front faces only: on a back face the reference's lobe value is
    // powf(h·F.z, α) of a negative base (cg.hpp:1830-1846)
    if ((gd.lobes & LAM_GUIDE_BLINN_SPEC) && m.w_t == 0.f && front) {
      pp = pp_spec;
      return kGuideBlinnSpec;
    }
    return kGuideNone;
  }
  if (xl < pp_spec + pp_d && (gd.lobes & LAM_GUIDE_BLINN_KD)) {
    pp = pp_d;
    return kGuideCosine;
  }
  return kGuideNone;
}

// power-heuristic weight a² / (a² + b²) of a strategy with pdf a against one
// with pdf b; 1 where b = 0 (a pdf that underflowed to 0 leaves 0/0 otherwise)
__device__ __forceinline__ float mis_w(float a, float b2) {
  const float a2 = a * a;
  return b2 > 0.f ? a2 / (a2 + b2) : 1.f;
}

// network inputs of a vertex, each in [0,1]: position in the scene box, the
// outgoing direction (towards the previous vertex, world space) as
// (θ/π, φ/2π), hero wavelength on the grid
template <int H>
__device__ inline void guide_inputs(const GuideDev &gd, const DevScene &sc, f3 x, f3 o, const Lam<H> &L,
                                    float *in) {
  in[0] = fminf(fmaxf((x.x - gd.lo.x) * gd.inv_ext.x, 0.f), 1.f);
  in[1] = fminf(fmaxf((x.y - gd.lo.y) * gd.inv_ext.y, 0.f), 1.f);
  in[2] = fminf(fmaxf((x.z - gd.lo.z) * gd.inv_ext.z, 0.f), 1.f);
  in[3] = acosf(fminf(fmaxf(o.z, -1.f), 1.f)) * kInvPi;
  in[4] = (atan2f(o.y, o.x) + kPi) * (.5f * kInvPi);
  in[5] = (L.l[0] - sc.grid_inf) / (sc.grid_sup - sc.grid_inf);
}

}  // namespace lam

// Guiding network kernels: parameter packing into MFMA fragments, batched
// inference (test hook), and online training.
//
// Training runs on batches of vertex records the tracer writes during the
// first passes of a guided render.  Each record holds the network inputs of a
// vertex, the continuation direction ω it sampled, and (after the pass) a
// weight w = (radiance gathered through ω) / (throughput at the vertex).  The
// loss is the weighted negative log-likelihood of the mixture,
//   loss = -Σ_i (w_i / Σw) log p_g(ω_i | x_i),
// whose minimiser is the mixture closest in KL divergence to the sampled
// distribution of radiance (times the plain continuation's weight).  Every
// GEMM of the forward and backward passes runs on bf16 MFMA with fp32
// accumulation (k_gemm_mfma): the forward reproduces the inference network
// (bf16 weights, features and activations); the backward rounds dZ to bf16
// where it enters a product and treats the forward's bf16 roundings as
// identity.  Adam updates fp32 master parameters, which are then repacked
// into bf16 fragments.
#include <algorithm>

#include "lam_guide.h"

namespace lam {

// ---- packing ------------------------------------------------------------------------
__global__ void k_guide_pack(const float *__restrict__ th, uint4 *__restrict__ frag, float *__restrict__ bias,
                             float *__restrict__ thq) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < uint32_t(kGFrags) * 64) {
    const uint32_t f = tid / 64, lane = tid % 64, n = lane % 16, g = lane / 16;
    uint16_t v[8];
    for (int j = 0; j < 8; j++) {
      float w = 0.f;
      if (f < uint32_t(kGFragL2)) {  // layer 1: natural feature order
        const uint32_t t = f / 5, c = f % 5, row = 16 * t + n, col = 32 * c + 8 * g + j;
        if (col < uint32_t(kGFeat)) w = th[kGW1 + row * kGFeat + col];
      } else if (f < uint32_t(kGFragL5)) {
        const uint32_t l = (f - kGFragL2) / 8, r = (f - kGFragL2) % 8, t = r / 2, c = r % 2;
        const uint32_t row = 16 * t + n, col = uint32_t(g_perm(int(32 * c + 8 * g + j)));
        const int base = l == 0 ? kGW2 : (l == 1 ? kGW3 : kGW4);
        w = th[base + row * kGHid + col];
      } else {
        const uint32_t r = f - kGFragL5, t = r / 2, c = r % 2, row = 16 * t + n;
        const uint32_t col = uint32_t(g_perm(int(32 * c + 8 * g + j)));
        if (row < uint32_t(kGOut)) w = th[kGW5 + row * kGHid + col];
      }
      v[j] = f2bf(w);
    }
    frag[tid] = make_uint4(v[0] | (uint32_t(v[1]) << 16), v[2] | (uint32_t(v[3]) << 16),
                           v[4] | (uint32_t(v[5]) << 16), v[6] | (uint32_t(v[7]) << 16));
  }
  if (tid < uint32_t(kGBiasWords)) {
    float b = 0.f;
    if (tid < 4u * kGHid) {
      const int l = tid / kGHid, i = tid % kGHid;
      const int base = l == 0 ? kGB1 : (l == 1 ? kGB2 : (l == 2 ? kGB3 : kGB4));
      b = th[base + i];
    } else if (tid - 4u * kGHid < uint32_t(kGOut)) {
      b = th[kGB5 + (tid - 4 * kGHid)];
    }
    bias[tid] = b;
  }
  // bf16-rounded copy of every parameter for the training passes, which take
  // the weights from it and the (fp32) biases from th.  (A per-range select
  // between th and its rounding here was compiled for gfx950 with the store
  // address of the bias ranges clobbered to the first lane's; a plain copy
  // has no such select.)
  if (tid < uint32_t(kGParams)) thq[tid] = bfround(th[tid]);
}

// ---- batched inference (test hook) ------------------------------------------------------
// the same staging and evaluation as the guided render's k_guide_queue
__global__ __launch_bounds__(256) void k_guide_infer(const uint4 *__restrict__ pack, const float *__restrict__ in,
                                                     uint32_t n, float *__restrict__ raw) {
  __shared__ uint4 frag[kGFrags * 64];
  __shared__ float bias[kGBiasWords];
  {
    const float *gb = reinterpret_cast<const float *>(pack + kGFrags * 64);
    for (uint32_t i = threadIdx.x; i < uint32_t(kGFrags) * 64; i += blockDim.x) frag[i] = pack[i];
    for (uint32_t i = threadIdx.x; i < uint32_t(kGBiasWords); i += blockDim.x) bias[i] = gb[i];
  }
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x % 64, ln = lane % 16, lg = lane / 16;
  for (uint32_t wb = first - lane; wb < n; wb += stride) {  // wave-uniform trips
    const uint32_t q = wb + lane;
    float x[kGIn];
    for (int d = 0; d < kGIn; d++) x[d] = q < n ? in[size_t(q) * kGIn + d] : 0.f;
    f32x4 o[4][3];
    mlp_wave(frag, bias, x, o);
#pragma unroll
    for (int G = 0; G < 4; G++) {
      const uint32_t qq = wb + 16 * G + ln;
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const uint32_t row = 16 * t + 4 * lg;
        if (qq < n && row < uint32_t(kGOut))
          *reinterpret_cast<float4 *>(raw + size_t(qq) * kGOut + row) =
              make_float4(o[G][t][0], o[G][t][1], o[G][t][2], o[G][t][3]);
      }
    }
  }
}

// ---- training ---------------------------------------------------------------------------
__device__ inline uint32_t hash32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return uint32_t(x);
}

// training batch from records: features (bf16-rounded) and (ω, w)
__global__ void k_guide_gather(const float4 *__restrict__ rec, const uint32_t *__restrict__ rec_count,
                               uint32_t rec_cap, uint32_t B, uint64_t seed, float *__restrict__ X0,
                               float4 *__restrict__ tgt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const uint32_t nr = min(*rec_count, rec_cap);
  float in[kGIn];
  float4 t = make_float4(0.f, 0.f, 1.f, 0.f);
  if (nr == 0) {
    for (int d = 0; d < kGIn; d++) in[d] = 0.f;
  } else {
    const uint32_t r = uint32_t((uint64_t(hash32(seed * 0x9E3779B97F4A7C15ull + i)) * nr) >> 32);
    const float4 a = rec[4 * size_t(r)], b = rec[4 * size_t(r) + 1], c = rec[4 * size_t(r) + 2],
                 d = rec[4 * size_t(r) + 3];
    in[0] = a.x; in[1] = a.y; in[2] = a.z; in[3] = a.w; in[4] = b.x; in[5] = b.y;
    t = make_float4(c.x, c.y, c.z, d.z);
  }
  for (int f = 0; f < kGFeatPad; f++) X0[size_t(i) * kGFeatPad + f] = bfround(g_feature(f, in));
  tgt[i] = t;
}

// training batch from host rows (in[6], ω[3], w)
__global__ void k_guide_feat_rows(const float *__restrict__ rows, uint32_t B, float *__restrict__ X0,
                                  float4 *__restrict__ tgt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const float *r = rows + size_t(i) * 10;
  float in[kGIn];
  for (int d = 0; d < kGIn; d++) in[d] = r[d];
  for (int f = 0; f < kGFeatPad; f++) X0[size_t(i) * kGFeatPad + f] = bfround(g_feature(f, in));
  tgt[i] = make_float4(r[6], r[7], r[8], r[9]);
}

// C[M x N] (+)= A[M x K] B[K x N] on v_mfma_f32_16x16x32_bf16: A and B are
// read as fp32, rounded to bf16 (nearest even) into LDS, and their products
// summed in fp32.  A(m,k) = TA ? A[k lda + m] : A[m lda + k],
// B(k,n) = TB ? B[n ldb + k] : B[k ldb + n].  64 x 64 block tiles, four waves
// of 32 x 32 (2 x 2 MFMA tiles), K steps of 32; blockIdx.z splits K into
// chunks of `kchunk` (a multiple of 32) whose partial sums are added with
// atomics.  The operands of the network's layers are bf16 already
// (weights, features, activations: exact); the backward's dZ is rounded.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_mfma(uint32_t M, uint32_t N, uint32_t K, const float *__restrict__ A,
                                                   uint32_t lda, const float *__restrict__ Bm, uint32_t ldb,
                                                   float *__restrict__ C, uint32_t ldc, uint32_t kchunk, int atomic) {
  // [row][k] bf16; rows of 40 halves (80 B) keep every 8-half group 16-byte aligned
  __shared__ __attribute__((aligned(16))) uint16_t As[64][40];
  __shared__ __attribute__((aligned(16))) uint16_t Bt[64][40];
  const uint32_t tid = threadIdx.x, wave = tid / 64, lane = tid % 64, ln = lane % 16, lg = lane / 16;
  const uint32_t wm = wave / 2, wn = wave % 2;
  const uint32_t m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const uint32_t k_begin = blockIdx.z * kchunk, k_end = min(K, k_begin + kchunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (uint32_t k0 = k_begin; k0 < k_end; k0 += 32) {
    // 64 x 32 elements of each operand, 8 per thread; the global index that
    // varies fastest follows the operand's storage order
    for (uint32_t e = tid; e < 64u * 32u; e += 256u) {
      const uint32_t am = TA ? e % 64u : e / 32u, ak = TA ? e / 64u : e % 32u;
      const uint32_t gm = m0 + am, gk = k0 + ak;
      float av = 0.f;
      if (gm < M && gk < k_end) av = TA ? A[size_t(gk) * lda + gm] : A[size_t(gm) * lda + gk];
      As[am][ak] = f2bf(av);
      const uint32_t bn = TB ? e / 32u : e % 64u, bk = TB ? e % 32u : e / 64u;
      const uint32_t gn = n0 + bn, gk2 = k0 + bk;
      float bv = 0.f;
      if (gn < N && gk2 < k_end) bv = TB ? Bm[size_t(gn) * ldb + gk2] : Bm[size_t(gk2) * ldb + gn];
      Bt[bn][bk] = f2bf(bv);
    }
    __syncthreads();
    // lane (n, g) supplies A[n][8g .. 8g+7] and B[8g .. 8g+7][n]
#pragma unroll
    for (int mi = 0; mi < 2; mi++) {
      const uint4 a = *reinterpret_cast<const uint4 *>(&As[32 * wm + 16 * mi + ln][8 * lg]);
#pragma unroll
      for (int ni = 0; ni < 2; ni++) {
        const uint4 b = *reinterpret_cast<const uint4 *>(&Bt[32 * wn + 16 * ni + ln][8 * lg]);
        acc[mi][ni] = mfma16(a, b, acc[mi][ni]);
      }
    }
    __syncthreads();
  }
  // lane (n, g) holds D[4g + i][n] of each tile
#pragma unroll
  for (int mi = 0; mi < 2; mi++)
#pragma unroll
    for (int ni = 0; ni < 2; ni++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t m = m0 + 32 * wm + 16 * mi + 4 * lg + i, n = n0 + 32 * wn + 16 * ni + ln;
        if (m < M && n < N) {
          if (atomic) atomicAdd(C + size_t(m) * ldc + n, acc[mi][ni][i]);
          else C[size_t(m) * ldc + n] = acc[mi][ni][i];
        }
      }
}

// Z[b][j] += bias[j]; X = bf16(relu(Z)) (hidden layers) or X = Z (output)
__global__ void k_bias_act(float *__restrict__ Z, const float *__restrict__ bias, uint32_t B, uint32_t N,
                           float *__restrict__ X, int relu) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B * N) return;
  const float z = Z[e] + bias[e % N];
  Z[e] = z;
  if (X) X[e] = relu ? bfround(fmaxf(z, 0.f)) : z;
}

// dZ = dX where Z > 0, else 0
__global__ void k_relu_grad(const float *__restrict__ dX, const float *__restrict__ Z, uint32_t n,
                            float *__restrict__ dZ) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) dZ[e] = Z[e] > 0.f ? dX[e] : 0.f;
}

// db[j] += Σ_b dZ[b][j]
__global__ void k_colsum(const float *__restrict__ dZ, uint32_t B, uint32_t N, float *__restrict__ db) {
  const uint32_t j = blockIdx.x, t = threadIdx.x;
  float s = 0.f;
  for (uint32_t b = blockIdx.y * 256 + t; b < B; b += 256 * gridDim.y) s += dZ[size_t(b) * N + j];
  __shared__ float red[256];
  red[t] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) atomicAdd(db + j, red[0]);
}

__global__ void k_sum_w(const float4 *__restrict__ tgt, uint32_t B, float *__restrict__ out) {
  __shared__ float red[256];
  const uint32_t t = threadIdx.x;
  float s = 0.f;
  for (uint32_t b = blockIdx.x * 256 + t; b < B; b += 256 * gridDim.x) s += tgt[b].w;
  red[t] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) atomicAdd(out, red[0]);
}

// per sample: loss term -w/Σw log p_g(ω) and its gradient with respect to the
// 40 raw outputs.  scal[0] = Σw (input), scal[1] += loss.
__global__ void k_guide_loss(const float *__restrict__ raw, const float4 *__restrict__ tgt, uint32_t B,
                             float *__restrict__ scal, float *__restrict__ dZ) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const float *r = raw + size_t(i) * kGOut;
  float *d = dZ + size_t(i) * kGOut;
  const float4 t = tgt[i];
  const float sw = scal[0];
  const float wn = sw > 0.f ? t.w / sw : 0.f;
  if (wn == 0.f) {
    for (int k = 0; k < kGOut; k++) d[k] = 0.f;
    return;
  }
  const f3 om = mk(t.x, t.y, t.z);
  float a_max = -3.4e38f;
  for (int L = 0; L < kGLobes; L++) a_max = fmaxf(a_max, r[5 * L + 4]);
  float ws[kGLobes], lv[kGLobes], kap[kGLobes], md[kGLobes], vlen[kGLobes];
  f3 mu[kGLobes];
  float wsum = 0.f;
  for (int L = 0; L < kGLobes; L++) {
    ws[L] = expf(r[5 * L + 4] - a_max);
    wsum += ws[L];
    const f3 v = mk(r[5 * L], r[5 * L + 1], r[5 * L + 2]);
    const float l2 = len2(v);
    vlen[L] = sqrtf(l2);
    mu[L] = l2 > 1e-24f ? mul(v, 1.f / vlen[L]) : mk(0.f, 0.f, 1.f);
    kap[L] = expf(fminf(fmaxf(r[5 * L + 3], kGLogKMin), kGLogKMax));
    md[L] = dot(mu[L], om);
    // log vMF = log κ - log 2π - log(1 - e^{-2κ}) + κ(μ·ω - 1)
    lv[L] = logf(kap[L]) - logf(2.f * kPi) - logf(-expm1f(-2.f * kap[L])) + kap[L] * (md[L] - 1.f);
  }
  // log p = logsumexp(log w_L + log vMF_L); responsibilities
  float zmax = -3.4e38f, z[kGLobes];
  for (int L = 0; L < kGLobes; L++) {
    z[L] = logf(ws[L] / wsum) + lv[L];
    zmax = fmaxf(zmax, z[L]);
  }
  float zs = 0.f;
  for (int L = 0; L < kGLobes; L++) zs += expf(z[L] - zmax);
  const float logp = zmax + logf(zs);
  atomicAdd(scal + 1, -wn * logp);
  for (int L = 0; L < kGLobes; L++) {
    const float resp = expf(z[L] - logp), wl = ws[L] / wsum;
    const float k = kap[L];
    // ∂ log vMF / ∂ log κ = 1 - 2κ/(e^{2κ}-1) + κ(μ·ω - 1)
    const float dlk = 1.f - 2.f * k / expm1f(2.f * k) + k * (md[L] - 1.f);
    const float rk = r[5 * L + 3];
    const float g_k = (rk > kGLogKMin && rk < kGLogKMax) ? resp * dlk : 0.f;
    // ∂(μ·ω)/∂v = (ω - (μ·ω) μ) / |v|
    f3 gv = mk(0.f, 0.f, 0.f);
    if (vlen[L] > 1e-12f) gv = mul(sub(om, mul(mu[L], md[L])), resp * k / vlen[L]);
    d[5 * L] = -wn * gv.x;
    d[5 * L + 1] = -wn * gv.y;
    d[5 * L + 2] = -wn * gv.z;
    d[5 * L + 3] = -wn * g_k;
    d[5 * L + 4] = -wn * (resp - wl);
  }
}

__global__ void k_adam(float *__restrict__ th, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- host-side drivers ------------------------------------------------------------------
static inline uint32_t cdiv(size_t a, uint32_t b) { return uint32_t((a + b - 1) / b); }

template <bool TA, bool TB>
static void gemm(uint32_t M, uint32_t N, uint32_t K, const float *A, uint32_t lda, const float *B, uint32_t ldb,
                 float *C, uint32_t ldc, bool split, hipStream_t st) {
  uint32_t zs = 1, kchunk = (K + 31) / 32 * 32;
  if (split) {  // K = batch: spread over about 256 blocks with atomics
    const uint32_t tiles = cdiv(M, 64) * cdiv(N, 64);
    zs = std::max(1u, std::min(cdiv(K, 256), 256u / std::max(tiles, 1u)));
    kchunk = (cdiv(K, zs) + 31) / 32 * 32;
    zs = cdiv(K, kchunk);
  }
  hipLaunchKernelGGL((k_gemm_mfma<TA, TB>), dim3(cdiv(N, 64), cdiv(M, 64), zs), dim3(256), 0, st, M, N, K, A, lda,
                     B, ldb, C, ldc, kchunk, split ? 1 : 0);
}

const float *guide_train_raw(const float *ws, uint32_t B) {
  return ws + size_t(B) * (kGFeatPad + 8 * kGHid);
}

size_t guide_train_bytes(uint32_t B) {
  return size_t(B) * (kGFeatPad + 9 * kGHid + 2 * kGOut + 4) * 4 + 64;
}

// One forward + backward over a prepared batch (X0, tgt in ws): grad
// (kGParams, zeroed here) and scal[1] (loss) are written.  Weights come from
// thq (bf16-rounded), biases from th (fp32), as in the MFMA network.
static void guide_fwd_bwd(const float *th, const float *thq, float *ws, uint32_t B, float *grad, float *scal,
                          hipStream_t st) {
  float *X0 = ws;
  float *Z[5], *X[5];
  float *p = X0 + size_t(B) * kGFeatPad;
  for (int l = 0; l < 4; l++) {
    Z[l] = p; p += size_t(B) * kGHid;
    X[l] = p; p += size_t(B) * kGHid;
  }
  Z[4] = p; p += size_t(B) * kGOut;
  float *dZo = p; p += size_t(B) * kGOut;
  float *dA = p; p += size_t(B) * kGHid;
  float4 *tgt = reinterpret_cast<float4 *>(p);
  const int Wo[5] = {kGW1, kGW2, kGW3, kGW4, kGW5}, Bo[5] = {kGB1, kGB2, kGB3, kGB4, kGB5};
  const uint32_t In[5] = {kGFeat, kGHid, kGHid, kGHid, kGHid}, Out[5] = {kGHid, kGHid, kGHid, kGHid, kGOut};
  const uint32_t ldx[5] = {kGFeatPad, kGHid, kGHid, kGHid, kGHid};
  // forward: Z_l = X_{l-1} W_lᵀ + b_l
  for (int l = 0; l < 5; l++) {
    const float *Xin = l == 0 ? X0 : X[l - 1];
    gemm<false, true>(B, Out[l], In[l], Xin, ldx[l], thq + Wo[l], In[l], Z[l], Out[l], false, st);
    hipLaunchKernelGGL(k_bias_act, dim3(cdiv(size_t(B) * Out[l], 256)), dim3(256), 0, st, Z[l], th + Bo[l], B,
                       Out[l], l < 4 ? X[l] : nullptr, l < 4 ? 1 : 0);
  }
  (void)launch_zero(scal, 16, st);  // whole 16-byte words
  (void)launch_zero(grad, size_t(kGParams) * 4, st);
  hipLaunchKernelGGL(k_sum_w, dim3(std::min(cdiv(B, 256), 64u)), dim3(256), 0, st, tgt, B, scal);
  hipLaunchKernelGGL(k_guide_loss, dim3(cdiv(B, 256)), dim3(256), 0, st, Z[4], tgt, B, scal, dZo);
  // backward
  const float *dZ = dZo;
  for (int l = 4; l >= 0; l--) {
    const float *Xin = l == 0 ? X0 : X[l - 1];
    // dW_l = dZ_lᵀ X_{l-1}
    gemm<true, false>(Out[l], In[l], B, dZ, Out[l], Xin, ldx[l], grad + Wo[l], In[l], true, st);
    hipLaunchKernelGGL(k_colsum, dim3(Out[l], std::min(cdiv(B, 256), 16u)), dim3(256), 0, st, dZ, B, Out[l],
                       grad + Bo[l]);
    if (l > 0) {
      // dX_{l-1} = dZ_l W_l, then through the ReLU of layer l-1
      gemm<false, false>(B, In[l], Out[l], dZ, Out[l], thq + Wo[l], In[l], dA, kGHid, false, st);
      // X_{l-1} is no longer needed once dW_l is formed: it takes dZ_{l-1}
      float *dZn = X[l - 1];
      hipLaunchKernelGGL(k_relu_grad, dim3(cdiv(size_t(B) * kGHid, 256)), dim3(256), 0, st, dA, Z[l - 1],
                         B * kGHid, dZn);
      dZ = dZn;
    }
  }
}

hipError_t guide_pack(const float *th, void *pack, float *thq, hipStream_t st) {
  uint4 *frag = static_cast<uint4 *>(pack);
  float *bias = reinterpret_cast<float *>(frag + kGFrags * 64);
  const uint32_t n = std::max<uint32_t>(kGFrags * 64, kGParams);
  hipLaunchKernelGGL(k_guide_pack, dim3(cdiv(n, 256)), dim3(256), 0, st, th, frag, bias, thq);
  return hipGetLastError();
}

hipError_t guide_infer(const void *pack, const float *in, uint32_t n, float *raw, uint32_t blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_guide_infer, dim3(blocks), dim3(256), 0, st, static_cast<const uint4 *>(pack), in, n, raw);
  return hipGetLastError();
}

hipError_t guide_grad_rows(const float *th, const float *thq, const float *rows, uint32_t B, float *ws,
                           float *grad, float *scal, hipStream_t st) {
  float4 *tgt = reinterpret_cast<float4 *>(ws + size_t(B) * (kGFeatPad + 8 * kGHid + 2 * kGOut + kGHid));
  hipLaunchKernelGGL(k_guide_feat_rows, dim3(cdiv(B, 256)), dim3(256), 0, st, rows, B, ws, tgt);
  guide_fwd_bwd(th, thq, ws, B, grad, scal, st);
  return hipGetLastError();
}

hipError_t guide_train_records(GuideTrain &gt, const float4 *rec, const uint32_t *rec_count, uint32_t rec_cap,
                               uint32_t steps, hipStream_t st) {
  const uint32_t B = gt.batch;
  float4 *tgt = reinterpret_cast<float4 *>(gt.ws + size_t(B) * (kGFeatPad + 8 * kGHid + 2 * kGOut + kGHid));
  for (uint32_t s = 0; s < steps; s++) {
    gt.step++;
    hipLaunchKernelGGL(k_guide_gather, dim3(cdiv(B, 256)), dim3(256), 0, st, rec, rec_count, rec_cap, B,
                       gt.seed ^ (uint64_t(gt.step) << 20), gt.ws, tgt);
    guide_fwd_bwd(gt.th, gt.thq, gt.ws, B, gt.grad, gt.scal, st);
    const float c1 = 1.f - powf(gt.b1, float(gt.step)), c2 = 1.f - powf(gt.b2, float(gt.step));
    hipLaunchKernelGGL(k_adam, dim3(cdiv(kGParams, 256)), dim3(256), 0, st, gt.th, gt.m, gt.v, gt.grad,
                       uint32_t(kGParams), gt.lr, gt.b1, gt.b2, gt.eps, c1, c2);
    (void)guide_pack(gt.th, gt.pack, gt.thq, st);
  }
  return hipGetLastError();
}

hipError_t guide_train_rows(GuideTrain &gt, const float *rows, uint32_t B, uint32_t steps, hipStream_t st) {
  for (uint32_t s = 0; s < steps; s++) {
    gt.step++;
    (void)guide_grad_rows(gt.th, gt.thq, rows, B, gt.ws, gt.grad, gt.scal, st);
    const float c1 = 1.f - powf(gt.b1, float(gt.step)), c2 = 1.f - powf(gt.b2, float(gt.step));
    hipLaunchKernelGGL(k_adam, dim3(cdiv(kGParams, 256)), dim3(256), 0, st, gt.th, gt.m, gt.v, gt.grad,
                       uint32_t(kGParams), gt.lr, gt.b1, gt.b2, gt.eps, c1, c2);
    (void)guide_pack(gt.th, gt.pack, gt.thq, st);
  }
  return hipGetLastError();
}

}  // namespace lam

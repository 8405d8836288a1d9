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
// distribution of radiance (times the plain continuation's weight).
//
// One Adam step is two kernels.  k_guide_fwdbwd runs the forward and the
// backward of 64 batch rows per wave, entirely on bf16 MFMA with fp32
// accumulation: the forward is the inference network's own evaluation (the
// same fragments, features and ReLU packing, so its outputs are the
// inference outputs bit for bit); the backward takes each layer's dZ,
// rounded to bf16, as an MFMA operand in registers for dX = dZ W (against
// fragments of the transposed weights) and, through a transpose in LDS, for
// dW = dZᵀ X over the wave's 64 rows; the ReLU masks are kept from the
// forward as bits.  Each block (one wave, weights and activations in LDS)
// writes its gradient sums to a partial buffer.  k_guide_update then sums the
// partials in a fixed order, applies Adam to the fp32 master parameters and
// writes each updated parameter into the packed bf16 fragments in place; its
// threads also draw the next step's batch from the records.  Sums are in a
// fixed order throughout, so training is deterministic.
#include <algorithm>

#include "lam_guide.h"

namespace lam {

// ---- packing ------------------------------------------------------------------------
// one parameter into its packed places: forward fragment (or bias) and, for
// the weights of layers 2-5, the backward's transposed fragment
__device__ inline void pack_param(uint4 *__restrict__ pack, int p, float w) {
  uint16_t *h = reinterpret_cast<uint16_t *>(pack);
  float *bias = reinterpret_cast<float *>(pack + kGFrags * 64);
  uint16_t *hb = reinterpret_cast<uint16_t *>(reinterpret_cast<char *>(pack) + kGPackFwdBytes);
  const uint16_t b = f2bf(w);
  auto put = [](uint16_t *base, int f, int g, int n, int j, uint16_t v) { base[(f * 64 + 16 * g + n) * 8 + j] = v; };
  if (p < kGB1) {
    const int row = p / kGFeat, col = p % kGFeat;
    put(h, kGFragL1 + (row / 16) * 5 + col / 32, (col % 32) / 8, row % 16, col % 8, b);
  } else if (p < kGW2) {
    bias[p - kGB1] = w;
  } else if (p < kGW5) {
    const int l = (p - kGW2) / (kGHid * kGHid + kGHid), idx = (p - kGW2) % (kGHid * kGHid + kGHid);
    if (idx >= kGHid * kGHid) {
      bias[kGHid * (l + 1) + idx - kGHid * kGHid] = w;
      return;
    }
    const int row = idx / kGHid, col = idx % kGHid;
    const int kf = g_perm_inv(col), kb = g_perm_inv(row);
    put(h, kGFragL2 + 8 * l + 2 * (row / 16) + kf / 32, (kf >> 3) & 3, row % 16, kf & 7, b);
    put(hb, kGBFragL2 + 8 * l + 2 * (col / 16) + kb / 32, (kb >> 3) & 3, col % 16, kb & 7, b);
  } else if (p < kGB5) {
    const int row = (p - kGW5) / kGHid, col = (p - kGW5) % kGHid;
    const int kf = g_perm_inv(col);
    put(h, kGFragL5 + 2 * (row / 16) + kf / 32, (kf >> 3) & 3, row % 16, kf & 7, b);
    put(hb, kGBFragL5 + 2 * (col / 16) + row / 32, (row % 32) / 8, col % 16, row % 8, b);
  } else {
    bias[4 * kGHid + p - kGB5] = w;
  }
}

// the whole pack from fp32 parameters, padding included (zero)
__global__ void k_guide_pack(const float *__restrict__ th, uint4 *__restrict__ pack) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  uint4 *bpack = reinterpret_cast<uint4 *>(reinterpret_cast<char *>(pack) + kGPackFwdBytes);
  if (tid < uint32_t(kGFrags + kGBFrags) * 64) {
    const bool fwd = tid < uint32_t(kGFrags) * 64;
    const uint32_t f = fwd ? tid / 64 : tid / 64 - kGFrags, lane = tid % 64, n = lane % 16, g = lane / 16;
    uint16_t v[8];
    for (int j = 0; j < 8; j++) {
      float w = 0.f;
      if (!fwd) {  // transposed weights: row tiles over inputs, K over outputs
        const uint32_t l = f / 8, t = (f % 8) / 2, c = f % 2, in = 16 * t + n;
        if (f < uint32_t(kGBFragL5)) {
          const int base = l == 0 ? kGW2 : (l == 1 ? kGW3 : kGW4);
          w = th[base + g_perm(int(32 * c + 8 * g + j)) * kGHid + in];
        } else {
          const uint32_t out = 32 * c + 8 * g + j;
          if (out < uint32_t(kGOut)) w = th[kGW5 + out * kGHid + in];
        }
      } else if (f < uint32_t(kGFragL2)) {  // layer 1: natural feature order
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
    (fwd ? pack : bpack)[fwd ? tid : tid - uint32_t(kGFrags) * 64] =
        make_uint4(v[0] | (uint32_t(v[1]) << 16), v[2] | (uint32_t(v[3]) << 16), v[4] | (uint32_t(v[5]) << 16),
                   v[6] | (uint32_t(v[7]) << 16));
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
    reinterpret_cast<float *>(pack + kGFrags * 64)[tid] = b;
  }
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
constexpr uint32_t kGTrainBlocks = 256;  // blocks (one wave each) of k_guide_fwdbwd
constexpr int kGParamsPad = (kGParams + 63) / 64 * 64;
constexpr int kGRow = 10;  // batch row: in[6], ω[3], w
constexpr int kXS = 72;    // LDS row stride (halves) of transposed activations
constexpr int kZS = 49;    // LDS row stride (floats) of the output exchange

__device__ inline uint32_t hash32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return uint32_t(x);
}

// Batch row i of a training step drawn from the records (uniform, with
// replacement), and the batch's Σw per block of 256 rows (a fixed-order tree
// in LDS).  Every thread of a block must call it.
struct Gather {
  const float4 *rec;
  const uint32_t *rec_count;
  uint32_t rec_cap, B;
  uint64_t seed;
  float *rows, *wpart;
};
__device__ inline void gather_row(const Gather &ga, uint32_t i, float *red) {
  float w = 0.f;
  if (i < ga.B) {
    const uint32_t nr = min(*ga.rec_count, ga.rec_cap);
    float v[kGRow] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f};
    if (nr > 0) {
      const uint32_t r = uint32_t((uint64_t(hash32(ga.seed * 0x9E3779B97F4A7C15ull + i)) * nr) >> 32);
      const float4 a = ga.rec[4 * size_t(r)], b = ga.rec[4 * size_t(r) + 1], c = ga.rec[4 * size_t(r) + 2],
                   d = ga.rec[4 * size_t(r) + 3];
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
      v[6] = c.x; v[7] = c.y; v[8] = c.z; v[9] = d.z;
    }
    for (int k = 0; k < kGRow; k++) ga.rows[size_t(i) * kGRow + k] = v[k];
    w = v[9];
  }
  const uint32_t t = threadIdx.x;
  red[t] = w;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  if (t == 0 && blockIdx.x * 256 < ga.B) ga.wpart[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_guide_gather(Gather ga) {
  __shared__ float red[256];
  gather_row(ga, blockIdx.x * 256 + threadIdx.x, red);
}

// Σw per block of 256 rows of a host batch
__global__ __launch_bounds__(256) void k_guide_rows_w(const float *__restrict__ rows, uint32_t B,
                                                      float *__restrict__ wpart) {
  __shared__ float red[256];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x, t = threadIdx.x;
  red[t] = i < B ? rows[size_t(i) * kGRow + 9] : 0.f;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  if (t == 0) wpart[blockIdx.x] = red[0];
}

// bf16 half j (0..7) of a packed fragment word
__device__ inline uint16_t half_of(const uint4 &v, int j) {
  const uint32_t w = j < 2 ? v.x : (j < 4 ? v.y : (j < 6 ? v.z : v.w));
  return uint16_t((j & 1) ? (w >> 16) : (w & 0xFFFFu));
}

// Loss term of one row and its gradient with respect to the 40 raw outputs
// (r in, d out), scaled by wn = w / Σw; returns -wn log p.
__device__ inline float row_loss(const float *r, float wn, f3 om, float *d) {
  float a_max = -3.4e38f;
#pragma unroll
  for (int L = 0; L < kGLobes; L++) a_max = fmaxf(a_max, r[5 * L + 4]);
  float ws[kGLobes], lv[kGLobes], kap[kGLobes], md[kGLobes], vlen[kGLobes];
  f3 mu[kGLobes];
  float wsum = 0.f;
#pragma unroll
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
#pragma unroll
  for (int L = 0; L < kGLobes; L++) {
    z[L] = logf(ws[L] / wsum) + lv[L];
    zmax = fmaxf(zmax, z[L]);
  }
  float zs = 0.f;
#pragma unroll
  for (int L = 0; L < kGLobes; L++) zs += expf(z[L] - zmax);
  const float logp = zmax + logf(zs);
#pragma unroll
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
  return -wn * logp;
}

// Store (the block's first chunk) or add a batch of gradient sums at
// bp[idx[u]] (idx < 0: none); a batch's loads are issued before its adds (a
// block is one wave, with nothing else to hide their latency).
template <int N>
__device__ inline void put_n(float *__restrict__ bp, bool first, const int (&idx)[N], const float (&val)[N]) {
  float old[N];
#pragma unroll
  for (int u = 0; u < N; u++) old[u] = (!first && idx[u] >= 0) ? bp[idx[u]] : 0.f;
#pragma unroll
  for (int u = 0; u < N; u++)
    if (idx[u] >= 0) bp[idx[u]] = first ? val[u] : old[u] + val[u];
}

// Forward and backward of the batch `rows` (B rows, kGRow floats each), 64
// rows per block of four waves: wave G carries rows 16G .. 16G+15 of the
// chunk through the forward and the dX products (its 16 MFMA columns), and
// output row tile G of every dW product (over all 64 rows, from LDS).
// Block b takes chunks b, b + grid, ... and writes its gradient sums to
// part[b][·] (stored by its first chunk, added by the others) and its loss
// sum to lpart[b].  Σw is the fixed-order sum of wpart[0 .. nw-1]; block 0
// stores it in scal[0].  raw (optional): the forward's outputs, B x 40.
__global__ __launch_bounds__(256) void k_guide_fwdbwd(const uint4 *__restrict__ pack, const float *__restrict__ rows,
                                                      uint32_t B, const float *__restrict__ wpart, uint32_t nw,
                                                      float *__restrict__ part, float *__restrict__ lpart,
                                                      float *__restrict__ scal, float *__restrict__ raw) {
  __shared__ uint4 sfrag[kGFrags * 64];
  __shared__ uint4 sbfrag[kGBFrags * 64];
  __shared__ float sbias[kGBiasWords];
  // X1ᵀ..X4ᵀ as [neuron][row] (bf16); X0ᵀ (160 features) once the others are used
  __shared__ __attribute__((aligned(16))) uint16_t xt[4 * 64 * kXS];
  __shared__ __attribute__((aligned(16))) uint16_t dzt[64 * kXS];  // bf16(dZ)ᵀ of the current layer
  __shared__ float zs[64 * kZS];                                    // layer-5 outputs, then dZ5, per row
  __shared__ float dbs[4 * 64];                                     // per-wave bias-gradient sums
  __shared__ float s_sw;
  const uint32_t tid = threadIdx.x, G = tid >> 6, lane = tid & 63u, n = lane & 15u, g = lane >> 4;
  {
    const float *gb = reinterpret_cast<const float *>(pack + kGFrags * 64);
    const uint4 *bf = reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(pack) + kGPackFwdBytes);
    auto stage = [&](const uint4 *src, uint4 *dst, uint32_t cnt) {  // 8 loads in flight per thread
      for (uint32_t i0 = 0; i0 < cnt; i0 += 8 * 256) {
        uint4 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const uint32_t i = i0 + 256 * u + tid;
          v[u] = i < cnt ? src[i] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const uint32_t i = i0 + 256 * u + tid;
          if (i < cnt) dst[i] = v[u];
        }
      }
    };
    stage(pack, sfrag, uint32_t(kGFrags) * 64);
    stage(bf, sbfrag, uint32_t(kGBFrags) * 64);
    for (uint32_t i = tid; i < uint32_t(kGBiasWords); i += 256) sbias[i] = gb[i];
    if (tid == 0) {
      float sw = 0.f;  // in index order; the loads 16 at a time
      for (uint32_t k0 = 0; k0 < nw; k0 += 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = k0 + u < nw ? wpart[k0 + u] : 0.f;
#pragma unroll
        for (int u = 0; u < 16; u++) sw += v[u];
      }
      s_sw = sw;
      if (blockIdx.x == 0) scal[0] = sw;
    }
  }
  __syncthreads();
  const float sw = s_sw;
  float *bp = part + size_t(blockIdx.x) * kGParamsPad;
  float lsum = 0.f;
  bool first = true;
  for (uint32_t chunk = blockIdx.x; chunk * 64u < B; chunk += gridDim.x, first = false) {
    const uint32_t rq = 16 * G + n;  // this lane's row in the chunk (shared by the lanes g)
    const uint32_t qi = chunk * 64u + rq;
    float in[kGIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    f3 om = mk(0.f, 0.f, 1.f);
    float w = 0.f;
    if (qi < B) {
      const float *r = rows + size_t(qi) * kGRow;
#pragma unroll
      for (int d = 0; d < kGIn; d++) in[d] = r[d];
      om = mk(r[6], r[7], r[8]);
      w = r[9];
    }
    const GIn q{in[0], in[1], in[2], in[3], in[4], in[5]};
    // ---- forward (mlp_wave for one column group, keeping masks and activations) ----
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int i = 0; i < 4; i++) acc[t][i] = sbias[16 * t + 4 * g + i];
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const uint4 xb = feature_chunk(c, g, q);
#pragma unroll
      for (int t = 0; t < 4; t++) acc[t] = mfma16(sfrag[(kGFragL1 + t * 5 + c) * 64 + lane], xb, acc[t]);
    }
    uint32_t msk[4];  // Z > 0 of hidden layers 1..4, bit 4 t + i
    f32x4 o[3];
#pragma unroll
    for (int L = 0; L < 4; L++) {
      uint32_t mb = 0;
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (acc[t][i] > 0.f) mb |= 1u << (4 * t + i);
      msk[L] = mb;
      uint4 xb[2];
      relu_pack(acc, xb);
      // X_{L+1}ᵀ: K position 32c + 8g + j of xb[c] is neuron g_perm of it
      uint16_t *X = xt + L * 64 * kXS;
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int j = 0; j < 8; j++) X[(32 * c + 16 * (j >> 2) + 4 * int(g) + (j & 3)) * kXS + rq] = half_of(xb[c], j);
      if (L < 3) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
#pragma unroll
          for (int i = 0; i < 4; i++) acc[t][i] = sbias[kGHid * (L + 1) + 16 * t + 4 * g + i];
#pragma unroll
          for (int c = 0; c < 2; c++) acc[t] = mfma16(sfrag[(kGFragL2 + 8 * L + t * 2 + c) * 64 + lane], xb[c], acc[t]);
        }
      } else {
#pragma unroll
        for (int t = 0; t < 3; t++) {
#pragma unroll
          for (int i = 0; i < 4; i++) o[t][i] = sbias[4 * kGHid + 16 * t + 4 * g + i];
#pragma unroll
          for (int c = 0; c < 2; c++) o[t] = mfma16(sfrag[(kGFragL5 + t * 2 + c) * 64 + lane], xb[c], o[t]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 3; t++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = 16 * t + 4 * int(g) + i;
        if (row < kGOut) zs[rq * kZS + row] = o[t][i];
      }
    __syncthreads();
    // ---- loss, one row per lane g = 0; dZ5 = ∂loss/∂raw ----
    if (g == 0) {
      float r[kGOut], d[kGOut];
#pragma unroll
      for (int k = 0; k < kGOut; k++) r[k] = zs[rq * kZS + k];
      if (raw && qi < B)
#pragma unroll
        for (int k = 0; k < kGOut; k++) raw[size_t(qi) * kGOut + k] = r[k];
      const float wn = sw > 0.f ? w / sw : 0.f;
      if (wn != 0.f) {
        lsum += row_loss(r, wn, om, d);
      } else {
#pragma unroll
        for (int k = 0; k < kGOut; k++) d[k] = 0.f;
      }
#pragma unroll
      for (int k = 0; k < kGOut; k++) zs[rq * kZS + k] = d[k];
      // bf16(dZ5)ᵀ, rows 40..47 zero
#pragma unroll
      for (int k = 0; k < kGOutPad; k += 2) {
        const uint32_t wd = pack_bf2(k < kGOut ? d[k] : 0.f, k + 1 < kGOut ? d[k + 1] : 0.f);
        dzt[k * kXS + rq] = uint16_t(wd & 0xFFFFu);
        dzt[(k + 1) * kXS + rq] = uint16_t(wd >> 16);
      }
    }
    __syncthreads();
    if (tid < uint32_t(kGOut)) {  // db5 from the unrounded dZ5
      float s = 0.f;
      for (int r = 0; r < 64; r++) s += zs[r * kZS + tid];
      put_n<1>(bp, first, {kGB5 + int(tid)}, {s});
    }
    if (G < 3) {  // dW5 = dZ5ᵀ X4, output rows 16G .. 16G+15 (< 40)
      f32x4 a4[4];
#pragma unroll
      for (int ti = 0; ti < 4; ti++) a4[ti] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const uint4 a = *reinterpret_cast<const uint4 *>(&dzt[(16 * G + n) * kXS + 32 * c + 8 * g]);
#pragma unroll
        for (int ti = 0; ti < 4; ti++) {
          const uint4 b = *reinterpret_cast<const uint4 *>(&xt[(3 * 64 + 16 * ti + n) * kXS + 32 * c + 8 * g]);
          a4[ti] = mfma16(a, b, a4[ti]);
        }
      }
      int idx[16];
      float val[16];
#pragma unroll
      for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int row = 16 * int(G) + 4 * int(g) + i;
          idx[4 * ti + i] = row < kGOut ? kGW5 + row * kGHid + 16 * ti + int(n) : -1;
          val[4 * ti + i] = a4[ti][i];
        }
      put_n(bp, first, idx, val);
    }
    // dX4ᵀ = W5ᵀ dZ5ᵀ: B operand bf16(dZ5) of this wave's rows, outputs 32c + 8g + j
    f32x4 dz[4];
    {
      uint4 zb[2];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int k = 32 * c + 8 * int(g) + j;
          v[j] = k < kGOut ? zs[rq * kZS + k] : 0.f;
        }
        zb[c] = make_uint4(pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]), pack_bf2(v[4], v[5]), pack_bf2(v[6], v[7]));
      }
#pragma unroll
      for (int t = 0; t < 4; t++) {
        dz[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 2; c++) dz[t] = mfma16(sbfrag[(kGBFragL5 + 2 * t + c) * 64 + lane], zb[c], dz[t]);
      }
    }
    // ---- hidden layers l = 4 .. 1: dz holds dX_l; through the ReLU to dZ_l ----
#pragma unroll 1
    for (int l = 4; l >= 1; l--) {
      const uint32_t mb = msk[l - 1];
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (!((mb >> (4 * t + i)) & 1u)) dz[t][i] = 0.f;
      // db_l (unrounded): this wave's 16 rows (lanes n), then the waves in order
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          float s = dz[t][i];
          s += __shfl_xor(s, 8, 64);
          s += __shfl_xor(s, 4, 64);
          s += __shfl_xor(s, 2, 64);
          s += __shfl_xor(s, 1, 64);
          if (n == 0) dbs[64 * G + 16 * t + 4 * g + i] = s;
        }
      // bf16(dZ_l): the B operand of dX_{l-1} (K permuted as the forward's)
      // and its transpose for dW_l
      uint4 zb[2];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const f32x4 a = dz[2 * c], b = dz[2 * c + 1];
        zb[c] = make_uint4(pack_bf2(a[0], a[1]), pack_bf2(a[2], a[3]), pack_bf2(b[0], b[1]), pack_bf2(b[2], b[3]));
#pragma unroll
        for (int j = 0; j < 8; j++)
          dzt[(32 * c + 16 * (j >> 2) + 4 * int(g) + (j & 3)) * kXS + rq] = half_of(zb[c], j);
      }
      if (l == 1) {
        // X0ᵀ (the forward's features) over the activations, which are no
        // longer needed
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 5; c++) {
          const uint4 f = feature_chunk(c, g, q);
#pragma unroll
          for (int j = 0; j < 8; j++) xt[(32 * c + 8 * int(g) + j) * kXS + rq] = half_of(f, j);
        }
      }
      __syncthreads();
      if (tid < 64) {
        const int bb = l == 1 ? kGB1 : (l == 2 ? kGB2 : (l == 3 ? kGB3 : kGB4));
        put_n<1>(bp, first, {bb + int(tid)}, {((dbs[tid] + dbs[64 + tid]) + dbs[128 + tid]) + dbs[192 + tid]});
      }
      // dW_l = dZ_lᵀ X_{l-1}, output rows 16G .. 16G+15
      const uint16_t *Xp = l == 1 ? xt : xt + (l - 2) * 64 * kXS;
      const int wbase = l == 1 ? kGW1 : (l == 2 ? kGW2 : (l == 3 ? kGW3 : kGW4));
      const int nti = l == 1 ? 10 : 4, ld = l == 1 ? kGFeat : kGHid;
      {
        f32x4 aw[10];
#pragma unroll
        for (int ti = 0; ti < 10; ti++) aw[ti] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const uint4 a = *reinterpret_cast<const uint4 *>(&dzt[(16 * G + n) * kXS + 32 * c + 8 * g]);
#pragma unroll
          for (int ti = 0; ti < 10; ti++)
            if (ti < nti) {
              const uint4 b = *reinterpret_cast<const uint4 *>(&Xp[(16 * ti + n) * kXS + 32 * c + 8 * g]);
              aw[ti] = mfma16(a, b, aw[ti]);
            }
        }
        int idx[40];
        float val[40];
#pragma unroll
        for (int ti = 0; ti < 10; ti++) {
          const int col = 16 * ti + int(n);
#pragma unroll
          for (int i = 0; i < 4; i++) {
            idx[4 * ti + i] = ti < nti && col < ld ? wbase + (16 * int(G) + 4 * int(g) + i) * ld + col : -1;
            val[4 * ti + i] = aw[ti][i];
          }
        }
        put_n(bp, first, idx, val);
      }
      if (l > 1) {
        // dX_{l-1}ᵀ = W_lᵀ dZ_lᵀ
#pragma unroll
        for (int t = 0; t < 4; t++) {
          f32x4 nx = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int c = 0; c < 2; c++) nx = mfma16(sbfrag[(kGBFragL2 + 8 * (l - 2) + 2 * t + c) * 64 + lane], zb[c], nx);
          dz[t] = nx;
        }
      }
      __syncthreads();  // dzt, dbs and xt are rewritten next
    }
  }
  // the block's loss sum: a fixed-order tree over each wave, then the waves in order
  lsum += __shfl_xor(lsum, 32, 64);
  lsum += __shfl_xor(lsum, 16, 64);
  lsum += __shfl_xor(lsum, 8, 64);
  lsum += __shfl_xor(lsum, 4, 64);
  lsum += __shfl_xor(lsum, 2, 64);
  lsum += __shfl_xor(lsum, 1, 64);
  if (lane == 0) dbs[G] = lsum;
  __syncthreads();
  if (tid == 0) lpart[blockIdx.x] = ((dbs[0] + dbs[1]) + dbs[2]) + dbs[3];
}

// The step's gradient (the partials summed in a fixed order) into grad
// (optional) and, with adam, the Adam update of each parameter written to
// th and into the pack; thread 0 stores the loss in scal[1].  A block takes
// 64 parameters, its four waves a quarter of the partial rows each (16 loads
// in flight per thread), then one sum over the quarters in LDS.  Threads
// also draw the next step's batch (ga.B > 0).
__global__ __launch_bounds__(256) void k_guide_update(const float *__restrict__ part, uint32_t nblk,
                                                      const float *__restrict__ lpart, float *__restrict__ grad,
                                                      float *__restrict__ th, float *__restrict__ m,
                                                      float *__restrict__ v, uint4 *__restrict__ pack, int adam,
                                                      float lr, float b1, float b2, float eps, float c1, float c2,
                                                      float *__restrict__ scal, Gather ga) {
  __shared__ float red[256];
  const uint32_t pl = threadIdx.x % 64, quarter = threadIdx.x / 64;
  const uint32_t j = blockIdx.x * 64 + pl;
  const uint32_t r_begin = quarter * ((nblk + 3) / 4), r_end = min(nblk, r_begin + (nblk + 3) / 4);
  float gs = 0.f;
  if (j < uint32_t(kGParams))
    for (uint32_t b0 = r_begin; b0 < r_end; b0 += 16) {
      float x[16];
#pragma unroll
      for (int u = 0; u < 16; u++) x[u] = b0 + u < r_end ? part[size_t(b0 + u) * kGParamsPad + j] : 0.f;
#pragma unroll
      for (int u = 0; u < 16; u++) gs += x[u];
    }
  red[threadIdx.x] = gs;
  __syncthreads();
  if (quarter == 0 && j < uint32_t(kGParams)) {
    gs = ((red[pl] + red[64 + pl]) + red[128 + pl]) + red[192 + pl];
    if (grad) grad[j] = gs;
    if (adam) {
      const float mi = b1 * m[j] + (1.f - b1) * gs;
      const float vi = b2 * v[j] + (1.f - b2) * gs * gs;
      m[j] = mi;
      v[j] = vi;
      const float t = th[j] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
      th[j] = t;
      pack_param(pack, int(j), t);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (uint32_t b = 0; b < nblk; b++) s += lpart[b];
    scal[1] = s;
  }
  __syncthreads();  // red is reused by the batch draw
  if (ga.B > 0 && blockIdx.x * 256 < ga.B) gather_row(ga, blockIdx.x * 256 + threadIdx.x, red);
}

// ---- host-side drivers ------------------------------------------------------------------
static inline uint32_t cdiv(size_t a, uint32_t b) { return uint32_t((a + b - 1) / b); }

// workspace: batch rows, Σw per 256 rows, gradient and loss partials
struct TrainWs {
  float *rows, *wpart, *part, *lpart;
};
static TrainWs train_ws(float *ws, uint32_t B) {
  TrainWs t;
  t.rows = ws;
  t.wpart = t.rows + (size_t(B) * kGRow + 63) / 64 * 64;
  t.part = t.wpart + (cdiv(B, 256) + 63) / 64 * 64;
  t.lpart = t.part + size_t(kGTrainBlocks) * kGParamsPad;
  return t;
}

size_t guide_train_bytes(uint32_t B) {
  return (size_t(B) * kGRow + 64 + cdiv(B, 256) + 64 + size_t(kGTrainBlocks) * kGParamsPad + kGTrainBlocks) * 4 +
         64;
}

static hipError_t fwdbwd(const GuideTrain &gt, const float *rows, uint32_t B, const float *wpart, float *part,
                         float *lpart, float *raw, hipStream_t st, uint32_t &nblk) {
  nblk = std::min(cdiv(B, 64), kGTrainBlocks);
  hipLaunchKernelGGL(k_guide_fwdbwd, dim3(nblk), dim3(256), 0, st, static_cast<const uint4 *>(gt.pack), rows, B,
                     wpart, cdiv(B, 256), part, lpart, gt.scal, raw);
  return hipGetLastError();
}

static hipError_t update(GuideTrain &gt, uint32_t nblk, const TrainWs &t, float *grad, bool adam, const Gather &ga,
                         hipStream_t st) {
  float c1 = 1.f, c2 = 1.f;
  if (adam) {
    gt.step++;
    c1 = 1.f - powf(gt.b1, float(gt.step));
    c2 = 1.f - powf(gt.b2, float(gt.step));
  }
  const uint32_t blocks = std::max<uint32_t>(cdiv(kGParams, 64), cdiv(ga.B, 256));
  hipLaunchKernelGGL(k_guide_update, dim3(blocks), dim3(256), 0, st, t.part, nblk, t.lpart, grad,
                     gt.th, gt.m, gt.v, static_cast<uint4 *>(gt.pack), adam ? 1 : 0, gt.lr, gt.b1, gt.b2, gt.eps,
                     c1, c2, gt.scal, ga);
  return hipGetLastError();
}

hipError_t guide_pack(const float *th, void *pack, hipStream_t st) {
  const uint32_t n = std::max<uint32_t>((kGFrags + kGBFrags) * 64, kGBiasWords);
  hipLaunchKernelGGL(k_guide_pack, dim3(cdiv(n, 256)), dim3(256), 0, st, th, static_cast<uint4 *>(pack));
  return hipGetLastError();
}

hipError_t guide_infer(const void *pack, const float *in, uint32_t n, float *raw, uint32_t blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_guide_infer, dim3(blocks), dim3(256), 0, st, static_cast<const uint4 *>(pack), in, n, raw);
  return hipGetLastError();
}

hipError_t guide_loss_grad_rows(GuideTrain &gt, const float *rows, uint32_t B, float *grad, float *raw,
                                hipStream_t st) {
  const TrainWs t = train_ws(gt.ws, B);
  hipLaunchKernelGGL(k_guide_rows_w, dim3(cdiv(B, 256)), dim3(256), 0, st, rows, B, t.wpart);
  uint32_t nblk = 0;
  hipError_t e = fwdbwd(gt, rows, B, t.wpart, t.part, t.lpart, raw, st, nblk);
  if (e != hipSuccess) return e;
  return update(gt, nblk, t, grad, false, Gather{}, st);
}

hipError_t guide_train_rows(GuideTrain &gt, const float *rows, uint32_t B, uint32_t steps, hipStream_t st) {
  const TrainWs t = train_ws(gt.ws, B);
  hipLaunchKernelGGL(k_guide_rows_w, dim3(cdiv(B, 256)), dim3(256), 0, st, rows, B, t.wpart);
  for (uint32_t s = 0; s < steps; s++) {
    uint32_t nblk = 0;
    hipError_t e = fwdbwd(gt, rows, B, t.wpart, t.part, t.lpart, nullptr, st, nblk);
    if (e == hipSuccess) e = update(gt, nblk, t, nullptr, true, Gather{}, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t guide_train_records(GuideTrain &gt, const float4 *rec, const uint32_t *rec_count, uint32_t rec_cap,
                               uint32_t steps, hipStream_t st) {
  const uint32_t B = gt.batch;
  const TrainWs t = train_ws(gt.ws, B);
  auto batch = [&](uint32_t step) {
    return Gather{rec, rec_count, rec_cap, B, gt.seed ^ (uint64_t(step) << 20), t.rows, t.wpart};
  };
  if (steps == 0) return hipSuccess;
  hipLaunchKernelGGL(k_guide_gather, dim3(cdiv(B, 256)), dim3(256), 0, st, batch(gt.step + 1));
  for (uint32_t s = 0; s < steps; s++) {
    uint32_t nblk = 0;
    hipError_t e = fwdbwd(gt, t.rows, B, t.wpart, t.part, t.lpart, nullptr, st, nblk);
    // the update kernel draws the batch of the step after it
    if (e == hipSuccess) e = update(gt, nblk, t, nullptr, true, s + 1 < steps ? batch(gt.step + 2) : Gather{}, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace lam

// gfx950 kernels: extend / occlusion, path tracing (one path per lane),
// film accumulation and resolve.  Launch wrappers are in lam_launch.h.
#include "lam_launch.h"
#include "lam_shade.h"

namespace lam {

// ---------------------------------------------------------------------------
// extend / occlusion over a ray batch: the persistent queue traversal of the
// wavefront's extend and shadow kernels (split stacks, per-lane refill)
// ---------------------------------------------------------------------------
template <bool ANY, bool COUNT>
__global__ __launch_bounds__(64, 6) void k_trace_rays(DevScene sc, const lam_ray *__restrict__ rays, uint32_t n,
                                                   lam_hit *__restrict__ hits, uint32_t *__restrict__ occl,
                                                   uint32_t *__restrict__ counts, uint32_t *__restrict__ err,
                                                   uint32_t *__restrict__ fetch, uint32_t *__restrict__ spill) {
  __shared__ uint32_t stack[kLdsStack * 64];
  auto load = [&](uint32_t i) {
    const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
    const float4 a = rp[0], b = rp[1];
    RayItem it;
    it.r.o = mk(a.x, a.y, a.z);
    it.r.tmax = a.w;
    it.r.d = mk(b.x, b.y, b.z);
    it.r.skip = __float_as_uint(b.w);
    return it;
  };
  auto done = [&](uint32_t i, const RayItem &, const Trav &tv, bool ovf, const TravCounts &c) {
    const Hit &h = tv.best;
    if (ovf) atomicOr(err, 1u);
    if (ANY) {
      occl[i] = h.obj != LAM_NONE ? 1u : 0u;
    } else {
      // input index of the hit triangle (kept by the traversal: no fetch)
      const uint32_t tri = h.slot != LAM_NONE ? tv.orig : LAM_NONE;
      float4 o;
      o.x = h.obj == LAM_NONE ? kFltMax : h.t;
      o.y = __uint_as_float(tri);
      o.z = __uint_as_float(h.obj);
      o.w = __uint_as_float(h.obj == LAM_NONE ? LAM_NONE : h.prim);
      reinterpret_cast<float4 *>(hits)[i] = o;
    }
    if (COUNT) {
      counts[2 * i + 0] = c.nodes;
      counts[2 * i + 1] = c.tris;
    }
  };
  const SplitStack stk = SplitStack::of(stack, spill, blockIdx.x, gridDim.x);
  trace_queue<ANY, COUNT>(sc, n, fetch, stk, load, done);
}

// ---------------------------------------------------------------------------
// one path per lane: Renderer::tracePath (src/renderer.cpp:64-146) unrolled
// into a forward loop.  The recursion is linear, so
//   Li_s = M_IS,s * w_mat,s + L_IS,s * w_L,s
// becomes L += β_s ⊙ (w_L,s L_IS,s) and β_{s+1} = β_s ⊙ mult_s / prob_s * w_mat,s,
// equal to the reference up to floating-point reassociation.
// ---------------------------------------------------------------------------
template <int H>
__device__ void trace_path(const DevScene &sc, const PathArgs &pa, uint32_t pixel, uint32_t sample, float *Lout,
                           float &inv_omega, Lam<H> &lam, const SplitStack &stk, bool &ovf) {
  const uint64_t key = lam_path_key(pa.seed, pixel, sample);
  const uint32_t w = sc.cam.width;
  const uint32_t row = pixel / w, colj = pixel - row * w;
  // render (src/renderer.cpp:35-40)
  const float u = float(colj) + lam_draw_f(key, LAM_DIM_U);
  const float v = float(row) + lam_draw_f(key, LAM_DIM_V);
  Ray ray;
  camera_ray(sc.cam, u, v, lam_draw_f(key, LAM_DIM_LENS_X), lam_draw_f(key, LAM_DIM_LENS_Y), ray.o, ray.d);
  ray.tmax = kFltMax;
  ray.skip = LAM_NONE;
  inv_omega = hero_lambdas<H>(sc, lam_draw_f(key, LAM_DIM_HERO), lam);
  float beta[H];
#pragma unroll
  for (int k = 0; k < H; k++) { beta[k] = 1.f; Lout[k] = 0.f; }
  path_loop<H>(sc, pa.max_scat, pa.sample_p, key, lam, 0u, ray, beta, Lout, stk, ovf);
}

// grid-stride over the paths; spill holds the traversal-stack spill
// columns of the grid
template <int H>
__global__ __launch_bounds__(64) void k_trace_paths(DevScene sc, PathArgs pa, const uint32_t *__restrict__ pixel,
                                                    const uint32_t *__restrict__ sample, uint32_t n,
                                                    float *__restrict__ out, uint32_t *__restrict__ err,
                                                    uint32_t *__restrict__ spill) {
  __shared__ uint32_t stack[kLdsStack * 64];
  const uint32_t lane = threadIdx.x;
  const SplitStack stk = SplitStack::of(stack, spill, blockIdx.x, gridDim.x);
  for (uint32_t i = blockIdx.x * 64u + lane; i < n; i += gridDim.x * 64u) {
    float L[H];
    float inv_om;
    Lam<H> lam;
    bool ovf = false;
    trace_path<H>(sc, pa, pixel[i], sample[i], L, inv_om, lam, stk, ovf);
    if (ovf) atomicOr(err, 1u);
#pragma unroll
    for (int k = 0; k < H; k++) out[size_t(i) * H + k] = L[k];
  }
}

template <int H>
__global__ __launch_bounds__(64) void k_render_mega(DevScene sc, PathArgs pa, uint32_t sample, uint32_t npix,
                                                    float *__restrict__ film, uint32_t *__restrict__ err,
                                                    uint32_t *__restrict__ spill) {
  __shared__ uint32_t stack[kLdsStack * 64];
  const uint32_t lane = threadIdx.x;
  const SplitStack stk = SplitStack::of(stack, spill, blockIdx.x, gridDim.x);
  for (uint32_t pixel = blockIdx.x * 64u + lane; pixel < npix; pixel += gridDim.x * 64u) {
    float L[H];
    float inv_om;
    Lam<H> lam;
    bool ovf = false;
    trace_path<H>(sc, pa, pixel, sample, L, inv_om, lam, stk, ovf);
    if (ovf) atomicOr(err, 1u);
    float X = 0.f, Y = 0.f, Z = 0.f;
#pragma unroll
    for (int k = 0; k < H; k++) splat_xyz(sc, lam.l[k], L[k] / inv_om, X, Y, Z, 1.f / sc.grid_delta);
    film_add(film, pixel, X, Y, Z);
  }
}

// coefλ2linRGB of the normalised film (src/renderer.cpp:52-56,
// cg.hpp:144-148, 340-341), written top row first
__global__ void k_resolve(DevScene sc, const float *__restrict__ film, uint32_t w, uint32_t h, float inv_n,
                          float *__restrict__ rgb, float *__restrict__ xyz) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w * h) return;
  const uint32_t i = p / w, j = p - i * w;
  const float4 s = reinterpret_cast<const float4 *>(film)[size_t(p) * 2];
  const float X = s.x * inv_n * sc.grid_delta;
  const float Y = s.y * inv_n * sc.grid_delta;
  const float Z = s.z * inv_n * sc.grid_delta;
  const size_t o = (size_t(h - i - 1) * w + j) * 3;
  rgb[o + 0] = 3.2404542f * X - 1.5371385f * Y - 0.4985314f * Z;
  rgb[o + 1] = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
  rgb[o + 2] = 0.0556434f * X - 0.2040259f * Y + 1.0572252f * Z;
  if (xyz) {
    xyz[o + 0] = s.x;
    xyz[o + 1] = s.y;
    xyz[o + 2] = s.z;
  }
}

// ---------------------------------------------------------------------------
// tone map + 8-bit sRGB of a linear-RGB image (Renderer::toneMap,
// src/renderer.cpp:149-176; rendering::rgb24, include/renderer.hpp:20-29)
// ---------------------------------------------------------------------------
constexpr uint32_t kToneBlocks = 1024;

// per-block sums of log(max(Y, FLT_EPSILON)), in double
__global__ __launch_bounds__(256) void k_logsum(const float *__restrict__ rgb, uint32_t n, double *__restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const float Y = 0.2126f * rgb[3 * size_t(i)] + 0.7152f * rgb[3 * size_t(i) + 1] + 0.0722f * rgb[3 * size_t(i) + 2];
    acc += double(logf(Y >= 1.1920929e-07f ? Y : 1.1920929e-07f));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// exposure = y_mid / exp(mean log Y), from the per-block sums (one block)
__global__ __launch_bounds__(256) void k_exposure(const double *__restrict__ partial, uint32_t nb, uint32_t n,
                                                  float y_mid, float *__restrict__ exposure) {
  __shared__ double red[256];
  double acc = 0.0;
  for (uint32_t b = threadIdx.x; b < nb; b += 256u) acc += partial[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *exposure = y_mid / float(exp(red[0] / double(n)));
}

// koiFilmic (filmicToneMap(10 L, .3, .14, .3, 2.8, 10), cg.hpp:181-194),
// then linRGB2sRGB + sRGB2rgb24 (cg.hpp:149-162, numli.hpp:105-106)
__global__ __launch_bounds__(256) void k_tonemap8(const float *__restrict__ rgb, uint32_t n,
                                                  const float *__restrict__ exposure, uint8_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= 3 * n) return;
  const float x = 10.f * (rgb[i] * *exposure);
  const float cl = (x * (0.3f * x + 0.14f)) / (x * (0.3f * x + 2.8f) + 10.f);
  const float s = cl > 0.0031308f ? 1.055f * lam_powf(cl, 1.f / 2.4f) - 0.055f : 12.92f * cl;
  const float v = 255.f * s + 0.5f;
  out[i] = !(v > 0.f) ? uint8_t(0) : (v >= 255.f ? uint8_t(255) : uint8_t(int(v)));
}

hipError_t launch_tonemap8(const float *rgb, uint32_t n, float y_mid, double *partial, float *exposure,
                           uint8_t *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const uint32_t nb = std::min(kToneBlocks, (n + 255) / 256);
  hipLaunchKernelGGL(k_logsum, dim3(nb), dim3(256), 0, st, rgb, n, partial);
  hipLaunchKernelGGL(k_exposure, dim3(1), dim3(256), 0, st, partial, nb, n, y_mid, exposure);
  hipLaunchKernelGGL(k_tonemap8, dim3(uint32_t((3 * size_t(n) + 255) / 256)), dim3(256), 0, st, rgb, n, exposure,
                     out);
  return hipGetLastError();
}

__global__ void k_math(int fn, const float *__restrict__ a, const float *__restrict__ b, uint32_t n,
                       float *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = a[i], y = b[i];
  float r;
  switch (fn) {
    case 0: r = lam_sinf(x); break;
    case 1: r = lam_cosf(x); break;
    case 2: r = lam_powf(x, y); break;
    case 3: r = lam_atan2f(x, y); break;
    case 4: r = sqrtf(x); break;
    default: r = x / y; break;
  }
  out[i] = r;
}

__global__ void k_zero(float4 *__restrict__ p, size_t n4) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n4) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks64(size_t n) { return uint32_t((n + 63) / 64); }

hipError_t launch_trace_rays(const DevScene &sc, const lam_ray *rays, size_t n, lam_hit *hits, uint32_t *occl,
                             uint32_t *counts, uint32_t *err, uint32_t *fetch, uint32_t *spill, uint32_t blocks,
                             hipStream_t st) {
  if (n == 0) return hipSuccess;
  const dim3 g(std::min(blocks64(n), blocks)), b(64);
  if (occl && counts) {
    hipLaunchKernelGGL((k_trace_rays<true, true>), g, b, 0, st, sc, rays, uint32_t(n), hits, occl, counts, err,
                       fetch, spill);
  } else if (occl) {
    hipLaunchKernelGGL((k_trace_rays<true, false>), g, b, 0, st, sc, rays, uint32_t(n), hits, occl, counts, err,
                       fetch, spill);
  } else if (counts) {
    hipLaunchKernelGGL((k_trace_rays<false, true>), g, b, 0, st, sc, rays, uint32_t(n), hits, occl, counts, err,
                       fetch, spill);
  } else {
    hipLaunchKernelGGL((k_trace_rays<false, false>), g, b, 0, st, sc, rays, uint32_t(n), hits, occl, counts, err,
                       fetch, spill);
  }
  return hipGetLastError();
}

hipError_t launch_trace_paths(const DevScene &sc, const PathArgs &pa, int hero, const uint32_t *pixel,
                              const uint32_t *sample, size_t n, float *out, uint32_t *err, uint32_t *spill,
                              uint32_t blocks, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const dim3 g(std::min(blocks64(n), blocks)), b(64);
  if (hero == 16)
    hipLaunchKernelGGL((k_trace_paths<16>), g, b, 0, st, sc, pa, pixel, sample, uint32_t(n), out, err, spill);
  else
    hipLaunchKernelGGL((k_trace_paths<4>), g, b, 0, st, sc, pa, pixel, sample, uint32_t(n), out, err, spill);
  return hipGetLastError();
}

hipError_t launch_render_mega(const DevScene &sc, const PathArgs &pa, int hero, uint32_t sample, uint32_t npix,
                              float *film, uint32_t *err, uint32_t *spill, uint32_t blocks, hipStream_t st) {
  const dim3 g(std::min(blocks64(npix), blocks)), b(64);
  if (hero == 16)
    hipLaunchKernelGGL((k_render_mega<16>), g, b, 0, st, sc, pa, sample, npix, film, err, spill);
  else
    hipLaunchKernelGGL((k_render_mega<4>), g, b, 0, st, sc, pa, sample, npix, film, err, spill);
  return hipGetLastError();
}

hipError_t launch_resolve(const DevScene &sc, const float *film, uint32_t w, uint32_t h, float inv_n, float *rgb,
                          float *xyz, hipStream_t st) {
  const size_t n = size_t(w) * h;
  hipLaunchKernelGGL(k_resolve, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, sc, film, w, h, inv_n, rgb, xyz);
  return hipGetLastError();
}

hipError_t launch_math(int fn, const float *a, const float *b, size_t n, float *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_math, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, fn, a, b, uint32_t(n), out);
  return hipGetLastError();
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t st) {
  const size_t n4 = bytes / 16;
  if (n4 == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero, dim3(uint32_t((n4 + 255) / 256)), dim3(256), 0, st, reinterpret_cast<float4 *>(p), n4);
  return hipGetLastError();
}

}  // namespace lam

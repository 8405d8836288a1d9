// Elementary functions and the counter-based RNG used on the device.
//
// The reference calls libm's cosf/sinf/powf/atan2f (lib/numli/cg.hpp,
// stoch.hpp).  Device libm and host libm round differently in the last
// bit, which would let a sampled direction or a lobe choice differ between
// the GPU and the CPU oracle.  These functions are built only from
// correctly rounded IEEE operations (+ - * /, sqrt, floor, frexp, ldexp) in
// double precision and rounded once to float, so with -ffp-contract=off
// they give the same bits on x86 and gfx950.  Accuracy: below 1 ulp of
// float over the argument ranges the renderer uses (checked against libm
// in tests/test_oracle.py::test_math_*, and bitwise against the oracle in tests/test_gpu_parity.py).
//
// Methods (textbook, no vendor code):
//   sin/cos  Cody-Waite reduction by pi/2 (two-part constant) + Taylor
//            series to degree 11 / 12 on [-pi/4, pi/4] (truncation below
//            1e-11 relative);
//   log2     frexp split, ln m = 2 atanh((m-1)/(m+1)) series to f^19,
//            |f| <= 0.172 (truncation below 1e-17: pow scales log2's error
//            by |y|, up to 2048 here);
//   exp2     2^y = 2^n * exp(r ln2), |r| <= 1/2, Taylor to degree 11
//            (truncation below 1e-14);
//   pow      exp2(y log2 x) with C99 special cases;
//   atan2    two half-angle reductions atan t = 2 atan(t/(1+sqrt(1+t^2)))
//            then the odd Taylor series to t^15, t <= tan(pi/16)
//            (truncation below 1e-13).
// Each truncation is far below half a float ulp, so the rounded results
// stay within 1 ulp of libm; shorter series cost fewer double operations.
// RNG: splitmix64 finalizer (Steele, Lea & Flood 2014) keyed by
// (seed, pixel, sample) and a per-draw dimension index.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LAM_HD __host__ __device__ inline
#else
#include <math.h>
#define LAM_HD static inline
#endif

#define LAM_PI_F 3.14159265358979323846f

LAM_HD double lam_sin_series(double r) {
  const double r2 = r * r;
  double p = 1.0 / 39916800.0;             /* 1/11! */
  p = p * r2 - 1.0 / 362880.0;             /* -1/9! */
  p = p * r2 + 1.0 / 5040.0;               /* 1/7! */
  p = p * r2 - 1.0 / 120.0;
  p = p * r2 + 1.0 / 6.0;
  return r - r * r2 * p;
}
LAM_HD double lam_cos_series(double r) {
  const double r2 = r * r;
  double p = -1.0 / 479001600.0;           /* -1/12! */
  p = p * r2 + 1.0 / 3628800.0;            /* 1/10! */
  p = p * r2 - 1.0 / 40320.0;              /* -1/8! */
  p = p * r2 + 1.0 / 720.0;
  p = p * r2 - 1.0 / 24.0;
  p = p * r2 + 0.5;
  return 1.0 - r2 * p;
}
/* x -> (r, quadrant) with x = q*pi/2 + r */
LAM_HD double lam_reduce_pio2(double x, int *q) {
  const double two_over_pi = 0.63661977236758134308;
  const double pio2_1 = 1.57079632673412561417e+00;  /* first 33 bits of pi/2 */
  const double pio2_1t = 6.07710050650619224932e-11; /* pi/2 - pio2_1 */
  const double k = floor(x * two_over_pi + 0.5);
  *q = (int)((long long)k & 3);
  return (x - k * pio2_1) - k * pio2_1t;
}
LAM_HD float lam_sinf(float xf) {
  const double x = (double)xf;
  if (!(x - x == 0.0)) return xf - xf; /* inf / nan -> nan */
  int q;
  const double r = lam_reduce_pio2(x, &q);
  double v;
  switch (q) {
    case 0: v = lam_sin_series(r); break;
    case 1: v = lam_cos_series(r); break;
    case 2: v = -lam_sin_series(r); break;
    default: v = -lam_cos_series(r); break;
  }
  return (float)v;
}
LAM_HD float lam_cosf(float xf) {
  const double x = (double)xf;
  if (!(x - x == 0.0)) return xf - xf;
  int q;
  const double r = lam_reduce_pio2(x, &q);
  double v;
  switch (q) {
    case 0: v = lam_cos_series(r); break;
    case 1: v = -lam_sin_series(r); break;
    case 2: v = -lam_cos_series(r); break;
    default: v = lam_sin_series(r); break;
  }
  return (float)v;
}

/* sin and cos of one argument: one reduction and each series once (the
 * same operations as lam_sinf and lam_cosf, so the same bits) */
LAM_HD void lam_sincosf(float xf, float *sf, float *cf) {
  const double x = (double)xf;
  if (!(x - x == 0.0)) {
    *sf = *cf = xf - xf;
    return;
  }
  int q;
  const double r = lam_reduce_pio2(x, &q);
  const double s = lam_sin_series(r), c = lam_cos_series(r);
  const double sv = (q & 1) ? c : s, cv = (q & 1) ? s : c;
  *sf = (float)((q & 2) ? -sv : sv);
  *cf = (float)(((q + 1) & 2) ? -cv : cv);
}

/* log2 of a positive finite double */
LAM_HD double lam_log2_pos(double x) {
  int e;
  double m = frexp(x, &e); /* m in [0.5, 1) */
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double f = (m - 1.0) / (m + 1.0);
  const double f2 = f * f;
  double p = 1.0 / 19.0;
  p = p * f2 + 1.0 / 17.0;
  p = p * f2 + 1.0 / 15.0;
  p = p * f2 + 1.0 / 13.0;
  p = p * f2 + 1.0 / 11.0;
  p = p * f2 + 1.0 / 9.0;
  p = p * f2 + 1.0 / 7.0;
  p = p * f2 + 1.0 / 5.0;
  p = p * f2 + 1.0 / 3.0;
  p = p * f2 + 1.0;
  const double ln_m = 2.0 * f * p;
  return (double)e + ln_m * 1.44269504088896340736; /* 1/ln 2 */
}
LAM_HD double lam_exp2_d(double y) {
  if (y > 1030.0) return 1.0 / 0.0;
  if (y < -1100.0) return 0.0;
  const double n = floor(y + 0.5);
  const double a = (y - n) * 0.69314718055994530942; /* r ln 2 */
  double p = 1.0 / 39916800.0;             /* 1/11! */
  p = p * a + 1.0 / 3628800.0;              /* 1/10! */
  p = p * a + 1.0 / 362880.0;
  p = p * a + 1.0 / 40320.0;
  p = p * a + 1.0 / 5040.0;
  p = p * a + 1.0 / 720.0;
  p = p * a + 1.0 / 120.0;
  p = p * a + 1.0 / 24.0;
  p = p * a + 1.0 / 6.0;
  p = p * a + 0.5;
  p = p * a + 1.0;
  p = p * a + 1.0;
  return ldexp(p, (int)n);
}
LAM_HD int lam_is_int(float y) { return floorf(y) == y; }
LAM_HD int lam_is_odd_int(float y) {
  if (!lam_is_int(y)) return 0;
  const double h = (double)y * 0.5;
  return floor(h) != h;
}
/* powf with the C99 special cases the renderer can reach */
LAM_HD float lam_powf(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (x == 1.0f) return 1.0f;
  if (x != x || y != y) return x + y;
  const float ax = x < 0.0f ? -x : x;
  const float ay = y < 0.0f ? -y : y;
  if (ay > 3.0e38f && ay - ay != 0.0f) { /* y = +-inf */
    if (ax == 1.0f) return 1.0f;
    return ((ax > 1.0f) == (y > 0.0f)) ? y * y : 0.0f;
  }
  if (x == 0.0f) {
    const int odd = lam_is_odd_int(y);
    if (y > 0.0f) return odd ? x : 0.0f;
    return odd ? 1.0f / x : 1.0f / 0.0f;
  }
  if (ax - ax != 0.0f) { /* x = +-inf */
    const int odd = lam_is_odd_int(y);
    const float r = y > 0.0f ? ax : 0.0f;
    return (x < 0.0f && odd) ? -r : r;
  }
  float sign = 1.0f;
  if (x < 0.0f) {
    if (!lam_is_int(y)) return (x - x) / (x - x); /* nan */
    if (lam_is_odd_int(y)) sign = -1.0f;
  }
  const double r = lam_exp2_d((double)y * lam_log2_pos((double)ax));
  return sign * (float)r;
}

LAM_HD double lam_atan_pos(double t) { /* t >= 0, finite */
  const double pio2 = 1.57079632679489661923;
  int inv = 0;
  if (t > 1.0) {
    t = 1.0 / t;
    inv = 1;
  }
  t = t / (1.0 + sqrt(1.0 + t * t));
  t = t / (1.0 + sqrt(1.0 + t * t));
  const double t2 = t * t;
  double p = -1.0 / 15.0;
  p = p * t2 + 1.0 / 13.0;
  p = p * t2 - 1.0 / 11.0;
  p = p * t2 + 1.0 / 9.0;
  p = p * t2 - 1.0 / 7.0;
  p = p * t2 + 1.0 / 5.0;
  p = p * t2 - 1.0 / 3.0;
  p = p * t2 + 1.0;
  const double a = 4.0 * (t * p);
  return inv ? pio2 - a : a;
}
LAM_HD float lam_atan2f(float yf, float xf) {
  const double pi = 3.14159265358979323846;
  if (xf != xf || yf != yf) return xf + yf;
  const double y = (double)yf, x = (double)xf;
  const int xneg = (xf < 0.0f) || (xf == 0.0f && 1.0f / xf < 0.0f);
  const int yneg = (yf < 0.0f) || (yf == 0.0f && 1.0f / yf < 0.0f);
  const double ay = yneg ? -y : y, ax = xneg ? -x : x;
  double a;
  if (ax == 0.0 && ay == 0.0) a = 0.0;
  else if (ax - ax != 0.0 && ay - ay != 0.0) a = pi / 4.0;
  else if (ax == 0.0 || ay - ay != 0.0) a = pi / 2.0;
  else if (ax - ax != 0.0) a = 0.0;
  else a = lam_atan_pos(ay / ax);
  if (xneg) a = pi - a;
  return (float)(yneg ? -a : a);
}

/* ---- counter-based RNG ------------------------------------------------ */
LAM_HD uint64_t lam_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
/* per-path key: one per (pixel, sample) */
LAM_HD uint64_t lam_path_key(uint64_t seed, uint32_t pixel, uint32_t sample) {
  return lam_mix64(lam_mix64(seed) ^ (((uint64_t)pixel << 32) | (uint64_t)sample));
}
LAM_HD uint32_t lam_draw_u32(uint64_t key, uint32_t dim) {
  return (uint32_t)(lam_mix64(key + (uint64_t)(dim + 1u) * 0x9E3779B97F4A7C15ull) >> 32);
}
/* uniform float in [0,1) with 24 random bits */
LAM_HD float lam_draw_f(uint64_t key, uint32_t dim) {
  return (float)(lam_draw_u32(key, dim) >> 8) * 5.9604644775390625e-8f;
}

/* RNG dimensions: fixed slots per sample and per path vertex, so every
 * draw is addressed by (pixel, sample, dim) independently of branch
 * history.  Order follows the reference's draw order
 * (src/renderer.cpp:35-40, :92-117). */
#define LAM_DIM_U 0u
#define LAM_DIM_V 1u
#define LAM_DIM_LENS_X 2u
#define LAM_DIM_LENS_Y 3u
#define LAM_DIM_HERO 4u
#define LAM_DIM_VERTEX0 8u
#define LAM_DIMS_PER_VERTEX 8u
#define LAM_VDIM_LIGHT_PICK 0u
#define LAM_VDIM_LIGHT_A 1u
#define LAM_VDIM_LIGHT_B 2u
#define LAM_VDIM_LOBE 3u
#define LAM_VDIM_DIR_A 4u
#define LAM_VDIM_DIR_B 5u
#define LAM_VDIM_GUIDE 6u
#define LAM_VDIM_SPARE 7u
LAM_HD uint32_t lam_vdim(uint32_t vertex, uint32_t slot) {
  return LAM_DIM_VERTEX0 + vertex * LAM_DIMS_PER_VERTEX + slot;
}

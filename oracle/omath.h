/* omath.h -- TEST INFRASTRUCTURE (oracle).  Elementary functions and the
 * counter RNG of the CPU oracle.
 *
 * The reference calls glibc's cosf/sinf/powf/atan2f.  So that the oracle
 * and the GPU draw identical directions, both sides use the same published
 * constructions, evaluated in double from correctly rounded IEEE operations
 * and rounded once to float (the oracle is compiled with -ffp-contract=off):
 *   sin/cos: Cody-Waite reduction by pi/2 with fdlibm's two-part pi/2,
 *            Taylor series on [-pi/4, pi/4] (to x^11 / x^12);
 *   log2:    x = m 2^e, ln m = 2 atanh((m-1)/(m+1)) (odd series to f^19);
 *   exp2:    2^n * exp(r ln 2), Taylor to r^11;
 *   atan:    two half-angle reductions then the odd series to t^15.
 * Each series is truncated where its remainder is far below half a float
 * ulp (1e-11 relative or less; log2 1e-17, as pow scales its error by |y|).
 * They are pinned against glibc in tests/test_oracle.py (test_math_*).
 * RNG: splitmix64 finalizer (Steele, Lea, Flood 2014), keyed by
 * (seed, pixel, sample) plus a per-draw dimension.
 */
#ifndef ORACLE_OMATH_H
#define ORACLE_OMATH_H
#include <math.h>
#include <stdint.h>

static inline double om_sin_poly(double r) {
  double r2 = r * r;
  double p = 1.0 / 39916800.0;
  p = p * r2 - 1.0 / 362880.0;
  p = p * r2 + 1.0 / 5040.0;
  p = p * r2 - 1.0 / 120.0;
  p = p * r2 + 1.0 / 6.0;
  return r - r * r2 * p;
}
static inline double om_cos_poly(double r) {
  double r2 = r * r;
  double p = -1.0 / 479001600.0;
  p = p * r2 + 1.0 / 3628800.0;
  p = p * r2 - 1.0 / 40320.0;
  p = p * r2 + 1.0 / 720.0;
  p = p * r2 - 1.0 / 24.0;
  p = p * r2 + 0.5;
  return 1.0 - r2 * p;
}
static inline double om_reduce(double x, int *q) {
  double k = floor(x * 0.63661977236758134308 + 0.5);
  *q = (int)((long long)k & 3);
  return (x - k * 1.57079632673412561417e+00) - k * 6.07710050650619224932e-11;
}
static inline float om_sinf(float xf) {
  double x = (double)xf, r, v;
  int q;
  if (!(x - x == 0.0)) return xf - xf;
  r = om_reduce(x, &q);
  if (q == 0) v = om_sin_poly(r);
  else if (q == 1) v = om_cos_poly(r);
  else if (q == 2) v = -om_sin_poly(r);
  else v = -om_cos_poly(r);
  return (float)v;
}
static inline float om_cosf(float xf) {
  double x = (double)xf, r, v;
  int q;
  if (!(x - x == 0.0)) return xf - xf;
  r = om_reduce(x, &q);
  if (q == 0) v = om_cos_poly(r);
  else if (q == 1) v = -om_sin_poly(r);
  else if (q == 2) v = -om_cos_poly(r);
  else v = om_sin_poly(r);
  return (float)v;
}
static inline double om_log2(double x) {
  int e;
  double m = frexp(x, &e), f, f2, p;
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  f = (m - 1.0) / (m + 1.0);
  f2 = f * f;
  p = 1.0 / 19.0;
  p = p * f2 + 1.0 / 17.0;
  p = p * f2 + 1.0 / 15.0;
  p = p * f2 + 1.0 / 13.0;
  p = p * f2 + 1.0 / 11.0;
  p = p * f2 + 1.0 / 9.0;
  p = p * f2 + 1.0 / 7.0;
  p = p * f2 + 1.0 / 5.0;
  p = p * f2 + 1.0 / 3.0;
  p = p * f2 + 1.0;
  return (double)e + (2.0 * f * p) * 1.44269504088896340736;
}
static inline double om_exp2(double y) {
  double n, a, p;
  if (y > 1030.0) return 1.0 / 0.0;
  if (y < -1100.0) return 0.0;
  n = floor(y + 0.5);
  a = (y - n) * 0.69314718055994530942;
  p = 1.0 / 39916800.0;
  p = p * a + 1.0 / 3628800.0;
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
static inline int om_is_int(float y) { return floorf(y) == y; }
static inline int om_is_odd(float y) {
  double h;
  if (!om_is_int(y)) return 0;
  h = (double)y * 0.5;
  return floor(h) != h;
}
static inline float om_powf(float x, float y) {
  float ax, ay, sign = 1.0f;
  if (y == 0.0f) return 1.0f;
  if (x == 1.0f) return 1.0f;
  if (x != x || y != y) return x + y;
  ax = x < 0.0f ? -x : x;
  ay = y < 0.0f ? -y : y;
  if (ay > 3.0e38f && ay - ay != 0.0f) {
    if (ax == 1.0f) return 1.0f;
    return ((ax > 1.0f) == (y > 0.0f)) ? y * y : 0.0f;
  }
  if (x == 0.0f) {
    int odd = om_is_odd(y);
    if (y > 0.0f) return odd ? x : 0.0f;
    return odd ? 1.0f / x : 1.0f / 0.0f;
  }
  if (ax - ax != 0.0f) {
    int odd = om_is_odd(y);
    float r = y > 0.0f ? ax : 0.0f;
    return (x < 0.0f && odd) ? -r : r;
  }
  if (x < 0.0f) {
    if (!om_is_int(y)) return (x - x) / (x - x);
    if (om_is_odd(y)) sign = -1.0f;
  }
  return sign * (float)om_exp2((double)y * om_log2((double)ax));
}
static inline double om_atan_pos(double t) {
  int inv = 0;
  double t2, p, a;
  if (t > 1.0) {
    t = 1.0 / t;
    inv = 1;
  }
  t = t / (1.0 + sqrt(1.0 + t * t));
  t = t / (1.0 + sqrt(1.0 + t * t));
  t2 = t * t;
  p = -1.0 / 15.0;
  p = p * t2 + 1.0 / 13.0;
  p = p * t2 - 1.0 / 11.0;
  p = p * t2 + 1.0 / 9.0;
  p = p * t2 - 1.0 / 7.0;
  p = p * t2 + 1.0 / 5.0;
  p = p * t2 - 1.0 / 3.0;
  p = p * t2 + 1.0;
  a = 4.0 * (t * p);
  return inv ? 1.57079632679489661923 - a : a;
}
static inline float om_atan2f(float yf, float xf) {
  const double pi = 3.14159265358979323846;
  double y = (double)yf, x = (double)xf, ay, ax, a;
  int xneg, yneg;
  if (xf != xf || yf != yf) return xf + yf;
  xneg = (xf < 0.0f) || (xf == 0.0f && 1.0f / xf < 0.0f);
  yneg = (yf < 0.0f) || (yf == 0.0f && 1.0f / yf < 0.0f);
  ay = yneg ? -y : y;
  ax = xneg ? -x : x;
  if (ax == 0.0 && ay == 0.0) a = 0.0;
  else if (ax - ax != 0.0 && ay - ay != 0.0) a = pi / 4.0;
  else if (ax == 0.0 || ay - ay != 0.0) a = pi / 2.0;
  else if (ax - ax != 0.0) a = 0.0;
  else a = om_atan_pos(ay / ax);
  if (xneg) a = pi - a;
  return (float)(yneg ? -a : a);
}

/* ---- counter RNG ------------------------------------------------------- */
static inline uint64_t om_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t om_path_key(uint64_t seed, uint32_t pixel, uint32_t sample) {
  return om_mix64(om_mix64(seed) ^ (((uint64_t)pixel << 32) | (uint64_t)sample));
}
static inline uint32_t om_u32(uint64_t key, uint32_t dim) {
  return (uint32_t)(om_mix64(key + (uint64_t)(dim + 1u) * 0x9E3779B97F4A7C15ull) >> 32);
}
static inline float om_flt(uint64_t key, uint32_t dim) {
  return (float)(om_u32(key, dim) >> 8) * 5.9604644775390625e-8f;
}
/* dimension layout: 5 per sample, then 8 per path vertex */
#define OM_DIM_U 0u
#define OM_DIM_V 1u
#define OM_DIM_LX 2u
#define OM_DIM_LY 3u
#define OM_DIM_HERO 4u
static inline uint32_t om_vdim(uint32_t vertex, uint32_t slot) { return 8u + vertex * 8u + slot; }

#endif

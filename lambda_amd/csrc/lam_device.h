// Device-side scene view, vector helpers and the closest-hit / any-hit
// traversal shared by the extend, shadow and path kernels.
//
// Evaluation order of every float expression restates the reference
// (lib/numli/cg.hpp, bra.hpp) and the whole library is compiled with
// -ffp-contract=off, so the CPU oracle computes the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lambda_amd.h"
#include "lam_math.h"

namespace lam {

// ---- constants ------------------------------------------------------------
// intersect::BIAS = constexprSqrt(ε<float>) (cg.hpp:1189, numli.hpp:69-70):
// the Newton iteration converges to this float.
constexpr float kBias = 0x1.6a09e6p-12f;
constexpr float kEps = 1.1920928955078125e-07f;  // ε<float>
constexpr float kFltMax = 3.40282346638528859812e+38f;
constexpr float kInvPi = 0.318309886183790671538f;  // inv_π<float>
constexpr float kPi = 3.14159265358979323846f;
constexpr int kStack = 128;  // traversal stack entries per lane (LDS part + HBM spill)
constexpr uint32_t kLeafBit = 0x80000000u;

// ---- small vector type with the reference's operator semantics -----------
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float a, float b, float c) { return f3{a, b, c}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 neg(f3 a) { return f3{a.x * -1.f, a.y * -1.f, a.z * -1.f}; }
__device__ __forceinline__ f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len2(f3 a) {
  float s = 0.f;
  s += a.x * a.x;
  s += a.y * a.y;
  s += a.z * a.z;
  return s;
}
__device__ __forceinline__ f3 normalized(f3 a) {
  const float l = sqrtf(len2(a));
  return f3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float fmax_ref(float a, float b) { return a >= b ? a : b; }
__device__ __forceinline__ float fmin_ref(float a, float b) { return a <= b ? a : b; }

// M (3x4 rows) * (v, w) with a sequential k-sum (bra.hpp:325-336)
__device__ __forceinline__ f3 xmul(const float *M, f3 v, float w) {
  float r[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float s = 0.f;
    s += M[i * 4 + 0] * v.x;
    s += M[i * 4 + 1] * v.y;
    s += M[i * 4 + 2] * v.z;
    s += M[i * 4 + 3] * w;
    r[i] = s;
  }
  return f3{r[0], r[1], r[2]};
}
// subMatT<3,3>(M_inv) * n (cg.hpp:442-443)
__device__ __forceinline__ f3 normal_to_world(const float *Minv, f3 n) {
  float r[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float s = 0.f;
    s += Minv[0 * 4 + i] * n.x;
    s += Minv[1 * 4 + i] * n.y;
    s += Minv[2 * 4 + i] * n.z;
    r[i] = s;
  }
  return f3{r[0], r[1], r[2]};
}

// ---- device scene -----------------------------------------------------------
struct Node8d {  // 128 B, see host/bvh8.hpp
  float ox, oy, oz;
  uint32_t e_n;  // ex | ey<<8 | ez<<16 | nchild<<24
  uint32_t child[8];
  uint32_t qlo[6];  // 24 B: x[8], y[8], z[8]
  uint32_t qhi[6];
  uint32_t pad[8];
};
static_assert(sizeof(Node8d) == 128, "node record");

struct DevScene {
  const Node8d *nodes;
  const float4 *tris;   // 3 float4 per triangle slot: v0 v1 v2 | obj prim orig
  const float4 *shade;  // 3 float4 per slot: n0 n1 n2 | gn
  const uint32_t *analytic;  // object ids of spheres and planes
  const lam_object *objects;
  const lam_material *materials;
  const lam_light *lights;
  const float *spectra;
  const float *cmf;
  uint32_t n_nodes, n_tris, n_analytic, n_objects, n_materials, n_lights;
  uint32_t n_spectra;
  uint32_t grid_n;
  float grid_inf, grid_sup, grid_delta;
  lam_camera cam;
};

struct Ray {
  f3 o, d;
  float tmax;
  uint32_t skip;
};

struct Hit {
  float t;
  uint32_t slot;  // triangle slot (BVH order) or LAM_NONE
  uint32_t obj;
  uint32_t prim;
};

__device__ __forceinline__ bool better(float t, uint32_t obj, uint32_t prim, const Hit &b) {
  return t < b.t || (t == b.t && (obj < b.obj || (obj == b.obj && prim < b.prim)));
}

// ---- primitive tests -------------------------------------------------------

// intersect::triangle, Plücker form (cg.hpp:1341-1379).  Returns t or a
// negative value on a miss; the t-window test is done by the caller.
__device__ __forceinline__ float tri_t(f3 v0, f3 v1, f3 v2, f3 o, f3 u, f3 mr) {
  const f3 e0 = sub(v1, v0);
  const f3 e2 = sub(v0, v2);
  const f3 n = cross(e2, e0);
  const float D = dot(u, n);
  if (fabsf(D) < kEps) return -1.f;
  const f3 m0 = cross(v0, e0);
  const float s0 = dot(u, m0) + dot(e0, mr);
  if (s0 * D < 0.f) return -1.f;
  const f3 e1 = sub(v2, v1);
  const f3 m1 = cross(v1, e1);
  const float s1 = dot(u, m1) + dot(e1, mr);
  if (s1 * D < 0.f) return -1.f;
  const f3 m2 = cross(v2, e2);
  const float s2 = dot(u, m2) + dot(e2, mr);
  if (s2 * D < 0.f) return -1.f;
  if (fabsf(s0 + s1 + s2) < kEps) return -1.f;
  const float inv_D = 1.f / D;
  const float t = dot(sub(v0, o), n) * inv_D;
  return t;
}
// barycentrics of the same test, b = (s1, s2, s0)/D (cg.hpp:1371)
__device__ __forceinline__ f3 tri_bary(f3 v0, f3 v1, f3 v2, f3 o, f3 u) {
  const f3 mr = cross(o, u);
  const f3 e0 = sub(v1, v0);
  const f3 e2 = sub(v0, v2);
  const f3 n = cross(e2, e0);
  const float D = dot(u, n);
  const float s0 = dot(u, cross(v0, e0)) + dot(e0, mr);
  const f3 e1 = sub(v2, v1);
  const float s1 = dot(u, cross(v1, e1)) + dot(e1, mr);
  const float s2 = dot(u, cross(v2, e2)) + dot(e2, mr);
  const float inv_D = 1.f / D;
  return f3{s1 * inv_D, s2 * inv_D, s0 * inv_D};
}

// ray.toLocal (cg.hpp:438-439)
__device__ __forceinline__ void to_local(const lam_object &ob, f3 o, f3 u, f3 &lo, f3 &lu) {
  lo = xmul(ob.T.Minv, o, 1.f);
  lu = xmul(ob.T.Minv, u, 0.f);
}

// intersect::sphere t (cg.hpp:1196-1214); negative on miss
__device__ __forceinline__ float sphere_t(f3 lo, f3 lu) {
  const float a = dot(lu, lu);
  const float b = 2 * dot(lo, lu);
  const float c = dot(lo, lo) - 1.f;
  const float disc = b * b - 4 * a * c;
  if (disc < .1f * kBias) return -1.f;
  const float inv_2a = .5f / a;
  const float tp = (-b + sqrtf(disc)) * inv_2a;
  const float tm = (-b - sqrtf(disc)) * inv_2a;
  float t = tm;
  if (tm < kBias) t = tp;
  if (t < kBias) return -1.f;
  return t;
}
// intersect::plane t (cg.hpp:1245-1252); negative on miss
__device__ __forceinline__ float plane_t(f3 lo, f3 lu) {
  if (lu.z == 0.f) return -1.f;
  const float t = -lo.z / lu.z;
  const f3 x = add(lo, mul(lu, t));
  if (x.x < -1.f || x.x > 1.f || x.y < -1.f || x.y > 1.f || t < kBias) return -1.f;
  return t;
}

__device__ __forceinline__ float analytic_t(const lam_object &ob, f3 o, f3 u) {
  f3 lo, lu;
  to_local(ob, o, u, lo, lu);
  return ob.type == LAM_OBJ_SPHERE ? sphere_t(lo, lu) : plane_t(lo, lu);
}

// ---- BVH8 traversal ---------------------------------------------------------

__device__ __forceinline__ float scale_of(uint32_t e) { return __uint_as_float(e << 23); }

// the relative widening of the quantized-frame slab test: 1 + 2^-21 (covers
// the relative roundings of tn and tf and of the widened far distance)
constexpr float kRobustQ = 1.000000476837158203125f;

__device__ __forceinline__ void cswap(float &ta, uint32_t &ca, float &tb, uint32_t &cb) {
  const bool s = tb < ta;
  const float t0 = s ? tb : ta, t1 = s ? ta : tb;
  const uint32_t c0 = s ? cb : ca, c1 = s ? ca : cb;
  ta = t0; tb = t1; ca = c0; cb = c1;
}

struct TravCounts {
  uint32_t nodes, tris;
};

// Traversal stack of one lane: entry k lives in the lane's LDS column,
// lds[k*64], for k < kLdsStack, else in the lane's column of a global spill
// buffer, spill[(k - kLdsStack) * stride] (stride = lanes of the grid).
// Deep entries are rare (a push needs several overlapping children at every
// level above), so the split stack trades a few global accesses for a block
// footprint of kLdsStack*256 B (6 KB) instead of kStack*256 B.
constexpr int kLdsStack = 24;
struct SplitStack {
  uint32_t *lds;    // this lane's LDS column: entry k at lds[k*64]
  uint32_t *spill;  // this lane's spill column: entry k at spill[(k - kLdsStack)*stride]
  uint32_t stride;  // lanes of the grid
  // the stack of lane threadIdx.x % 64 of block `block` in a grid of
  // `blocks` 64-lane blocks (the spill buffer holds kStack - kLdsStack
  // words per grid lane)
  __device__ __forceinline__ static SplitStack of(uint32_t *lds_base, uint32_t *spill_base, uint32_t block,
                                                  uint32_t blocks) {
    const uint32_t lane = threadIdx.x % 64;
    return SplitStack{lds_base + lane, spill_base + size_t(block) * 64 + lane, blocks * 64};
  }
  __device__ __forceinline__ bool in_lds(int top) const { return top <= kLdsStack; }
  __device__ __forceinline__ void put_lds(int k, uint32_t v) const { lds[k * 64] = v; }
  __device__ __forceinline__ void put(int k, uint32_t v) const {
    if (k < kLdsStack)
      lds[k * 64] = v;
    else
      spill[size_t(k - kLdsStack) * stride] = v;
  }
  __device__ __forceinline__ uint32_t get(int k) const {
    return k < kLdsStack ? lds[k * 64] : spill[size_t(k - kLdsStack) * stride];
  }
};

// Traversal state of one ray: the ray, its slab-test constants, the best
// hit so far, the current node and the stack depth (the stack itself is
// the lane's SplitStack).  Closest hit (ANY=false) or any hit (ANY=true)
// for t in [BIAS, tmax], objects `skip` excluded; ties go to the lowest
// (obj, prim).
struct Trav {
  Ray r;
  float ix, iy, iz;
  f3 mr;
  Hit best;
  uint32_t orig;  // input index of best's triangle (LAM_NONE otherwise)
  uint32_t cur;
  int sp;
};

// Starts a ray: the analytic objects (spheres, planes) as the linear loop
// of intersect::scene (cg.hpp:1477-1487), then the BVH root.  Returns false
// when the ray is already decided (any-hit found, or no BVH).
template <bool ANY>
__device__ __forceinline__ bool trav_begin(const DevScene &sc, const Ray &r, Trav &t) {
  t.r = r;
  t.best.t = r.tmax;
  t.best.slot = LAM_NONE;
  t.best.obj = LAM_NONE;
  t.best.prim = LAM_NONE;
  t.orig = LAM_NONE;
  for (uint32_t a = 0; a < sc.n_analytic; a++) {
    const uint32_t oi = sc.analytic[a];
    if (oi == r.skip) continue;
    const float th = analytic_t(sc.objects[oi], r.o, r.d);
    if (th >= kBias && better(th, oi, 0u, t.best)) {
      t.best.t = th;
      t.best.slot = LAM_NONE;
      t.best.obj = oi;
      t.best.prim = 0;
      if (ANY) return false;
    }
  }
  if (sc.n_nodes == 0) return false;
  // ray constructor inv_u (cg.hpp:357-361), clamped to +-2^64 for the
  // quantized-frame slab test (trav_step): the slab distances then stay
  // finite, and an axis the ray does not move along tests o_j against the
  // planes up to the test's widening
  t.ix = fminf(fmaxf(r.d.x != 0.f ? 1.f / r.d.x : kFltMax, -0x1p64f), 0x1p64f);
  t.iy = fminf(fmaxf(r.d.y != 0.f ? 1.f / r.d.y : kFltMax, -0x1p64f), 0x1p64f);
  t.iz = fminf(fmaxf(r.d.z != 0.f ? 1.f / r.d.z : kFltMax, -0x1p64f), 0x1p64f);
  t.mr = cross(r.o, r.d);
  t.sp = 0;
  t.cur = 0;  // root is always an interior record
  return true;
}

// One traversal step: an interior node (test the 8 children, push the hit
// ones but the nearest, farthest first, descend into the nearest) or a leaf
// (test its triangles), then pop if nothing was descended into.  Returns
// false when the ray is finished (stack empty, any-hit found or overflow).
template <bool ANY, bool COUNT, class Stack>
__device__ __forceinline__ bool trav_step(const DevScene &sc, Trav &t, const Stack &stk, TravCounts &cnt,
                                          bool &overflow) {
  const uint32_t cur = t.cur;
  if (!(cur & kLeafBit)) {
    const Node8d *nd = sc.nodes + cur;
    const float4 h0 = *reinterpret_cast<const float4 *>(&nd->ox);
    const uint4 c0 = *reinterpret_cast<const uint4 *>(&nd->child[0]);
    const uint4 c1 = *reinterpret_cast<const uint4 *>(&nd->child[4]);
    const uint4 q0 = *reinterpret_cast<const uint4 *>(&nd->qlo[0]);
    const uint4 q1 = *reinterpret_cast<const uint4 *>(&nd->qlo[4]);  // qlo[4..5], qhi[0..1]
    const uint4 q2 = *reinterpret_cast<const uint4 *>(&nd->qhi[2]);  // qhi[2..5]
    if (COUNT) cnt.nodes++;
    const uint32_t en = __float_as_uint(h0.w);
    const float sx = scale_of(en & 0xffu), sy = scale_of((en >> 8) & 0xffu), sz = scale_of((en >> 16) & 0xffu);
    const uint32_t nchild = en >> 24;
    const uint32_t qlo[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
    const uint32_t qhi[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
    const uint32_t ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    // The slab test in the node's quantized frame: the plane of byte q on
    // axis j is origin_j + q s_j, reached at t = q (s_j ix_j) + (origin_j -
    // o_j) ix_j: one fused multiply-add per plane (s_j ix_j is exact, s_j a
    // power of two).  The rounding of b_j = (origin_j - o_j) ix_j is absolute
    // in t, and so is the rounding the decode origin_j + q s_j had where the
    // builder checked the boxes conservative: E_j = 2^-22 (|b_j| + (|origin_j|
    // + 255 s_j) |ix_j|) bounds both (and the rounding of b_j -+ E_j), so the
    // near planes start from b_j - E_j and the far planes from b_j + E_j; the
    // relative roundings of the fused multiply-adds and of the comparison
    // take the factor 1 + 2^-21.  No child box the ray enters is culled
    // (ix is clamped to +-2^64: for a vanishing direction component the
    // slab distances shrink, which culls less).  Near and far planes per
    // axis by the sign of ix (t is monotone in q).  The host traversal
    // (host/trace.cpp) computes the same values.
    const f3 o = t.r.o;
    const float ixv[3] = {t.ix, t.iy, t.iz}, orv[3] = {h0.x, h0.y, h0.z}, sv[3] = {sx, sy, sz};
    const float rov[3] = {o.x, o.y, o.z};
    const float kInf = __builtin_inff();
    float av[3], bn[3], bf[3];
    uint32_t qn[6], qf[6];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const float ix = ixv[j];
      const float b = (orv[j] - rov[j]) * ix;
      const float e = (fabsf(b) + (fabsf(orv[j]) + 255.f * sv[j]) * fabsf(ix)) * 0x1p-22f;
      av[j] = sv[j] * ix;
      bn[j] = b - e;
      bf[j] = b + e;
      const bool pos = ix > 0.f;
#pragma unroll
      for (int w = 0; w < 2; w++) {
        qn[2 * j + w] = pos ? qlo[2 * j + w] : qhi[2 * j + w];
        qf[2 * j + w] = pos ? qhi[2 * j + w] : qlo[2 * j + w];
      }
    }
    float tk[8];
    uint32_t ck[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int w = k >> 2, sh = (k & 3) * 8;
      const float tnx = __builtin_fmaf(float((qn[0 + w] >> sh) & 0xffu), av[0], bn[0]);
      const float tny = __builtin_fmaf(float((qn[2 + w] >> sh) & 0xffu), av[1], bn[1]);
      const float tnz = __builtin_fmaf(float((qn[4 + w] >> sh) & 0xffu), av[2], bn[2]);
      const float tfx = __builtin_fmaf(float((qf[0 + w] >> sh) & 0xffu), av[0], bf[0]);
      const float tfy = __builtin_fmaf(float((qf[2 + w] >> sh) & 0xffu), av[1], bf[1]);
      const float tfz = __builtin_fmaf(float((qf[4 + w] >> sh) & 0xffu), av[2], bf[2]);
      const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.f));
      const float tf = fminf(fminf(tfx, tfy), fminf(tfz, t.best.t));
      const bool hit = (uint32_t(k) < nchild) && (tn <= tf * kRobustQ);
      tk[k] = hit ? fminf(tn, kFltMax) : kInf;
      ck[k] = ch[k];
    }
    if constexpr (ANY) {
      // any hit: the order of the children does not matter, so no sort:
      // descend into the first hit child in slot order, push the others
      int nh = 0;
#pragma unroll
      for (int k = 0; k < 8; k++) nh += tk[k] != kInf ? 1 : 0;
      if (nh > 0) {
        const int top = t.sp + nh - 1;
        if (top > kStack) {
          overflow = true;
          return false;
        }
        uint32_t first = 0;
        int j = 0;
        const bool lds = stk.in_lds(top);
#pragma unroll
        for (int k = 0; k < 8; k++) {
          if (tk[k] != kInf) {
            if (j == 0)
              first = ck[k];
            else if (lds)
              stk.put_lds(t.sp + j - 1, ck[k]);
            else
              stk.put(t.sp + j - 1, ck[k]);
            j++;
          }
        }
        t.sp = top;
        t.cur = first;
        return true;
      }
    } else {
    // sorting network for 8 (19 comparators), nearest first
    cswap(tk[0], ck[0], tk[1], ck[1]); cswap(tk[2], ck[2], tk[3], ck[3]);
    cswap(tk[4], ck[4], tk[5], ck[5]); cswap(tk[6], ck[6], tk[7], ck[7]);
    cswap(tk[0], ck[0], tk[2], ck[2]); cswap(tk[1], ck[1], tk[3], ck[3]);
    cswap(tk[4], ck[4], tk[6], ck[6]); cswap(tk[5], ck[5], tk[7], ck[7]);
    cswap(tk[1], ck[1], tk[2], ck[2]); cswap(tk[5], ck[5], tk[6], ck[6]);
    cswap(tk[0], ck[0], tk[4], ck[4]); cswap(tk[3], ck[3], tk[7], ck[7]);
    cswap(tk[1], ck[1], tk[5], ck[5]); cswap(tk[2], ck[2], tk[6], ck[6]);
    cswap(tk[1], ck[1], tk[4], ck[4]); cswap(tk[3], ck[3], tk[6], ck[6]);
    cswap(tk[2], ck[2], tk[4], ck[4]); cswap(tk[3], ck[3], tk[5], ck[5]);
    cswap(tk[3], ck[3], tk[4], ck[4]);
    if (tk[0] != kInf) {
      // the hit children are ck[0 .. nh-1] (misses sorted last as +inf):
      // descend into ck[0], push the others farthest first (ck[nh-1] at
      // sp, ..., ck[1] at sp + nh - 2)
      int nh = 1;
#pragma unroll
      for (int k = 1; k < 8; k++) nh += tk[k] != kInf ? 1 : 0;
      const int top = t.sp + nh - 1;
      if (top > kStack) {
        overflow = true;
        return false;
      }
      if (stk.in_lds(top)) {
#pragma unroll
        for (int k = 1; k < 8; k++)
          if (k < nh) stk.put_lds(top - k, ck[k]);
      } else {
#pragma unroll
        for (int k = 1; k < 8; k++)
          if (k < nh) stk.put(top - k, ck[k]);
      }
      t.sp = top;
      t.cur = ck[0];
      return true;
    }
    }
  } else {
    const uint32_t first = cur & 0x07FFFFFFu;
    const uint32_t count = ((cur >> 27) & 7u) + 1u;
    for (uint32_t i = 0; i < count; i++) {
      const uint32_t s = first + i;
      const float4 a = sc.tris[3 * s + 0];
      const float4 b = sc.tris[3 * s + 1];
      const float4 c = sc.tris[3 * s + 2];
      if (COUNT) cnt.tris++;
      const uint32_t obj = __float_as_uint(c.y);
      if (obj == t.r.skip) continue;
      const f3 v0 = mk(a.x, a.y, a.z), v1 = mk(a.w, b.x, b.y), v2 = mk(b.z, b.w, c.x);
      const float th = tri_t(v0, v1, v2, t.r.o, t.r.d, t.mr);
      const uint32_t prim = __float_as_uint(c.z);
      // selects rather than a branch: hipcc for gfx950 lowered the branch
      // form with the new input index lost on ties won by a lower id (the
      // register it sat in was reset for every lane of the t >= best.t arm)
      const bool bt = th >= kBias && better(th, obj, prim, t.best);
      t.best.t = bt ? th : t.best.t;
      t.best.slot = bt ? s : t.best.slot;
      t.best.obj = bt ? obj : t.best.obj;
      t.best.prim = bt ? prim : t.best.prim;
      t.orig = bt ? __float_as_uint(c.w) : t.orig;
      if (ANY && bt) return false;
    }
  }
  if (t.sp == 0) return false;
  t.sp--;
  t.cur = stk.get(t.sp);
  return true;
}

// The whole traversal of one ray on the lane's stack.
template <bool ANY, bool COUNT, class Stack>
__device__ __forceinline__ Hit trace(const DevScene &sc, const Ray &r, const Stack &stk, TravCounts &cnt,
                                     bool &overflow) {
  Trav t;
  if (trav_begin<ANY>(sc, r, t))
    while (trav_step<ANY, COUNT>(sc, t, stk, cnt, overflow)) {
    }
  return t.best;
}

// Persistent traversal of a queue of n rays with per-lane refill: a lane
// whose ray has ended takes the next ray of the queue once at least
// kRefill lanes of its wave are idle, so a wave keeps most lanes tracing
// until the queue runs dry instead of waiting for its longest ray.  Ray
// indices come from a wave-local pool: block b starts with rays
// [64 b, 64 b + 64) (no atomic), and an exhausted pool is refilled with
// `psize` consecutive indices from *fetch (ray count past 64 gridDim.x,
// zero at launch), one device atomic per reservation.  The reservation
// size gives each wave about kPoolRounds reservations of the queue's
// remainder (32..kPool rays): large enough that the atomic's latency is
// rare, small enough that the last reservations do not leave a few waves
// with most of the work.  load(i) returns queue entry i (an item with the
// ray as member r, and whatever done needs: loaded with the ray, in one
// round trip); done(i, item, trav, overflow, counts) consumes its result
// (trav.best, trav.orig; counts: nodes and triangles of this ray when
// COUNT).
// (Measured: holding a second, prefetched entry per lane and starting it
// without waiting was 13-20% slower on the extend and shadow batches: the
// prefetch's HBM round trip then sits in front of every following node
// fetch on the wave's in-order memory counter.)
constexpr uint32_t kRefill = 16;
constexpr uint32_t kPool = 256;
constexpr uint32_t kPoolRounds = 2;

// queue entry of a plain ray batch
struct RayItem {
  Ray r;
};

// Wave-uniform pool of queue indices [pool, pool_end) (see trace_queue).
// dry: the queue has no unreserved entries left.
struct QueuePool {
  uint32_t pool, pool_end, psize, n;
  bool dry;
  __device__ __forceinline__ void init(uint32_t n_) {
    n = n_;
    pool = blockIdx.x * 64u;
    pool_end = pool < n ? min(pool + 64u, n) : pool;
    dry = pool >= n;
    const uint32_t first_round = 64u * gridDim.x;
    psize = n > first_round ? min(kPool, max(32u, (n - first_round) / (gridDim.x * kPoolRounds) + 1u)) : 32u;
  }
  // indices left to hand out (in the pool or still reservable)
  __device__ __forceinline__ bool more() const { return pool_end > pool || !dry; }
  // k_finish's refill rule: at least `least` idle lanes and indices left
  __device__ __forceinline__ bool done(uint32_t nidle) const { return nidle == 64 && !more(); }
  __device__ __forceinline__ bool can_refill(uint32_t nidle, uint32_t least = 16) const {
    return nidle >= least && more();
  }
  // Hands the lanes of ballot `idle` (nidle of them) the next indices: the
  // rest of the pool, then a new reservation from *fetch.  Returns this
  // lane's index (n for none; only meaningful where `is_idle`).
  __device__ __forceinline__ uint32_t take(uint64_t idle, uint32_t nidle, bool is_idle, uint32_t *fetch) {
    const uint32_t lane = threadIdx.x % 64;
    const uint32_t k1 = min(nidle, pool_end - pool);
    uint32_t nb = 0, nb_cnt = 0;
    if (k1 < nidle && !dry) {
      uint32_t r = 0;
      if (lane == 0) r = atomicAdd(fetch, psize);
      nb = 64u * gridDim.x + __shfl(r, 0, 64);
      if (nb >= n)
        dry = true;
      else
        nb_cnt = min(psize, n - nb);
    }
    uint32_t idx = n;
    if (is_idle) {
      const uint32_t rank = __popcll(idle & ((uint64_t(1) << lane) - 1));
      idx = rank < k1 ? pool + rank : (rank - k1 < nb_cnt ? nb + (rank - k1) : n);
    }
    if (nb_cnt > 0) {
      pool = nb + min(nidle - k1, nb_cnt);
      pool_end = nb + nb_cnt;
    } else {
      pool += k1;
    }
    return idx;
  }
};

template <bool ANY, bool COUNT, class Stack, class Load, class Done>
__device__ __forceinline__ void trace_queue(const DevScene &sc, uint32_t n, uint32_t *fetch, const Stack &stk,
                                            Load load, Done done) {
  using Item = decltype(load(0u));
  TravCounts cnt{0, 0};
  Trav t;
  bool active = false, ovf = false;
  uint32_t idx = 0;
  Item cur{};
  QueuePool qp;
  qp.init(n);
  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = __popcll(idle);
    if (qp.done(nidle)) break;
    if (qp.can_refill(nidle, kRefill)) {
      const uint32_t i = qp.take(idle, nidle, !active, fetch);
      if (!active && i < n) {
        idx = i;
        ovf = false;
        cnt = TravCounts{0, 0};
        cur = load(idx);
        active = trav_begin<ANY>(sc, cur.r, t);
        if (!active) done(idx, cur, t, false, cnt);
      }
    }
    if (active) {
      active = trav_step<ANY, COUNT>(sc, t, stk, cnt, ovf);
      if (!active) done(idx, cur, t, ovf, cnt);
    }
  }
}

}  // namespace lam

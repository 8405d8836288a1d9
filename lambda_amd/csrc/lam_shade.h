// Hero-wavelength shading on the device: hit records -> surface frames,
// BSDFs, light sampling, material sampling, MIS weights, camera and hero
// sampling.  Each function restates the reference function named in its
// comment, with the same float evaluation order; H is the number of hero
// wavelengths carried by a path (16 in the reference, cg.hpp:41).
#pragma once
#include "lam_device.h"

namespace lam {

// per-path spectral lookup: λ_k -> (bin i_k, fraction t_k) of the
// coefficient grid, reused by every spectrum evaluation on the path
template <int H>
struct Lam {
  float l[H];
  uint32_t i[H];
  float t[H];
};

// coefficientλ::operator()(float) split into index/fraction (cg.hpp:289-297).
// The clamped cases map to (0,0) and (n-2,1), which reproduce v[0] and
// v[n-1] exactly in the interpolation formula.
__device__ __forceinline__ void lookup(const DevScene &sc, float l, uint32_t &i, float &t) {
  if (l <= sc.grid_inf) { i = 0; t = 0.f; return; }
  if (l >= sc.grid_sup) { i = sc.grid_n - 2; t = 1.f; return; }
  const float x = (l - sc.grid_inf) / sc.grid_delta;
  uint32_t ii = static_cast<uint32_t>(x);
  if (ii > sc.grid_n - 2) ii = sc.grid_n - 2;
  i = ii;
  t = x - float(ii);
}

__device__ __forceinline__ const float *spec_ptr(const DevScene &sc, int idx) {
  return sc.spectra + size_t(idx) * sc.grid_n;
}
__device__ __forceinline__ float spec_at(const float *v, uint32_t i, float t) {
  return (1.f - t) * v[i] + t * v[i + 1];
}
template <int H>
__device__ __forceinline__ void spec_eval(const DevScene &sc, int idx, const Lam<H> &L, float *out) {
  const float *v = spec_ptr(sc, idx);
#pragma unroll
  for (int k = 0; k < H; k++) out[k] = spec_at(v, L.i[k], L.t[k]);
}
__device__ __forceinline__ float spec_eval1(const DevScene &sc, int idx, float l) {
  uint32_t i;
  float t;
  lookup(sc, l, i, t);
  return spec_at(spec_ptr(sc, idx), i, t);
}

// ---- frames -------------------------------------------------------------------

struct Frame {
  f3 x, y, z;
};
// orthonormalBasisOf (cg.hpp:384-405)
__device__ __forceinline__ Frame onb(f3 v) {
  const f3 e2 = normalized(v);
  const float x = e2.x, y = e2.y, z = e2.z;
  Frame F;
  if (z >= y) {
    const float a = 1.f / (1.f + z);
    const float b = -x * y * a;
    F.x = mk(1 - x * x * a, b, -x);
    F.y = mk(b, 1 - y * y * a, -y);
  } else {
    const float a = 1.f / (1.f + y);
    const float b = -x * z * a;
    F.x = mk(b, -z, 1 - z * z * a);
    F.y = mk(1 - x * x * a, -x, b);
  }
  F.z = e2;
  return F;
}
// basis::toBasis (cg.hpp:378)
__device__ __forceinline__ f3 to_basis(const Frame &F, f3 v) {
  return add(add(mul(F.x, v.x), mul(F.y, v.y)), mul(F.z, v.z));
}
// reflect (bra.hpp:397-398)
__device__ __forceinline__ f3 reflect(f3 x, f3 a) { return add(neg(x), mul(a, 2 * dot(x, a))); }

// hitinfo (cg.hpp:1138-1149)
struct Surf {
  f3 p;
  Frame F;
  f3 gn;
  bool front;
  uint32_t mat, obj;
  __device__ __forceinline__ f3 n() const { return front ? F.z : neg(F.z); }
};

__device__ __forceinline__ f3 col(const float *M, int j) { return mk(M[j], M[4 + j], M[8 + j]); }

// sphere hit frame (cg.hpp:1216-1240)
__device__ __forceinline__ void sphere_surf(const lam_object &ob, f3 lp, f3 wu, Surf &s) {
  const f3 n = normalized(normal_to_world(ob.T.Minv, lp));
  s.front = dot(n, wu) < 0.f;
  const float ct = n.z;
  const float st = sqrtf(1.f - ct * ct);
  const float phi = lam_atan2f(n.y, n.x);
  float spp, cp;
  lam_sincosf(phi, &spp, &cp);
  s.F.x = mk(-st * spp, st * cp, 0.f);
  s.F.y = mk(ct * cp, ct * spp, -st);
  s.F.z = n;
  s.gn = n;
  s.p = xmul(ob.T.M, lp, 1.f);
}
// plane hit frame (cg.hpp:1254-1261)
__device__ __forceinline__ void plane_surf(const lam_object &ob, f3 lp, f3 wu, Surf &s) {
  s.p = xmul(ob.T.M, lp, 1.f);
  s.F.z = normalized(normal_to_world(ob.T.Minv, mk(0.f, 0.f, 1.f)));
  s.F.x = normalized(normal_to_world(ob.T.Minv, mk(0.f, 1.f, 0.f)));
  s.F.y = cross(s.F.z, s.F.x);
  s.gn = s.F.z;
  s.front = dot(wu, s.F.z) < 0.f;
}

// Surface record of a closest hit.  Triangles: world hit point o + t d,
// shading normal = barycentric blend of the world vertex normals, frame =
// orthonormalBasisOf of it (cg.hpp:1371-1377, 1450-1452).
__device__ __forceinline__ Surf surface(const DevScene &sc, const Hit &h, f3 o, f3 d) {
  Surf s;
  s.obj = h.obj;
  const lam_object &ob = sc.objects[h.obj];
  s.mat = ob.material;
  if (h.slot != LAM_NONE) {
    // both records requested before either is used: one round trip
    const float4 a = sc.tris[3 * h.slot + 0];
    const float4 b = sc.tris[3 * h.slot + 1];
    const float4 c = sc.tris[3 * h.slot + 2];
    const float4 n0 = sc.shade[3 * h.slot + 0];
    const float4 n1 = sc.shade[3 * h.slot + 1];
    const float4 n2 = sc.shade[3 * h.slot + 2];
    __builtin_amdgcn_sched_barrier(0);
    const f3 bb = tri_bary(mk(a.x, a.y, a.z), mk(a.w, b.x, b.y), mk(b.z, b.w, c.x), o, d);
    const f3 N0 = mk(n0.x, n0.y, n0.z), N1 = mk(n0.w, n1.x, n1.y), N2 = mk(n1.z, n1.w, n2.x);
    s.gn = mk(n2.y, n2.z, n2.w);
    s.p = add(o, mul(d, h.t));
    const f3 ns = add(add(mul(N0, bb.x), mul(N1, bb.y)), mul(N2, bb.z));
    s.F = onb(ns);
    s.front = dot(s.gn, d) < 0.f;
  } else {
    f3 lo, lu;
    to_local(ob, o, d, lo, lu);
    const f3 lp = add(lo, mul(lu, h.t));
    if (ob.type == LAM_OBJ_SPHERE) sphere_surf(ob, lp, d, s);
    else plane_surf(ob, lp, d, s);
  }
  return s;
}

// ---- sampling helpers (stoch.hpp:53-68, cg.hpp:1790-1796) --------------------
__device__ __forceinline__ f3 cos_hemi(float x0, float x1) {
  const float cc = 1.f - x0;
  const float ct = sqrtf(cc);
  const float st = sqrtf(1.f - cc);
  const float phi = 2.f * kPi * x1;
  float sp, cp;
  lam_sincosf(phi, &sp, &cp);
  return mk(st * cp, st * sp, ct);
}
__device__ __forceinline__ f3 unif_hemi(float x0, float x1) {
  const float ct = 1.f - x0;
  const float st = sqrtf(1.f - ct * ct);
  const float phi = 2.f * kPi * x1;
  float sp, cp;
  lam_sincosf(phi, &sp, &cp);
  return mk(st * cp, st * sp, ct);
}
__device__ __forceinline__ f3 blinn_h(float alpha, float x0, float x1) {
  const float ct = lam_powf(1.f - x0, 1.f / (alpha + 1.f));
  const float st = sqrtf(1.f - ct * ct);
  const float phi = 2.f * kPi * x1;
  float sp, cp;
  lam_sincosf(phi, &sp, &cp);
  return mk(st * cp, st * sp, ct);
}

// ---- materials ------------------------------------------------------------------
__device__ __forceinline__ float F0s(float n) { return (n * n - 2.f * n + 1.f) / (n * n + 2.f * n + 1.f); }
__device__ __forceinline__ float fresnel_s(float F0, float c) {
  return F0 + (1.f - F0) * lam_powf(1.f - c, 5.f);
}
// blinn::fresnelSplit (cg.hpp:950-959)
__device__ __forceinline__ void fresnel_split(const lam_material &m, float F, float &pr, float &pt) {
  float p_r = m.w_r + F * m.w_t;
  float p_t = (1.f - F) * m.w_t;
  const float tot = p_r + p_t;
  if (tot == 0.f) { pr = 1.f; pt = 0.f; return; }
  pr = p_r / tot;
  pt = p_t / tot;
}
__device__ __forceinline__ float blinnh_prob(float alpha, float ch) {
  return lam_powf(fabsf(ch), alpha + 1) * (alpha + 1) * .5f * kInvPi;
}
// h = normalize(i+o) with the degenerate fallback (cg.hpp:976-979)
__device__ __forceinline__ f3 half_vec(f3 i, f3 o, f3 n) {
  f3 h = add(i, o);
  const float l2 = len2(h);
  if (l2 < kEps) h = normalized(sub(n, mul(o, dot(n, o))));
  else h = divs(h, sqrtf(l2));
  return h;
}

// blinn::BSDFcosθ (cg.hpp:967-986)
template <int H>
__device__ void blinn_f(const DevScene &sc, const lam_material &m, const Lam<H> &L, f3 i, f3 o, f3 n,
                        bool front, float *out) {
  float ior[H], eta[H], F[H];
  spec_eval<H>(sc, m.spec[4], L, ior);
#pragma unroll
  for (int k = 0; k < H; k++) eta[k] = front ? 1.f / ior[k] : ior[k];
  const float ci = dot(i, n), co = dot(o, n);
  const float pw5 = lam_powf(1.f - co, 5.f);
#pragma unroll
  for (int k = 0; k < H; k++) {
    const float f0 = F0s(eta[k]);
    F[k] = f0 + (1.f - f0) * pw5;
  }
  const float a2 = (m.alpha + 2) * .125f;
  if (ci * co > 0.f) {
    const f3 h = half_vec(i, o, n);
    const float pw = lam_powf(dot(h, n), m.alpha);
    const float md = fmax_ref(0.f, ci);
    float kd[H], ks[H], kt[H];
    spec_eval<H>(sc, m.spec[0], L, kd);
    spec_eval<H>(sc, m.spec[1], L, ks);
    spec_eval<H>(sc, m.spec[2], L, kt);
#pragma unroll
    for (int k = 0; k < H; k++) {
      const float fd = kd[k] * kInvPi * md;
      const float fr = (ks[k] + F[k] * kt[k]) * (m.alpha + 2) * .125f * kInvPi * pw;
      out[k] = fd + fr;
    }
  } else {
    const f3 ht = neg(normalized(add(i, mul(o, eta[0]))));
    const float pw = lam_powf(fabsf(dot(ht, n)), m.alpha);
    float kt[H];
    spec_eval<H>(sc, m.spec[2], L, kt);
#pragma unroll
    for (int k = 0; k < H; k++) out[k] = kt[k] * (1.f - F[k]) * (m.alpha + 2) * .125f * kInvPi * pw;
  }
  (void)a2;
}

// thinfilm::BRDFcosθ (cg.hpp:1031-1051), Airy two-interface reflectance with
// the reference's cos_l expression (cg.hpp:1007)
template <int H>
__device__ void thinfilm_f(const DevScene &sc, const lam_material &m, const Lam<H> &L, f3 i, f3 o, f3 n,
                           bool front, float *out) {
  if (!front || dot(i, n) < 0.f || dot(o, n) < 0.f) {
#pragma unroll
    for (int k = 0; k < H; k++) out[k] = 0.f;
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
#pragma unroll
  for (int k = 0; k < H; k++) {
    const float n01 = 1.f / n1[k];
    const float cos_l = sqrtf(1.f - n01 * n01) * (1.f - ci * ci);
    const float cos_d = lam_cosf(4 * kPi * n1[k] * m.d * cos_l / L.l[k]);
    const float a01 = 1.f - n1[k], b01 = 1.f + n1[k];
    const float r0_01 = a01 * a01 / (b01 * b01);
    const float a12 = n1[k] - n2[k], b12 = n1[k] + n2[k];
    const float r0_12 = a12 * a12 / (b12 * b12);
    const float R01 = r0_01 + (1.f - r0_01) * pw5;
    const float R12 = r0_12 + (1.f - r0_12) * pw5;
    const float ravg = 2 * sqrtf(R01 * R12);
    const float rt = (R01 + R12 + ravg * cos_d) / (R01 * R12 + 1.f + ravg * cos_d);
    out[k] = rt * c * pw;
  }
}

// BxDFcosθ dispatch (cg.hpp:1088-1102)
template <int H>
__device__ void bxdf(const DevScene &sc, const lam_material &m, const Lam<H> &L, f3 i, f3 o, f3 n, bool front,
                     float *out) {
  if (m.type == LAM_MAT_LAMBERTIAN) {
    float a[H];
    spec_eval<H>(sc, m.spec[0], L, a);
    const float md = fmax_ref(0.f, dot(i, n));
#pragma unroll
    for (int k = 0; k < H; k++) out[k] = a[k] * kInvPi * md;
  } else if (m.type == LAM_MAT_BLINN) {
    blinn_f<H>(sc, m, L, i, o, n, front, out);
  } else if (m.type == LAM_MAT_THINFILM) {
    thinfilm_f<H>(sc, m, L, i, o, n, front, out);
  } else {
#pragma unroll
    for (int k = 0; k < H; k++) out[k] = 0.f;
  }
}

// ---- material sampling (cg.hpp:1755-1968) --------------------------------------
template <int H>
struct MatSample {
  bool ok;
  f3 dir;
  float prob;
  float mult[H];
};

template <int H>
__device__ void sample_material(const DevScene &sc, const lam_material &m, const Lam<H> &L, const Surf &s, f3 o,
                                float xl, float xa, float xb, float p, MatSample<H> &ms) {
  ms.ok = false;
  if (m.type == LAM_MAT_LAMBERTIAN) {  // lambertiani (cg.hpp:1755-1778)
    if (xl > p) return;
    f3 dir = cos_hemi(xa, xb);
    if (!s.front) dir.z *= -1;
    const f3 i = to_basis(s.F, dir);
    ms.prob = p * fabsf(dir.z) * kInvPi;
    ms.dir = i;
    bxdf<H>(sc, m, L, i, o, s.n(), s.front, ms.mult);
    ms.ok = true;
  } else if (m.type == LAM_MAT_BLINN) {  // blinni (cg.hpp:1799-1863)
    if (xl >= p) return;
    const float pp_d = p * m.p_d;
    const float pp_spec = p * m.p_spec;
    if (xl < pp_spec) {
      float ior[H], eta[H], F[H];
      spec_eval<H>(sc, m.spec[4], L, ior);
#pragma unroll
      for (int k = 0; k < H; k++) eta[k] = s.front ? 1.f / ior[k] : ior[k];
      f3 w = blinn_h(m.alpha, xa, xb);
      if (!s.front) w.z *= -1;
      const f3 h = to_basis(s.F, w);
      const f3 iR = reflect(o, h);
      // transmit (cg.hpp:500-513): TIR <=> zero vector
      const float cos_i = dot(o, h);
      const float pw5 = lam_powf(1.f - cos_i, 5.f);
      f3 iT0 = mk(0.f, 0.f, 0.f);
#pragma unroll
      for (int k = 0; k < H; k++) {
        const float ct2 = 1.f - eta[k] * eta[k] * (1.f - cos_i * cos_i);
        f3 td = mk(0.f, 0.f, 0.f);
        if (!(ct2 < 0.f)) td = sub(mul(o, -eta[k]), mul(h, sqrtf(ct2) - eta[k] * cos_i));
        if (k == 0) iT0 = td;
        const float f0 = F0s(eta[k]);
        F[k] = f0 + (1.f - f0) * pw5;
        if (td.x == 0.f && td.y == 0.f && td.z == 0.f) F[k] = 1.f;
      }
      float p_r, p_t;
      fresnel_split(m, F[0], p_r, p_t);
      const float bp = blinnh_prob(m.alpha, w.z);
      const float pw = lam_powf(w.z, m.alpha);
      if (xl < pp_spec * p_r) {
        float ks[H], kt[H];
        spec_eval<H>(sc, m.spec[1], L, ks);
        spec_eval<H>(sc, m.spec[2], L, kt);
        ms.prob = pp_spec * p_r * bp * .25f;
        ms.dir = iR;
#pragma unroll
        for (int k = 0; k < H; k++) ms.mult[k] = (ks[k] + F[k] * kt[k]) * (m.alpha + 2) * .125f * kInvPi * pw;
      } else {
        float kt[H];
        spec_eval<H>(sc, m.spec[2], L, kt);
        ms.prob = pp_spec * p_t * bp * .25f;
        ms.dir = iT0;
#pragma unroll
        for (int k = 0; k < H; k++) ms.mult[k] = kt[k] * (1.f - F[k]) * (m.alpha + 2) * .125f * kInvPi * pw;
      }
      ms.ok = true;
    } else if (xl < pp_spec + pp_d) {
      f3 dir = cos_hemi(xa, xb);
      if (!s.front) dir.z *= -1;
      const f3 i = to_basis(s.F, dir);
      ms.prob = pp_d * (fabsf(dir.z) * kInvPi);
      ms.dir = i;
      float kd[H];
      spec_eval<H>(sc, m.spec[0], L, kd);
      const float md = fmax_ref(0.f, dir.z);
#pragma unroll
      for (int k = 0; k < H; k++) ms.mult[k] = kd[k] * kInvPi * md;
      ms.ok = true;
    }
  } else if (m.type == LAM_MAT_THINFILM) {  // thinfilmi (cg.hpp:1905-1931)
    const f3 n = s.F.z;
    if (!s.front || dot(o, n) < 0.f) return;
    if (xl >= p) return;
    const f3 w = blinn_h(m.alpha, xa, xb);
    const f3 h = to_basis(s.F, w);
    const f3 i = reflect(o, h);
    ms.prob = p * lam_powf(dot(n, h), m.alpha + 1) * (m.alpha + 1) * .5f * kInvPi * .25f;
    ms.dir = i;
    thinfilm_f<H>(sc, m, L, i, o, n, s.front, ms.mult);
    ms.ok = true;
  }
}

// probForMateriali (cg.hpp:1971-1985): pdf of the material sampler for
// direction `wl` (the light sample), evaluated at the hero wavelength l0.
// The reference passes (i, o) into probForBlinn/probForThinFilm's (o, i)
// parameters; the swap is kept.
__device__ float prob_material(const DevScene &sc, const lam_material &m, const Surf &s, f3 wl, f3 o, float l0,
                               float p) {
  if (m.type == LAM_MAT_LAMBERTIAN) {
    const f3 n = s.n();
    const float cos_i = dot(wl, n);
    if (dot(wl, n) < 0.f || dot(o, n) < 0.f) return 0.f;
    return p * cos_i * kInvPi;
  }
  if (m.type == LAM_MAT_BLINN) {
    const f3 oo = wl, ii = o;  // probForBlinn's (o, i)
    const f3 n = s.n();
    const float iorl = spec_eval1(sc, m.spec[4], l0);
    const float ior = s.front ? 1.f / iorl : iorl;
    const float cin = dot(ii, n), con = dot(oo, n);
    f3 h;
    if (cin * con > 0.f) h = half_vec(ii, oo, n);
    else h = neg(normalized(add(ii, mul(oo, ior))));
    const float cos_o = fabsf(dot(oo, h));
    const float F = fresnel_s(F0s(ior), cos_o);
    float p_r, p_t;
    fresnel_split(m, F, p_r, p_t);
    if (cin * con > 0.f) {
      return p * (m.p_d * (fabsf(cin) * kInvPi) + m.p_spec * p_r * blinnh_prob(m.alpha, dot(h, n)) * .25f);
    }
    const float c2 = 1.f - ior * ior * (1.f - cos_o * cos_o);
    if (c2 < 0.f) return 0.f;
    return p * m.p_spec * p_t * blinnh_prob(m.alpha, dot(h, n)) * .25f;
  }
  if (m.type == LAM_MAT_THINFILM) {
    const f3 oo = wl, ii = o;  // probForThinFilm's (o, i)
    const f3 n = s.F.z;
    if (!s.front || dot(ii, n) < 0.f || dot(oo, n) < 0.f) return 0.f;
    const f3 h = half_vec(ii, oo, n);
    return p * lam_powf(dot(oo, h), m.alpha + 1) * (m.alpha + 1) * .5f * kInvPi * .25f;
  }
  return 0.f;
}

// ---- lights (cg.hpp:1548-1747) -------------------------------------------------------
template <int H>
struct LightSample {
  f3 dir;
  float prob;     // p(ωi | light)
  float mult[H];  // unshadowed L(ωi) factor
  bool shadow;    // needs a shadow query
  Ray sray;
};

template <int H>
__device__ void sample_light(const DevScene &sc, const lam_light &lt, const Lam<H> &L, const Surf &s, float xa,
                             float xb, LightSample<H> &ls) {
  float rad[H];
  spec_eval<H>(sc, lt.spec, L, rad);
  ls.shadow = true;
  ls.sray.o = s.p;
  ls.sray.skip = LAM_NONE;
  if (lt.type == LAM_LIGHT_SPHERE) {  // spherelight (cg.hpp:1671-1700)
    const lam_object &ob = sc.objects[lt.object];
    const f3 Lv = sub(col(ob.T.M, 3), s.p);
    const float dist2 = dot(Lv, Lv);
    const float r2 = lt.size * lt.size;
    const float sr = 1.f - sqrtf(fmax_ref(0.f, 1.f - r2 / dist2));
    ls.prob = 0.5f / (kPi * sr);
    const Frame B = onb(Lv);
    const float ct = fmax_ref(0.f, 1.f - xa * sr);
    const float st = sqrtf(1.f - ct * ct);
    const float phi = 2 * kPi * xb;
    float sp, cp;
    lam_sincosf(phi, &sp, &cp);
    const f3 w = add(add(mul(mul(B.x, st), cp), mul(mul(B.y, st), sp)), mul(B.z, ct));
    ls.dir = w;
    ls.sray.tmax = sqrtf(dist2);
    ls.sray.skip = lt.object;
#pragma unroll
    for (int k = 0; k < H; k++) ls.mult[k] = rad[k];
  } else if (lt.type == LAM_LIGHT_QUADDIR) {  // quaddirlight (cg.hpp:1625-1655)
    const lam_object &ob = sc.objects[lt.object];
    const float sx = 2.f * xa - 1.f;
    const float sy = 2.f * xb - 1.f;
    const f3 x = add(add(col(ob.T.M, 3), mul(col(ob.T.M, 0), sx)), mul(col(ob.T.M, 1), sy));
    const f3 r = sub(x, s.p);
    const f3 w = normalized(r);
    ls.dir = w;
    const float dist2 = dot(r, r);
    const float cs = lam_powf(dot(neg(w), normalized(col(ob.T.M, 2))), lt.specular);
    ls.prob = fmax_ref(1e-6f, (lt.specular + 1) * .5f * kInvPi * cs * lt.inv_area);
    ls.sray.tmax = sqrtf(dist2);
    ls.sray.skip = lt.object;
#pragma unroll
    for (int k = 0; k < H; k++) ls.mult[k] = rad[k] * cs;
  } else if (lt.type == LAM_LIGHT_POINT) {  // pointlight (cg.hpp:1580-1600)
    const f3 Lv = sub(mk(lt.pos[0], lt.pos[1], lt.pos[2]), s.p);
    const f3 w = normalized(Lv);
    ls.dir = w;
    ls.prob = 1.f;
    const float dist = sqrtf(len2(Lv));
    ls.sray.tmax = dist;
#pragma unroll
    for (int k = 0; k < H; k++) ls.mult[k] = rad[k] / (dist * dist);
  } else if (lt.type == LAM_LIGHT_DIR) {  // dirlight (cg.hpp:1604-1621)
    ls.dir = neg(mk(lt.dir[0], lt.dir[1], lt.dir[2]));
    ls.prob = 1.f;
    ls.sray.tmax = kFltMax;
#pragma unroll
    for (int k = 0; k < H; k++) ls.mult[k] = rad[k];
  } else {  // ambientlight (cg.hpp:1560-1576): no shadow query
    f3 dir = unif_hemi(xa, xb);
    if (!s.front) dir.z *= -1.f;
    ls.dir = to_basis(s.F, dir);
    ls.prob = .5f * kInvPi;
    ls.shadow = false;
#pragma unroll
    for (int k = 0; k < H; k++) ls.mult[k] = rad[k];
  }
  ls.sray.d = ls.dir;
}

// probForLight (cg.hpp:1735-1747) for the ray (p, dir)
__device__ float prob_light(const DevScene &sc, const lam_light &lt, f3 p, f3 dir) {
  if (lt.type == LAM_LIGHT_SPHERE) {  // probForSphereLight (cg.hpp:1702-1714)
    const lam_object &ob = sc.objects[lt.object];
    f3 lo, lu;
    to_local(ob, p, dir, lo, lu);
    const float t = sphere_t(lo, lu);
    if (!(t >= kBias)) return 0.f;
    const f3 lp = add(lo, mul(lu, t));
    const f3 n = normalized(normal_to_world(ob.T.Minv, lp));
    if (!(dot(n, dir) < 0.f)) return 0.f;
    const f3 Lv = sub(col(ob.T.M, 3), p);
    const float dist2 = dot(Lv, Lv);
    const float sr = 1.f - sqrtf(fmax_ref(0.f, 1.f - lt.size * lt.size / dist2));
    return .5f / (kPi * sr);
  }
  if (lt.type == LAM_LIGHT_QUADDIR) {  // probForQuadDirLight (cg.hpp:1657-1668)
    const lam_object &ob = sc.objects[lt.object];
    f3 lo, lu;
    to_local(ob, p, dir, lo, lu);
    const float t = plane_t(lo, lu);
    if (!(t >= kBias)) return 0.f;
    const f3 nz = normalized(normal_to_world(ob.T.Minv, mk(0.f, 0.f, 1.f)));
    if (!(dot(dir, nz) < 0.f)) return 0.f;
    const float ct = dot(neg(normalized(dir)), nz);
    return fmax_ref(1e-6f, (lt.specular + 1) * .5f * kInvPi * lam_powf(ct, lt.specular) * lt.inv_area);
  }
  if (lt.type == LAM_LIGHT_AMBIENT) return .5f * kInvPi;
  return 0.f;
}

// ---- camera + hero wavelengths ----------------------------------------------------
// sample::camera (cg.hpp:1534-1542); uv in pixels
__device__ __forceinline__ void camera_ray(const lam_camera &c, float u, float v, float lx, float ly, f3 &o, f3 &d) {
  const f3 pos = mk(c.pos[0], c.pos[1], c.pos[2]);
  const f3 X = mk(c.x[0], c.x[1], c.x[2]), Y = mk(c.y[0], c.y[1], c.y[2]), Z = mk(c.z[0], c.z[1], c.z[2]);
  const f3 wij = sub(add(add(pos, mul(X, -0.5f * c.w + c.delta * u)), mul(Y, -0.5f * c.h + c.delta * v)), mul(Z, c.D));
  const f3 wkl = add(pos, mul(add(mul(X, lx - .5f), mul(Y, ly - .5f)), 2 * c.dof));
  o = wkl;
  d = normalized(sub(wij, wkl));
}

// sample::spectrum, hero form (cg.hpp:1517-1531); returns 1/Ω
template <int H>
__device__ __forceinline__ float hero_lambdas(const DevScene &sc, float xi, Lam<H> &L) {
  const float om = sc.grid_sup - sc.grid_inf;
  const float hero = sc.grid_inf + om * xi;
  const float step = om / float(H);
#pragma unroll
  for (int k = 0; k < H; k++) {
    float l = hero + float(k) * step;
    if (l > sc.grid_sup) l -= om;
    L.l[k] = l;
    lookup(sc, l, L.i[k], L.t[k]);
  }
  return 1.f / om;
}

// ---- one path, vertex s0 onwards -------------------------------------------------
// Renderer::tracePath (src/renderer.cpp:64-146) unrolled into a forward
// loop.  The recursion is linear, so
//   Li_s = M_IS,s * w_mat,s + L_IS,s * w_L,s
// becomes L += β_s ⊙ (w_L,s L_IS,s) and β_{s+1} = β_s ⊙ mult_s / prob_s * w_mat,s,
// equal to the reference up to floating-point reassociation.  Starting at
// vertex s0 with (ray, β, L) lets the wavefront hand its tail of long paths
// to one thread each.
// One vertex of the path: intersect, emission, light IS, material IS and
// MIS (src/renderer.cpp:64-146), updating (ray, β, L).  Returns false when
// the path ends (L final), true with the continuation ray in `ray`.
// n_shadow (nullable) counts the shadow rays traced.
template <int H, class Stack>
__device__ __forceinline__ bool path_vertex(const DevScene &sc, uint32_t max_scat, float sample_p, uint64_t key,
                                            const Lam<H> &lam, uint32_t s, Ray &ray, float *beta, float *L,
                                            const Stack &stk, bool &ovf, uint32_t *n_shadow = nullptr) {
  TravCounts tc{0, 0};
  const Hit hit = trace<false, false>(sc, ray, stk, tc, ovf);
  if (hit.obj == LAM_NONE || ovf) return false;
  const Surf sf = surface(sc, hit, ray.o, ray.d);
  const f3 o = neg(normalized(ray.d));
  const f3 n = sf.n();
  const lam_material &mat = sc.materials[sf.mat];
  if (mat.type == LAM_MAT_EMITTER || mat.type == LAM_MAT_DIREMITTER) {
    float le[H];
    spec_eval<H>(sc, mat.spec[0], lam, le);
    float f = 1.f;
    if (mat.type == LAM_MAT_DIREMITTER) f = lam_powf(fmax_ref(dot(o, sf.F.z), 0.f), mat.specular);
#pragma unroll
    for (int k = 0; k < H; k++) L[k] += beta[k] * (mat.type == LAM_MAT_DIREMITTER ? le[k] * f : le[k]);
    return false;
  }
  if (s > max_scat) return false;
  // light IS (src/renderer.cpp:91-106)
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
    if (n_shadow) (*n_shadow)++;
  }
  float lis[H];
#pragma unroll
  for (int k = 0; k < H; k++) lis[k] = ls.mult[k] * coef[k] / (ls.prob * pl);
  // material IS (src/renderer.cpp:116-120)
  MatSample<H> ms;
  sample_material<H>(sc, mat, lam, sf, o, lam_draw_f(key, lam_vdim(s, LAM_VDIM_LOBE)),
                     lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_A)), lam_draw_f(key, lam_vdim(s, LAM_VDIM_DIR_B)),
                     sample_p, ms);
  if (!ms.ok) {
    if (vis) {
#pragma unroll
      for (int k = 0; k < H; k++) L[k] += beta[k] * lis[k];
    }
    return false;
  }
  // MIS (src/renderer.cpp:135-145), power heuristic β=2
  const float p_L_mati = prob_light(sc, lt, sf.p, ms.dir) * pl;
  const float p_mat_Li = prob_material(sc, mat, sf, ls.dir, o, lam.l[0], sample_p);
  const float p_L_i = ls.prob * pl;
  const float w_L = p_L_i * p_L_i / (p_L_i * p_L_i + p_mat_Li * p_mat_Li);
  const float w_mat = ms.prob * ms.prob / (ms.prob * ms.prob + p_L_mati * p_L_mati);
  if (vis) {
#pragma unroll
    for (int k = 0; k < H; k++) L[k] += beta[k] * (lis[k] * w_L);
  }
#pragma unroll
  for (int k = 0; k < H; k++) beta[k] = beta[k] * (ms.mult[k] / ms.prob * w_mat);
  ray.o = sf.p;
  ray.d = ms.dir;
  ray.tmax = kFltMax;
  ray.skip = LAM_NONE;
  return true;
}

template <int H, class Stack>
__device__ void path_loop(const DevScene &sc, uint32_t max_scat, float sample_p, uint64_t key, const Lam<H> &lam,
                          uint32_t s0, Ray ray, float *beta, float *L, const Stack &stk, bool &ovf) {
  for (uint32_t s = s0; path_vertex<H>(sc, max_scat, sample_p, key, lam, s, ray, beta, L, stk, ovf); s++) {
  }
}

// ---- film -------------------------------------------------------------------------
// addHeroλ (cg.hpp:306-320) projected onto the CMFs: the splat weights of
// wavelength l on the coefficient grid, with the reference's ordering of
// (1-t) and t, folded into XYZ.
// i = the number of grid points p_c = inf + c Δ (float arithmetic) with
// p_c <= l, capped at n.  p_c does not decrease with c, so i is the one
// count with "i = 0 or p_{i-1} <= l" (not A) and "i = n or p_i > l" (not B),
// whatever the start of the search.  The start is a guess from
// (l - inf) / Δ (inv_d: any approximation of 1/Δ); one step fixes a guess
// that is off by one, and the loops, which stop only where neither A nor B
// holds, run for the rest.  A NaN l counts 0 grid points.
__device__ __forceinline__ uint32_t splat_index(const DevScene &sc, float l, float inv_d) {
  const int n = int(sc.grid_n);
  const float inf = sc.grid_inf, D = sc.grid_delta;
  const float g = (l - inf) * inv_d;
  int ci = g >= 0.f ? int(fminf(g, float(n))) + 1 : 0;
  if (ci > n) ci = n;
  auto A = [&](int c) { return c > 0 && !(inf + float(c - 1) * D <= l); };
  auto B = [&](int c) { return c < n && inf + float(c) * D <= l; };
  if (A(ci))
    ci--;
  else if (B(ci))
    ci++;
  if (A(ci) || B(ci)) {
    while (A(ci)) ci--;
    while (B(ci)) ci++;
  }
  return uint32_t(ci);
}

__device__ __forceinline__ void splat_xyz(const DevScene &sc, float l, float s, float &X, float &Y, float &Z,
                                          float inv_d) {
  const uint32_t n = sc.grid_n;
  const float inf = sc.grid_inf, D = sc.grid_delta;
  const uint32_t i = splat_index(sc, l, inv_d);
  const float l1 = inf + float(i) * D;
  float t = (l1 - l) / D;
  if (i == 0) t = 1.f;
  if (i == n) t = 0.f;
  if (i > 0) {
    const float a = (1.f - t) * s;
    X += a * sc.cmf[i - 1];
    Y += a * sc.cmf[n + i - 1];
    Z += a * sc.cmf[2 * n + i - 1];
  }
  if (i < n) {
    const float b = t * s;
    X += b * sc.cmf[i];
    Y += b * sc.cmf[n + i];
    Z += b * sc.cmf[2 * n + i];
  }
}

// Kahan-compensated accumulation of one sample per pixel (6 floats/px:
// sums s.xyz, compensations c.xyz, in 8 floats)
__device__ __forceinline__ void kahan_add(float4 &s, float4 &c, float X, float Y, float Z) {
  float v[3] = {X, Y, Z};
  float *sp[3] = {&s.x, &s.y, &s.z};
  float *cp[3] = {&c.x, &c.y, &c.z};
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const float y = v[q] - *cp[q];
    const float t = *sp[q] + y;
    *cp[q] = (t - *sp[q]) - y;
    *sp[q] = t;
  }
}
__device__ __forceinline__ void film_add(float *__restrict__ film, uint32_t pixel, float X, float Y, float Z) {
  float4 *f = reinterpret_cast<float4 *>(film) + size_t(pixel) * 2;
  float4 s = f[0], c = f[1];
  kahan_add(s, c, X, Y, Z);
  f[0] = s;
  f[1] = c;
}

}  // namespace lam

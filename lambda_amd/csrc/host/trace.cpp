// Host traversal of the product's BVH8 records: the closest-hit and any-hit
// queries of the extend / shadow kernels (lam_device.h trav_begin /
// trav_step) restated in C++ over the same 128-byte nodes and the same
// leaf-order triangle slots that lam_scene_upload builds, with the same
// float operations in the same order (-ffp-contract=off), counting the
// interior nodes fetched and the triangles tested per ray.  It is the
// independent count behind the extend roofline's algorithmic bytes (no
// counterpart in the reference; its rays and hits mean what
// intersect::scene's do, lib/numli/cg.hpp:1469-1489).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "bvh8.hpp"
#include "lambda_host.h"

namespace {

struct V3 {
  float x, y, z;
};
inline V3 mk(float a, float b, float c) { return V3{a, b, c}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

constexpr float kBias = 0x1.6a09e6p-12f;
constexpr float kEps = 1.1920928955078125e-07f;
constexpr float kRobustQ = 1.000000476837158203125f;  // 1 + 2^-21
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct Hit {
  float t;
  uint32_t slot, obj, prim;
};
inline bool better(float t, uint32_t obj, uint32_t prim, const Hit &b) {
  return t < b.t || (t == b.t && (obj < b.obj || (obj == b.obj && prim < b.prim)));
}

// intersect::triangle, Plücker form, as tri_t in lam_device.h
float tri_t(V3 v0, V3 v1, V3 v2, V3 o, V3 u, V3 mr) {
  const V3 e0 = sub(v1, v0);
  const V3 e2 = sub(v0, v2);
  const V3 n = cross(e2, e0);
  const float D = dot(u, n);
  if (std::fabs(D) < kEps) return -1.f;
  const V3 m0 = cross(v0, e0);
  const float s0 = dot(u, m0) + dot(e0, mr);
  if (s0 * D < 0.f) return -1.f;
  const V3 e1 = sub(v2, v1);
  const V3 m1 = cross(v1, e1);
  const float s1 = dot(u, m1) + dot(e1, mr);
  if (s1 * D < 0.f) return -1.f;
  const V3 m2 = cross(v2, e2);
  const float s2 = dot(u, m2) + dot(e2, mr);
  if (s2 * D < 0.f) return -1.f;
  if (std::fabs(s0 + s1 + s2) < kEps) return -1.f;
  const float inv_D = 1.f / D;
  return dot(sub(v0, o), n) * inv_D;
}

V3 xmul(const float *M, V3 v, float w) {
  float r[3];
  for (int i = 0; i < 3; i++) {
    float s = 0.f;
    s += M[i * 4 + 0] * v.x;
    s += M[i * 4 + 1] * v.y;
    s += M[i * 4 + 2] * v.z;
    s += M[i * 4 + 3] * w;
    r[i] = s;
  }
  return V3{r[0], r[1], r[2]};
}

float sphere_t(V3 lo, V3 lu) {
  const float a = dot(lu, lu);
  const float b = 2 * dot(lo, lu);
  const float c = dot(lo, lo) - 1.f;
  const float disc = b * b - 4 * a * c;
  if (disc < .1f * kBias) return -1.f;
  const float inv_2a = .5f / a;
  const float tp = (-b + std::sqrt(disc)) * inv_2a;
  const float tm = (-b - std::sqrt(disc)) * inv_2a;
  float t = tm;
  if (tm < kBias) t = tp;
  if (t < kBias) return -1.f;
  return t;
}
float plane_t(V3 lo, V3 lu) {
  if (lu.z == 0.f) return -1.f;
  const float t = -lo.z / lu.z;
  const V3 x = add(lo, mul(lu, t));
  if (x.x < -1.f || x.x > 1.f || x.y < -1.f || x.y > 1.f || t < kBias) return -1.f;
  return t;
}

struct Scene {
  const lam_scene_desc *d;
  lamh::Bvh8 bvh;
  std::vector<float> tri;  // 12 floats per slot: v0 v1 v2, then obj prim input (as bits)
  std::vector<uint32_t> analytic;
};

inline void cswap(float &ta, uint32_t &ca, float &tb, uint32_t &cb) {
  const bool s = tb < ta;
  const float t0 = s ? tb : ta, t1 = s ? ta : tb;
  const uint32_t c0 = s ? cb : ca, c1 = s ? ca : cb;
  ta = t0; tb = t1; ca = c0; cb = c1;
}

// one ray; returns false on stack overflow (more than 128 entries)
bool trace(const Scene &S, const lam_ray &ray, bool any, Hit &best, uint32_t &nodes, uint32_t &tris) {
  const lam_scene_desc *d = S.d;
  const V3 o = mk(ray.o[0], ray.o[1], ray.o[2]), u = mk(ray.d[0], ray.d[1], ray.d[2]);
  best = Hit{ray.tmax, kNone, kNone, kNone};
  nodes = tris = 0;
  for (uint32_t oi : S.analytic) {
    if (oi == ray.skip) continue;
    const lam_object &ob = d->objects[oi];
    const V3 lo = xmul(ob.T.Minv, o, 1.f), lu = xmul(ob.T.Minv, u, 0.f);
    const float th = ob.type == LAM_OBJ_SPHERE ? sphere_t(lo, lu) : plane_t(lo, lu);
    if (th >= kBias && better(th, oi, 0u, best)) {
      best = Hit{th, kNone, oi, 0u};
      if (any) return true;
    }
  }
  if (S.bvh.nodes.empty()) return true;
  // inverse direction clamped to +-2^64, as trav_begin does
  const float ix = std::fmin(std::fmax(u.x != 0.f ? 1.f / u.x : FLT_MAX, -0x1p64f), 0x1p64f);
  const float iy = std::fmin(std::fmax(u.y != 0.f ? 1.f / u.y : FLT_MAX, -0x1p64f), 0x1p64f);
  const float iz = std::fmin(std::fmax(u.z != 0.f ? 1.f / u.z : FLT_MAX, -0x1p64f), 0x1p64f);
  const V3 mr = cross(o, u);
  uint32_t stack[128];
  int sp = 0;
  uint32_t cur = 0;
  const float kInf = std::numeric_limits<float>::infinity();
  for (;;) {
    if (!(cur & lamh::kLeafBit)) {
      const lamh::Node8 &nd = S.bvh.nodes[cur];
      nodes++;
      float tk[8];
      uint32_t ck[8];
      // the kernel's slab test in the node's quantized frame (trav_step in
      // lam_device.h): t = q (s ix) + (origin - o) ix per plane, one fused
      // multiply-add from b - E (near) or b + E (far), the far distance
      // widened by 1 + 2^-21
      const float ixv[3] = {ix, iy, iz}, rov[3] = {o.x, o.y, o.z};
      const uint8_t ev[3] = {nd.ex, nd.ey, nd.ez};
      float av[3], bn[3], bf[3];
      bool pos[3];
      for (int j = 0; j < 3; j++) {
        // scale 2^(e-127) as the kernel reads it: the exponent field alone
        // (e = 0 reads as 0)
        const uint32_t sbits = uint32_t(ev[j]) << 23;
        float sj;
        std::memcpy(&sj, &sbits, 4);
        const float b = (nd.origin[j] - rov[j]) * ixv[j];
        const float e = (std::fabs(b) + (std::fabs(nd.origin[j]) + 255.f * sj) * std::fabs(ixv[j])) * 0x1p-22f;
        av[j] = sj * ixv[j];
        bn[j] = b - e;
        bf[j] = b + e;
        pos[j] = ixv[j] > 0.f;
      }
      for (int k = 0; k < 8; k++) {
        float tnj[3], tfj[3];
        for (int j = 0; j < 3; j++) {
          const float qn = float(pos[j] ? nd.qlo[j][k] : nd.qhi[j][k]);
          const float qf = float(pos[j] ? nd.qhi[j][k] : nd.qlo[j][k]);
          tnj[j] = std::fma(qn, av[j], bn[j]);
          tfj[j] = std::fma(qf, av[j], bf[j]);
        }
        const float tn = std::fmax(std::fmax(tnj[0], tnj[1]), std::fmax(tnj[2], 0.f));
        const float tf = std::fmin(std::fmin(tfj[0], tfj[1]), std::fmin(tfj[2], best.t));
        const bool hit = k < nd.nchild && tn <= tf * kRobustQ;
        tk[k] = hit ? std::fmin(tn, FLT_MAX) : kInf;
        ck[k] = nd.child[k];
      }
      if (any) {
        int nh = 0;
        for (int k = 0; k < 8; k++) nh += tk[k] != kInf ? 1 : 0;
        if (nh > 0) {
          const int top = sp + nh - 1;
          if (top > 128) return false;
          uint32_t first = 0;
          int j = 0;
          for (int k = 0; k < 8; k++)
            if (tk[k] != kInf) {
              if (j == 0) first = ck[k];
              else stack[sp + j - 1] = ck[k];
              j++;
            }
          sp = top;
          cur = first;
          continue;
        }
      } else {
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
          int nh = 1;
          for (int k = 1; k < 8; k++) nh += tk[k] != kInf ? 1 : 0;
          const int top = sp + nh - 1;
          if (top > 128) return false;
          for (int k = 1; k < nh; k++) stack[top - k] = ck[k];
          sp = top;
          cur = ck[0];
          continue;
        }
      }
    } else {
      const uint32_t first = cur & lamh::kLeafFirstMask;
      const uint32_t count = ((cur >> lamh::kLeafCountShift) & 7u) + 1u;
      for (uint32_t i = 0; i < count; i++) {
        const uint32_t s = first + i;
        const float *r = &S.tri[size_t(s) * 12];
        tris++;
        uint32_t obj, prim;
        std::memcpy(&obj, r + 9, 4);
        std::memcpy(&prim, r + 10, 4);
        if (obj == ray.skip) continue;
        const float th = tri_t(mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]), mk(r[6], r[7], r[8]), o, u, mr);
        if (th >= kBias && better(th, obj, prim, best)) {
          best = Hit{th, s, obj, prim};
          if (any) return true;
        }
      }
    }
    if (sp == 0) return true;
    sp--;
    cur = stack[sp];
  }
}

}  // namespace

extern "C" int lamh_trace_counts(const lam_scene_desc *desc, const lam_ray *rays, size_t n, int any,
                                 lam_hit *hits, uint32_t *occluded, uint32_t *counts) {
  if (!desc || (n && !rays)) return LAM_E_INVALID;
  if (desc->n_tris > lamh::kLeafFirstMask) return LAM_E_LIMIT;
  Scene S;
  S.d = desc;
  if (desc->n_tris && !lamh::build_bvh8(desc->tri_v, desc->n_tris, S.bvh)) return LAM_E_LIMIT;
  S.tri.resize(size_t(desc->n_tris) * 12);
  for (uint32_t s = 0; s < desc->n_tris; s++) {
    const uint32_t o = S.bvh.order.empty() ? s : S.bvh.order[s];
    std::memcpy(&S.tri[size_t(s) * 12], desc->tri_v + size_t(o) * 9, 36);
    const uint32_t ids[3] = {desc->tri_obj[o], desc->tri_prim[o], o};
    std::memcpy(&S.tri[size_t(s) * 12 + 9], ids, 12);
  }
  for (uint32_t i = 0; i < desc->n_objects; i++)
    if (desc->objects[i].type != LAM_OBJ_TRIMESH) S.analytic.push_back(i);
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<int> ok(nth, 1);
  std::vector<std::thread> th;
  for (unsigned w = 0; w < nth; w++)
    th.emplace_back([&, w] {
      for (size_t i = w; i < n; i += nth) {
        Hit h;
        uint32_t nn, nt;
        if (!trace(S, rays[i], any != 0, h, nn, nt)) ok[w] = 0;
        if (hits) {
          const bool miss = h.obj == kNone;
          hits[i].t = miss ? FLT_MAX : h.t;
          hits[i].tri = h.slot == kNone ? kNone : S.bvh.order[h.slot];
          hits[i].obj = h.obj;
          hits[i].prim = miss ? kNone : h.prim;
        }
        if (occluded) occluded[i] = h.obj != kNone ? 1u : 0u;
        if (counts) {
          counts[2 * i] = nn;
          counts[2 * i + 1] = nt;
        }
      }
    });
  for (auto &t : th) t.join();
  for (int v : ok)
    if (!v) return LAM_E_BVH;
  return LAM_OK;
}

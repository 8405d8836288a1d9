/* oracle.c -- TEST INFRASTRUCTURE ONLY.  CPU restatement of the reference's
 * render loop (webstek/lambda), used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker.  The product never links or
 * calls this code.
 *
 * Built twice (oracle/Makefile):
 *   liboracle32.so  FP32, -ffp-contract=off: bit-level parity target of the
 *                   GPU extend/shadow kernels and per-path radiance;
 *   liboracle64.so  FP64 "truth mode": same estimator in double with libm,
 *                   the reference for image-level error.
 *
 * What it restates (reference file:line):
 *   render loop + film              src/renderer.cpp:14-61, cg.hpp:306-341
 *   tracePath (recursive, MIS)      src/renderer.cpp:64-146
 *   intersect::sphere/plane         cg.hpp:1196-1265
 *   intersect::triangle (Plücker)   cg.hpp:1341-1379
 *   trimesh BVH traversal           cg.hpp:1382-1460 (closest-first stack)
 *   bvh<>::build (8 equal bins)     cg.hpp:632-749, leaf <= 4 prims
 *   sampling + pdfs                 cg.hpp:1508-1985
 *   materials                       cg.hpp:921-1102
 *   orthonormalBasisOf              cg.hpp:384-405
 * Deliberate differences, shared with the GPU path: triangles are the
 * flattened world-space instances of lam_scene_desc; ties in t go to the
 * lowest (object, prim) id (the reference keeps the last equal-t hit); the
 * BVH box test is widened (Ize 2013) so BVH traversal equals brute force;
 * random numbers come from the counter RNG of omath.h.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "lambda_amd.h"
#include "omath.h"

#ifdef ORC_F64
typedef double real;
#define RSQRT sqrt
#define RSIN(x) sin(x)
#define RCOS(x) cos(x)
#define RPOW(x, y) pow(x, y)
#define RATAN2(y, x) atan2(y, x)
#define RFABS fabs
#else
typedef float real;
#define RSQRT sqrtf
#define RSIN(x) om_sinf(x)
#define RCOS(x) om_cosf(x)
#define RPOW(x, y) om_powf(x, y)
#define RATAN2(y, x) om_atan2f(y, x)
#define RFABS fabsf
#endif

#define BIAS ((real)0x1.6a09e6p-12f) /* constexprSqrt(ε<float>), cg.hpp:1189 */
#define EPS ((real)FLT_EPSILON)
#define UBF ((real)FLT_MAX)
#define PI ((real)3.14159265358979323846)
#define INV_PI ((real)0.318309886183790671538)
#define HMAX 16
#define NONE 0xFFFFFFFFu

/* ---- vectors (bra.hpp operator order) ------------------------------------ */
typedef struct { real x, y, z; } v3;
static inline v3 V(real a, real b, real c) { v3 r; r.x = a; r.y = b; r.z = c; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return V(a.x * (real)-1, a.y * (real)-1, a.z * (real)-1); }
static inline v3 vmul(v3 a, real s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, real s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vcross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline real vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline real vlen2(v3 a) { real s = 0; s += a.x * a.x; s += a.y * a.y; s += a.z * a.z; return s; }
static inline v3 vnorm(v3 a) { real l = RSQRT(vlen2(a)); return V(a.x / l, a.y / l, a.z / l); }
static inline real rmax(real a, real b) { return a >= b ? a : b; } /* numli.hpp:64 */
static inline v3 fv(const float *p) { return V((real)p[0], (real)p[1], (real)p[2]); }

/* M (3x4 rows) * (v, w), sequential k-sum (bra.hpp:325-336) */
static inline v3 xmul(const float *M, v3 v, real w) {
  real r[3];
  int i;
  for (i = 0; i < 3; i++) {
    real s = 0;
    s += (real)M[i * 4 + 0] * v.x;
    s += (real)M[i * 4 + 1] * v.y;
    s += (real)M[i * 4 + 2] * v.z;
    s += (real)M[i * 4 + 3] * w;
    r[i] = s;
  }
  return V(r[0], r[1], r[2]);
}
/* subMatT<3,3>(M_inv) * n (cg.hpp:442-443) */
static inline v3 ntw(const float *Mi, v3 n) {
  real r[3];
  int i;
  for (i = 0; i < 3; i++) {
    real s = 0;
    s += (real)Mi[0 * 4 + i] * n.x;
    s += (real)Mi[1 * 4 + i] * n.y;
    s += (real)Mi[2 * 4 + i] * n.z;
    r[i] = s;
  }
  return V(r[0], r[1], r[2]);
}
static inline v3 colv(const float *M, int j) { return V((real)M[j], (real)M[4 + j], (real)M[8 + j]); }

/* ---- the reference's BVH (cg.hpp:568-749) over the flat triangles -------- */
typedef struct {
  int leaf;
  int count;       /* children (interior) or prims (leaf) */
  int first;       /* first child node (interior) or offset into prim_idx */
  float lo[8][3], hi[8][3];
} onode;

typedef struct {
  const lam_scene_desc *d;
  onode *nodes;
  int n_nodes, cap_nodes;
  uint32_t *prim_idx;
  int n_prim_idx;
  float root_lo[3], root_hi[3];
  int depth;
  uint32_t *analytic;
  int n_analytic;
  real delta; /* coefficient grid Δ */
} orc_scene;

typedef struct { float lo[3], hi[3]; } obox;

static void tri_box(const lam_scene_desc *d, uint32_t t, obox *b) {
  const float *v = d->tri_v + (size_t)t * 9;
  int a, k;
  for (a = 0; a < 3; a++) { b->lo[a] = 1e30f; b->hi[a] = -1e30f; }
  for (k = 0; k < 3; k++)
    for (a = 0; a < 3; a++) {
      float x = v[3 * k + a];
      b->lo[a] = b->lo[a] <= x ? b->lo[a] : x; /* nl::min */
      b->hi[a] = b->hi[a] >= x ? b->hi[a] : x; /* nl::max */
    }
}

/* build node: recursive split into 8 equal-width bins of the centroid on
 * the longest axis (splitbuildnode, cg.hpp:632-683) */
typedef struct bnode {
  obox bb;
  uint32_t *idx;
  int n;
  struct bnode *child[8];
  int leaf;
} bnode;

static void grow(obox *b, const obox *c) {
  int a;
  for (a = 0; a < 3; a++) {
    b->lo[a] = b->lo[a] <= c->lo[a] ? b->lo[a] : c->lo[a];
    b->hi[a] = b->hi[a] >= c->hi[a] ? b->hi[a] : c->hi[a];
  }
}

static void split(const lam_scene_desc *d, bnode *node, int *n_nodes, int *n_leaves) {
  float l[3];
  int i = 0, a, b, k;
  uint32_t *bins[8];
  int bc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  (*n_nodes)++;
  if (node->n <= 4) { node->leaf = 1; (*n_leaves)++; return; }
  for (a = 0; a < 3; a++) l[a] = node->bb.hi[a] - node->bb.lo[a];
  if (l[1] > l[0]) i = 1;
  if (l[2] > l[1] && l[2] > l[0]) i = 2;
  for (b = 0; b < 8; b++) bins[b] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)node->n);
  for (k = 0; k < node->n; k++) {
    const float *v = d->tri_v + (size_t)node->idx[k] * 9;
    float c = (v[i] + v[3 + i] + v[6 + i]) / 3.f;
    int bin;
    if (l[i] > 0.f) {
      float f = 8 * (c - node->bb.lo[i]) / l[i];
      bin = (int)f;
      if (bin > 7) bin = 7;
      if (bin < 0) bin = 0;
    } else {
      bin = 0;
    }
    bins[bin][bc[bin]++] = node->idx[k];
  }
  for (b = 0; b < 8; b++) {
    if (bc[b] != node->n) continue; /* all in one bin: equal chunks (cg.hpp:658-672) */
    {
      uint32_t *tmp = bins[b];
      int chunk = (node->n + 7) / 8, q;
      bins[b] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)node->n);
      for (q = 0; q < 8; q++) bc[q] = 0;
      for (q = 0; q < 8; q++) {
        int s = q * chunk, e = s + chunk < node->n ? s + chunk : node->n, m;
        for (m = s; m < e; m++) bins[q][bc[q]++] = tmp[m];
      }
      free(tmp);
    }
    break;
  }
  for (b = 0; b < 8; b++) {
    bnode *c;
    if (bc[b] == 0) { free(bins[b]); continue; }
    c = (bnode *)calloc(1, sizeof(bnode));
    c->idx = bins[b];
    c->n = bc[b];
    for (a = 0; a < 3; a++) { c->bb.lo[a] = 1e30f; c->bb.hi[a] = -1e30f; }
    for (k = 0; k < c->n; k++) {
      obox tb;
      tri_box(d, c->idx[k], &tb);
      grow(&c->bb, &tb);
    }
    node->child[b] = c;
    split(d, c, n_nodes, n_leaves);
  }
}

static void free_b(bnode *n) {
  int b;
  if (!n) return;
  for (b = 0; b < 8; b++) free_b(n->child[b]);
  free(n->idx);
  free(n);
}

/* convertbuildnodes: breadth-first flatten, children contiguous (cg.hpp:687-728) */
static int flatten_b(orc_scene *s, bnode *root, int n_nodes) {
  bnode **q = (bnode **)malloc(sizeof(bnode *) * (size_t)n_nodes);
  int *qi = (int *)malloc(sizeof(int) * (size_t)n_nodes), *qd = (int *)malloc(sizeof(int) * (size_t)n_nodes);
  int head = 0, tail = 0, next_free = 1, k, b, a;
  s->nodes = (onode *)calloc((size_t)n_nodes, sizeof(onode));
  s->n_nodes = n_nodes;
  s->prim_idx = (uint32_t *)malloc(sizeof(uint32_t) * (s->d->n_tris ? s->d->n_tris : 1));
  s->n_prim_idx = 0;
  q[tail] = root; qi[tail] = 0; qd[tail] = 1; tail++;
  while (head < tail) {
    bnode *n = q[head];
    int idx = qi[head], dep = qd[head];
    onode *o = &s->nodes[idx];
    head++;
    if (dep > s->depth) s->depth = dep;
    if (n->leaf) {
      o->leaf = 1;
      o->count = n->n;
      o->first = s->n_prim_idx;
      for (k = 0; k < n->n; k++) s->prim_idx[s->n_prim_idx++] = n->idx[k];
    } else {
      int nc = 0;
      o->leaf = 0;
      o->first = next_free;
      for (b = 0; b < 8; b++) {
        if (!n->child[b]) continue;
        for (a = 0; a < 3; a++) { o->lo[nc][a] = n->child[b]->bb.lo[a]; o->hi[nc][a] = n->child[b]->bb.hi[a]; }
        q[tail] = n->child[b]; qi[tail] = next_free + nc; qd[tail] = dep + 1; tail++;
        nc++;
      }
      o->count = nc;
      next_free += nc;
    }
  }
  free(q); free(qi); free(qd);
  return 0;
}

orc_scene *orc_scene_new(const lam_scene_desc *d) {
  orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
  uint32_t i;
  int a;
  s->d = d;
  s->delta = (real)((d->grid.sup - d->grid.inf) / (float)(d->grid.n - 1));
  s->analytic = (uint32_t *)malloc(sizeof(uint32_t) * (d->n_objects ? d->n_objects : 1));
  for (i = 0; i < d->n_objects; i++)
    if (d->objects[i].type != LAM_OBJ_TRIMESH) s->analytic[s->n_analytic++] = i;
  if (d->n_tris) {
    bnode *root = (bnode *)calloc(1, sizeof(bnode));
    int n_nodes = 0, n_leaves = 0;
    root->idx = (uint32_t *)malloc(sizeof(uint32_t) * d->n_tris);
    root->n = (int)d->n_tris;
    for (a = 0; a < 3; a++) { root->bb.lo[a] = 1e30f; root->bb.hi[a] = -1e30f; }
    for (i = 0; i < d->n_tris; i++) {
      obox tb;
      root->idx[i] = i;
      tri_box(d, i, &tb);
      grow(&root->bb, &tb);
    }
    split(d, root, &n_nodes, &n_leaves);
    for (a = 0; a < 3; a++) { s->root_lo[a] = root->bb.lo[a]; s->root_hi[a] = root->bb.hi[a]; }
    flatten_b(s, root, n_nodes);
    free_b(root);
  }
  return s;
}

void orc_scene_free(orc_scene *s) {
  if (!s) return;
  free(s->nodes);
  free(s->prim_idx);
  free(s->analytic);
  free(s);
}

/* nodes, interior nodes, leaves, depth of the restated reference BVH */
void orc_bvh_stats(const orc_scene *s, int *out4) {
  int i, interior = 0;
  for (i = 0; i < s->n_nodes; i++) interior += !s->nodes[i].leaf;
  out4[0] = s->n_nodes;
  out4[1] = interior;
  out4[2] = s->n_nodes - interior;
  out4[3] = s->depth;
}

/* ---- primitive tests --------------------------------------------------------- */
typedef struct { v3 o, d; real tmax; uint32_t skip; } oray;
typedef struct { real t; uint32_t tri, obj, prim; } ohit;

static inline int better(real t, uint32_t obj, uint32_t prim, const ohit *b) {
  return t < b->t || (t == b->t && (obj < b->obj || (obj == b->obj && prim < b->prim)));
}

/* intersect::triangle (cg.hpp:1341-1379): t or -1 */
static inline real tri_t(v3 v0, v3 v1, v3 v2, v3 p, v3 u) {
  v3 e0 = vsub(v1, v0), e2 = vsub(v0, v2), n = vcross(e2, e0), mr, m0, e1, m1, m2;
  real D = vdot(u, n), s0, s1, s2, inv_D;
  if (RFABS(D) < EPS) return -1;
  mr = vcross(p, u);
  m0 = vcross(v0, e0);
  s0 = vdot(u, m0) + vdot(e0, mr);
  if (s0 * D < 0) return -1;
  e1 = vsub(v2, v1);
  m1 = vcross(v1, e1);
  s1 = vdot(u, m1) + vdot(e1, mr);
  if (s1 * D < 0) return -1;
  m2 = vcross(v2, e2);
  s2 = vdot(u, m2) + vdot(e2, mr);
  if (s2 * D < 0) return -1;
  if (RFABS(s0 + s1 + s2) < EPS) return -1;
  inv_D = (real)1 / D;
  return vdot(vsub(v0, p), n) * inv_D;
}
static inline v3 tri_bary(v3 v0, v3 v1, v3 v2, v3 p, v3 u) {
  v3 mr = vcross(p, u), e0 = vsub(v1, v0), e2 = vsub(v0, v2), n = vcross(e2, e0), e1 = vsub(v2, v1);
  real D = vdot(u, n);
  real s0 = vdot(u, vcross(v0, e0)) + vdot(e0, mr);
  real s1 = vdot(u, vcross(v1, e1)) + vdot(e1, mr);
  real s2 = vdot(u, vcross(v2, e2)) + vdot(e2, mr);
  real inv_D = (real)1 / D;
  return V(s1 * inv_D, s2 * inv_D, s0 * inv_D);
}
/* intersect::sphere (cg.hpp:1196-1214) */
static inline real sphere_t(v3 lo, v3 lu) {
  real a = vdot(lu, lu), b = 2 * vdot(lo, lu), c = vdot(lo, lo) - (real)1;
  real disc = b * b - 4 * a * c, inv_2a, tp, tm, t;
  if (disc < (real)0.1f * BIAS) return -1;
  inv_2a = (real)0.5 / a;
  tp = (-b + RSQRT(disc)) * inv_2a;
  tm = (-b - RSQRT(disc)) * inv_2a;
  t = tm;
  if (tm < BIAS) t = tp;
  if (t < BIAS) return -1;
  return t;
}
/* intersect::plane (cg.hpp:1245-1252) */
static inline real plane_t(v3 lo, v3 lu) {
  real t;
  v3 x;
  if (lu.z == 0) return -1;
  t = -lo.z / lu.z;
  x = vadd(lo, vmul(lu, t));
  if (x.x < -1 || x.x > 1 || x.y < -1 || x.y > 1 || t < BIAS) return -1;
  return t;
}
static inline real analytic_t(const lam_object *ob, v3 o, v3 u) {
  v3 lo = xmul(ob->T.Minv, o, 1), lu = xmul(ob->T.Minv, u, 0);
  return ob->type == LAM_OBJ_SPHERE ? sphere_t(lo, lu) : plane_t(lo, lu);
}

#define ROBUST ((real)1.0000003576278687)

/* closest (any=0) or any (any=1) hit, t in [BIAS, tmax], skip excluded.
 * mode 0: brute force over every triangle; mode 1: the BVH. */
static int trace(const orc_scene *s, const oray *r, int any, int mode, ohit *best) {
  const lam_scene_desc *d = s->d;
  uint32_t a;
  best->t = r->tmax;
  best->tri = NONE; best->obj = NONE; best->prim = NONE;
  for (a = 0; a < (uint32_t)s->n_analytic; a++) {
    uint32_t oi = s->analytic[a];
    real t;
    if (oi == r->skip) continue;
    t = analytic_t(&d->objects[oi], r->o, r->d);
    if (t >= BIAS && better(t, oi, 0, best)) {
      best->t = t; best->tri = NONE; best->obj = oi; best->prim = 0;
      if (any) return 0;
    }
  }
  if (d->n_tris == 0) return 0;
  if (mode == 0) {
    uint32_t t;
    for (t = 0; t < d->n_tris; t++) {
      const float *v = d->tri_v + (size_t)t * 9;
      uint32_t obj = d->tri_obj[t], prim = d->tri_prim[t];
      real th;
      if (obj == r->skip) continue;
      th = tri_t(fv(v), fv(v + 3), fv(v + 6), r->o, r->d);
      if (th >= BIAS && better(th, obj, prim, best)) {
        best->t = th; best->tri = t; best->obj = obj; best->prim = prim;
        if (any) return 0;
      }
    }
    return 0;
  }
  {
    /* trimesh traversal (cg.hpp:1405-1446): stack of (node, t), closest on top */
    enum { STACK = 4096 };
    int sn[STACK];
    real st[STACK];
    int sp = 0;
    real inv[3];
    inv[0] = r->d.x != 0 ? (real)1 / r->d.x : UBF; /* ray ctor, cg.hpp:357-361 */
    inv[1] = r->d.y != 0 ? (real)1 / r->d.y : UBF;
    inv[2] = r->d.z != 0 ? (real)1 / r->d.z : UBF;
    sn[sp] = 0; st[sp] = 0; sp++;
    while (sp > 0) {
      int node;
      real tn0;
      const onode *nd;
      sp--;
      node = sn[sp]; tn0 = st[sp];
      if (tn0 > best->t * ROBUST) continue;
      nd = &s->nodes[node];
      if (!nd->leaf) {
        int hits = 0, k, m, hc[8];
        real ht[8];
        for (k = 0; k < nd->count; k++) {
          real o3[3] = {r->o.x, r->o.y, r->o.z}, tn = 0, tf = best->t;
          int q;
          for (q = 0; q < 3; q++) {
            real t0 = ((real)nd->lo[k][q] - o3[q]) * inv[q];
            real t1 = ((real)nd->hi[k][q] - o3[q]) * inv[q];
            real lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            if (t0 != t0 || t1 != t1) { lo = 0; hi = UBF; }
            tn = tn > lo ? tn : lo;
            tf = tf < hi ? tf : hi;
          }
          if (tn <= tf * ROBUST) { ht[hits] = tn; hc[hits] = nd->first + k; hits++; }
        }
        /* sort far -> near, push so the closest is on top (cg.hpp:1427-1437) */
        for (k = 1; k < hits; k++)
          for (m = k; m > 0 && (ht[m] > ht[m - 1] || (ht[m] == ht[m - 1] && hc[m] > hc[m - 1])); m--) {
            real tt = ht[m]; int cc = hc[m];
            ht[m] = ht[m - 1]; hc[m] = hc[m - 1]; ht[m - 1] = tt; hc[m - 1] = cc;
          }
        for (k = 0; k < hits; k++) {
          if (sp >= STACK) return -1;
          sn[sp] = hc[k]; st[sp] = ht[k]; sp++;
        }
      } else {
        int k;
        for (k = 0; k < nd->count; k++) {
          uint32_t t = s->prim_idx[nd->first + k];
          const float *v = d->tri_v + (size_t)t * 9;
          uint32_t obj = d->tri_obj[t], prim = d->tri_prim[t];
          real th;
          if (obj == r->skip) continue;
          th = tri_t(fv(v), fv(v + 3), fv(v + 6), r->o, r->d);
          if (th >= BIAS && better(th, obj, prim, best)) {
            best->t = th; best->tri = t; best->obj = obj; best->prim = prim;
            if (any) return 0;
          }
        }
      }
    }
  }
  return 0;
}

/* intersection mode of the path tracer's queries: 1 BVH (default), 0 brute
 * force over every triangle (orc_set_path_brute, a test hook) */
static int path_trace_mode = 1;
void orc_set_path_brute(int brute) { path_trace_mode = brute ? 0 : 1; }

static oray ray_from(const lam_ray *r) {
  oray o;
  o.o = fv(r->o);
  o.d = fv(r->d);
  o.tmax = (real)r->tmax;
  o.skip = r->skip;
  return o;
}

int orc_closest(const orc_scene *s, const lam_ray *rays, size_t n, lam_hit *hits, int mode) {
  long i;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (i = 0; i < (long)n; i++) {
    oray r = ray_from(&rays[i]);
    ohit h;
    bad |= trace(s, &r, 0, mode, &h) != 0;
    hits[i].t = h.obj == NONE ? FLT_MAX : (float)h.t;
    hits[i].tri = h.tri;
    hits[i].obj = h.obj;
    hits[i].prim = h.obj == NONE ? NONE : h.prim;
  }
  return bad ? -1 : 0;
}

int orc_occluded(const orc_scene *s, const lam_ray *rays, size_t n, uint32_t *out, int mode) {
  long i;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (i = 0; i < (long)n; i++) {
    oray r = ray_from(&rays[i]);
    ohit h;
    bad |= trace(s, &r, 1, mode, &h) != 0;
    out[i] = h.obj != NONE;
  }
  return bad ? -1 : 0;
}

/* ---- spectra ------------------------------------------------------------------ */
typedef struct {
  int H;
  real l[HMAX];
} olam;

/* coefficientλ::operator()(float) (cg.hpp:289-297) */
static inline real spec1(const orc_scene *s, int idx, real l) {
  const lam_scene_desc *d = s->d;
  const float *v = d->spectra + (size_t)idx * d->grid.n;
  real x, t;
  uint32_t i;
  if (l <= (real)d->grid.inf) return (real)v[0];
  if (l >= (real)d->grid.sup) return (real)v[d->grid.n - 1];
  x = (l - (real)d->grid.inf) / s->delta;
  i = (uint32_t)x;
  if (i > d->grid.n - 2) i = d->grid.n - 2;
  t = x - (real)i;
  return ((real)1 - t) * (real)v[i] + t * (real)v[i + 1];
}
static inline void specH(const orc_scene *s, int idx, const olam *L, real *out) {
  int k;
  for (k = 0; k < L->H; k++) out[k] = spec1(s, idx, L->l[k]);
}

/* ---- frames and surfaces -------------------------------------------------------- */
typedef struct { v3 x, y, z; } oframe;
/* orthonormalBasisOf (cg.hpp:384-405) */
static oframe onb(v3 v) {
  oframe F;
  v3 e2 = vnorm(v);
  real x = e2.x, y = e2.y, z = e2.z;
  if (z >= y) {
    real a = (real)1 / ((real)1 + z), b = -x * y * a;
    F.x = V(1 - x * x * a, b, -x);
    F.y = V(b, 1 - y * y * a, -y);
  } else {
    real a = (real)1 / ((real)1 + y), b = -x * z * a;
    F.x = V(b, -z, 1 - z * z * a);
    F.y = V(1 - x * x * a, -x, b);
  }
  F.z = e2;
  return F;
}
static inline v3 to_basis(const oframe *F, v3 v) { return vadd(vadd(vmul(F->x, v.x), vmul(F->y, v.y)), vmul(F->z, v.z)); }
static inline v3 reflect(v3 x, v3 a) { return vadd(vneg(x), vmul(a, 2 * vdot(x, a))); } /* bra.hpp:397 */

typedef struct {
  v3 p, gn;
  oframe F;
  int front;
  uint32_t mat, obj;
} osurf;
static inline v3 surf_n(const osurf *s) { return s->front ? s->F.z : vneg(s->F.z); }

static void surface(const orc_scene *s, const ohit *h, v3 o, v3 d, osurf *sf) {
  const lam_scene_desc *sd = s->d;
  const lam_object *ob = &sd->objects[h->obj];
  sf->obj = h->obj;
  sf->mat = ob->material;
  if (h->tri != NONE) {
    const float *v = sd->tri_v + (size_t)h->tri * 9, *nn = sd->tri_n + (size_t)h->tri * 9;
    v3 b = tri_bary(fv(v), fv(v + 3), fv(v + 6), o, d);
    v3 ns = vadd(vadd(vmul(fv(nn), b.x), vmul(fv(nn + 3), b.y)), vmul(fv(nn + 6), b.z));
    sf->gn = fv(sd->tri_gn + (size_t)h->tri * 3);
    sf->p = vadd(o, vmul(d, h->t));
    sf->F = onb(ns);
    sf->front = vdot(sf->gn, d) < 0;
  } else {
    v3 lo = xmul(ob->T.Minv, o, 1), lu = xmul(ob->T.Minv, d, 0), lp = vadd(lo, vmul(lu, h->t));
    if (ob->type == LAM_OBJ_SPHERE) { /* cg.hpp:1216-1240 */
      v3 n = vnorm(ntw(ob->T.Minv, lp));
      real ct = n.z, st = RSQRT(1 - ct * ct), phi = RATAN2(n.y, n.x), cp = RCOS(phi), spp = RSIN(phi);
      sf->front = vdot(n, d) < 0;
      sf->F.x = V(-st * spp, st * cp, 0);
      sf->F.y = V(ct * cp, ct * spp, -st);
      sf->F.z = n;
      sf->gn = n;
      sf->p = xmul(ob->T.M, lp, 1);
    } else { /* cg.hpp:1254-1261 */
      sf->p = xmul(ob->T.M, lp, 1);
      sf->F.z = vnorm(ntw(ob->T.Minv, V(0, 0, 1)));
      sf->F.x = vnorm(ntw(ob->T.Minv, V(0, 1, 0)));
      sf->F.y = vcross(sf->F.z, sf->F.x);
      sf->gn = sf->F.z;
      sf->front = vdot(d, sf->F.z) < 0;
    }
  }
}

/* ---- sampling helpers (stoch.hpp:53-68, cg.hpp:1790-1796) ------------------------- */
static inline v3 cos_hemi(real x0, real x1) {
  real cc = 1 - x0, ct = RSQRT(cc), st = RSQRT(1 - cc), phi = (2 * PI) * x1;
  return V(st * RCOS(phi), st * RSIN(phi), ct);
}
static inline v3 unif_hemi(real x0, real x1) {
  real ct = 1 - x0, st = RSQRT(1 - ct * ct), phi = (2 * PI) * x1;
  return V(st * RCOS(phi), st * RSIN(phi), ct);
}
static inline v3 blinn_h(real a, real x0, real x1) {
  real ct = RPOW(1 - x0, (real)1 / (a + 1)), st = RSQRT(1 - ct * ct), phi = (2 * PI) * x1;
  return V(st * RCOS(phi), st * RSIN(phi), ct);
}

/* ---- materials (cg.hpp:921-1052) ---------------------------------------------------- */
static inline real F0(real n) { return (n * n - 2 * n + 1) / (n * n + 2 * n + 1); }
static inline real fresnel1(real f0, real c) { return f0 + (1 - f0) * RPOW(1 - c, 5); }
static void fsplit(const lam_material *m, real F, real *pr, real *pt) {
  real p_r = (real)m->w_r + F * (real)m->w_t, p_t = (1 - F) * (real)m->w_t, tot = p_r + p_t;
  if (tot == 0) { *pr = 1; *pt = 0; return; }
  *pr = p_r / tot;
  *pt = p_t / tot;
}
static inline real blinnh_prob(real a, real ch) { return RPOW(RFABS(ch), a + 1) * (a + 1) * (real)0.5 * INV_PI; }
static v3 half_vec(v3 i, v3 o, v3 n) {
  v3 h = vadd(i, o);
  real l2 = vlen2(h);
  if (l2 < EPS) return vnorm(vsub(n, vmul(o, vdot(n, o))));
  return vdiv(h, RSQRT(l2));
}

/* BxDFcosθ (cg.hpp:1088-1102) */
static void bxdf(const orc_scene *s, const lam_material *m, const olam *L, v3 i, v3 o, v3 n, int front, real *out) {
  int k, H = L->H;
  real alpha = (real)m->alpha;
  if (m->type == LAM_MAT_LAMBERTIAN) { /* cg.hpp:924-925 */
    real a[HMAX], md = rmax(0, vdot(i, n));
    specH(s, m->spec[0], L, a);
    for (k = 0; k < H; k++) out[k] = a[k] * INV_PI * md;
  } else if (m->type == LAM_MAT_BLINN) { /* cg.hpp:967-986 */
    real ior[HMAX], eta[HMAX], F[HMAX], ci = vdot(i, n), co = vdot(o, n), pw5;
    specH(s, m->spec[4], L, ior);
    for (k = 0; k < H; k++) eta[k] = front ? 1 / ior[k] : ior[k];
    pw5 = RPOW(1 - co, 5);
    for (k = 0; k < H; k++) { real f0 = F0(eta[k]); F[k] = f0 + (1 - f0) * pw5; }
    if (ci * co > 0) {
      v3 h = half_vec(i, o, n);
      real pw = RPOW(vdot(h, n), alpha), md = rmax(0, ci), kd[HMAX], ks[HMAX], kt[HMAX];
      specH(s, m->spec[0], L, kd);
      specH(s, m->spec[1], L, ks);
      specH(s, m->spec[2], L, kt);
      for (k = 0; k < H; k++)
        out[k] = kd[k] * INV_PI * md + (ks[k] + F[k] * kt[k]) * (alpha + 2) * (real)0.125 * INV_PI * pw;
    } else {
      v3 ht = vneg(vnorm(vadd(i, vmul(o, eta[0]))));
      real pw = RPOW(RFABS(vdot(ht, n)), alpha), kt[HMAX];
      specH(s, m->spec[2], L, kt);
      for (k = 0; k < H; k++) out[k] = kt[k] * (1 - F[k]) * (alpha + 2) * (real)0.125 * INV_PI * pw;
    }
  } else if (m->type == LAM_MAT_THINFILM) { /* cg.hpp:1031-1051 */
    v3 h;
    real ci, n1[HMAX], n2[HMAX], pw5, c, pw;
    if (!front || vdot(i, n) < 0 || vdot(o, n) < 0) {
      for (k = 0; k < H; k++) out[k] = 0;
      return;
    }
    h = half_vec(i, o, n);
    ci = vdot(i, h);
    specH(s, m->spec[0], L, n1);
    specH(s, m->spec[1], L, n2);
    pw5 = RPOW(1 - ci, 5);
    c = (alpha + 2) * (real)0.125 * INV_PI;
    pw = RPOW(vdot(h, n), alpha);
    for (k = 0; k < H; k++) {
      real n01 = 1 / n1[k];
      real cos_l = RSQRT(1 - n01 * n01) * (1 - ci * ci); /* cg.hpp:1007 */
      real cos_d = RCOS(4 * PI * n1[k] * (real)m->d * cos_l / L->l[k]);
      real a01 = 1 - n1[k], b01 = 1 + n1[k], r001 = a01 * a01 / (b01 * b01);
      real a12 = n1[k] - n2[k], b12 = n1[k] + n2[k], r012 = a12 * a12 / (b12 * b12);
      real R01 = r001 + (1 - r001) * pw5, R12 = r012 + (1 - r012) * pw5;
      real ravg = 2 * RSQRT(R01 * R12);
      real rt = (R01 + R12 + ravg * cos_d) / (R01 * R12 + 1 + ravg * cos_d);
      out[k] = rt * c * pw;
    }
  } else {
    for (k = 0; k < H; k++) out[k] = 0;
  }
}

typedef struct {
  int ok;
  v3 dir;
  real prob;
  real mult[HMAX];
} omsample;

/* materiali (cg.hpp:1953-1968) */
static void sample_material(const orc_scene *s, const lam_material *m, const olam *L, const osurf *sf, v3 o, real xl,
                            real xa, real xb, real p, omsample *ms) {
  int k, H = L->H;
  real alpha = (real)m->alpha;
  ms->ok = 0;
  if (m->type == LAM_MAT_LAMBERTIAN) { /* cg.hpp:1755-1778 */
    v3 dir, i;
    if (xl > p) return;
    dir = cos_hemi(xa, xb);
    if (!sf->front) dir.z *= -1;
    i = to_basis(&sf->F, dir);
    ms->prob = p * RFABS(dir.z) * INV_PI;
    ms->dir = i;
    bxdf(s, m, L, i, o, surf_n(sf), sf->front, ms->mult);
    ms->ok = 1;
  } else if (m->type == LAM_MAT_BLINN) { /* cg.hpp:1799-1863 */
    real pp_d, pp_spec;
    if (xl >= p) return;
    pp_d = p * (real)m->p_d;
    pp_spec = p * (real)m->p_spec;
    if (xl < pp_spec) {
      real ior[HMAX], eta[HMAX], F[HMAX], cos_i, pw5, p_r, p_t, bp, pw;
      v3 w, h, iR, iT0 = V(0, 0, 0);
      specH(s, m->spec[4], L, ior);
      for (k = 0; k < H; k++) eta[k] = sf->front ? 1 / ior[k] : ior[k];
      w = blinn_h(alpha, xa, xb);
      if (!sf->front) w.z *= -1;
      h = to_basis(&sf->F, w);
      iR = reflect(o, h);
      cos_i = vdot(o, h);
      pw5 = RPOW(1 - cos_i, 5);
      for (k = 0; k < H; k++) { /* transmit (cg.hpp:500-513) */
        real ct2 = 1 - eta[k] * eta[k] * (1 - cos_i * cos_i), f0;
        v3 td = V(0, 0, 0);
        if (!(ct2 < 0)) td = vsub(vmul(o, -eta[k]), vmul(h, RSQRT(ct2) - eta[k] * cos_i));
        if (k == 0) iT0 = td;
        f0 = F0(eta[k]);
        F[k] = f0 + (1 - f0) * pw5;
        if (td.x == 0 && td.y == 0 && td.z == 0) F[k] = 1;
      }
      fsplit(m, F[0], &p_r, &p_t);
      bp = blinnh_prob(alpha, w.z);
      pw = RPOW(w.z, alpha);
      if (xl < pp_spec * p_r) {
        real ks[HMAX], kt[HMAX];
        specH(s, m->spec[1], L, ks);
        specH(s, m->spec[2], L, kt);
        ms->prob = pp_spec * p_r * bp * (real)0.25;
        ms->dir = iR;
        for (k = 0; k < H; k++) ms->mult[k] = (ks[k] + F[k] * kt[k]) * (alpha + 2) * (real)0.125 * INV_PI * pw;
      } else {
        real kt[HMAX];
        specH(s, m->spec[2], L, kt);
        ms->prob = pp_spec * p_t * bp * (real)0.25;
        ms->dir = iT0;
        for (k = 0; k < H; k++) ms->mult[k] = kt[k] * (1 - F[k]) * (alpha + 2) * (real)0.125 * INV_PI * pw;
      }
      ms->ok = 1;
    } else if (xl < pp_spec + pp_d) {
      v3 dir = cos_hemi(xa, xb), i;
      real kd[HMAX], md;
      if (!sf->front) dir.z *= -1;
      i = to_basis(&sf->F, dir);
      ms->prob = pp_d * (RFABS(dir.z) * INV_PI);
      ms->dir = i;
      specH(s, m->spec[0], L, kd);
      md = rmax(0, dir.z);
      for (k = 0; k < H; k++) ms->mult[k] = kd[k] * INV_PI * md;
      ms->ok = 1;
    }
  } else if (m->type == LAM_MAT_THINFILM) { /* cg.hpp:1905-1931 */
    v3 n = sf->F.z, w, h, i;
    if (!sf->front || vdot(o, n) < 0) return;
    if (xl >= p) return;
    w = blinn_h(alpha, xa, xb);
    h = to_basis(&sf->F, w);
    i = reflect(o, h);
    ms->prob = p * RPOW(vdot(n, h), alpha + 1) * (alpha + 1) * (real)0.5 * INV_PI * (real)0.25;
    ms->dir = i;
    bxdf(s, m, L, i, o, n, sf->front, ms->mult);
    ms->ok = 1;
  }
}

/* probForMateriali (cg.hpp:1971-1985); the reference passes (i, o) into the
 * (o, i) parameters of probForBlinn / probForThinFilm, kept here */
static real prob_material(const orc_scene *s, const lam_material *m, const osurf *sf, v3 wl, v3 o, real l0, real p) {
  real alpha = (real)m->alpha;
  if (m->type == LAM_MAT_LAMBERTIAN) { /* cg.hpp:1780-1787 */
    v3 n = surf_n(sf);
    real cos_i = vdot(wl, n);
    if (vdot(wl, n) < 0 || vdot(o, n) < 0) return 0;
    return p * cos_i * INV_PI;
  }
  if (m->type == LAM_MAT_BLINN) { /* cg.hpp:1865-1902 */
    v3 oo = wl, ii = o, n = surf_n(sf), h;
    real iorl = spec1(s, m->spec[4], l0), ior = sf->front ? 1 / iorl : iorl;
    real cin = vdot(ii, n), con = vdot(oo, n), cos_o, F, p_r, p_t, c2;
    if (cin * con > 0) h = half_vec(ii, oo, n);
    else h = vneg(vnorm(vadd(ii, vmul(oo, ior))));
    cos_o = RFABS(vdot(oo, h));
    F = fresnel1(F0(ior), cos_o);
    fsplit(m, F, &p_r, &p_t);
    if (cin * con > 0)
      return p * ((real)m->p_d * (RFABS(cin) * INV_PI) + (real)m->p_spec * p_r * blinnh_prob(alpha, vdot(h, n)) * (real)0.25);
    c2 = 1 - ior * ior * (1 - cos_o * cos_o);
    if (c2 < 0) return 0;
    return p * (real)m->p_spec * p_t * blinnh_prob(alpha, vdot(h, n)) * (real)0.25;
  }
  if (m->type == LAM_MAT_THINFILM) { /* cg.hpp:1932-1949 */
    v3 oo = wl, ii = o, n = sf->F.z, h;
    if (!sf->front || vdot(ii, n) < 0 || vdot(oo, n) < 0) return 0;
    h = half_vec(ii, oo, n);
    return p * RPOW(vdot(oo, h), alpha + 1) * (alpha + 1) * (real)0.5 * INV_PI * (real)0.25;
  }
  return 0;
}

/* ---- lights (cg.hpp:1548-1747) ------------------------------------------------------- */
typedef struct {
  v3 dir;
  real prob;
  real mult[HMAX]; /* includes shadowing, as the reference */
} olsample;

/* rays traced by this thread's paths (extend queries of trace_path, shadow
 * queries of the light samples), summed by orc_count_rays */
static __thread uint64_t n_ext_rays, n_shd_rays;

/* ray log of orc_path_rays: every query of one path, in order */
static __thread lam_ray *rec_rays;
static __thread uint32_t *rec_kind;
static __thread size_t rec_n, rec_cap;
static void rec_ray(const oray *r, uint32_t kind) {
  if (rec_rays && rec_n < rec_cap) {
    lam_ray *q = &rec_rays[rec_n];
    q->o[0] = (float)r->o.x; q->o[1] = (float)r->o.y; q->o[2] = (float)r->o.z;
    q->d[0] = (float)r->d.x; q->d[1] = (float)r->d.y; q->d[2] = (float)r->d.z;
    q->tmax = (float)r->tmax;
    q->skip = r->skip;
    rec_kind[rec_n] = kind;
  }
  if (rec_rays) rec_n++;
}

static int shadowed(const orc_scene *s, v3 p, v3 dir, real tmax, uint32_t skip) {
  oray r;
  ohit h;
  r.o = p; r.d = dir; r.tmax = tmax; r.skip = skip;
  rec_ray(&r, 1);
  n_shd_rays++;
  trace(s, &r, 1, path_trace_mode, &h);
  return h.obj != NONE;
}

static void sample_light(const orc_scene *s, const lam_light *lt, const olam *L, const osurf *sf, real xa, real xb,
                         olsample *ls) {
  const lam_scene_desc *d = s->d;
  real rad[HMAX], sh;
  int k, H = L->H;
  specH(s, lt->spec, L, rad);
  if (lt->type == LAM_LIGHT_SPHERE) { /* cg.hpp:1671-1700 */
    const lam_object *ob = &d->objects[lt->object];
    v3 Lv = vsub(colv(ob->T.M, 3), sf->p), w;
    real dist2 = vdot(Lv, Lv), r2 = (real)lt->size * (real)lt->size, sr = 1 - RSQRT(rmax(0, 1 - r2 / dist2));
    real ct, st, phi;
    oframe B;
    ls->prob = (real)0.5 / (PI * sr);
    B = onb(Lv);
    ct = rmax(0, 1 - xa * sr);
    st = RSQRT(1 - ct * ct);
    phi = 2 * PI * xb;
    w = vadd(vadd(vmul(vmul(B.x, st), RCOS(phi)), vmul(vmul(B.y, st), RSIN(phi))), vmul(B.z, ct));
    ls->dir = w;
    sh = shadowed(s, sf->p, w, RSQRT(dist2), lt->object) ? 0 : 1;
    for (k = 0; k < H; k++) ls->mult[k] = rad[k] * sh;
  } else if (lt->type == LAM_LIGHT_QUADDIR) { /* cg.hpp:1625-1655 */
    const lam_object *ob = &d->objects[lt->object];
    real sx = 2 * xa - 1, sy = 2 * xb - 1, dist2, cs;
    v3 x = vadd(vadd(colv(ob->T.M, 3), vmul(colv(ob->T.M, 0), sx)), vmul(colv(ob->T.M, 1), sy));
    v3 r = vsub(x, sf->p), w = vnorm(r);
    ls->dir = w;
    dist2 = vdot(r, r);
    cs = RPOW(vdot(vneg(w), vnorm(colv(ob->T.M, 2))), (real)lt->specular);
    ls->prob = rmax((real)1e-6f, ((real)lt->specular + 1) * (real)0.5 * INV_PI * cs * (real)lt->inv_area);
    sh = shadowed(s, sf->p, w, RSQRT(dist2), lt->object) ? 0 : 1;
    for (k = 0; k < H; k++) ls->mult[k] = rad[k] * cs * sh;
  } else if (lt->type == LAM_LIGHT_POINT) { /* cg.hpp:1580-1600 */
    v3 Lv = vsub(fv(lt->pos), sf->p), w = vnorm(Lv);
    real dist = RSQRT(vlen2(Lv));
    ls->dir = w;
    ls->prob = 1;
    sh = shadowed(s, sf->p, w, dist, NONE) ? 0 : 1;
    for (k = 0; k < H; k++) ls->mult[k] = rad[k] * sh / (dist * dist);
  } else if (lt->type == LAM_LIGHT_DIR) { /* cg.hpp:1604-1621 */
    v3 w = vneg(fv(lt->dir));
    ls->dir = w;
    ls->prob = 1;
    sh = shadowed(s, sf->p, w, UBF, NONE) ? 0 : 1;
    for (k = 0; k < H; k++) ls->mult[k] = rad[k] * sh;
  } else { /* ambient (cg.hpp:1560-1576) */
    v3 dir = unif_hemi(xa, xb);
    if (!sf->front) dir.z *= -1;
    ls->dir = to_basis(&sf->F, dir);
    ls->prob = (real)0.5 * INV_PI;
    for (k = 0; k < H; k++) ls->mult[k] = rad[k];
  }
}

/* probForLight (cg.hpp:1735-1747) */
static real prob_light(const orc_scene *s, const lam_light *lt, v3 p, v3 dir) {
  const lam_scene_desc *d = s->d;
  if (lt->type == LAM_LIGHT_SPHERE) {
    const lam_object *ob = &d->objects[lt->object];
    v3 lo = xmul(ob->T.Minv, p, 1), lu = xmul(ob->T.Minv, dir, 0), lp, n, Lv;
    real t = sphere_t(lo, lu), dist2, sr;
    if (!(t >= BIAS)) return 0;
    lp = vadd(lo, vmul(lu, t));
    n = vnorm(ntw(ob->T.Minv, lp));
    if (!(vdot(n, dir) < 0)) return 0;
    Lv = vsub(colv(ob->T.M, 3), p);
    dist2 = vdot(Lv, Lv);
    sr = 1 - RSQRT(rmax(0, 1 - (real)lt->size * (real)lt->size / dist2));
    return (real)0.5 / (PI * sr);
  }
  if (lt->type == LAM_LIGHT_QUADDIR) {
    const lam_object *ob = &d->objects[lt->object];
    v3 lo = xmul(ob->T.Minv, p, 1), lu = xmul(ob->T.Minv, dir, 0), nz;
    real t = plane_t(lo, lu), ct;
    if (!(t >= BIAS)) return 0;
    nz = vnorm(ntw(ob->T.Minv, V(0, 0, 1)));
    if (!(vdot(dir, nz) < 0)) return 0;
    ct = vdot(vneg(vnorm(dir)), nz);
    return rmax((real)1e-6f, ((real)lt->specular + 1) * (real)0.5 * INV_PI * RPOW(ct, (real)lt->specular) * (real)lt->inv_area);
  }
  if (lt->type == LAM_LIGHT_AMBIENT) return (real)0.5 * INV_PI;
  return 0;
}

/* ---- the path (src/renderer.cpp:64-146) ----------------------------------------------- */
typedef struct {
  uint64_t key;
  uint32_t max_scat;
  real p;
} opath;

/* ---- forward-mode tangents: the checker of the path-replay adjoint ------------------
 * Directional derivatives of the path radiance along K tangent directions of
 * the differentiated parameters (thin-film d, film IOR coefficients,
 * Lambertian albedo coefficients; the layout of lam_diff_params).  Sampling
 * decisions do not depend on these parameters, so the tangent of the
 * recursive estimator is the recursion of the tangents of its factors. */
#define KMAX 16
#define DPI 3.14159265358979323846
typedef struct {
  const int32_t *mat_d, *mat_ior, *mat_albedo; /* per material; -1: not a parameter */
  const double *v;                             /* K directions of n_params entries */
  uint32_t n_params;
  int K;
} otan;

/* spec1's interpolation applied to tangent coefficients */
static double spec1_tan(const orc_scene *s, const double *row, real l) {
  const lam_scene_desc *d = s->d;
  real x, t;
  uint32_t i;
  if (l <= (real)d->grid.inf) return row[0];
  if (l >= (real)d->grid.sup) return row[d->grid.n - 1];
  x = (l - (real)d->grid.inf) / s->delta;
  i = (uint32_t)x;
  if (i > d->grid.n - 2) i = d->grid.n - 2;
  t = x - (real)i;
  return (1 - (double)t) * row[i] + (double)t * row[i + 1];
}

/* thinfilm::BRDFcosθ (cg.hpp:1031-1051) in double with libm as a function of
 * (d, n1) at fixed geometry (ci = i·h, hn = h·n) */
static double tf_val(double alpha, double d, double n1, double n2, double l, double ci, double hn) {
  double n01 = 1 / n1, cos_l = sqrt(1 - n01 * n01) * (1 - ci * ci);
  double cos_d = cos(4 * DPI * n1 * d * cos_l / l);
  double r001 = (1 - n1) * (1 - n1) / ((1 + n1) * (1 + n1));
  double r012 = (n1 - n2) * (n1 - n2) / ((n1 + n2) * (n1 + n2));
  double pw5 = pow(1 - ci, 5);
  double R01 = r001 + (1 - r001) * pw5, R12 = r012 + (1 - r012) * pw5, ravg = 2 * sqrt(R01 * R12);
  return (R01 + R12 + ravg * cos_d) / (R01 * R12 + 1 + ravg * cos_d) * (alpha + 2) * 0.125 / DPI * pow(hn, alpha);
}

/* tangents of BxDFcosθ of material mi along the K directions: exact for the
 * albedo (linear), fp64 central differences for the thin film */
static void bxdf_tan(const orc_scene *s, uint32_t mi, const lam_material *m, const olam *L, v3 i, v3 o, v3 n,
                     int front, const otan *tn, double *out) {
  int j, k, H = L->H;
  for (j = 0; j < tn->K; j++)
    for (k = 0; k < H; k++) out[j * HMAX + k] = 0;
  if (m->type == LAM_MAT_LAMBERTIAN && tn->mat_albedo[mi] >= 0) {
    double md = (double)rmax(0, vdot(i, n)) / DPI;
    for (j = 0; j < tn->K; j++) {
      const double *row = tn->v + (size_t)j * tn->n_params + tn->mat_albedo[mi];
      for (k = 0; k < H; k++) out[j * HMAX + k] = spec1_tan(s, row, L->l[k]) * md;
    }
  } else if (m->type == LAM_MAT_THINFILM && (tn->mat_d[mi] >= 0 || tn->mat_ior[mi] >= 0)) {
    v3 h;
    double ci, hn, d = m->d;
    if (!front || vdot(i, n) < 0 || vdot(o, n) < 0) return;
    h = half_vec(i, o, n);
    ci = (double)vdot(i, h);
    hn = (double)vdot(h, n);
    for (k = 0; k < H; k++) {
      double n1 = (double)spec1(s, m->spec[0], L->l[k]), n2 = (double)spec1(s, m->spec[1], L->l[k]);
      double l = (double)L->l[k];
      for (j = 0; j < tn->K; j++) {
        const double *vj = tn->v + (size_t)j * tn->n_params;
        double dd = tn->mat_d[mi] >= 0 ? vj[tn->mat_d[mi]] : 0;
        double dn = tn->mat_ior[mi] >= 0 ? spec1_tan(s, vj + tn->mat_ior[mi], L->l[k]) : 0;
        double sc = fabs(dd) / (fabs(d) > 1 ? fabs(d) : 1), e;
        if (fabs(dn) / n1 > sc) sc = fabs(dn) / n1;
        if (sc == 0) continue;
        e = 1e-6 / sc;
        out[j * HMAX + k] = (tf_val(m->alpha, d + e * dd, n1 + e * dn, n2, l, ci, hn) -
                             tf_val(m->alpha, d - e * dd, n1 - e * dn, n2, l, ci, hn)) / (2 * e);
      }
    }
  }
}

/* tn, dLi: optional tangents (dLi[j*HMAX+k] for direction j) */
static void trace_path(const orc_scene *s, const opath *pa, const olam *L, v3 ro, v3 rd, uint32_t sct, real *Li,
                       const otan *tn, double *dLi) {
  const lam_scene_desc *d = s->d;
  oray r;
  ohit h;
  osurf sf;
  v3 o, n;
  const lam_material *mat;
  const lam_light *lt;
  olsample ls;
  omsample ms;
  real coef[HMAX], lis[HMAX], nxt[HMAX], pl, p_L_mati, p_mat_Li, p_L_i, w_L, w_mat;
  uint32_t li;
  int j, k, H = L->H, K = tn ? tn->K : 0;
  double dcoef[KMAX * HMAX], dnxt[KMAX * HMAX], dmult[KMAX * HMAX];
  for (j = 0; j < K; j++)
    for (k = 0; k < H; k++) dLi[j * HMAX + k] = 0;
  r.o = ro; r.d = rd; r.tmax = UBF; r.skip = NONE;
  rec_ray(&r, 0);
  n_ext_rays++;
  trace(s, &r, 0, path_trace_mode, &h);
  if (h.obj == NONE) { for (k = 0; k < H; k++) Li[k] = 0; return; }
  surface(s, &h, ro, rd, &sf);
  o = vneg(vnorm(rd));
  n = surf_n(&sf);
  mat = &d->materials[sf.mat];
  if (mat->type == LAM_MAT_EMITTER) { specH(s, mat->spec[0], L, Li); return; }
  if (mat->type == LAM_MAT_DIREMITTER) {
    real f = RPOW(rmax(vdot(o, sf.F.z), 0), (real)mat->specular);
    specH(s, mat->spec[0], L, Li);
    for (k = 0; k < H; k++) Li[k] = Li[k] * f;
    return;
  }
  if (sct > pa->max_scat) { for (k = 0; k < H; k++) Li[k] = 0; return; }
  li = om_u32(pa->key, om_vdim(sct, 0)) % d->n_lights;
  pl = (real)(1.f / (float)d->n_lights);
  lt = &d->lights[li];
  sample_light(s, lt, L, &sf, (real)om_flt(pa->key, om_vdim(sct, 1)), (real)om_flt(pa->key, om_vdim(sct, 2)), &ls);
  bxdf(s, mat, L, ls.dir, o, n, sf.front, coef);
  for (k = 0; k < H; k++) lis[k] = ls.mult[k] * coef[k] / (ls.prob * pl);
  if (K) bxdf_tan(s, sf.mat, mat, L, ls.dir, o, n, sf.front, tn, dcoef);
  sample_material(s, mat, L, &sf, o, (real)om_flt(pa->key, om_vdim(sct, 3)), (real)om_flt(pa->key, om_vdim(sct, 4)),
                  (real)om_flt(pa->key, om_vdim(sct, 5)), pa->p, &ms);
  if (!ms.ok) {
    for (k = 0; k < H; k++) Li[k] = lis[k];
    for (j = 0; j < K; j++)
      for (k = 0; k < H; k++) dLi[j * HMAX + k] = (double)ls.mult[k] * dcoef[j * HMAX + k] / ((double)ls.prob * pl);
    return;
  }
  trace_path(s, pa, L, sf.p, ms.dir, sct + 1, nxt, tn, dnxt);
  if (K) bxdf_tan(s, sf.mat, mat, L, ms.dir, o, n, sf.front, tn, dmult);
  p_L_mati = prob_light(s, lt, sf.p, ms.dir) * pl;
  p_mat_Li = prob_material(s, mat, &sf, ls.dir, o, L->l[0], pa->p);
  p_L_i = ls.prob * pl;
  w_L = p_L_i * p_L_i / (p_L_i * p_L_i + p_mat_Li * p_mat_Li);
  w_mat = ms.prob * ms.prob / (ms.prob * ms.prob + p_L_mati * p_L_mati);
  for (k = 0; k < H; k++) Li[k] = nxt[k] * ms.mult[k] / ms.prob * w_mat + lis[k] * w_L;
  for (j = 0; j < K; j++)
    for (k = 0; k < H; k++) {
      const int q = j * HMAX + k;
      dLi[q] = (dnxt[q] * (double)ms.mult[k] + (double)nxt[k] * dmult[q]) / (double)ms.prob * (double)w_mat +
               (double)ls.mult[k] * dcoef[q] / ((double)ls.prob * pl) * (double)w_L;
    }
}

/* camera ray + hero wavelengths (src/renderer.cpp:35-40, cg.hpp:1517-1542)
 * for film position (u, v) in pixels, lens numbers (lx, ly) and hero ξ;
 * returns prob_λ = 1/Ω */
static real camera_ray(const orc_scene *s, real u, real v, real lx, real ly, real xh, int H, olam *L, v3 *ro,
                       v3 *rd) {
  const lam_camera *c = &s->d->camera;
  v3 pos = fv(c->pos), X = fv(c->x), Y = fv(c->y), Z = fv(c->z), wij, wkl;
  real om, hero, step;
  int k;
  wij = vsub(vadd(vadd(pos, vmul(X, (real)-0.5f * (real)c->w + (real)c->delta * u)),
                  vmul(Y, (real)-0.5f * (real)c->h + (real)c->delta * v)), vmul(Z, (real)c->D));
  wkl = vadd(pos, vmul(vadd(vmul(X, lx - (real)0.5), vmul(Y, ly - (real)0.5)), 2 * (real)c->dof));
  *ro = wkl;
  *rd = vnorm(vsub(wij, wkl));
  om = (real)s->d->grid.sup - (real)s->d->grid.inf;
  hero = (real)s->d->grid.inf + om * xh;
  step = om / (real)H;
  L->H = H;
  for (k = 0; k < H; k++) {
    real l = hero + (real)k * step;
    if (l > (real)s->d->grid.sup) l -= om;
    L->l[k] = l;
  }
  return (real)1 / om;
}

/* the camera ray of (pixel, sample) from its counter-RNG key */
static real start_path(const orc_scene *s, uint64_t key, uint32_t pixel, int H, olam *L, v3 *ro, v3 *rd) {
  uint32_t w = s->d->camera.width, row = pixel / w, col = pixel - row * w;
  real u = (real)((float)col + om_flt(key, OM_DIM_U)), v = (real)((float)row + om_flt(key, OM_DIM_V));
  return camera_ray(s, u, v, (real)om_flt(key, OM_DIM_LX), (real)om_flt(key, OM_DIM_LY),
                    (real)om_flt(key, OM_DIM_HERO), H, L, ro, rd);
}

int orc_trace_paths(const orc_scene *s, const lam_render_params *p, const uint32_t *pixel, const uint32_t *sample,
                    size_t n, float *out) {
  long i;
  int H = (int)p->hero;
  if (H != 4 && H != 16) return -1;
#pragma omp parallel for schedule(dynamic, 16)
  for (i = 0; i < (long)n; i++) {
    opath pa;
    olam L;
    v3 ro, rd;
    real Li[HMAX];
    int k;
    pa.key = om_path_key(p->seed, pixel[i], sample[i]);
    pa.max_scat = p->max_scatterings;
    pa.p = (real)p->sample_p;
    start_path(s, pa.key, pixel[i], H, &L, &ro, &rd);
    trace_path(s, &pa, &L, ro, rd, 0, Li, NULL, NULL);
    for (k = 0; k < H; k++) out[i * H + k] = (float)Li[k];
  }
  return 0;
}

/* The queries of one path in the order tracePath makes them: rays (up to
 * cap) and kind (0 closest hit, 1 shadow); returns the number of queries
 * (may exceed cap).  A test hook for locating where two tracers part. */
long orc_path_rays(const orc_scene *s, const lam_render_params *p, uint32_t pixel, uint32_t sample, lam_ray *rays,
                   uint32_t *kind, size_t cap) {
  opath pa;
  olam L;
  v3 ro, rd;
  real Li[HMAX];
  int H = (int)p->hero;
  if (H != 4 && H != 16) return -1;
  pa.key = om_path_key(p->seed, pixel, sample);
  pa.max_scat = p->max_scatterings;
  pa.p = (real)p->sample_p;
  start_path(s, pa.key, pixel, H, &L, &ro, &rd);
  rec_rays = rays; rec_kind = kind; rec_n = 0; rec_cap = cap;
  trace_path(s, &pa, &L, ro, rd, 0, Li, NULL, NULL);
  rec_rays = NULL; rec_kind = NULL;
  return (long)rec_n;
}

/* The extend and shadow rays tracePath traces for these (pixel, sample)
 * paths (src/renderer.cpp:64-146: one closest-hit query per vertex, one
 * shadow query per light sample): counts[0] extend, counts[1] shadow. */
int orc_count_rays(const orc_scene *s, const lam_render_params *p, const uint32_t *pixel, const uint32_t *sample,
                   size_t n, uint64_t *counts) {
  long i;
  int H = (int)p->hero;
  uint64_t ext = 0, shd = 0;
  if (H != 4 && H != 16) return -1;
#pragma omp parallel reduction(+ : ext, shd)
  {
    n_ext_rays = 0;
    n_shd_rays = 0;
#pragma omp for schedule(dynamic, 16)
    for (i = 0; i < (long)n; i++) {
      opath pa;
      olam L;
      v3 ro, rd;
      real Li[HMAX];
      pa.key = om_path_key(p->seed, pixel[i], sample[i]);
      pa.max_scat = p->max_scatterings;
      pa.p = (real)p->sample_p;
      start_path(s, pa.key, pixel[i], H, &L, &ro, &rd);
      trace_path(s, &pa, &L, ro, rd, 0, Li, NULL, NULL);
    }
    ext += n_ext_rays;
    shd += n_shd_rays;
  }
  counts[0] = ext;
  counts[1] = shd;
  return 0;
}

/* coefficientλ::addHeroλ (cg.hpp:306-320), weights exactly as the reference:
 * wavelength l adds (1-t)·s to bin i-1 (if i > 0) and t·s to bin i (if i < n) */
static void hero_weights(const orc_scene *s, real l, uint32_t *bin, real *tw) {
  const lam_scene_desc *d = s->d;
  uint32_t n = d->grid.n, i = 0;
  real l1, t;
  for (; (real)((float)d->grid.inf + (float)i * (float)s->delta) <= l && i < n; i++) {
  }
  l1 = (real)((float)d->grid.inf + (float)i * (float)s->delta);
  t = (l1 - l) / s->delta;
  if (i == 0) t = 1;
  if (i == n) t = 0;
  *bin = i;
  *tw = t;
}
static void add_hero(const orc_scene *s, real *v, const olam *L, const real *sv) {
  uint32_t n = s->d->grid.n, i;
  real t;
  int k;
  for (k = 0; k < L->H; k++) {
    hero_weights(s, L->l[k], &i, &t);
    if (i > 0) v[i - 1] += (1 - t) * sv[k];
    if (i < n) v[i] += t * sv[k];
  }
}

/* per-pixel spectral accumulator of orc_render: grids of up to this many
 * bins (32 KB of stack per thread in the FP64 build) */
#define ORC_RENDER_MAX_BINS 4096

/* Renderer::render (src/renderer.cpp:14-61): OpenMP over pixels, a spectral
 * film per pixel, coefλ2linRGB with the row flip.  rows: render only rows
 * i with i % row_step == row_phase (bounded CPU-baseline samples); others
 * are left untouched.  xyz (optional) receives Σ v_i·CMF_i before the
 * 1/(SPP·H) and Δ factors, same orientation as linrgb. */
int orc_render(const orc_scene *s, const lam_render_params *p, float *linrgb, float *xyz, int threads, int row_step,
               int row_phase) {
  const lam_scene_desc *d = s->d;
  long w = d->camera.width, h = d->camera.height, i, j;
  int H = (int)p->hero;
  uint32_t nb = d->grid.n;
  if (H != 4 && H != 16) return -1;
  if (nb > ORC_RENDER_MAX_BINS) return -1;
  if (row_step < 1) row_step = 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
  for (i = 0; i < h; i++)
    for (j = 0; j < w; j++) {
      real acc[ORC_RENDER_MAX_BINS], X = 0, Y = 0, Z = 0, N = (real)(float)((uint64_t)p->spp * (uint64_t)H), r, g, b;
      uint32_t k, q;
      size_t o;
      if (i % row_step != row_phase) continue;
      for (q = 0; q < nb; q++) acc[q] = 0;
      for (k = 0; k < p->spp; k++) {
        opath pa;
        olam L;
        v3 ro, rd;
        real Li[HMAX], sv[HMAX], prob;
        int m;
        pa.key = om_path_key(p->seed, (uint32_t)(i * w + j), p->spp_offset + k);
        pa.max_scat = p->max_scatterings;
        pa.p = (real)p->sample_p;
        prob = start_path(s, pa.key, (uint32_t)(i * w + j), H, &L, &ro, &rd);
        trace_path(s, &pa, &L, ro, rd, 0, Li, NULL, NULL);
        for (m = 0; m < H; m++) sv[m] = Li[m] / prob;
        add_hero(s, acc, &L, sv);
      }
      for (q = 0; q < nb; q++) {
        X += acc[q] * (real)d->cmf[q];
        Y += acc[q] * (real)d->cmf[nb + q];
        Z += acc[q] * (real)d->cmf[2 * nb + q];
      }
      o = (size_t)((h - i - 1) * w + j) * 3;
      if (xyz) { xyz[o] = (float)X; xyz[o + 1] = (float)Y; xyz[o + 2] = (float)Z; }
      /* (v/N)|CMF = Σ (v_i/N) CMF_i · Δ, then XYZ2linRGB (cg.hpp:144-148) */
      X = 0; Y = 0; Z = 0;
      for (q = 0; q < nb; q++) {
        real vn = acc[q] / N;
        X += vn * (real)d->cmf[q];
        Y += vn * (real)d->cmf[nb + q];
        Z += vn * (real)d->cmf[2 * nb + q];
      }
      X = X * s->delta; Y = Y * s->delta; Z = Z * s->delta;
      r = (real)3.2404542f * X - (real)1.5371385f * Y - (real)0.4985314f * Z;
      g = (real)-0.9692660f * X + (real)1.8760108f * Y + (real)0.0415560f * Z;
      b = (real)0.0556434f * X - (real)0.2040259f * Y + (real)1.0572252f * Z;
      linrgb[o] = (float)r; linrgb[o + 1] = (float)g; linrgb[o + 2] = (float)b;
    }
  return 0;
}

/* Directional derivatives of the image loss: out[j] = Σ_px Σ_c dloss[px][c] ·
 * d rgb[px][c] / dε along direction j (v + j*n_params) for the image orc_render
 * produces, 1 <= K <= 16.  The film is linear in the path radiance, so each
 * path's radiance tangent is splatted, projected and converted like its
 * radiance (src/renderer.cpp:14-61, cg.hpp:144-148, 306-341).  Per-pixel sums
 * are added in pixel order, so the result does not depend on threading. */
int orc_jvp(const orc_scene *s, const lam_render_params *p, const int32_t *mat_d, const int32_t *mat_ior,
            const int32_t *mat_albedo, uint32_t n_params, int K, const double *v, const float *dloss, int threads,
            double *out) {
  const lam_scene_desc *d = s->d;
  long w = d->camera.width, h = d->camera.height, px;
  int H = (int)p->hero, j;
  uint32_t nb = d->grid.n, m;
  double *part;
  otan tn;
  if (H != 4 && H != 16) return -1;
  if (nb > 512 || K < 1 || K > KMAX) return -1;
  for (m = 0; m < d->n_materials; m++) {
    if (mat_d[m] >= 0 && (uint32_t)mat_d[m] >= n_params) return -1;
    if (mat_ior[m] >= 0 && (uint64_t)mat_ior[m] + nb > n_params) return -1;
    if (mat_albedo[m] >= 0 && (uint64_t)mat_albedo[m] + nb > n_params) return -1;
  }
  part = (double *)malloc(sizeof(double) * (size_t)(w * h) * (size_t)K);
  if (!part) return -1;
  tn.mat_d = mat_d; tn.mat_ior = mat_ior; tn.mat_albedo = mat_albedo; tn.v = v; tn.n_params = n_params; tn.K = K;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
  for (px = 0; px < w * h; px++) {
    long i = px / w, jj = px - i * w;
    double dacc[KMAX][512], N = (double)(float)((uint64_t)p->spp * (uint64_t)H);
    size_t o = (size_t)((h - i - 1) * w + jj) * 3;
    uint32_t k, q, bin;
    int a;
    for (a = 0; a < K; a++)
      for (q = 0; q < nb; q++) dacc[a][q] = 0;
    for (k = 0; k < p->spp; k++) {
      opath pa;
      olam L;
      v3 ro, rd;
      real Li[HMAX], prob, t;
      double dLi[KMAX * HMAX];
      int mm;
      pa.key = om_path_key(p->seed, (uint32_t)px, p->spp_offset + k);
      pa.max_scat = p->max_scatterings;
      pa.p = (real)p->sample_p;
      prob = start_path(s, pa.key, (uint32_t)px, H, &L, &ro, &rd);
      trace_path(s, &pa, &L, ro, rd, 0, Li, &tn, dLi);
      for (mm = 0; mm < H; mm++) {
        hero_weights(s, L.l[mm], &bin, &t);
        for (a = 0; a < K; a++) {
          double sv = dLi[a * HMAX + mm] / (double)prob;
          if (bin > 0) dacc[a][bin - 1] += (1 - (double)t) * sv;
          if (bin < nb) dacc[a][bin] += (double)t * sv;
        }
      }
    }
    for (a = 0; a < K; a++) {
      double X = 0, Y = 0, Z = 0, r, g, b;
      for (q = 0; q < nb; q++) {
        double vn = dacc[a][q] / N;
        X += vn * (double)d->cmf[q];
        Y += vn * (double)d->cmf[nb + q];
        Z += vn * (double)d->cmf[2 * nb + q];
      }
      X *= (double)s->delta; Y *= (double)s->delta; Z *= (double)s->delta;
      r = (double)3.2404542f * X - (double)1.5371385f * Y - (double)0.4985314f * Z;
      g = (double)-0.9692660f * X + (double)1.8760108f * Y + (double)0.0415560f * Z;
      b = (double)0.0556434f * X - (double)0.2040259f * Y + (double)1.0572252f * Z;
      part[(size_t)px * K + a] = (double)dloss[o] * r + (double)dloss[o + 1] * g + (double)dloss[o + 2] * b;
    }
  }
  for (j = 0; j < K; j++) out[j] = 0;
  for (px = 0; px < w * h; px++)
    for (j = 0; j < K; j++) out[j] += part[(size_t)px * K + j];
  free(part);
  return 0;
}

/* ---- unit probes: one material, light, film or camera function per call ------
 * Each probe runs the oracle's own function (the one trace_path and
 * orc_render use) on caller-given geometry and random numbers, so tests can
 * pin it against closed forms and histograms of the mathematics.  Arrays
 * are float regardless of the build's `real`.
 *   frame: 9 floats per query (x, y, z of the shading frame, z = normal)
 *   lam:   H floats per query (hero wavelengths) */
static void probe_frame(const float *f, int front, osurf *sf) {
  sf->F.x = fv(f);
  sf->F.y = fv(f + 3);
  sf->F.z = fv(f + 6);
  sf->front = front;
  sf->p = V(0, 0, 0);
  sf->gn = sf->F.z;
  sf->mat = sf->obj = NONE;
}
static void probe_lam(int H, const float *lam, olam *L) {
  int k;
  L->H = H;
  for (k = 0; k < H; k++) L->l[k] = (real)lam[k];
}

/* BxDFcosθ of material `mat` (bxdf): out = cnt*H */
int orc_probe_bxdf(const orc_scene *s, uint32_t mat, int H, const float *lam, const float *i, const float *o,
                   const float *n, const int32_t *front, size_t cnt, float *out) {
  size_t q;
  int k;
  if (mat >= s->d->n_materials || H < 1 || H > HMAX) return -1;
  for (q = 0; q < cnt; q++) {
    olam L;
    real v[HMAX];
    probe_lam(H, lam + q * H, &L);
    bxdf(s, &s->d->materials[mat], &L, fv(i + 3 * q), fv(o + 3 * q), fv(n + 3 * q), front[q], v);
    for (k = 0; k < H; k++) out[q * H + k] = (float)v[k];
  }
  return 0;
}

/* sample::materiali (sample_material) with random numbers xi = (xl, xa, xb)
 * per query: ok, dir (3), prob, mult (H) */
int orc_probe_sample_material(const orc_scene *s, uint32_t mat, int H, const float *lam, const float *frame,
                              const int32_t *front, const float *o, const float *xi, float p, size_t cnt,
                              int32_t *ok, float *dir, float *prob, float *mult) {
  size_t q;
  int k;
  if (mat >= s->d->n_materials || H < 1 || H > HMAX) return -1;
  for (q = 0; q < cnt; q++) {
    olam L;
    osurf sf;
    omsample ms;
    probe_lam(H, lam + q * H, &L);
    probe_frame(frame + 9 * q, front[q], &sf);
    memset(&ms, 0, sizeof ms);
    sample_material(s, &s->d->materials[mat], &L, &sf, fv(o + 3 * q), (real)xi[3 * q], (real)xi[3 * q + 1],
                    (real)xi[3 * q + 2], (real)p, &ms);
    ok[q] = ms.ok;
    dir[3 * q] = (float)ms.dir.x;
    dir[3 * q + 1] = (float)ms.dir.y;
    dir[3 * q + 2] = (float)ms.dir.z;
    prob[q] = (float)ms.prob;
    for (k = 0; k < H; k++) mult[q * H + k] = ms.ok ? (float)ms.mult[k] : 0.f;
  }
  return 0;
}

/* probForMateriali (prob_material), called as trace_path calls it:
 * wl = the light sample's direction, o = the outgoing direction */
int orc_probe_prob_material(const orc_scene *s, uint32_t mat, const float *frame, const int32_t *front,
                            const float *wl, const float *o, const float *l0, float p, size_t cnt, float *out) {
  size_t q;
  if (mat >= s->d->n_materials) return -1;
  for (q = 0; q < cnt; q++) {
    osurf sf;
    probe_frame(frame + 9 * q, front[q], &sf);
    out[q] = (float)prob_material(s, &s->d->materials[mat], &sf, fv(wl + 3 * q), fv(o + 3 * q), (real)l0[q],
                                  (real)p);
  }
  return 0;
}

/* sample::light (sample_light, including its shadow query) at surface point
 * pos with shading frame `frame`: dir (3), prob, mult (H) */
int orc_probe_sample_light(const orc_scene *s, uint32_t light, int H, const float *lam, const float *pos,
                           const float *frame, const int32_t *front, const float *xi, size_t cnt, float *dir,
                           float *prob, float *mult) {
  size_t q;
  int k;
  if (light >= s->d->n_lights || H < 1 || H > HMAX) return -1;
  for (q = 0; q < cnt; q++) {
    olam L;
    osurf sf;
    olsample ls;
    probe_lam(H, lam + q * H, &L);
    probe_frame(frame + 9 * q, front[q], &sf);
    sf.p = fv(pos + 3 * q);
    sample_light(s, &s->d->lights[light], &L, &sf, (real)xi[2 * q], (real)xi[2 * q + 1], &ls);
    dir[3 * q] = (float)ls.dir.x;
    dir[3 * q + 1] = (float)ls.dir.y;
    dir[3 * q + 2] = (float)ls.dir.z;
    prob[q] = (float)ls.prob;
    for (k = 0; k < H; k++) mult[q * H + k] = (float)ls.mult[k];
  }
  return 0;
}

/* probForLight (prob_light) */
int orc_probe_prob_light(const orc_scene *s, uint32_t light, const float *pos, const float *dir, size_t cnt,
                         float *out) {
  size_t q;
  if (light >= s->d->n_lights) return -1;
  for (q = 0; q < cnt; q++) out[q] = (float)prob_light(s, &s->d->lights[light], fv(pos + 3 * q), fv(dir + 3 * q));
  return 0;
}

/* addHeroλ (add_hero): splats the (lam[q], val[q]) pairs one at a time into
 * the grid.n-bin spectrum `bins` (accumulated, not cleared) */
int orc_probe_splat(const orc_scene *s, const float *lam, const float *val, size_t cnt, float *bins) {
  uint32_t n = s->d->grid.n, b;
  size_t q;
  real acc[512];
  if (n > 512) return -1;
  for (b = 0; b < n; b++) acc[b] = (real)bins[b];
  for (q = 0; q < cnt; q++) {
    olam L;
    real v = (real)val[q];
    L.H = 1;
    L.l[0] = (real)lam[q];
    add_hero(s, acc, &L, &v);
  }
  for (b = 0; b < n; b++) bins[b] = (float)acc[b];
  return 0;
}

/* camera_ray for explicit sample numbers: u = (u, v, lens x, lens y,
 * hero ξ) per query with u, v film positions in pixels (column + fraction,
 * row from the bottom + fraction) -> ro (3), rd (3), lam (H) */
int orc_probe_camera(const orc_scene *s, int H, const float *u, size_t cnt, float *ro, float *rd, float *lam) {
  size_t q;
  int k;
  if (H < 1 || H > HMAX) return -1;
  for (q = 0; q < cnt; q++) {
    const float *x = u + 5 * q;
    olam L;
    v3 o, d;
    camera_ray(s, (real)x[0], (real)x[1], (real)x[2], (real)x[3], (real)x[4], H, &L, &o, &d);
    ro[3 * q] = (float)o.x; ro[3 * q + 1] = (float)o.y; ro[3 * q + 2] = (float)o.z;
    rd[3 * q] = (float)d.x; rd[3 * q + 1] = (float)d.y; rd[3 * q + 2] = (float)d.z;
    for (k = 0; k < H; k++) lam[q * H + k] = (float)L.l[k];
  }
  return 0;
}

/* math hooks for tests: fn 0 sin, 1 cos, 2 pow, 3 atan2 */
void orc_math(int fn, const float *a, const float *b, size_t n, float *out) {
  size_t i;
  for (i = 0; i < n; i++) {
    switch (fn) {
      case 0: out[i] = om_sinf(a[i]); break;
      case 1: out[i] = om_cosf(a[i]); break;
      case 2: out[i] = om_powf(a[i], b[i]); break;
      default: out[i] = om_atan2f(a[i], b[i]); break;
    }
  }
}
void orc_rng(uint64_t seed, const uint32_t *pixel, const uint32_t *sample, const uint32_t *dim, size_t n,
             uint32_t *out) {
  size_t i;
  for (i = 0; i < n; i++) out[i] = om_u32(om_path_key(seed, pixel[i], sample[i]), dim[i]);
}
int orc_is_f64(void) {
#ifdef ORC_F64
  return 1;
#else
  return 0;
#endif
}

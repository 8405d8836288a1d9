// Binned-SAH BVH2 -> BVH8 collapse -> 8-bit quantized 128-B nodes.
#include "bvh8.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>

namespace lamh {

namespace {

struct Box {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(const Box &b) {
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  void grow(const float p[3]) {
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  double area() const {
    double d[3];
    for (int a = 0; a < 3; a++) d[a] = std::max(0.0, double(hi[a]) - double(lo[a]));
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }
};

struct BNode {
  Box box;
  int left = -1, right = -1;
  uint32_t first = 0, count = 0;  // leaf when count > 0
};

constexpr int kBins = 16;
constexpr double kCostTrav = 1.0, kCostTri = 1.0;

struct Builder {
  const std::vector<Box> &tbox;
  std::vector<float> cent;  // 3 per tri
  std::vector<uint32_t> idx;
  std::vector<BNode> nodes;

  explicit Builder(const std::vector<Box> &b) : tbox(b) {
    const size_t n = b.size();
    cent.resize(3 * n);
    idx.resize(n);
    for (size_t i = 0; i < n; i++) {
      idx[i] = uint32_t(i);
      for (int a = 0; a < 3; a++) cent[3 * i + a] = 0.5f * (b[i].lo[a] + b[i].hi[a]);
    }
  }

  int make_leaf(const Box &box, uint32_t first, uint32_t count) {
    BNode n;
    n.box = box;
    n.first = first;
    n.count = count;
    nodes.push_back(n);
    return int(nodes.size() - 1);
  }

  int build(uint32_t first, uint32_t count) {
    Box box, cbox;
    for (uint32_t i = first; i < first + count; i++) {
      box.grow(tbox[idx[i]]);
      cbox.grow(&cent[3 * idx[i]]);
    }
    if (count <= 1) return make_leaf(box, first, count);

    // binned SAH over the centroid bounds
    double best_cost = DBL_MAX;
    int best_axis = -1, best_split = -1;
    for (int a = 0; a < 3; a++) {
      const float cmin = cbox.lo[a], cmax = cbox.hi[a];
      if (!(cmax > cmin)) continue;
      Box bb[kBins];
      uint32_t bc[kBins] = {0};
      const float scale = float(kBins) / (cmax - cmin);
      for (uint32_t i = first; i < first + count; i++) {
        int b = int((cent[3 * idx[i] + a] - cmin) * scale);
        b = std::min(std::max(b, 0), kBins - 1);
        bb[b].grow(tbox[idx[i]]);
        bc[b]++;
      }
      double rarea[kBins];
      uint32_t rcount[kBins];
      Box r;
      uint32_t rc = 0;
      for (int b = kBins - 1; b > 0; b--) {
        r.grow(bb[b]);
        rc += bc[b];
        rarea[b] = r.area();
        rcount[b] = rc;
      }
      Box l;
      uint32_t lc = 0;
      for (int b = 0; b < kBins - 1; b++) {
        l.grow(bb[b]);
        lc += bc[b];
        if (lc == 0 || rcount[b + 1] == 0) continue;
        double c = l.area() * lc + rarea[b + 1] * rcount[b + 1];
        if (c < best_cost) { best_cost = c; best_axis = a; best_split = b; }
      }
    }
    const double area = box.area();
    const double split_cost = best_axis < 0 ? DBL_MAX
        : kCostTrav + (area > 0 ? best_cost / area : best_cost) * kCostTri;
    const double leaf_cost = double(count) * kCostTri;
    if (count <= uint32_t(kMaxLeafTris) && leaf_cost <= split_cost) return make_leaf(box, first, count);

    uint32_t mid;
    if (best_axis < 0) {
      mid = first + count / 2;  // coincident centroids: split by index
    } else {
      const float cmin = cbox.lo[best_axis], cmax = cbox.hi[best_axis];
      const float scale = float(kBins) / (cmax - cmin);
      uint32_t *b = idx.data() + first, *e = b + count;
      uint32_t *m = std::partition(b, e, [&](uint32_t t) {
        int bin = int((cent[3 * t + best_axis] - cmin) * scale);
        bin = std::min(std::max(bin, 0), kBins - 1);
        return bin <= best_split;
      });
      mid = uint32_t(m - idx.data());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    const int self = int(nodes.size());
    nodes.emplace_back();
    nodes[self].box = box;
    const int l = build(first, mid - first);
    const int r = build(mid, first + count - mid);
    nodes[self].left = l;
    nodes[self].right = r;
    return self;
  }
};

// decode as the kernel: o + float(q) * 2^(e-127), no contraction
inline float decode(float o, uint8_t q, uint8_t e) {
  uint32_t bits = uint32_t(e) << 23;
  float s;
  std::memcpy(&s, &bits, 4);
  float qs = float(q) * s;
  return o + qs;
}

void quantize_node(Node8 &n, const Box *cb, int nc) {
  Box nb;
  for (int k = 0; k < nc; k++) nb.grow(cb[k]);
  for (int a = 0; a < 3; a++) {
    n.origin[a] = nb.lo[a];
    const double ext = double(nb.hi[a]) - double(nb.lo[a]);
    int e = -126;
    if (ext > 0) {
      e = int(std::ceil(std::log2(ext / 255.0)));
      e = std::max(e, -126);
    }
    for (;;) {
      bool ok = true;
      uint8_t be = uint8_t(e + 127);
      const double s = std::ldexp(1.0, e);
      for (int k = 0; k < 8 && ok; k++) {
        if (k >= nc) { n.qlo[a][k] = 255; n.qhi[a][k] = 0; continue; }
        long ql = long(std::floor((double(cb[k].lo[a]) - double(n.origin[a])) / s));
        long qh = long(std::ceil((double(cb[k].hi[a]) - double(n.origin[a])) / s));
        ql = std::max(0L, std::min(255L, ql));
        qh = std::max(0L, std::min(255L, qh));
        while (ql > 0 && decode(n.origin[a], uint8_t(ql), be) > cb[k].lo[a]) ql--;
        while (qh < 255 && decode(n.origin[a], uint8_t(qh), be) < cb[k].hi[a]) qh++;
        if (decode(n.origin[a], uint8_t(ql), be) > cb[k].lo[a] ||
            decode(n.origin[a], uint8_t(qh), be) < cb[k].hi[a]) { ok = false; break; }
        n.qlo[a][k] = uint8_t(ql);
        n.qhi[a][k] = uint8_t(qh);
      }
      if (ok) {
        if (a == 0) n.ex = be; else if (a == 1) n.ey = be; else n.ez = be;
        break;
      }
      e++;
    }
  }
}

}  // namespace

void decode_child_box(const Node8 &n, int k, float lo[3], float hi[3]) {
  const uint8_t e[3] = {n.ex, n.ey, n.ez};
  for (int a = 0; a < 3; a++) {
    lo[a] = decode(n.origin[a], n.qlo[a][k], e[a]);
    hi[a] = decode(n.origin[a], n.qhi[a][k], e[a]);
  }
}

bool build_bvh8(const float *tri_v, uint32_t n_tris, Bvh8 &out) {
  out = Bvh8();
  if (n_tris == 0) return true;
  if (n_tris > kLeafFirstMask) return false;
  std::vector<Box> tb(n_tris);
  for (uint32_t i = 0; i < n_tris; i++)
    for (int v = 0; v < 3; v++) tb[i].grow(tri_v + 9 * i + 3 * v);
  Builder B(tb);
  const int root = B.build(0, n_tris);
  out.order = B.idx;

  // collapse: BFS over BVH2 interior nodes
  struct Item { int bnode; uint32_t n8; uint32_t depth; };
  std::vector<Item> queue;
  out.nodes.emplace_back();
  std::memset(&out.nodes[0], 0, sizeof(Node8));
  queue.push_back({root, 0, 1});
  const double root_area = B.nodes[root].box.area();
  double sah = 0.0;
  for (size_t qi = 0; qi < queue.size(); qi++) {
    const Item it = queue[qi];
    out.depth = std::max(out.depth, it.depth);
    int c[8];
    int nc = 0;
    const BNode &bn = B.nodes[it.bnode];
    if (bn.count > 0) {
      c[nc++] = it.bnode;  // whole tree is one leaf
    } else {
      c[nc++] = bn.left;
      c[nc++] = bn.right;
      while (nc < 8) {
        int best = -1;
        double ba = -1.0;
        for (int k = 0; k < nc; k++) {
          const BNode &x = B.nodes[c[k]];
          if (x.count == 0 && x.box.area() > ba) { ba = x.box.area(); best = k; }
        }
        if (best < 0) break;
        const BNode &x = B.nodes[c[best]];
        c[best] = x.left;
        c[nc++] = x.right;
      }
    }
    if (root_area > 0) sah += kCostTrav * bn.box.area() / root_area;
    Box cb[8];
    Node8 node;
    std::memset(&node, 0, sizeof(Node8));
    node.nchild = uint8_t(nc);
    for (int k = 0; k < 8; k++) node.child[k] = 0xFFFFFFFFu;
    for (int k = 0; k < nc; k++) {
      const BNode &x = B.nodes[c[k]];
      cb[k] = x.box;
      if (x.count > 0) {
        node.child[k] = kLeafBit | ((x.count - 1) << kLeafCountShift) | x.first;
        out.n_leaves++;
        if (root_area > 0) sah += kCostTri * x.count * x.box.area() / root_area;
      } else {
        const uint32_t idx = uint32_t(out.nodes.size());
        out.nodes.emplace_back();
        node.child[k] = idx;
        queue.push_back({c[k], idx, it.depth + 1});
      }
    }
    quantize_node(node, cb, nc);
    out.nodes[it.n8] = node;
  }
  out.sah_cost = sah;

  // worst-case stack: max over root-to-leaf paths of sum(nchild-1), + 1
  std::vector<uint32_t> bound(out.nodes.size(), 0);
  for (size_t i = out.nodes.size(); i-- > 0;) {
    const Node8 &n = out.nodes[i];
    uint32_t m = 0;
    for (int k = 0; k < n.nchild; k++)
      if (!(n.child[k] & kLeafBit)) m = std::max(m, bound[n.child[k]]);
    bound[i] = m + (n.nchild > 0 ? n.nchild - 1u : 0u);
  }
  out.stack_bound = bound[0] + 1;
  return true;
}

}  // namespace lamh

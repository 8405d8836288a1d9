// BVH8 built on the GPU (LAM_UPLOAD_GPU_BVH): a linear BVH over the
// triangle centroids' 30-bit Morton codes (Karras 2012, "Maximizing
// parallelism in the construction of BVHs, octrees and k-d trees"), boxes
// propagated bottom-up, then collapsed to 8-wide nodes the way the host
// builder collapses its SAH tree (repeatedly open the largest-area child
// until 8 children; ranges of at most 4 triangles become leaves), with the
// child boxes quantized to 8 bits against a power-of-two scale and checked
// conservative with the traversal's own decode.  The node record is the
// host builder's (host/bvh8.hpp), so the traversal kernels take either.
//
// The host builder (binned SAH) gives better trees; this one trades tree
// quality for build time on large scenes (10 M triangles: seconds on the
// host, milliseconds here).  Images do not depend on the tree: the closest
// hit and its tie-break are exact for any conservative BVH.
#include <cfloat>

#include "lam_launch.h"

namespace lam {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kPrimRef = 0x80000000u;  // BVH2 child reference to sorted primitive i
constexpr uint32_t kLeafMax = 4;            // triangles per leaf, as the host builder
constexpr uint32_t kRedBlocks = 1024;       // blocks of the two-pass reductions

// float <-> unsigned with the float order (min/max of floats as integers)
__device__ __forceinline__ uint32_t f2o(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// triangle record: (v0.xyz v1.x) (v1.yz v2.xy) (v2.z obj prim orig)
__device__ __forceinline__ void tri_box(const float4 *__restrict__ tri, uint32_t t, float4 &lo, float4 &hi) {
  const float4 a = tri[3 * size_t(t)], b = tri[3 * size_t(t) + 1], c = tri[3 * size_t(t) + 2];
  lo = make_float4(fminf(fminf(a.x, a.w), b.z), fminf(fminf(a.y, b.x), b.w), fminf(fminf(a.z, b.y), c.x), 0.f);
  hi = make_float4(fmaxf(fmaxf(a.x, a.w), b.z), fmaxf(fmaxf(a.y, b.x), b.w), fmaxf(fmaxf(a.z, b.y), c.x), 0.f);
}

// per-primitive boxes and per-block centroid bounds (6 ordered words per block)
__global__ __launch_bounds__(256) void k_prim_boxes(const float4 *__restrict__ tri, uint32_t n,
                                                    float4 *__restrict__ pbox, uint32_t *__restrict__ partial) {
  __shared__ uint32_t red[6][256];
  uint32_t m[6] = {kNone, kNone, kNone, 0u, 0u, 0u};
  for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < n; t += gridDim.x * 256u) {
    float4 lo, hi;
    tri_box(tri, t, lo, hi);
    pbox[2 * size_t(t)] = lo;
    pbox[2 * size_t(t) + 1] = hi;
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      m[a] = min(m[a], f2o(c[a]));
      m[3 + a] = max(m[3 + a], f2o(c[a]));
    }
  }
#pragma unroll
  for (int a = 0; a < 6; a++) red[a][threadIdx.x] = m[a];
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int a = 0; a < 6; a++)
        red[a][threadIdx.x] = a < 3 ? min(red[a][threadIdx.x], red[a][threadIdx.x + s])
                                    : max(red[a][threadIdx.x], red[a][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// the per-block bounds reduced by one block: bounds[0..5] = centroid min xyz, max xyz (floats)
__global__ __launch_bounds__(256) void k_bounds_final(const uint32_t *__restrict__ partial, uint32_t nb,
                                                      float *__restrict__ bounds) {
  __shared__ uint32_t red[6][256];
  uint32_t m[6] = {kNone, kNone, kNone, 0u, 0u, 0u};
  for (uint32_t b = threadIdx.x; b < nb; b += 256u)
#pragma unroll
    for (int a = 0; a < 6; a++) m[a] = a < 3 ? min(m[a], partial[6 * b + a]) : max(m[a], partial[6 * b + a]);
#pragma unroll
  for (int a = 0; a < 6; a++) red[a][threadIdx.x] = m[a];
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int a = 0; a < 6; a++)
        red[a][threadIdx.x] = a < 3 ? min(red[a][threadIdx.x], red[a][threadIdx.x + s])
                                    : max(red[a][threadIdx.x], red[a][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 6) bounds[threadIdx.x] = o2f(red[threadIdx.x][0]);
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {  // 10 bits -> every third bit
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ __launch_bounds__(256) void k_morton(const float4 *__restrict__ pbox, uint32_t n,
                                                const float *__restrict__ bounds, uint32_t *__restrict__ keys,
                                                uint32_t *__restrict__ vals) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= n) return;
  const float4 lo = pbox[2 * size_t(t)], hi = pbox[2 * size_t(t) + 1];
  const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  uint32_t q[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float ext = bounds[3 + a] - bounds[a];
    const float u = ext > 0.f ? (c[a] - bounds[a]) / ext : 0.f;
    q[a] = uint32_t(fminf(fmaxf(u * 1024.f, 0.f), 1023.f));
  }
  keys[t] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  vals[t] = t;
}

// common-prefix length of sorted keys i and j; -1 outside [0, n); equal
// keys are told apart by their indices
__device__ __forceinline__ int delta(const uint32_t *__restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  return a != b ? __clz(a ^ b) : 32 + __clz(uint32_t(i) ^ uint32_t(j));
}

// internal node i of the radix tree over n >= 2 sorted keys: its children
// (kPrimRef | leaf index, or internal index) and key range; parents of all
// internal nodes and leaves
__global__ __launch_bounds__(256) void k_karras(const uint32_t *__restrict__ k, int n, uint2 *__restrict__ child,
                                                uint2 *__restrict__ range, uint32_t *__restrict__ parent_in,
                                                uint32_t *__restrict__ parent_leaf) {
  const int i = int(blockIdx.x * 256u + threadIdx.x);
  if (i >= n - 1) return;
  const int d = delta(k, n, i, i + 1) - delta(k, n, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = delta(k, n, i, i - d);
  int lmax = 2;
  while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(k, n, i, j);
  int s = 0;
  for (int t = l;;) {
    t = (t + 1) / 2;
    if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int g = i + s * d + min(d, 0);
  const int first = min(i, j), last = max(i, j);
  const uint32_t left = first == g ? (kPrimRef | uint32_t(g)) : uint32_t(g);
  const uint32_t right = last == g + 1 ? (kPrimRef | uint32_t(g + 1)) : uint32_t(g + 1);
  child[i] = make_uint2(left, right);
  range[i] = make_uint2(uint32_t(first), uint32_t(last));
  if (left & kPrimRef) parent_leaf[left & ~kPrimRef] = uint32_t(i); else parent_in[left] = uint32_t(i);
  if (right & kPrimRef) parent_leaf[right & ~kPrimRef] = uint32_t(i); else parent_in[right] = uint32_t(i);
  if (i == 0) parent_in[0] = kNone;
}

__device__ __forceinline__ void ref_box(uint32_t ref, const float4 *__restrict__ pbox, const uint32_t *__restrict__ perm,
                                        const float4 *ibox, float4 &lo, float4 &hi) {
  if (ref & kPrimRef) {
    const uint32_t t = perm[ref & ~kPrimRef];
    lo = pbox[2 * size_t(t)];
    hi = pbox[2 * size_t(t) + 1];
  } else {
    lo = ibox[2 * size_t(ref)];
    hi = ibox[2 * size_t(ref) + 1];
  }
}

// Boxes bottom-up: the second child to arrive at a node writes its box.
// The climb is split so that nearly all of it needs only workgroup-scope
// ordering (a device-scope fence writes back and invalidates the L2 of the
// wave's XCD; one per climb step cost milliseconds).  Phase 1: block b owns
// the kBuBlock consecutive sorted leaves [kBuBlock b, kBuBlock (b+1)) and
// climbs only through nodes whose key range lies inside them, with block-
// scope atomics and fences; a climber whose parent's range crosses a block
// boundary appends that parent to `top` and stops.  Phase 2 (one block,
// after the kernel boundary has made phase 1's boxes visible) climbs from
// the appended nodes through the crossing nodes, which form the top of the
// tree.  Each crossing node gets exactly two arrivals: one per child, from
// the phase-1 completer of a non-crossing child (or the leaf's own thread)
// and from the phase-2 completer of a crossing child.
constexpr uint32_t kBuBlock = 1024;

__device__ __forceinline__ void node_box(uint32_t node, const uint2 *__restrict__ child,
                                         const uint32_t *__restrict__ perm, const float4 *__restrict__ pbox,
                                         float4 *ibox) {
  const uint2 c = child[node];
  float4 la, ha, lb, hb;
  ref_box(c.x, pbox, perm, ibox, la, ha);
  ref_box(c.y, pbox, perm, ibox, lb, hb);
  ibox[2 * size_t(node)] = make_float4(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z), 0.f);
  ibox[2 * size_t(node) + 1] = make_float4(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z), 0.f);
}

__global__ __launch_bounds__(kBuBlock) void k_bottom_up_local(
    const uint2 *__restrict__ child, const uint2 *__restrict__ range, const uint32_t *__restrict__ parent_in,
    const uint32_t *__restrict__ parent_leaf, const uint32_t *__restrict__ perm, const float4 *__restrict__ pbox,
    uint32_t n, float4 *ibox, uint32_t *arrive, uint32_t *top, uint32_t *top_count, uint32_t top_cap) {
  const uint32_t t = blockIdx.x * kBuBlock + threadIdx.x;
  if (t >= n) return;
  uint32_t node = parent_leaf[t];
  while (node != kNone) {
    const uint2 r = range[node];
    if (r.x / kBuBlock != r.y / kBuBlock) {
      const uint32_t k = atomicAdd(top_count, 1u);
      if (k < top_cap) top[k] = node;  // at most two per crossing node: k < 2(n-1)
      return;
    }
    __threadfence_block();
    if (__hip_atomic_fetch_add(arrive + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u) return;
    __threadfence_block();
    node_box(node, child, perm, pbox, ibox);
    node = parent_in[node];
  }
}

__global__ __launch_bounds__(kBuBlock) void k_bottom_up_top(const uint2 *__restrict__ child,
                                                            const uint32_t *__restrict__ parent_in,
                                                            const uint32_t *__restrict__ perm,
                                                            const float4 *__restrict__ pbox, float4 *ibox,
                                                            uint32_t *arrive, const uint32_t *__restrict__ top,
                                                            const uint32_t *__restrict__ top_count,
                                                            uint32_t top_cap) {
  const uint32_t cnt = min(*top_count, top_cap);
  for (uint32_t e = threadIdx.x; e < cnt; e += kBuBlock) {
    uint32_t node = top[e];
    while (node != kNone) {
      __threadfence_block();
      if (__hip_atomic_fetch_add(arrive + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u) break;
      __threadfence_block();
      node_box(node, child, perm, pbox, ibox);
      node = parent_in[node];
    }
  }
}

// decode as the traversal: o + float(q) * 2^(e-127)
__device__ __forceinline__ float qdecode(float o, uint32_t q, uint32_t e) {
  const float s = __uint_as_float(e << 23);
  const float qs = float(q) * s;
  return o + qs;
}

// 8-bit child boxes against the node's origin and a power-of-two scale per
// axis, conservative under qdecode (host/bvh8.cpp quantize_node)
__device__ void quantize(const float4 *clo, const float4 *chi, int nc, Node8d &nd) {
  float o[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int k = 0; k < nc; k++) {
    const float l3[3] = {clo[k].x, clo[k].y, clo[k].z}, h3[3] = {chi[k].x, chi[k].y, chi[k].z};
    for (int a = 0; a < 3; a++) {
      o[a] = fminf(o[a], l3[a]);
      h[a] = fmaxf(h[a], h3[a]);
    }
  }
  nd.ox = o[0];
  nd.oy = o[1];
  nd.oz = o[2];
  uint32_t ebits = 0;
  uint8_t *ql8 = reinterpret_cast<uint8_t *>(nd.qlo), *qh8 = reinterpret_cast<uint8_t *>(nd.qhi);
  for (int a = 0; a < 3; a++) {
    const double ext = double(h[a]) - double(o[a]);
    int e = -126;
    if (ext > 0.0) e = max(int(ceil(log2(ext / 255.0))), -126);
    for (;;) {
      const uint32_t be = uint32_t(e + 127);
      const double s = ldexp(1.0, e);
      bool ok = true;
      for (int k = 0; k < 8 && ok; k++) {
        if (k >= nc) {
          ql8[8 * a + k] = 255;
          qh8[8 * a + k] = 0;
          continue;
        }
        const float lo = a == 0 ? clo[k].x : a == 1 ? clo[k].y : clo[k].z;
        const float hi = a == 0 ? chi[k].x : a == 1 ? chi[k].y : chi[k].z;
        long lq = long(floor((double(lo) - double(o[a])) / s));
        long hq = long(ceil((double(hi) - double(o[a])) / s));
        lq = lq < 0 ? 0 : (lq > 255 ? 255 : lq);
        hq = hq < 0 ? 0 : (hq > 255 ? 255 : hq);
        while (lq > 0 && qdecode(o[a], uint32_t(lq), be) > lo) lq--;
        while (hq < 255 && qdecode(o[a], uint32_t(hq), be) < hi) hq++;
        if (qdecode(o[a], uint32_t(lq), be) > lo || qdecode(o[a], uint32_t(hq), be) < hi) {
          ok = false;
          break;
        }
        ql8[8 * a + k] = uint8_t(lq);
        qh8[8 * a + k] = uint8_t(hq);
      }
      if (ok) {
        ebits |= be << (8 * a);
        break;
      }
      e++;
    }
  }
  nd.e_n = ebits | (uint32_t(nc) << 24);
}

__device__ __forceinline__ uint32_t ref_count(uint32_t ref, const uint2 *__restrict__ range) {
  if (ref & kPrimRef) return 1u;
  const uint2 r = range[ref];
  return r.y - r.x + 1u;
}
__device__ __forceinline__ uint32_t ref_first(uint32_t ref, const uint2 *__restrict__ range) {
  return (ref & kPrimRef) ? (ref & ~kPrimRef) : range[ref].x;
}

// One level of the collapse: BVH8 nodes [a, b), each made from the radix-
// tree node node_ref[i] (its acc: Σ(children-1) over its ancestors).  New
// interior children are numbered from *ctr (wave-aggregated), so the next
// level is [b, *ctr).  stats: [0] ctr, [1] stack bound, [2] leaves, [3] set
// when a child would be numbered at or past cap (then it is not written).
__global__ __launch_bounds__(256) void k_collapse(const uint2 *__restrict__ child, const uint2 *__restrict__ range,
                                                  const float4 *__restrict__ pbox, const uint32_t *__restrict__ perm,
                                                  const float4 *__restrict__ ibox, uint32_t a, uint32_t b,
                                                  uint32_t cap, uint32_t *__restrict__ node_ref,
                                                  uint32_t *__restrict__ node_acc, Node8d *__restrict__ nodes,
                                                  uint32_t *__restrict__ stats) {
  const uint32_t i = a + blockIdx.x * 256u + threadIdx.x;
  const bool live = i < b;
  uint32_t c[8];
  int nc = 0;
  uint32_t acc = 0;
  if (live) {
    const uint32_t root = node_ref[i];
    acc = node_acc[i];
    if ((root & kPrimRef) || ref_count(root, range) <= kLeafMax) {
      c[nc++] = root;  // the whole tree is one leaf
    } else {
      const uint2 cc = child[root];
      c[nc++] = cc.x;
      c[nc++] = cc.y;
      while (nc < 8) {
        int best = -1;
        float ba = -1.f;
        for (int k = 0; k < nc; k++) {
          if (ref_count(c[k], range) <= kLeafMax) continue;
          float4 lo, hi;
          ref_box(c[k], pbox, perm, ibox, lo, hi);
          const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
          const float area = dx * dy + dy * dz + dz * dx;
          if (area > ba) {
            ba = area;
            best = k;
          }
        }
        if (best < 0) break;
        const uint2 cc2 = child[c[best]];
        c[best] = cc2.x;
        c[nc++] = cc2.y;
      }
    }
  }
  // interior children of this node, numbered by a wave-aggregated atomic
  uint32_t ninner = 0, nleaf = 0;
  for (int k = 0; k < nc; k++) {
    if (ref_count(c[k], range) > kLeafMax) ninner++;
    else nleaf++;
  }
  const uint32_t lane = threadIdx.x % 64;
  uint32_t incl = ninner;  // inclusive prefix over the wave
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  uint32_t base = 0;
  if (lane == 63 && total) base = atomicAdd(stats, total);
  base = __shfl(base, 63, 64) + incl - ninner;
  // the stack bound (Σ(children-1) along the path + 1) and the leaf count
  uint32_t bound = live ? acc + uint32_t(nc) : 0u, leaves = nleaf;
  for (uint32_t o = 32; o > 0; o >>= 1) {
    bound = max(bound, __shfl_xor(bound, o, 64));
    leaves += __shfl_xor(leaves, o, 64);
  }
  if (lane == 0) {
    atomicMax(stats + 1, bound);
    if (leaves) atomicAdd(stats + 2, leaves);
  }
  if (!live) return;
  Node8d nd;
  float4 clo[8], chi[8];
  uint32_t next = base;
  for (int k = 0; k < 8; k++) nd.child[k] = kNone;
  for (int k = 0; k < 8; k++) nd.pad[k] = 0;
  for (int k = 0; k < nc; k++) {
    ref_box(c[k], pbox, perm, ibox, clo[k], chi[k]);
    const uint32_t cnt = ref_count(c[k], range);
    if (cnt <= kLeafMax) {
      nd.child[k] = kLeafBit | ((cnt - 1u) << 27) | ref_first(c[k], range);
    } else {
      nd.child[k] = next;
      if (next < cap) {
        node_ref[next] = c[k];
        node_acc[next] = acc + uint32_t(nc) - 1u;
      } else {
        atomicOr(stats + 3, 1u);
      }
      next++;
    }
  }
  quantize(clo, chi, nc, nd);
  nodes[i] = nd;
}

// triangle and shading records in the sorted (leaf) order
__global__ __launch_bounds__(256) void k_permute(const uint32_t *__restrict__ perm, uint32_t n,
                                                 const float4 *__restrict__ tri_in, const float4 *__restrict__ shd_in,
                                                 float4 *__restrict__ tri_out, float4 *__restrict__ shd_out) {
  const uint32_t s = blockIdx.x * 256u + threadIdx.x;
  if (s >= n) return;
  const uint32_t t = perm[s];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    tri_out[3 * size_t(s) + q] = tri_in[3 * size_t(t) + q];
    shd_out[3 * size_t(s) + q] = shd_in[3 * size_t(t) + q];
  }
}

__global__ void k_root(uint32_t n, uint32_t *node_ref, uint32_t *node_acc, uint32_t *stats) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    // a single primitive has no radix-tree node: the root refers to it
    node_ref[0] = n >= 2 ? 0u : kPrimRef;
    node_acc[0] = 0;
    stats[0] = 1;  // node 0 is the root
    stats[1] = 0;
    stats[2] = 0;
    stats[3] = 0;
  }
}

// ---- stable LSD radix sort of (key, value) pairs, 8 bits per pass ----------
// Per pass over a tile partition of kRsTile keys per block: k_rs_hist counts
// the tile's digits (LDS) into hist[digit * tiles + tile]; an exclusive
// scan over that digit-major array (k_scan_local / k_scan_top /
// k_scan_add) gives every (digit, tile) its first output index; k_rs_scatter
// re-reads the tile in rounds of 256 keys and writes each key to its
// digit's next slot in tile order: within a wave the rank among equal
// digits comes from eight ballots (the lanes whose digit matches bit for
// bit), across the block's four waves from per-wave counts in LDS.  Keys
// keep their input order within a digit, so the sort is stable and its
// output is the one stable order of (key, index) pairs.
constexpr uint32_t kRsTile = 2048;  // keys per block (8 rounds of 256)
constexpr uint32_t kScanBlock = 1024;  // elements per block of the scan (256 threads x 4)

__device__ __forceinline__ uint32_t rs_digit(uint32_t key, uint32_t shift) { return (key >> shift) & 0xFFu; }

__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t *__restrict__ keys, uint32_t n, uint32_t shift,
                                                 uint32_t tiles, uint32_t *__restrict__ hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kRsTile;
  for (uint32_t i = base + threadIdx.x; i < min(base + kRsTile, n); i += 256u)
    atomicAdd(&cnt[rs_digit(keys[i], shift)], 1u);
  __syncthreads();
  hist[size_t(threadIdx.x) * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// exclusive scan, level 1: each block scans kScanBlock elements in place
// and leaves their total in sums[block]
__global__ __launch_bounds__(256) void k_scan_local(uint32_t *__restrict__ a, uint32_t n, uint32_t *__restrict__ sums) {
  __shared__ uint32_t part[256];
  const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4u;
  uint32_t v[4], t = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v[k] = base + k < n ? a[base + k] : 0u;
    t += v[k];
  }
  part[threadIdx.x] = t;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off *= 2) {  // inclusive Hillis-Steele over the 256 thread totals
    const uint32_t x = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - t;  // exclusive prefix of this thread's four
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < n) a[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) sums[blockIdx.x] = part[255];
}
// level 2: one block turns the block totals into exclusive offsets
__global__ __launch_bounds__(256) void k_scan_top(uint32_t *__restrict__ sums, uint32_t m) {
  __shared__ uint32_t part[256];
  uint32_t carry = 0;
  for (uint32_t c = 0; c < m; c += 256) {
    const uint32_t i = c + threadIdx.x;
    const uint32_t v = i < m ? sums[i] : 0u;
    part[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off *= 2) {
      const uint32_t x = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += x;
      __syncthreads();
    }
    if (i < m) sums[i] = carry + part[threadIdx.x] - v;
    const uint32_t tot = part[255];
    __syncthreads();
    carry += tot;
  }
}
// level 3: add each block's offset
__global__ __launch_bounds__(256) void k_scan_add(uint32_t *__restrict__ a, uint32_t n,
                                                  const uint32_t *__restrict__ sums) {
  const uint32_t add = sums[blockIdx.x];
  const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4u;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (base + k < n) a[base + k] += add;
}

__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                    uint32_t n, uint32_t shift, uint32_t tiles,
                                                    const uint32_t *__restrict__ offs, uint32_t *__restrict__ kout,
                                                    uint32_t *__restrict__ vout) {
  __shared__ uint32_t next[256];     // next output index of each digit for this tile
  __shared__ uint32_t wcnt[4][256];  // this round's digit counts per wave
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  next[threadIdx.x] = offs[size_t(threadIdx.x) * tiles + blockIdx.x];
  const uint32_t base = blockIdx.x * kRsTile;
  for (uint32_t r = 0; r < kRsTile; r += 256) {
    const uint32_t i = base + r + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
    const uint32_t d = rs_digit(key, shift);
#pragma unroll
    for (int k = 0; k < 4; k++) wcnt[wave][lane * 4 + k] = 0;
    __syncthreads();
    // the lanes of this wave with the same digit: match all eight bits
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t lrank = __popcll(peers & below);
    if (valid && lrank == 0) wcnt[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = next[d] + lrank;
      for (uint32_t w = 0; w < wave; w++) pos += wcnt[w][d];
      if (pos < n) {  // always, when the counts are consistent
        kout[pos] = key;
        vout[pos] = val;
      }
    }
    __syncthreads();
    next[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
    __syncthreads();  // the counts are read before the next round clears them
  }
}

}  // namespace

// scratch of radix_sort_pairs for n keys
size_t radix_sort_scratch(uint32_t n) {
  const size_t tiles = (size_t(n) + kRsTile - 1) / kRsTile;
  const size_t hn = 256 * tiles;
  return (hn * 4 + 255) / 256 * 256 + ((hn + kScanBlock - 1) / kScanBlock * 4 + 255) / 256 * 256 +
         2 * ((size_t(n) * 4 + 255) / 256 * 256);
}
// Sorts (keys, vals) by the low `bits` bits of the keys, stable; the result
// is in (kout, vout); kin and vin are left unchanged.
hipError_t radix_sort_pairs(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, uint32_t n,
                            uint32_t bits, void *scratch, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const uint32_t tiles = (n + kRsTile - 1) / kRsTile;
  const uint32_t hn = 256u * tiles;
  const uint32_t sb = (hn + kScanBlock - 1) / kScanBlock;
  char *p = static_cast<char *>(scratch);
  auto take = [&](size_t bytes) {
    void *r = p;
    p += (bytes + 255) / 256 * 256;
    return static_cast<uint32_t *>(r);
  };
  uint32_t *hist = take(size_t(hn) * 4), *sums = take(size_t(sb) * 4);
  uint32_t *tk = take(size_t(n) * 4), *tv = take(size_t(n) * 4);
  const uint32_t passes = (bits + 7) / 8;
  // ping-pong so that the last pass writes (kout, vout)
  const uint32_t *ik = kin, *iv = vin;
  for (uint32_t ps = 0; ps < passes; ps++) {
    const bool to_out = ((passes - 1 - ps) % 2) == 0;
    uint32_t *ok = to_out ? kout : tk, *ov = to_out ? vout : tv;
    hipLaunchKernelGGL(k_rs_hist, dim3(tiles), dim3(256), 0, st, ik, n, 8u * ps, tiles, hist);
    hipLaunchKernelGGL(k_scan_local, dim3(sb), dim3(256), 0, st, hist, hn, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, sums, sb);
    hipLaunchKernelGGL(k_scan_add, dim3(sb), dim3(256), 0, st, hist, hn, sums);
    hipLaunchKernelGGL(k_rs_scatter, dim3(tiles), dim3(256), 0, st, ik, iv, n, 8u * ps, tiles, hist, ok, ov);
    ik = ok;
    iv = ov;
  }
  return hipGetLastError();
}

size_t gpu_bvh_scratch_bytes(uint32_t n) {
  const size_t sort_bytes = radix_sort_scratch(n);
  const size_t m = n > 1 ? n - 1 : 1;
  return (size_t(n) * 32)           // pbox
         + size_t(n) * 16           // keys, vals (in and out)
         + m * 8 * 2                // child, range
         + m * 4 + size_t(n) * 4    // parents
         + m * 32                   // internal boxes
         + m * 4                    // arrival counters
         + (2 * m + 1) * 4          // crossing nodes for the top of the bottom-up pass
         + size_t(n) * 8 + 8        // node_ref, node_acc (one per BVH8 node, at most n)
         + kRedBlocks * 24 + 64     // reduction partials, bounds, stats
         + sort_bytes + 32 * 256;   // sort scratch, alignment of each piece
}

hipError_t gpu_build_bvh8(const float4 *tri, const float4 *shd, uint32_t n, void *scratch, Node8d *nodes,
                          float4 *tri_out, float4 *shd_out, GpuBvhStats *st, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  char *p = static_cast<char *>(scratch);
  auto take = [&](size_t bytes) {
    void *r = p;
    p += (bytes + 255) / 256 * 256;
    return r;
  };
  const size_t m = n > 1 ? n - 1 : 1;
  float4 *pbox = static_cast<float4 *>(take(size_t(n) * 32));
  uint32_t *keys = static_cast<uint32_t *>(take(size_t(n) * 4));
  uint32_t *vals = static_cast<uint32_t *>(take(size_t(n) * 4));
  uint32_t *skeys = static_cast<uint32_t *>(take(size_t(n) * 4));
  uint32_t *perm = static_cast<uint32_t *>(take(size_t(n) * 4));
  uint2 *child = static_cast<uint2 *>(take(m * 8));
  uint2 *range = static_cast<uint2 *>(take(m * 8));
  uint32_t *parent_in = static_cast<uint32_t *>(take(m * 4));
  uint32_t *parent_leaf = static_cast<uint32_t *>(take(size_t(n) * 4));
  float4 *ibox = static_cast<float4 *>(take(m * 32));
  uint32_t *arrive = static_cast<uint32_t *>(take(m * 4));
  // at most two entries per crossing node (one per child), then the count
  uint32_t *top = static_cast<uint32_t *>(take((2 * m + 1) * 4));
  uint32_t *top_count = top + 2 * m;
  uint32_t *node_ref = static_cast<uint32_t *>(take(size_t(n) * 4 + 4));
  uint32_t *node_acc = static_cast<uint32_t *>(take(size_t(n) * 4 + 4));
  uint32_t *partial = static_cast<uint32_t *>(take(kRedBlocks * 24));
  float *bounds = static_cast<float *>(take(32));
  uint32_t *stats = static_cast<uint32_t *>(take(32));
  void *sort_tmp = take(radix_sort_scratch(n));
  hipError_t e = hipSuccess;

  const uint32_t b256 = (n + 255) / 256;
  const uint32_t nb = std::min(kRedBlocks, b256);
  hipLaunchKernelGGL(k_prim_boxes, dim3(nb), dim3(256), 0, stream, tri, n, pbox, partial);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, stream, partial, nb, bounds);
  hipLaunchKernelGGL(k_morton, dim3(b256), dim3(256), 0, stream, pbox, n, bounds, keys, vals);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // Morton codes in leaf order, the triangle of each leaf beside
  e = radix_sort_pairs(keys, vals, skeys, perm, n, 30, sort_tmp, stream);
  if (e != hipSuccess) return e;
  if (n >= 2) {
    if ((e = launch_zero(arrive, m * 4, stream)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(top_count, 0, 4, stream)) != hipSuccess) return e;
    // a parent the radix tree does not set ends the upward walk
    if ((e = hipMemsetAsync(parent_in, 0xFF, m * 4, stream)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(parent_leaf, 0xFF, size_t(n) * 4, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_karras, dim3(uint32_t((m + 255) / 256)), dim3(256), 0, stream, skeys, int(n), child, range,
                       parent_in, parent_leaf);
    hipLaunchKernelGGL(k_bottom_up_local, dim3((n + kBuBlock - 1) / kBuBlock), dim3(kBuBlock), 0, stream, child,
                       range, parent_in, parent_leaf, perm, pbox, n, ibox, arrive, top, top_count, uint32_t(2 * m));
    hipLaunchKernelGGL(k_bottom_up_top, dim3(1), dim3(kBuBlock), 0, stream, child, parent_in, perm, pbox, ibox, arrive,
                       top, top_count, uint32_t(2 * m));
  }
  hipLaunchKernelGGL(k_root, dim3(1), dim3(64), 0, stream, n, node_ref, node_acc, stats);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // level by level: nodes [a, b) make the next level [b, stats[0])
  // (node records and node_ref hold n entries; more would be a builder fault)
  uint32_t a = 0, b = 1, levels = 0;
  while (a < b) {
    hipLaunchKernelGGL(k_collapse, dim3((b - a + 255) / 256), dim3(256), 0, stream, child, range, pbox, perm, ibox,
                       a, b, n, node_ref, node_acc, nodes, stats);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t next[4] = {0, 0, 0, 0};
    if ((e = hipMemcpyAsync(next, stats, 16, hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    if (next[3] || next[0] > n || ++levels > 1000) return hipErrorUnknown;
    a = b;
    b = next[0];
  }
  hipLaunchKernelGGL(k_permute, dim3(b256), dim3(256), 0, stream, perm, n, tri, shd, tri_out, shd_out);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  uint32_t s3[3] = {0, 0, 0};
  if ((e = hipMemcpyAsync(s3, stats, 12, hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  st->n_nodes = s3[0];
  st->stack_bound = s3[1];
  st->n_leaves = s3[2];
  st->depth = levels;
  return hipSuccess;
}

}  // namespace lam

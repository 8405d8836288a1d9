// BVH8 over world-space triangles, laid out for the gfx950 extend kernel.
//
// Replaces bvh<>::build / convertbuildnodes (lib/numli/cg.hpp:632-749),
// which builds one BVH per mesh with 8 equal-width centroid bins and stores
// child boxes as six __m256 (192 B).  Here every instance is flattened into
// one BVH: a binned-SAH binary tree collapsed to 8-wide nodes whose child
// boxes are quantized to 8 bits against a per-node origin and power-of-two
// scale, one 128-byte (cache-line) record per node.  Quantized boxes are
// verified conservative with the decode arithmetic the kernel uses.
#pragma once
#include <cstdint>
#include <vector>

namespace lamh {

// 128-byte node record (see lam_device.h for the device view)
struct alignas(128) Node8 {
  float origin[3];
  uint8_t ex, ey, ez;  // scale = 2^(e-127) per axis
  uint8_t nchild;
  uint32_t child[8];   // interior: node index; leaf: 0x80000000|(cnt-1)<<27|first
  uint8_t qlo[3][8];   // [axis][child]
  uint8_t qhi[3][8];
  uint8_t pad[32];
};
static_assert(sizeof(Node8) == 128, "Node8 must be one 128-byte line");

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafCountShift = 27;
constexpr uint32_t kLeafFirstMask = (1u << 27) - 1;
constexpr int kMaxLeafTris = 4;

struct Bvh8 {
  std::vector<Node8> nodes;
  std::vector<uint32_t> order;  // leaf order -> input triangle index
  uint32_t depth = 0, n_leaves = 0, stack_bound = 0;
  double sah_cost = 0.0;
};

// tri_v: 9 floats per triangle.  Returns false if the triangle count
// exceeds the 27-bit leaf index range.
bool build_bvh8(const float *tri_v, uint32_t n_tris, Bvh8 &out);

// Decode a node's child box exactly as the kernel does (for tests).
void decode_child_box(const Node8 &n, int k, float lo[3], float hi[3]);

}  // namespace lamh

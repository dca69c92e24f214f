// Host-side BVH builder of the oxr path tracer: binned SAH (Wald 2007)
// over the prepared triangle soup, collapsed to 4-wide nodes laid out
// for coalesced 16-byte fetches on the device.
//
// Replaces the BVH of the reference's trace layer (oxideav-render
// src/trace.rs:320-351, which builds oxideav_mesh3d::Bvh with binned SAH).
#pragma once

#include <stdint.h>

#include <vector>

namespace oxr {

// 128-byte wide node, the builder's output (the device traverses its
// quantised copy, Node4Q below): child boxes as six float4 (lo/hi per
// axis, one lane per child), then child references and leaf triangle
// counts.
// count[i] == 0: inner child, child[i] is a node index;
// count[i]  > 0: leaf of count[i] triangles starting at tris[child[i]];
// child[i] == EMPTY: unused slot (its box is empty: lo = +inf, hi = -inf).
struct alignas(16) Node4 {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  uint32_t child[4];
  uint32_t count[4];
};
static_assert(sizeof(Node4) == 128, "Node4 is 128 bytes");

// 48-byte triangle record in leaf order: v0, global id; v1, flags; v2,
// alpha (BLEND coverage).
struct alignas(16) TriRec {
  float v0[3];
  uint32_t gid;
  float v1[3];
  uint32_t flags;
  float v2[3];
  float alpha;
};
static_assert(sizeof(TriRec) == 48, "TriRec is 48 bytes");

// TriRec.flags bits (acceptance rules of Integrator::accept,
// src/pathtrace.rs:1924-1934, resolved per material).
constexpr uint32_t TRI_DOUBLE_SIDED = 1u;
constexpr uint32_t TRI_MASK_REJECT = 2u;  // MASK with alpha < cutoff
constexpr uint32_t TRI_BLEND = 4u;
constexpr uint32_t TRI_ALPHA_TEX = 8u;  // alpha varies over the triangle (texture / vertex colours)
constexpr uint32_t TRI_MASK_TEX = 16u;  // MASK with varying alpha: the slot holds the cutoff

constexpr uint32_t NODE_EMPTY = 0xffffffffu;

struct Bvh4 {
  std::vector<Node4> nodes;
  std::vector<TriRec> tris;
  uint32_t num_leaves = 0;
  uint32_t max_depth = 0;  // levels of 4-wide nodes (root = 1)
  double sah_cost = 0.0;
  double build_ms = 0.0;
  float pad = 0.0f;  // absolute padding applied to every box
};

// pos: n x 9 floats (v0, v1, v2). flags/alpha: per global triangle.
// max_leaf: largest leaf (1..8). threads: 0 = hardware concurrency.
// fast: a linear BVH (Morton order, radix splits) instead of binned SAH:
// about 20x faster to build, somewhat slower to trace; for scenes built
// per request.
Bvh4 build_bvh4(const float* pos, uint32_t n, const uint32_t* flags, const float* alpha, int max_leaf, int threads,
                bool fast = false);

// 64-byte wide node, the device's traversal format: the child boxes of a
// Node4 quantised to 8 bits per plane (after Ylitie, Karras and Laine,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs", HPG 2017). Planes are steps of 2^e[k] on axis k from the corner
// org, rounded outwards and widened by one more step, so each decoded box
// holds the exact box with a margin of a step; the slab test then reads 64
// bytes a node instead of 128. Four 16-byte quads:
//   q0: org.xyz (f32), then e.x, e.y, e.z (int8) and a zero byte;
//   q1: lo.x, hi.x, lo.y, hi.y: one byte per child (child c in byte c);
//   q2: lo.z, hi.z, count (one byte per child), zero;
//   q3: child[4] as in Node4 (NODE_EMPTY: unused slot).
struct alignas(64) Node4Q {
  float org[3];
  uint32_t exps;
  uint32_t qlo_x, qhi_x, qlo_y, qhi_y;
  uint32_t qlo_z, qhi_z, counts, pad;
  uint32_t child[4];
};
static_assert(sizeof(Node4Q) == 64, "Node4Q is 64 bytes");

// The quantised copy of a Node4 array (same indices, same children).
std::vector<Node4Q> quantize_nodes(const std::vector<Node4>& nodes);

}  // namespace oxr

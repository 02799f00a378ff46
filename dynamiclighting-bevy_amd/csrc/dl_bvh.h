// Occluder BVH in device layout: wide nodes with 8-bit quantized child boxes (CNode), and
// triangles in leaf order as three float4 each (v0 | occluder index, e1 = v1 - v0, e2 = v2 - v0).
//
// Boxes are grown by a relative 2^-15 of their coordinate magnitude, so the slab test, itself
// widened by a relative 2^-16, never culls a triangle that Moller-Trumbore reports as hit: any-hit
// results equal a scan over all occluders and closest-hit results equal the lowest (t, index)
// over all occluders (occluder.rs:180-223 with ties resolved as at occluder.rs:205-209).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <memory>
#include <utility>
#include <vector>

namespace dl {

// Children per node (4 or 8, fixed at build time by DL_BVH_WIDTH).
#ifndef DL_BVH_WIDTH
#define DL_BVH_WIDTH 4
#endif
constexpr int kBvhWidth = DL_BVH_WIDTH;
static_assert(kBvhWidth == 4 || kBvhWidth == 8, "BVH nodes are 4 or 8 wide");

// A collapsed wide node with float child boxes (builder intermediate, quantized to CNode).
struct NodeW {
  float lox[kBvhWidth], hix[kBvhWidth], loy[kBvhWidth], hiy[kBvhWidth], loz[kBvhWidth], hiz[kBvhWidth];
  int32_t child[kBvhWidth];  // > 0 inner node; < 0 leaf: -1 - ((first << 2) | (count - 1)); 0 unused slot
  int32_t count;             // children in use (1..kBvhWidth); kernels read the child slots instead
};

// The node the kernels read: 64 bytes when 4-wide, 128 when 8-wide.  Child bounds are 8-bit steps
// from the node's origin (its padded minimum corner) with a power-of-two step per axis, rounded
// outward by one extra step, so each decoded box (origin + q * step, one float rounding) contains
// the padded box of the child it stands for.  q holds lo x, hi x, lo y, hi y, lo z, hi z in that
// order, kBvhWidth / 4 words each, child k in byte k % 4 of word k / 4; the child slots start on a
// 16-byte boundary.
constexpr int kNodeQWords = 6 * kBvhWidth / 4;
constexpr int kNodeQPad = (4 - (4 + kNodeQWords) % 4) % 4;
struct alignas(64) CNode {
  float ox, oy, oz;
  uint32_t exps;  // biased float exponents of the x, y, z steps in bits 0-7, 8-15, 16-23
  uint32_t q[kNodeQWords + kNodeQPad];
  int32_t child[kBvhWidth];  // as NodeW::child; 0 marks an unused slot
};
static_assert(sizeof(CNode) == 16 * kBvhWidth, "CNode is 64 bytes when 4-wide, 128 when 8-wide");
static_assert(offsetof(CNode, child) % 16 == 0, "child slots are 16-byte aligned");

// An allocator that leaves value-initialized elements uninitialized: for large buffers that are
// written in full right after they are sized.
template <class T>
struct NoInit : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <class U>
  NoInit(const NoInit<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... Args>
  void construct(U* p, Args&&... args) {
    ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};

struct Bvh {
  std::vector<CNode> nodes;
  std::vector<float, NoInit<float>> tris;      // 12 floats per reference, leaf order
  std::vector<uint32_t, NoInit<uint32_t>> order;  // leaf position -> occluder index (a split triangle appears more than once)
  uint32_t depth = 0;           // node levels on the longest root-to-leaf path
};

// Traversal stack entries a BVH of `depth` node levels needs: a node visit pushes its hit inner
// children except the nearest (at most kBvhWidth - 1) on top of what its ancestors pushed.
inline uint32_t bvh_stack_entries(uint32_t depth) { return (uint32_t)(kBvhWidth - 1) * (depth > 1u ? depth : 1u) + 1u; }

// Builds over `n` triangles given as 9 floats each (v0, v1, v2).
void build_bvh(const float* tris9, uint32_t n, unsigned threads, Bvh& out);

}  // namespace dl

// Occluder BVH in device layout: 4-wide nodes of 128 bytes (one cache line pair), child boxes in
// SoA order so one node is eight float4 loads, and triangles in leaf order as three float4 each
// (v0 | occluder index, e1 = v1 - v0, e2 = v2 - v0).
//
// Boxes are grown by a relative 2^-15 of their coordinate magnitude, so the slab test, itself
// widened by a relative 2^-16, never culls a triangle that Moller-Trumbore reports as hit: any-hit
// results equal a scan over all occluders and closest-hit results equal the lowest (t, index)
// over all occluders (occluder.rs:180-223 with ties resolved as at occluder.rs:205-209).
#pragma once

#include <stdint.h>

#include <memory>
#include <utility>
#include <vector>

namespace dl {

// A collapsed 4-wide node with float child boxes (builder intermediate, quantized to CNode).
struct alignas(16) Node4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];  // > 0 inner node; < 0 leaf: -1 - ((first << 2) | (count - 1)); 0 unused slot
  int32_t count;     // children in use (1..4); kernels read the child slots instead
  int32_t pad[3];
};
static_assert(sizeof(Node4) == 128, "Node4 is 128 bytes");

// The 4-wide node the kernels read: 64 bytes.  Child bounds are 8-bit steps from the node's
// origin (its padded minimum corner) with a power-of-two step per axis, rounded outward by one
// extra step, so each decoded box (origin + q * step, one float rounding) contains the padded
// box of the child it stands for.
struct alignas(64) CNode {
  float ox, oy, oz;
  uint32_t exps;        // biased float exponents of the x, y, z steps in bits 0-7, 8-15, 16-23
  uint32_t lox, hix, loy, hiy;  // byte k: child k
  uint32_t loz, hiz, pad0, pad1;
  int32_t child[4];     // as Node4::child; 0 marks an unused slot
};
static_assert(sizeof(CNode) == 64, "CNode is 64 bytes");

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

struct Bvh4 {
  std::vector<CNode> nodes;
  std::vector<float, NoInit<float>> tris;      // 12 floats per reference, leaf order
  std::vector<uint32_t, NoInit<uint32_t>> order;  // leaf position -> occluder index (a split triangle appears more than once)
  uint32_t depth = 0;           // node levels on the longest root-to-leaf path
};

// Builds over `n` triangles given as 9 floats each (v0, v1, v2).
void build_bvh4(const float* tris9, uint32_t n, unsigned threads, Bvh4& out);

}  // namespace dl

// Host build of the occluder BVH: binned SAH over centroids into a binary tree with leaves of at
// most four triangles, then collapsed to 4-wide nodes by opening the largest children first.
// Subtrees below a size threshold are built on worker threads.
#include "dl_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

namespace dl {
namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
  }
  float area() const {
    float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
    if (e0 < 0 || e1 < 0 || e2 < 0) return 0.0f;
    return e0 * e1 + e1 * e2 + e2 * e0;
  }
};

struct Node2 {
  Box box;
  int32_t left = -1, right = -1;  // children (inner); -1 for leaves
  uint32_t first = 0, count = 0;  // leaf range in `idx`
};

struct Builder {
  const float* tri;
  std::vector<Box> tb;
  std::vector<float> cen;
  std::vector<uint32_t> idx;
  std::vector<Node2> nodes;
  std::mutex mu;

  uint32_t alloc2() {
    std::lock_guard<std::mutex> g(mu);
    nodes.emplace_back();
    nodes.emplace_back();
    return (uint32_t)nodes.size() - 2;
  }

  // Splits [first, first+count) of idx; returns the split point or `first` for a leaf.
  uint32_t split(uint32_t first, uint32_t count, Box& box) {
    Box cb;
    box.reset();
    cb.reset();
    for (uint32_t k = first; k < first + count; ++k) {
      uint32_t p = idx[k];
      box.grow(tb[p]);
      for (int a = 0; a < 3; ++a) {
        cb.lo[a] = std::min(cb.lo[a], cen[p * 3 + a]);
        cb.hi[a] = std::max(cb.hi[a], cen[p * 3 + a]);
      }
    }
    if (count <= 4) return first;
    const int B = 32;
    float best = INFINITY;
    int best_axis = -1, best_bin = -1;
    for (int axis = 0; axis < 3; ++axis) {
      float ext = cb.hi[axis] - cb.lo[axis];
      if (!(ext > 0.0f)) continue;
      float scale = (float)B / ext;
      Box bins[B];
      uint32_t cnt[B] = {};
      for (int b = 0; b < B; ++b) bins[b].reset();
      for (uint32_t k = first; k < first + count; ++k) {
        uint32_t p = idx[k];
        int b = std::min(B - 1, std::max(0, (int)((cen[p * 3 + axis] - cb.lo[axis]) * scale)));
        cnt[b]++;
        bins[b].grow(tb[p]);
      }
      float right_area[B];
      uint32_t right_cnt[B];
      Box acc;
      acc.reset();
      uint32_t c = 0;
      for (int b = B - 1; b > 0; --b) {
        acc.grow(bins[b]);
        c += cnt[b];
        right_area[b] = acc.area();
        right_cnt[b] = c;
      }
      acc.reset();
      c = 0;
      for (int b = 0; b < B - 1; ++b) {
        acc.grow(bins[b]);
        c += cnt[b];
        if (c == 0 || right_cnt[b + 1] == 0) continue;
        float cost = (float)c * acc.area() + (float)right_cnt[b + 1] * right_area[b + 1];
        if (cost < best) { best = cost; best_axis = axis; best_bin = b + 1; }
      }
    }
    uint32_t mid;
    if (best_axis < 0) {
      mid = first + count / 2;  // coincident centroids: split the list
    } else {
      float ext = cb.hi[best_axis] - cb.lo[best_axis];
      float scale = (float)B / ext;
      int axis = best_axis, bin = best_bin;
      auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](uint32_t p) {
        int b = std::min(B - 1, std::max(0, (int)((cen[p * 3 + axis] - cb.lo[axis]) * scale)));
        return b < bin;
      });
      mid = (uint32_t)(it - idx.begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    return mid;
  }

  void build_subtree(uint32_t node, uint32_t first, uint32_t count, std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>>* spill,
                     uint32_t spill_below) {
    struct Item { uint32_t node, first, count; };
    std::vector<Item> work{{node, first, count}};
    while (!work.empty()) {
      Item it = work.back();
      work.pop_back();
      if (spill && it.count < spill_below && it.count > 4) {
        spill->push_back({it.node, {it.first, it.count}});
        continue;
      }
      Box box;
      uint32_t mid = split(it.first, it.count, box);
      if (mid == it.first) {
        std::lock_guard<std::mutex> g(mu);
        Node2& n = nodes[it.node];
        n.box = box;
        n.first = it.first;
        n.count = it.count;
        continue;
      }
      uint32_t l = alloc2();
      {
        std::lock_guard<std::mutex> g(mu);
        Node2& n = nodes[it.node];
        n.box = box;
        n.left = (int32_t)l;
        n.right = (int32_t)l + 1;
      }
      work.push_back({l, it.first, mid - it.first});
      work.push_back({l + 1, mid, it.first + it.count - mid});
    }
  }
};

void pad(const Box& b, float* lo, float* hi) {
  for (int a = 0; a < 3; ++a) {
    float p = (std::fabs(b.lo[a]) + std::fabs(b.hi[a]) + (b.hi[a] - b.lo[a])) * (1.0f / 32768.0f) + 1e-30f;
    lo[a] = b.lo[a] - p;
    hi[a] = b.hi[a] + p;
  }
}

}  // namespace

void build_bvh4(const float* tris9, uint32_t n, unsigned threads, Bvh4& out) {
  out.nodes.clear();
  out.tris.clear();
  out.order.clear();
  out.depth = 0;
  if (n == 0) return;
  Builder b;
  b.tri = tris9;
  b.tb.resize(n);
  b.cen.resize((size_t)n * 3);
  b.idx.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* t = tris9 + (size_t)i * 9;
    for (int a = 0; a < 3; ++a) {
      float v0 = t[a], v1 = t[3 + a], v2 = t[6 + a];
      b.tb[i].lo[a] = std::min(v0, std::min(v1, v2));
      b.tb[i].hi[a] = std::max(v0, std::max(v1, v2));
      b.cen[(size_t)i * 3 + a] = 0.5f * (b.tb[i].lo[a] + b.tb[i].hi[a]);
    }
    b.idx[i] = i;
  }
  b.nodes.reserve((size_t)n * 2 / 3 + 16);
  b.nodes.emplace_back();
  // Top of the tree on this thread until subtrees are small, then the rest in parallel.
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  uint32_t spill_below = threads > 1 ? std::max<uint32_t>(4096, n / (threads * 8)) : 0;
  std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>> spill;
  b.build_subtree(0, 0, n, threads > 1 ? &spill : nullptr, spill_below);
  if (!spill.empty()) {
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (;;) {
        size_t i = next.fetch_add(1);
        if (i >= spill.size()) break;
        b.build_subtree(spill[i].first, spill[i].second.first, spill[i].second.second, nullptr, 0);
      }
    };
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < threads; ++t) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
  }
  // Leaf-ordered triangles.
  out.order = b.idx;
  out.tris.resize((size_t)n * 12);
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t p = b.idx[k];
    const float* t = tris9 + (size_t)p * 9;
    float* o = out.tris.data() + (size_t)k * 12;
    o[0] = t[0]; o[1] = t[1]; o[2] = t[2];
    std::memcpy(&o[3], &p, 4);
    o[4] = t[3] - t[0]; o[5] = t[4] - t[1]; o[6] = t[5] - t[2]; o[7] = 0.0f;
    o[8] = t[6] - t[0]; o[9] = t[7] - t[1]; o[10] = t[8] - t[2]; o[11] = 0.0f;
  }
  // Collapse to 4-wide nodes.
  const std::vector<Node2>& n2 = b.nodes;
  struct Pending { uint32_t n2, n4, depth; };
  out.nodes.emplace_back();
  std::vector<Pending> work{{0, 0, 1}};
  while (!work.empty()) {
    Pending p = work.back();
    work.pop_back();
    out.depth = std::max(out.depth, p.depth);
    uint32_t kids[4];
    int nk = 0;
    if (n2[p.n2].left < 0) {
      kids[nk++] = p.n2;  // whole tree is one leaf
    } else {
      kids[nk++] = (uint32_t)n2[p.n2].left;
      kids[nk++] = (uint32_t)n2[p.n2].right;
      while (nk < 4) {
        int open = -1;
        float best = -1.0f;
        for (int k = 0; k < nk; ++k)
          if (n2[kids[k]].left >= 0 && n2[kids[k]].box.area() > best) { best = n2[kids[k]].box.area(); open = k; }
        if (open < 0) break;
        uint32_t o = kids[open];
        kids[open] = (uint32_t)n2[o].left;
        kids[nk++] = (uint32_t)n2[o].right;
      }
    }
    Node4 nd;
    std::memset(&nd, 0, sizeof nd);
    nd.count = nk;
    for (int k = 0; k < 4; ++k) {
      if (k >= nk) {
        nd.lox[k] = nd.loy[k] = nd.loz[k] = 0.0f;
        nd.hix[k] = nd.hiy[k] = nd.hiz[k] = 0.0f;
        nd.child[k] = 0;
        continue;
      }
      float lo[3], hi[3];
      pad(n2[kids[k]].box, lo, hi);
      nd.lox[k] = lo[0]; nd.loy[k] = lo[1]; nd.loz[k] = lo[2];
      nd.hix[k] = hi[0]; nd.hiy[k] = hi[1]; nd.hiz[k] = hi[2];
      if (n2[kids[k]].left < 0) {
        uint32_t first = n2[kids[k]].first, count = n2[kids[k]].count;
        nd.child[k] = -1 - (int32_t)((first << 2) | (count - 1));
      } else {
        uint32_t idx = (uint32_t)out.nodes.size();
        out.nodes.emplace_back();
        nd.child[k] = (int32_t)idx;
        work.push_back({kids[k], idx, p.depth + 1});
      }
    }
    out.nodes[p.n4] = nd;
  }
}

}  // namespace dl

// Host build of the occluder BVH: large triangles are first split into several references, each
// bounded by the part of the triangle inside half of its parent reference's box (clipped in
// double precision, bounds rounded outward); then binned SAH over reference centroids into a
// binary tree with leaves of at most four references, collapsed to kBvhWidth-wide nodes by
// opening the largest children first.  Reference cutting, subtrees above a size threshold, leaf records and
// quantization run on worker threads.  A triangle split into several references appears in
// several leaves; any-hit and closest-hit (lowest t, then lowest occluder index) results do not
// depend on that.
#include "dl_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <queue>
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

// One reference: the bounds of a triangle or of a part of it, their centre and the triangle.
struct Ref {
  Box box;
  float c[3];
  uint32_t tri;
};

// Binary build node.  Plain data: the node array is allocated once, uninitialized, for the
// largest possible tree (2 references - 1 nodes) and handed out by an atomic counter, so threads
// building different subtrees never wait for one another.
struct Node2 {
  Box box;
  int32_t left, right;    // children (inner); -1 for leaves
  uint32_t first, count;  // leaf range in the reference array
};

// fn(i) for i in [0, n), in contiguous blocks on up to `threads` threads.
template <class F>
void parallel_for(uint64_t n, unsigned threads, const F& fn) {
  const unsigned th = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(threads, (n + 8191) / 8192));
  if (th <= 1) {
    for (uint64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  const uint64_t per = (n + th - 1) / th;
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < th; ++t)
    ts.emplace_back([&fn, t, per, n]() {
      for (uint64_t i = t * per, e = std::min(n, (t + 1) * per); i < e; ++i) fn(i);
    });
  for (uint64_t i = 0; i < std::min(n, per); ++i) fn(i);
  for (auto& t : ts) t.join();
}

// fn(k) for blocks k in [0, blocks), handed out one at a time to up to `threads` threads (for
// blocks of uneven cost).
template <class F>
void parallel_blocks(uint64_t blocks, unsigned threads, const F& fn) {
  const unsigned th = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(threads, blocks));
  std::atomic<uint64_t> next{0};
  auto work = [&]() {
    for (uint64_t k; (k = next.fetch_add(1, std::memory_order_relaxed)) < blocks;) fn(k);
  };
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < th; ++t) ts.emplace_back(work);
  work();
  for (auto& t : ts) t.join();
}

struct Builder {
  Ref* refs = nullptr;
  Node2* nodes = nullptr;
  std::atomic<uint32_t> used{1};  // node 0 is the root

  // Work shared between threads: subtrees of at least `share` references.
  struct Item {
    uint32_t node, first, count;
  };
  std::mutex mu;
  std::condition_variable cv;
  std::vector<Item> shared;
  uint32_t pending = 0;  // shared items not yet finished (under mu)
  uint32_t share = 0;

  // Bounds of [first, first + count) in `box`; returns the split point of a binned SAH split
  // over reference centres (32 bins per axis), or `first` for a leaf of at most four.  The
  // references are partitioned in place.
  static constexpr int B = 32;
  struct Bins {
    Box box[3][B];
    uint32_t cnt[3][B];
  };
  unsigned threads = 1;
  static constexpr uint32_t kWide = 1u << 17;  // ranges split by several threads at once

  static void bounds(const Ref* r, uint32_t count, Box& box_out, Box& cb_out) {
    Box box, cb;  // locals: the caller's boxes may share cache lines with other threads'
    box.reset();
    cb.reset();
    for (uint32_t k = 0; k < count; ++k) {
      box.grow(r[k].box);
      for (int a = 0; a < 3; ++a) {
        cb.lo[a] = std::min(cb.lo[a], r[k].c[a]);
        cb.hi[a] = std::max(cb.hi[a], r[k].c[a]);
      }
    }
    box_out = box;
    cb_out = cb;
  }
  static void bin(const Ref* r, uint32_t count, const Box& cb, const float* scale, Bins& out) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < B; ++b) {
        out.box[a][b].reset();
        out.cnt[a][b] = 0;
      }
    for (uint32_t k = 0; k < count; ++k)
      for (int a = 0; a < 3; ++a) {
        const int b = std::min(B - 1, std::max(0, (int)((r[k].c[a] - cb.lo[a]) * scale[a])));
        out.cnt[a][b]++;
        out.box[a][b].grow(r[k].box);
      }
  }
  // fn(part, begin, count) over `parts` slices of [0, count) on their own threads
  template <class F>
  static void slices(uint32_t count, unsigned parts, const F& fn) {
    const uint32_t per = (count + parts - 1) / parts;
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < parts; ++t) {
      const uint32_t a = std::min(count, t * per);
      ts.emplace_back([&fn, t, a, per, count]() { fn(t, a, std::min(per, count - a)); });
    }
    fn(0, 0, std::min(per, count));
    for (auto& t : ts) t.join();
  }

  uint32_t split(uint32_t first, uint32_t count, Box& box) {
    Ref* const r = refs + first;
    const unsigned parts = count >= kWide ? std::min<unsigned>(threads, count / (kWide / 2)) : 1u;
    Box cb;
    if (parts > 1) {
      std::vector<Box> pb(parts), pc(parts);
      slices(count, parts, [&](unsigned t, uint32_t a, uint32_t c) { bounds(r + a, c, pb[t], pc[t]); });
      box = pb[0];
      cb = pc[0];
      for (unsigned t = 1; t < parts; ++t) {
        box.grow(pb[t]);
        cb.grow(pc[t]);
      }
    } else {
      bounds(r, count, box, cb);
    }
    if (count <= 4) return first;
    float scale[3];
    for (int a = 0; a < 3; ++a) {
      const float ext = cb.hi[a] - cb.lo[a];
      scale[a] = ext > 0.0f ? (float)B / ext : 0.0f;
    }
    Bins bins;
    if (parts > 1) {
      std::vector<Bins> pbins(parts);
      slices(count, parts, [&](unsigned t, uint32_t a, uint32_t c) { bin(r + a, c, cb, scale, pbins[t]); });
      bins = pbins[0];
      for (unsigned t = 1; t < parts; ++t)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < B; ++b) {
            bins.box[a][b].grow(pbins[t].box[a][b]);
            bins.cnt[a][b] += pbins[t].cnt[a][b];
          }
    } else {
      bin(r, count, cb, scale, bins);
    }
    float best = INFINITY;
    int best_axis = -1, best_bin = -1;
    for (int axis = 0; axis < 3; ++axis) {
      if (!(cb.hi[axis] - cb.lo[axis] > 0.0f)) continue;
      float right_area[B];
      uint32_t right_cnt[B];
      Box acc;
      acc.reset();
      uint32_t c = 0;
      for (int b = B - 1; b > 0; --b) {
        acc.grow(bins.box[axis][b]);
        c += bins.cnt[axis][b];
        right_area[b] = acc.area();
        right_cnt[b] = c;
      }
      acc.reset();
      c = 0;
      for (int b = 0; b < B - 1; ++b) {
        acc.grow(bins.box[axis][b]);
        c += bins.cnt[axis][b];
        if (c == 0 || right_cnt[b + 1] == 0) continue;
        const float cost = (float)c * acc.area() + (float)right_cnt[b + 1] * right_area[b + 1];
        if (cost < best) {
          best = cost;
          best_axis = axis;
          best_bin = b + 1;
        }
      }
    }
    if (best_axis < 0) return first + count / 2;  // coincident centres: split the list
    const int axis = best_axis, bin = best_bin;
    const float lo = cb.lo[axis], sc = scale[axis];
    Ref* const mid = std::partition(r, r + count, [&](const Ref& x) {
      return std::min(B - 1, std::max(0, (int)((x.c[axis] - lo) * sc))) < bin;
    });
    const uint32_t m = first + (uint32_t)(mid - r);
    return m == first || m == first + count ? first + count / 2 : m;
  }

  // The subtree of `it`, depth first; children of at least `share` references go to the shared
  // list while the thread carries on with their sibling.
  void run(Item it) {
    std::vector<Item> local{it};
    while (!local.empty()) {
      const Item w = local.back();
      local.pop_back();
      Node2& n = nodes[w.node];
      const uint32_t mid = split(w.first, w.count, n.box);
      if (mid == w.first) {
        n.left = n.right = -1;
        n.first = w.first;
        n.count = w.count;
        continue;
      }
      const uint32_t l = used.fetch_add(2, std::memory_order_relaxed);
      n.left = (int32_t)l;
      n.right = (int32_t)l + 1;
      n.first = w.first;  // inner nodes keep their range too: the collapse uses its size
      n.count = w.count;
      const Item a{l, w.first, mid - w.first}, b{l + 1, mid, w.first + w.count - mid};
      if (share && a.count >= share) {
        {
          std::lock_guard<std::mutex> g(mu);
          shared.push_back(a);
          pending++;
        }
        cv.notify_one();
      } else {
        local.push_back(a);
      }
      local.push_back(b);
    }
  }

  void worker() {
    for (;;) {
      Item it;
      {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&]() { return !shared.empty() || pending == 0; });
        if (shared.empty()) return;
        it = shared.back();
        shared.pop_back();
      }
      run(it);
      std::lock_guard<std::mutex> g(mu);
      if (--pending == 0) cv.notify_all();
    }
  }
};

// Bounds of the part of triangle `t` (9 floats) inside the box [lo, hi], clipped in double;
// false when the part is empty.  The result is rounded outward to float.
bool clip_bounds(const float* t, const double lo[3], const double hi[3], Box* out) {
  double poly[12][3], tmp[12][3];
  int n = 3;
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) poly[k][a] = t[3 * k + a];
  for (int plane = 0; plane < 6 && n > 0; ++plane) {
    const int axis = plane >> 1;
    const bool upper = plane & 1;
    const double bound = upper ? hi[axis] : lo[axis];
    auto inside = [&](const double* v) { return upper ? v[axis] <= bound : v[axis] >= bound; };
    int m = 0;
    for (int k = 0; k < n; ++k) {
      const double* a = poly[k];
      const double* b = poly[(k + 1) % n];
      const bool ia = inside(a), ib = inside(b);
      if (ia) {
        for (int c = 0; c < 3; ++c) tmp[m][c] = a[c];
        m++;
      }
      if (ia != ib) {
        const double f = (bound - a[axis]) / (b[axis] - a[axis]);
        for (int c = 0; c < 3; ++c) tmp[m][c] = a[c] + (b[c] - a[c]) * f;
        tmp[m][axis] = bound;
        m++;
      }
    }
    n = m;
    for (int k = 0; k < n; ++k)
      for (int c = 0; c < 3; ++c) poly[k][c] = tmp[k][c];
  }
  if (n == 0) return false;
  for (int a = 0; a < 3; ++a) {
    double mn = poly[0][a], mx = poly[0][a];
    for (int k = 1; k < n; ++k) {
      mn = std::min(mn, poly[k][a]);
      mx = std::max(mx, poly[k][a]);
    }
    float fl = (float)mn, fh = (float)mx;
    if ((double)fl > mn) fl = std::nextafter(fl, -INFINITY);
    if ((double)fh < mx) fh = std::nextafter(fh, INFINITY);
    out->lo[a] = fl;
    out->hi[a] = fh;
  }
  return true;
}

// Cuts the part of triangle `t` inside `b` at the middle of the box's longest axis: the bounds
// of the triangle's part on either side.  False when one side holds none of it.
bool cut(const float* t, const Box& b, Box* left, Box* right) {
  int axis = 0;
  for (int a = 1; a < 3; ++a)
    if (b.hi[a] - b.lo[a] > b.hi[axis] - b.lo[axis]) axis = a;
  const double mid = 0.5 * ((double)b.lo[axis] + (double)b.hi[axis]);
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = b.lo[a];
    hi[a] = b.hi[a];
  }
  hi[axis] = mid;
  const bool has_l = clip_bounds(t, lo, hi, left);
  hi[axis] = b.hi[axis];
  lo[axis] = mid;
  const bool has_r = clip_bounds(t, lo, hi, right);
  return has_l && has_r;
}

// Splits references of triangles whose box has a surface area above `factor` times the mean
// triangle box area into parts: a reference above that area is cut at the middle of its box's
// longest axis into the bounds of the triangle's part on either side, and both halves are cut
// again while above it; a cut that leaves one side empty keeps the reference whole.  At most
// `budget` extra references: when the triangles would need more, the largest references are
// cut first until the budget is spent.  Per triangle, its references in cut order.
void split_refs_heap(const float* tris9, uint32_t n, double limit, uint64_t budget, std::vector<uint32_t>& ref_tri,
                     std::vector<Box>& ref_box) {
  using Item = std::pair<float, uint32_t>;  // area, reference
  std::priority_queue<Item> heap;
  for (uint32_t i = 0; i < n; ++i)
    if (ref_box[i].area() > limit) heap.push({ref_box[i].area(), i});
  while (!heap.empty() && budget > 0) {
    const uint32_t r = heap.top().second;
    heap.pop();
    Box left, right;
    if (!cut(tris9 + (size_t)ref_tri[r] * 9, ref_box[r], &left, &right)) continue;  // a sliver: keep it whole
    ref_box[r] = left;
    ref_tri.push_back(ref_tri[r]);
    ref_box.push_back(right);
    budget--;
    if (left.area() > limit) heap.push({left.area(), r});
    if (right.area() > limit) heap.push({right.area(), (uint32_t)ref_box.size() - 1});
  }
}

void make_refs(const float* tris9, uint32_t n, float factor, uint64_t budget, unsigned threads, std::vector<Ref>& refs) {
  std::vector<Box> tb(n);
  parallel_for(n, threads, [&](uint64_t i) {
    const float* t = tris9 + i * 9;
    for (int a = 0; a < 3; ++a) {
      const float v0 = t[a], v1 = t[3 + a], v2 = t[6 + a];
      tb[i].lo[a] = std::min(v0, std::min(v1, v2));
      tb[i].hi[a] = std::max(v0, std::max(v1, v2));
    }
  });
  // mean box area, summed in fixed blocks so the value does not depend on the thread count
  constexpr uint64_t kBlock = 1 << 16;
  const uint64_t blocks = (n + kBlock - 1) / kBlock;
  std::vector<double> part(blocks, 0.0);
  parallel_for(blocks, threads, [&](uint64_t k) {
    double sum = 0.0;
    for (uint64_t i = k * kBlock, e = std::min<uint64_t>(n, (k + 1) * kBlock); i < e; ++i) sum += tb[i].area();
    part[k] = sum;
  });
  double mean = 0.0;
  for (double v : part) mean += v;
  mean /= std::max<uint32_t>(1, n);
  const double limit = mean * factor;
  auto ref_of = [&](const Box& b, uint32_t tri) {
    Ref r;
    r.box = b;
    for (int a = 0; a < 3; ++a) r.c[a] = 0.5f * (b.lo[a] + b.hi[a]);
    r.tri = tri;
    return r;
  };
  if (!(limit > 0.0) || budget == 0) {
    refs.resize(n);
    parallel_for(n, threads, [&](uint64_t i) { refs[i] = ref_of(tb[i], (uint32_t)i); });
    return;
  }
  // every triangle cut as far as the limit asks, small blocks of triangles in parallel (the
  // triangles that need cuts tend to sit together)
  constexpr uint64_t kCutBlock = 1 << 11;
  std::vector<std::vector<Ref>> out((n + kCutBlock - 1) / kCutBlock);
  parallel_blocks(out.size(), threads, [&](uint64_t k) {
    std::vector<Ref>& o = out[k];
    std::vector<Box> stack;
    for (uint64_t i = k * kCutBlock, e = std::min<uint64_t>(n, (k + 1) * kCutBlock); i < e; ++i) {
      stack.assign(1, tb[i]);
      while (!stack.empty()) {
        const Box b = stack.back();
        stack.pop_back();
        Box left, right;
        if (b.area() > limit && cut(tris9 + i * 9, b, &left, &right)) {
          stack.push_back(right);
          stack.push_back(left);
        } else {
          o.push_back(ref_of(b, (uint32_t)i));
        }
      }
    }
  });
  uint64_t total = 0;
  for (const auto& o : out) total += o.size();
  if (total - n <= budget) {
    refs.resize(total);
    uint64_t at = 0;
    for (auto& o : out) {
      std::copy(o.begin(), o.end(), refs.begin() + at);
      at += o.size();
      std::vector<Ref>().swap(o);
    }
    return;
  }
  // over budget: the largest references first
  std::vector<uint32_t> ref_tri(n);
  for (uint32_t i = 0; i < n; ++i) ref_tri[i] = i;
  split_refs_heap(tris9, n, limit, budget, ref_tri, tb);
  refs.resize(tb.size());
  for (size_t i = 0; i < tb.size(); ++i) refs[i] = ref_of(tb[i], ref_tri[i]);
}

void pad(const Box& b, float* lo, float* hi) {
  for (int a = 0; a < 3; ++a) {
    float p = (std::fabs(b.lo[a]) + std::fabs(b.hi[a]) + (b.hi[a] - b.lo[a])) * (1.0f / 32768.0f) + 1e-30f;
    lo[a] = b.lo[a] - p;
    hi[a] = b.hi[a] + p;
  }
}

// The step of one axis: a power of two with 254 steps covering the node's extent and no finer
// than the float spacing at the node's coordinates.
int step_exponent(float lo, float hi) {
  const double ext = (double)hi - (double)lo;
  int e = -126;
  if (ext > 0.0) e = std::max(e, (int)std::ceil(std::log2(ext / 254.0)));
  const float big = std::max(std::fabs(lo), std::fabs(hi));
  if (big > 0.0f) {
    int be;
    std::frexp(big, &be);     // big in [2^(be-1), 2^be)
    e = std::max(e, be - 24);  // the float spacing there is 2^(be-24)
  }
  while (std::ldexp(254.0, e) < ext) ++e;
  return std::min(e, 127);
}

// NodeW (padded float child boxes) -> CNode; checks that every decoded box contains its child's.
void quantize(const NodeW& nd, CNode* out) {
  std::memset(out, 0, sizeof *out);
  const float* los[3] = {nd.lox, nd.loy, nd.loz};
  const float* his[3] = {nd.hix, nd.hiy, nd.hiz};
  float origin[3];
  int exps[3];
  constexpr int kw = kBvhWidth / 4;  // words per bound row
  for (int a = 0; a < 3; ++a) {
    float mn = INFINITY, mx = -INFINITY;
    for (int k = 0; k < nd.count; ++k) {
      mn = std::min(mn, los[a][k]);
      mx = std::max(mx, his[a][k]);
    }
    if (nd.count == 0) mn = mx = 0.0f;
    origin[a] = mn;
    exps[a] = step_exponent(mn, mx);
    const double step = std::ldexp(1.0, exps[a]);
    for (int k = 0; k < kBvhWidth; ++k) {
      uint32_t ql = 255, qh = 0;
      if (k < nd.count) {
        const double l = std::floor(((double)los[a][k] - (double)mn) / step) - 1.0;
        const double h = std::ceil(((double)his[a][k] - (double)mn) / step) + 1.0;
        ql = (uint32_t)std::max(0.0, std::min(255.0, l));
        qh = (uint32_t)std::max(0.0, std::min(255.0, h));
        // the decode as the kernels do it: one float fma of the exact q * step and the origin
        const float dl = std::fma((float)ql, (float)step, mn), dh = std::fma((float)qh, (float)step, mn);
        if (!(dl <= los[a][k] && dh >= his[a][k])) {
          // cannot happen with the step bounds above; fall back to the whole node range
          ql = 0;
          qh = 255;
        }
      }
      out->q[(2 * a) * kw + k / 4] |= ql << (8 * (k % 4));
      out->q[(2 * a + 1) * kw + k / 4] |= qh << (8 * (k % 4));
    }
  }
  out->ox = origin[0];
  out->oy = origin[1];
  out->oz = origin[2];
  out->exps = (uint32_t)(exps[0] + 127) | ((uint32_t)(exps[1] + 127) << 8) | ((uint32_t)(exps[2] + 127) << 16);
  for (int k = 0; k < kBvhWidth; ++k) out->child[k] = nd.child[k];
}

struct Pending {
  uint32_t n2, n4, depth;
};

// The wide node `p.n4` from binary node `p.n2`: its two children, then the inner child with the
// largest box opened into its two, up to kBvhWidth children; inner children get the next indices
// of `wide` and are queued in `work`.
void collapse_node(const Node2* n2, const Pending& p, std::vector<NodeW>& wide, std::vector<Pending>& work) {
  uint32_t kids[kBvhWidth];
  int nk = 0;
  if (n2[p.n2].left < 0) {
    kids[nk++] = p.n2;  // whole tree is one leaf
  } else {
    kids[nk++] = (uint32_t)n2[p.n2].left;
    kids[nk++] = (uint32_t)n2[p.n2].right;
    while (nk < kBvhWidth) {
      int open = -1;
      float best = -1.0f;
      for (int k = 0; k < nk; ++k)
        if (n2[kids[k]].left >= 0 && n2[kids[k]].box.area() > best) {
          best = n2[kids[k]].box.area();
          open = k;
        }
      if (open < 0) break;
      const uint32_t o = kids[open];
      kids[open] = (uint32_t)n2[o].left;
      kids[nk++] = (uint32_t)n2[o].right;
    }
  }
  NodeW nd;
  std::memset(&nd, 0, sizeof nd);
  nd.count = nk;
  for (int k = 0; k < nk; ++k) {
    float lo[3], hi[3];
    pad(n2[kids[k]].box, lo, hi);
    nd.lox[k] = lo[0];
    nd.loy[k] = lo[1];
    nd.loz[k] = lo[2];
    nd.hix[k] = hi[0];
    nd.hiy[k] = hi[1];
    nd.hiz[k] = hi[2];
    if (n2[kids[k]].left < 0) {
      const uint32_t first = n2[kids[k]].first, count = n2[kids[k]].count;
      nd.child[k] = -1 - (int32_t)((first << 2) | (count - 1));
    } else {
      const uint32_t idx = (uint32_t)wide.size();
      wide.emplace_back();
      nd.child[k] = (int32_t)idx;
      work.push_back({kids[k], idx, p.depth + 1});
    }
  }
  wide[p.n4] = nd;
}

}  // namespace

void build_bvh(const float* tris9, uint32_t n, unsigned threads, Bvh& out) {
  out.nodes.clear();
  out.tris.clear();
  out.order.clear();
  out.depth = 0;
  if (n == 0) return;
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, 16u);
  float factor = 4.0f, budget_share = 1.0f;
  if (const char* e = std::getenv("DL_BVH_SPLIT")) factor = (float)std::atof(e);
  if (const char* e = std::getenv("DL_BVH_SPLIT_BUDGET")) budget_share = (float)std::atof(e);
  std::vector<Ref> refs;
  make_refs(tris9, n, factor, factor > 0.0f ? (uint64_t)(budget_share * (float)n) : 0, threads, refs);
  const uint32_t nr = (uint32_t)refs.size();

  Builder b;
  b.refs = refs.data();
  std::unique_ptr<Node2[]> node_store(new Node2[2 * (size_t)nr]);
  b.nodes = node_store.get();
  b.threads = threads;
  b.share = threads > 1 ? std::max<uint32_t>(2048, nr / (threads * 16)) : 0;
  if (b.share == 0) {
    b.run({0, 0, nr});
  } else {
    b.shared.push_back({0, 0, nr});
    b.pending = 1;
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < threads; ++t) ts.emplace_back([&b]() { b.worker(); });
    b.worker();
    for (auto& t : ts) t.join();
  }
  const Node2* n2 = b.nodes;

  // Leaf-ordered triangles (one record per reference).
  out.order.resize(nr);
  out.tris.resize((size_t)nr * 12);
  parallel_for(nr, threads, [&](uint64_t k) {
    const uint32_t p = refs[k].tri;
    out.order[k] = p;
    const float* t = tris9 + (size_t)p * 9;
    float* o = out.tris.data() + k * 12;
    o[0] = t[0]; o[1] = t[1]; o[2] = t[2];
    std::memcpy(&o[3], &p, 4);
    o[4] = t[3] - t[0]; o[5] = t[4] - t[1]; o[6] = t[5] - t[2]; o[7] = 0.0f;
    o[8] = t[6] - t[0]; o[9] = t[7] - t[1]; o[10] = t[8] - t[2]; o[11] = 0.0f;
  });
  // Collapse to wide nodes: the top of the tree here, subtrees below `cut` references on
  // worker threads; the nodes of each subtree follow the top's nodes in one block.
  const uint32_t cut = threads > 1 ? std::max<uint32_t>(4096, nr / (threads * 8)) : 0u;
  std::vector<NodeW> wide(1);
  std::vector<Pending> work{{0, 0, 1}}, deferred;
  while (!work.empty()) {
    const Pending p = work.back();
    work.pop_back();
    if (p.n2 != 0 && n2[p.n2].left >= 0 && n2[p.n2].count < cut) {
      deferred.push_back(p);
      continue;
    }
    out.depth = std::max(out.depth, p.depth);
    collapse_node(n2, p, wide, work);
  }
  std::vector<std::vector<NodeW>> sub(deferred.size());
  std::vector<uint32_t> sub_depth(deferred.size(), 0);
  parallel_blocks(deferred.size(), threads, [&](uint64_t i) {
    std::vector<NodeW>& local = sub[i];
    local.resize(1);
    std::vector<Pending> w{{deferred[i].n2, 0, deferred[i].depth}};
    while (!w.empty()) {
      const Pending p = w.back();
      w.pop_back();
      sub_depth[i] = std::max(sub_depth[i], p.depth);
      collapse_node(n2, p, local, w);
    }
  });
  std::vector<size_t> base(deferred.size());
  size_t total = wide.size();
  for (size_t i = 0; i < sub.size(); ++i) {
    base[i] = total;
    total += sub[i].size() - 1;
    out.depth = std::max(out.depth, sub_depth[i]);
  }
  out.nodes.resize(total);
  parallel_for(wide.size(), threads, [&](uint64_t i) { quantize(wide[i], &out.nodes[i]); });
  // a subtree's root takes the slot the top reserved for it; its other nodes follow at base[i]
  parallel_blocks(sub.size(), threads, [&](uint64_t i) {
    for (size_t j = 0; j < sub[i].size(); ++j) {
      NodeW nd = sub[i][j];
      for (int k = 0; k < kBvhWidth; ++k)
        if (nd.child[k] > 0) nd.child[k] = (int32_t)(base[i] + (size_t)nd.child[k] - 1);
      quantize(nd, &out.nodes[j == 0 ? deferred[i].n4 : base[i] + j - 1]);
    }
  });
}

}  // namespace dl

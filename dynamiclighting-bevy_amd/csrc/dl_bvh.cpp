// Host build of the occluder BVH: large triangles are first split into several references, each
// bounded by the part of the triangle inside half of its parent reference's box (clipped in
// double precision, bounds rounded outward); then binned SAH over reference centroids into a
// binary tree with leaves of at most four references, collapsed to 4-wide nodes by opening the
// largest children first.  Subtrees below a size threshold are built on worker threads.  A
// triangle split into several references appears in several leaves; any-hit and closest-hit
// (lowest t, then lowest occluder index) results do not depend on that.
#include "dl_bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
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

// Splits references of triangles whose box has a surface area above `factor` times the mean
// triangle box area, largest first, into at most `budget` extra references.
void split_refs(const float* tris9, uint32_t n, float factor, uint64_t budget, std::vector<uint32_t>& ref_tri, std::vector<Box>& ref_box) {
  double mean = 0.0;
  for (uint32_t i = 0; i < n; ++i) mean += ref_box[i].area();
  mean /= std::max<uint32_t>(1, n);
  const double limit = mean * factor;
  if (!(limit > 0.0) || budget == 0) return;
  using Item = std::pair<float, uint32_t>;  // area, reference
  std::priority_queue<Item> heap;
  for (uint32_t i = 0; i < n; ++i)
    if (ref_box[i].area() > limit) heap.push({ref_box[i].area(), i});
  while (!heap.empty() && budget > 0) {
    const uint32_t r = heap.top().second;
    heap.pop();
    const Box b = ref_box[r];
    int axis = 0;
    for (int a = 1; a < 3; ++a)
      if (b.hi[a] - b.lo[a] > b.hi[axis] - b.lo[axis]) axis = a;
    const double mid = 0.5 * ((double)b.lo[axis] + (double)b.hi[axis]);
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = b.lo[a]; hi[a] = b.hi[a]; }
    const float* t = tris9 + (size_t)ref_tri[r] * 9;
    Box left, right;
    hi[axis] = mid;
    const bool has_l = clip_bounds(t, lo, hi, &left);
    hi[axis] = b.hi[axis];
    lo[axis] = mid;
    const bool has_r = clip_bounds(t, lo, hi, &right);
    if (!has_l || !has_r) continue;  // a sliver: keep the reference whole
    ref_box[r] = left;
    ref_tri.push_back(ref_tri[r]);
    ref_box.push_back(right);
    budget--;
    if (left.area() > limit) heap.push({left.area(), r});
    if (right.area() > limit) heap.push({right.area(), (uint32_t)ref_box.size() - 1});
  }
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

// Node4 (padded float child boxes) -> CNode; checks that every decoded box contains its child's.
void quantize(const Node4& nd, CNode* out) {
  std::memset(out, 0, sizeof *out);
  const float* los[3] = {nd.lox, nd.loy, nd.loz};
  const float* his[3] = {nd.hix, nd.hiy, nd.hiz};
  float origin[3];
  uint32_t lo_q[3] = {0, 0, 0}, hi_q[3] = {0, 0, 0};
  int exps[3];
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
    for (int k = 0; k < 4; ++k) {
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
      lo_q[a] |= ql << (8 * k);
      hi_q[a] |= qh << (8 * k);
    }
  }
  out->ox = origin[0];
  out->oy = origin[1];
  out->oz = origin[2];
  out->exps = (uint32_t)(exps[0] + 127) | ((uint32_t)(exps[1] + 127) << 8) | ((uint32_t)(exps[2] + 127) << 16);
  out->lox = lo_q[0];
  out->hix = hi_q[0];
  out->loy = lo_q[1];
  out->hiy = hi_q[1];
  out->loz = lo_q[2];
  out->hiz = hi_q[2];
  for (int k = 0; k < 4; ++k) out->child[k] = nd.child[k];
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
  std::vector<uint32_t> ref_tri(n);
  b.tb.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* t = tris9 + (size_t)i * 9;
    for (int a = 0; a < 3; ++a) {
      float v0 = t[a], v1 = t[3 + a], v2 = t[6 + a];
      b.tb[i].lo[a] = std::min(v0, std::min(v1, v2));
      b.tb[i].hi[a] = std::max(v0, std::max(v1, v2));
    }
    ref_tri[i] = i;
  }
  float factor = 4.0f, budget_share = 1.0f;
  if (const char* e = std::getenv("DL_BVH_SPLIT")) factor = (float)std::atof(e);
  if (const char* e = std::getenv("DL_BVH_SPLIT_BUDGET")) budget_share = (float)std::atof(e);
  if (factor > 0.0f) split_refs(tris9, n, factor, (uint64_t)(budget_share * (float)n), ref_tri, b.tb);
  const uint32_t nr = (uint32_t)ref_tri.size();
  b.cen.resize((size_t)nr * 3);
  b.idx.resize(nr);
  for (uint32_t i = 0; i < nr; ++i) {
    for (int a = 0; a < 3; ++a) b.cen[(size_t)i * 3 + a] = 0.5f * (b.tb[i].lo[a] + b.tb[i].hi[a]);
    b.idx[i] = i;
  }
  b.nodes.reserve((size_t)n * 2 / 3 + 16);
  b.nodes.emplace_back();
  // Top of the tree on this thread until subtrees are small, then the rest in parallel.
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  uint32_t spill_below = threads > 1 ? std::max<uint32_t>(4096, nr / (threads * 8)) : 0;
  std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>> spill;
  b.build_subtree(0, 0, nr, threads > 1 ? &spill : nullptr, spill_below);
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
  // Leaf-ordered triangles (one record per reference).
  out.order.resize(nr);
  for (uint32_t k = 0; k < nr; ++k) out.order[k] = ref_tri[b.idx[k]];
  out.tris.resize((size_t)nr * 12);
  for (uint32_t k = 0; k < nr; ++k) {
    uint32_t p = out.order[k];
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
  std::vector<Node4> wide;
  wide.emplace_back();
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
        uint32_t idx = (uint32_t)wide.size();
        wide.emplace_back();
        nd.child[k] = (int32_t)idx;
        work.push_back({kids[k], idx, p.depth + 1});
      }
    }
    wide[p.n4] = nd;
  }
  out.nodes.resize(wide.size());
  for (size_t i = 0; i < wide.size(); ++i) quantize(wide[i], &out.nodes[i]);
}

}  // namespace dl

// Host orchestration of the bake: validation and ordering exactly as the reference
// (baker.rs:27-233, input.rs:233-373), per-light and per-mesh preparation, then the GPU passes:
//
//   setup/association -> owned texels -> traced texels -> burial -> compaction -> shadow rays
//   -> optimize -> seam padding -> crop/dilate/pack + serialization -> distance cubes
//
// Host work between passes is O(triangles + light entries) integer bookkeeping (regions and word
// offsets); every ray and every texel is processed on the device.
#include "dl_host.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

namespace dl {

// ---- trig exception tables ---------------------------------------------------------------
void build_trig_tables(TrigHost& out, unsigned threads) {
  // Every float with 2^-16 <= |x| (bits 0x37800000) and x in [-0.8, 8]: the arguments of the
  // burial rays, the concentric disk, on_unit_sphere and slerp.
  struct Range { uint32_t lo, hi; };
  const Range ranges[2] = {{0x37800000u, 0x41000000u}, {0xB7800000u, 0xBF4CCCCDu}};
  const uint32_t chunk = 1u << 20;
  std::vector<std::pair<uint32_t, uint32_t>> jobs;
  for (const Range& r : ranges)
    for (uint64_t b = r.lo; b <= r.hi; b += chunk) jobs.push_back({(uint32_t)b, (uint32_t)std::min<uint64_t>(r.hi, b + chunk - 1)});
  std::vector<std::vector<uint32_t>> sk(jobs.size()), sv(jobs.size()), ck(jobs.size()), cv(jobs.size());
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t j = next.fetch_add(1);
      if (j >= jobs.size()) break;
      for (uint64_t b = jobs[j].first; b <= jobs[j].second; ++b) {
        float x = u2f((uint32_t)b);
        float s = sinf(x), c = cosf(x);
        if (f2u(s) != f2u(trig_core(x, 0))) { sk[j].push_back((uint32_t)b); sv[j].push_back(f2u(s)); }
        if (f2u(c) != f2u(trig_core(x, 1))) { ck[j].push_back((uint32_t)b); cv[j].push_back(f2u(c)); }
      }
    }
  };
  if (threads == 0) threads = 1;
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < threads; ++t) ts.emplace_back(work);
  work();
  for (auto& t : ts) t.join();
  auto gather = [&](std::vector<std::vector<uint32_t>>& k, std::vector<std::vector<uint32_t>>& v, std::vector<uint32_t>& keys,
                    std::vector<uint32_t>& vals, std::vector<uint32_t>& runs) {
    keys.clear();
    vals.clear();
    // jobs are in ascending bit order within each range, and the positive range precedes the
    // negative one, so concatenation is sorted by bit pattern and by domain index.
    for (size_t j = 0; j < k.size(); ++j) {
      keys.insert(keys.end(), k[j].begin(), k[j].end());
      vals.insert(vals.end(), v[j].begin(), v[j].end());
    }
    runs.assign(2 * (size_t)kTrigRuns, 0u);
    for (uint32_t key : keys) {
      const uint32_t di = trig_domain_index(key);
      runs[2 * (size_t)(di >> 5)] |= 1u << (di & 31u);
    }
    uint32_t before = 0;
    for (size_t r = 0; r < kTrigRuns; ++r) {
      runs[2 * r + 1] = before;
      before += (uint32_t)__builtin_popcount(runs[2 * r]);
    }
  };
  gather(sk, sv, out.sin_keys, out.sin_vals, out.sin_runs);
  gather(ck, cv, out.cos_keys, out.cos_vals, out.cos_runs);
  for (int w = 0; w < 2; ++w) {  // one bit per run: the run holds an exception
    const std::vector<uint32_t>& runs = w ? out.cos_runs : out.sin_runs;
    std::vector<uint32_t>& any = w ? out.cos_any : out.sin_any;
    any.assign(((size_t)kTrigRuns + 31) / 32, 0u);
    for (size_t r = 0; r < kTrigRuns; ++r)
      if (runs[2 * r]) any[r / 32] |= 1u << (r % 32);
  }
  // The distance: every exception's double value lies within it of a rounding midpoint (see
  // trig_midpoint_distance), so the kernels search exactly when an exception is possible.
  uint32_t tol = 0;
  for (int w = 0; w < 2; ++w)
    for (uint32_t k : w ? out.cos_keys : out.sin_keys) tol = std::max(tol, trig_midpoint_distance(trig_value(u2f(k), w)));
  out.tol = tol;
}

// ---- light BVH: SAH split at light positions, leaves of at most two (light_bvh.rs:59-189) --------
namespace {
struct Aabb3 {
  f3 lo, hi;
};
inline Aabb3 empty_box() { return {mk3(INFINITY, INFINITY, INFINITY), mk3(-INFINITY, -INFINITY, -INFINITY)}; }
inline Aabb3 sphere_box(f3 c, float r) { return {sub(c, mk3(r, r, r)), add(c, mk3(r, r, r))}; }
inline Aabb3 unite(Aabb3 a, Aabb3 b) { return {vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
inline float half_area_x2(Aabb3 a) {
  f3 e = sub(a.hi, a.lo);
  if (min_elem(e) < 0.0f) return 0.0f;
  return 2.0f * ((e.x * e.y + e.y * e.z) + e.z * e.x);
}
inline float axis(f3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

struct LightBvhBuild {
  const std::vector<f3>& pos;
  const std::vector<float>& radii;
  std::vector<uint32_t> idx;
  std::vector<dl_light_bvh_node> nodes;
  LightBvhBuild(const std::vector<f3>& p, const std::vector<float>& r) : pos(p), radii(r) {}
  Aabb3 bounds(uint32_t l) const { return sphere_box(pos[l], radii[l]); }
  void update(size_t n) {
    Aabb3 a = empty_box();
    for (uint32_t k = nodes[n].left_first; k < nodes[n].left_first + nodes[n].count; ++k) a = unite(a, bounds(idx[k]));
    nodes[n].aabb_min[0] = a.lo.x; nodes[n].aabb_min[1] = a.lo.y; nodes[n].aabb_min[2] = a.lo.z;
    nodes[n].aabb_max[0] = a.hi.x; nodes[n].aabb_max[1] = a.hi.y; nodes[n].aabb_max[2] = a.hi.z;
  }
  void subdivide(size_t n) {
    uint32_t first = nodes[n].left_first, count = nodes[n].count;
    if (count <= 2) return;
    Aabb3 nb = {mk3(nodes[n].aabb_min[0], nodes[n].aabb_min[1], nodes[n].aabb_min[2]),
                mk3(nodes[n].aabb_max[0], nodes[n].aabb_max[1], nodes[n].aabb_max[2])};
    float parent = (float)count * half_area_x2(nb);
    bool found = false;
    int best_axis = 0;
    float best_split = 0.0f, best_cost = 3.40282347e+38f;
    // Every light's coordinate is a candidate split, tried axis by axis in the node's order (the
    // first lowest cost wins).  A split at v puts the lights whose coordinate is below v left, and
    // costs lc * area(left bounds) + rc * area(right bounds) (infinite with an empty side): with the
    // lights sorted along the axis the left ones are a prefix, so prefix and suffix bounds give
    // every candidate's cost in O(log n) instead of a pass over the node.  Bounds are unions by
    // min/max, exact in any order, so the costs are those of the pass bit for bit.
    std::vector<std::pair<float, uint32_t>> sv(count);
    std::vector<Aabb3> pre(count + 1), suf(count + 1);
    for (int a = 0; a < 3; ++a) {
      for (uint32_t k = 0; k < count; ++k) sv[k] = {axis(pos[idx[first + k]], a), idx[first + k]};
      std::sort(sv.begin(), sv.end(), [](const std::pair<float, uint32_t>& x, const std::pair<float, uint32_t>& y) {
        return x.first < y.first;
      });
      pre[0] = empty_box();
      for (uint32_t k = 0; k < count; ++k) pre[k + 1] = unite(pre[k], bounds(sv[k].second));
      suf[count] = empty_box();
      for (uint32_t k = count; k-- > 0;) suf[k] = unite(suf[k + 1], bounds(sv[k].second));
      for (uint32_t k = 0; k < count; ++k) {
        const float cand = axis(pos[idx[first + k]], a);
        const uint32_t lc = (uint32_t)(std::lower_bound(sv.begin(), sv.end(), cand,
                                                        [](const std::pair<float, uint32_t>& x, float v) { return x.first < v; }) -
                                       sv.begin());
        const uint32_t rc = count - lc;
        const float c = (lc == 0 || rc == 0) ? 3.40282347e+38f : (float)lc * half_area_x2(pre[lc]) + (float)rc * half_area_x2(suf[lc]);
        if (c < best_cost) { best_cost = c; found = true; best_axis = a; best_split = cand; }
      }
    }
    if (!found || best_cost >= parent) return;
    int64_t i = first, j = (int64_t)first + count - 1;
    while (i <= j) {
      if (axis(pos[idx[(size_t)i]], best_axis) < best_split) i++;
      else { std::swap(idx[(size_t)i], idx[(size_t)j]); j--; }
    }
    uint32_t left_count = (uint32_t)(i - first);
    if (left_count == 0 || left_count == count) return;
    size_t left = nodes.size();
    dl_light_bvh_node a{}, b{};
    a.left_first = first;
    a.count = left_count;
    b.left_first = (uint32_t)i;
    b.count = count - left_count;
    nodes.push_back(a);
    nodes.push_back(b);
    nodes[n].left_first = (uint32_t)left;
    nodes[n].count = 0;
    update(left);
    update(left + 1);
    subdivide(left);
    subdivide(left + 1);
  }
};
}  // namespace

void build_light_bvh(const std::vector<f3>& pos, const std::vector<float>& radii, LightBvhOut& out) {
  out.nodes.clear();
  out.order.clear();
  if (pos.empty()) return;
  LightBvhBuild b(pos, radii);
  for (uint32_t i = 0; i < pos.size(); ++i) b.idx.push_back(i);
  dl_light_bvh_node root{};
  root.count = (uint32_t)pos.size();
  b.nodes.push_back(root);
  b.update(0);
  b.subdivide(0);
  out.nodes = std::move(b.nodes);
  out.order = std::move(b.idx);
}

}  // namespace dl

// ---- validation: input.rs:233-373, baker.rs:205-233 ---------------------------------------------
namespace dl {

std::string hex_id(uint64_t id) {
  char b[32];
  snprintf(b, sizeof b, "0x%016llx", (unsigned long long)id);
  return b;
}
bool finite3p(const float* p) { return finite_(p[0]) && finite_(p[1]) && finite_(p[2]); }

int check_options(dl_ctx* ctx, const dl_bake_options& o) {
  auto bad = [&](const char* r) { ctx->err = std::string("invalid bake options: ") + r; return (int)DL_E_INVALID_OPTIONS; };
  if (!(finite_(o.pixel_density) && o.pixel_density > 0.0f)) return bad("pixel density must be positive");
  if (o.max_lightmap_size < 2) return bad("the maximum shadow map size must be at least 2");
  if (o.photon_cube_resolution == 0) return bad("the photon cube resolution must be positive");
  if (o.area_light_samples == 0) return bad("area lights need at least one sample");
  if (!(o.sun_map_resolution >= 1 && o.sun_map_resolution <= 8192)) return bad("the sun map resolution must be 1 to 8192");
  if (o.sky_samples == 0) return bad("skies need at least one sample");
  if (!(finite_(o.sky_probe_spacing) && o.sky_probe_spacing > 0.0f)) return bad("the sky probe spacing must be positive");
  return DL_OK;
}

int check_mesh(dl_ctx* ctx, const dl_mesh& m) {
  const uint32_t v = m.vertex_count;
  auto bad = [&](const std::string& r) { ctx->err = "mesh " + hex_id(m.id) + ": " + r; return (int)DL_E_INVALID_MESH; };
  if (m.index_count % 3 != 0) return bad(std::to_string(m.index_count) + " indices is not a multiple of 3");
  if (m.index_count && !m.indices) return bad("null index buffer");
  for (uint32_t i = 0; i < m.index_count; ++i)
    if (m.indices[i] >= v) return bad("index " + std::to_string(m.indices[i]) + " is out of range for " + std::to_string(v) + " vertices");
  if (m.uv1_count != v) return bad(std::to_string(m.uv1_count) + " UV1 coordinates for " + std::to_string(v) + " vertices");
  uint32_t nn = m.normals ? m.normal_count : 0;
  if (nn != 0 && nn != v) return bad(std::to_string(nn) + " normals for " + std::to_string(v) + " vertices");
  for (uint32_t i = 0; i < v; ++i)
    if (!finite3p(m.positions + 3 * (size_t)i)) return bad("non-finite vertex position");
  if (m.uv0 && m.uv0_count != v) return bad(std::to_string(m.uv0_count) + " UV0 coordinates for " + std::to_string(v) + " vertices");
  if (m.alpha) {
    if (!m.uv0) return bad("an alpha texture needs UV0");
    if (m.alpha->rgba8_len != (uint64_t)m.alpha->width * m.alpha->height * 4) return bad("alpha texture size doesn't match its data");
  }
  return DL_OK;
}

bool unit_quat(const float* q) {
  for (int i = 0; i < 4; ++i)
    if (!finite_(q[i])) return false;
  float len2 = (q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]);
  return fabsf(len2 - 1.0f) <= 2e-4f;
}

int check_light(dl_ctx* ctx, const dl_light& l) {
  auto bad = [&](const char* r) { ctx->err = "light " + hex_id(l.id) + ": " + r; return (int)DL_E_INVALID_LIGHT; };
  if (!finite3p(l.position)) return bad("non-finite position");
  bool distant = l.shape == DL_SHAPE_SUN || l.shape == DL_SHAPE_SKY;
  if (!distant && !(finite_(l.radius) && l.radius > 0.0f)) return bad("radius must be positive and finite");
  if (l.shape == DL_SHAPE_RECTANGLE || l.shape == DL_SHAPE_TUBE) {
    if (!(finite_(l.size[0]) && finite_(l.size[1])) || l.size[0] < 0.0f || l.size[1] < 0.0f)
      return bad("area light size must be finite and not negative");
    if (!unit_quat(l.rotation)) return bad("area light rotation must be a unit quaternion");
  } else if (l.shape == DL_SHAPE_SUN) {
    if (!unit_quat(l.rotation)) return bad("sun rotation must be a unit quaternion");
    if (!(l.angular_diameter >= 0.0f && l.angular_diameter <= 10.0f)) return bad("sun angular diameter must be 0 to 10 degrees");
  }
  return DL_OK;
}

DevLight make_dev_light(const dl_light& l, const dl_bake_options& o) {
  DevLight d;
  std::memset(&d, 0, sizeof d);
  f3 p = mk3(l.position[0], l.position[1], l.position[2]);
  d.pos[0] = p.x; d.pos[1] = p.y; d.pos[2] = p.z;
  d.radius = l.radius;
  f3 r3 = mk3(l.radius, l.radius, l.radius);
  f3 lo = sub(p, r3), hi = add(p, r3);
  d.bmin[0] = lo.x; d.bmin[1] = lo.y; d.bmin[2] = lo.z;
  d.bmax[0] = hi.x; d.bmax[1] = hi.y; d.bmax[2] = hi.z;
  d.shape = l.shape;
  bool area = l.shape == DL_SHAPE_RECTANGLE || l.shape == DL_SHAPE_TUBE;
  bool sun = l.shape == DL_SHAPE_SUN, sky = l.shape == DL_SHAPE_SKY;
  d.distant = sun || sky;
  d.alpha = l.transparency != DL_TRANSPARENCY_DISABLED;
  d.bakes_bounce = l.illumination == DL_ILLUMINATION_SINGLE_BOUNCE && !o.skip_bounce && !sky;
  d.bits = l.shape == DL_SHAPE_POINT ? 1u : area ? 4u : sun ? (l.angular_diameter > 0.0f ? 4u : 1u) : 8u;
  d.reach = l.radius;
  if (area) {
    f3 x = quat_rotate(l.rotation, mk3(1, 0, 0)), y = quat_rotate(l.rotation, mk3(0, 1, 0));
    f3 fw = quat_rotate(l.rotation, mk3(0, 0, -1));
    d.ax[0] = x.x; d.ax[1] = x.y; d.ax[2] = x.z;
    d.ay[0] = y.x; d.ay[1] = y.y; d.ay[2] = y.z;
    d.af[0] = fw.x; d.af[1] = fw.y; d.af[2] = fw.z;
    d.hx = l.size[0] * 0.5f;
    d.hy = l.size[1] * 0.5f;
    d.rect = l.shape == DL_SHAPE_RECTANGLE;
    d.reach = l.radius - sqrtf(d.hx * d.hx + d.hy * d.hy);
  }
  if (sun) {
    f3 t = normalize(quat_rotate(l.rotation, mk3(0, 0, 1)));
    d.toward[0] = t.x; d.toward[1] = t.y; d.toward[2] = t.z;
    d.srad = (l.angular_diameter * 0.5f) * (kPi / 180.0f);
    d.stan = tanf(d.srad);
    d.ssin = sinf(d.srad);
  }
  d.bounce_samples = l.bounce_samples;
  d.bounce_intensity = l.bounce_intensity;
  d.id = l.id;
  return d;
}

bool sun_map_setup(const float* corners, size_t ncorners, const float toward_in[3], uint32_t resolution, SunMapParams* p,
                   float rows[12]) {
  const f3 toward = mk3(toward_in[0], toward_in[1], toward_in[2]);
  f3 side, up;
  orthonormal_pair(toward, &side, &up);
  const f3 along = neg(toward);
  f3 mn = mk3(INFINITY, INFINITY, INFINITY), mx = mk3(-INFINITY, -INFINITY, -INFINITY);
  for (size_t i = 0; i < ncorners; ++i) {
    const f3 v = mk3(corners[3 * i], corners[3 * i + 1], corners[3 * i + 2]);
    const f3 l = mk3(dot(v, side), dot(v, up), dot(v, along));
    mn = vmin(mn, l);
    mx = vmax(mx, l);
  }
  if (!(finite_(mn.x) && finite_(mn.y) && finite_(mn.z) && finite_(mx.x) && finite_(mx.y) && finite_(mx.z))) return false;
  const f3 size = sub(mx, mn);
  const float margin = size.z * 0.01f + 0.05f;  // DEPTH_MARGIN (sun_map.rs:22)
  const float extent = maxf_(size.x, size.y) + 2.0f * margin;
  const f3 center = mul(add(mn, mx), 0.5f);
  const f3 bmin = mk3(center.x - extent * 0.5f, center.y - extent * 0.5f, mn.z - margin);
  const float depth = size.z + 2.0f * margin;
  const f3 axes[3] = {side, up, along};
  const float lo[3] = {bmin.x, bmin.y, bmin.z}, length[3] = {extent, extent, depth};
  for (int r = 0; r < 3; ++r) {
    rows[4 * r] = axes[r].x / length[r];
    rows[4 * r + 1] = axes[r].y / length[r];
    rows[4 * r + 2] = axes[r].z / length[r];
    rows[4 * r + 3] = -lo[r] / length[r];
  }
  for (int k = 0; k < 3; ++k) {
    p->side[k] = k == 0 ? side.x : k == 1 ? side.y : side.z;
    p->up[k] = k == 0 ? up.x : k == 1 ? up.y : up.z;
    p->along[k] = k == 0 ? along.x : k == 1 ? along.y : along.z;
    p->min[k] = lo[k];
  }
  p->texel = extent / (float)resolution;
  p->depth = depth;
  p->res = resolution;
  return true;
}

uint32_t pack_sun_depth(float lower, float upper) {
  auto unorm = [](float d) { return sat_u32(roundf(clampf_(d, 0.0f, 1.0f) * 65535.0f)); };
  return unorm(upper) << 16 | unorm(lower);
}

void probe_grid(f3 mn, f3 mx, float spacing, f3* origin, float* spacing_out, uint32_t counts[3]) {
  const uint64_t cap = 65536;  // MAX_SKY_PROBES (sky_probes.rs:18)
  const f3 extent = add(vmax(sub(mx, mn), mk3(0.0f, 0.0f, 0.0f)), mk3(spacing, spacing, spacing));
  auto counts_at = [&](float s) {
    counts[0] = sat_u32(ceilf(extent.x / s)) + 1;
    counts[1] = sat_u32(ceilf(extent.y / s)) + 1;
    counts[2] = sat_u32(ceilf(extent.z / s)) + 1;
  };
  auto count = [&]() { return (uint64_t)counts[0] * counts[1] * counts[2]; };
  counts_at(spacing);
  while (count() > cap) {
    const float ratio = (float)count() / (float)cap;
    spacing *= maxf_(cbrtf(ratio), 1.01f);
    counts_at(spacing);
  }
  const f3 span = mul(mk3((float)(counts[0] - 1), (float)(counts[1] - 1), (float)(counts[2] - 1)), spacing);
  *origin = sub(mul(add(mn, mx), 0.5f), mul(span, 0.5f));
  *spacing_out = spacing;
}

void fill_inside(const uint32_t counts[3], std::vector<float>& open, std::vector<uint8_t>& valid) {
  static const int step[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  const int64_t cx = counts[0], cy = counts[1], cz = counts[2];
  std::vector<float> before;
  std::vector<uint8_t> before_valid;
  for (bool changed = true; changed;) {
    changed = false;
    before = open;
    before_valid = valid;
    for (size_t i = 0; i < valid.size(); ++i) {
      if (valid[i]) continue;
      const int64_t x = (int64_t)i % cx, y = (int64_t)i % (cx * cy) / cx, z = (int64_t)i / (cx * cy);
      float sum[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      float found = 0.0f;
      for (const auto& st : step) {
        const int64_t nx = x + st[0], ny = y + st[1], nz = z + st[2];
        if (nx < 0 || ny < 0 || nz < 0 || nx >= cx || ny >= cy || nz >= cz) continue;
        const size_t j = (size_t)((nz * cy + ny) * cx + nx);
        if (!before_valid[j]) continue;
        for (int a = 0; a < 6; ++a) sum[a] += before[j * 6 + a];
        found += 1.0f;
      }
      if (found > 0.0f) {
        for (int a = 0; a < 6; ++a) open[i * 6 + a] = sum[a] / found;
        valid[i] = 1;
        changed = true;
      }
    }
  }
}

void pack_probe(const float* open6, uint32_t* words2) {
  words2[0] = words2[1] = 0;
  for (int a = 0; a < 6; ++a) words2[a / 4] |= sat_u32(roundf(clampf_(open6[a], 0.0f, 1.0f) * 255.0f)) << (8 * (a % 4));
}

namespace {
// Room for `bytes` in the staging buffer (allocated on first use), waiting for the stream when the
// buffer is full; null when staging is not possible (too large, or no pinned memory).
char* staging_slot(dl_ctx* ctx, size_t bytes, hipError_t* err) {
  Staging& st = ctx->staging;
  *err = hipSuccess;
  if (bytes > kStagedMax) return nullptr;
  if (!st.host) {
    void* h = nullptr;
    if (hipHostMalloc(&h, kStagingBytes, hipHostMallocDefault) != hipSuccess || !h) return nullptr;
    st.host = static_cast<char*>(h);
    st.cap = kStagingBytes;
    st.off = 0;
  }
  const size_t need = (bytes + 255) & ~size_t(255);
  if (st.off + need > st.cap) {  // every earlier staged copy must be done before the space is reused
    *err = hipStreamSynchronize(ctx->stream);
    if (*err != hipSuccess) return nullptr;
    st.off = 0;
  }
  char* slot = st.host + st.off;
  st.off += need;
  return slot;
}
}  // namespace

hipError_t copy_to_device(dl_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  hipError_t e = hipSuccess;
  if (char* slot = staging_slot(ctx, bytes, &e)) {
    std::memcpy(slot, src, bytes);
    return hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, ctx->stream);
  }
  if (e != hipSuccess) return e;
  // A copy from pageable memory may still read `src` after hipMemcpyAsync returns; the caller's
  // buffer is only borrowed for this call, so wait for the copy before giving it back.
  e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  return e == hipSuccess ? hipStreamSynchronize(ctx->stream) : e;
}

// memcpy on up to `threads` threads in 4 MiB parts: a fresh destination's pages fault in parallel.
static void parallel_copy(char* dst, const char* src, size_t bytes, unsigned threads) {
  constexpr size_t kPart = 4u << 20;
  const size_t parts = (bytes + kPart - 1) / kPart;
  const unsigned th = (unsigned)std::min<size_t>(std::min(threads, 8u), parts);
  if (th <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t k; (k = next.fetch_add(1)) < parts;) {
      const size_t a = k * kPart;
      std::memcpy(dst + a, src + a, std::min(kPart, bytes - a));
    }
  };
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < th; ++t) ts.emplace_back(work);
  work();
  for (auto& t : ts) t.join();
}

hipError_t copy_to_host(dl_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  hipError_t e = hipSuccess;
  if (bytes > kStagedMax) {
    // through the staging buffer in two alternating halves: the device copy of one piece runs
    // while the host copies the previous piece out
    if (!staging_slot(ctx, 1, &e)) {  // no pinned memory: one copy into pageable memory
      if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
      return e == hipSuccess ? hipStreamSynchronize(ctx->stream) : e;
    }
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;  // earlier staged copies
    Staging& st = ctx->staging;
    const size_t half = st.cap / 2;
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& v : ev)
      if (e == hipSuccess) e = hipEventCreateWithFlags(&v, hipEventDisableTiming);
    const char* from = static_cast<const char*>(src);
    char* to = static_cast<char*>(dst);
    size_t prev_off = 0, prev_len = 0;
    for (size_t off = 0, k = 0; e == hipSuccess && off < bytes; off += half, ++k) {
      const size_t len = std::min(half, bytes - off);
      char* slot = st.host + (k & 1u) * half;
      e = hipMemcpyAsync(slot, from + off, len, hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipEventRecord(ev[k & 1u], ctx->stream);
      if (e == hipSuccess && k > 0) {
        e = hipEventSynchronize(ev[(k - 1) & 1u]);
        if (e == hipSuccess) parallel_copy(to + prev_off, st.host + ((k - 1) & 1u) * half, prev_len, ctx->host_threads);
      }
      prev_off = off;
      prev_len = len;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
      parallel_copy(to + prev_off, st.host + ((((bytes + half - 1) / half) - 1) & 1u) * half, prev_len, ctx->host_threads);
    for (auto& v : ev)
      if (v) (void)hipEventDestroy(v);
    st.off = 0;
    return e;
  }
  char* slot = staging_slot(ctx, bytes, &e);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(slot ? slot : dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return e;
  if (slot) std::memcpy(dst, slot, bytes);
  ctx->staging.off = 0;  // the stream is idle: every staged copy is done
  return hipSuccess;
}

uint32_t lightmap_size(float area, float density, uint32_t max_size) {
  float size = ceilf(sqrtf(density * density * area));
  if (size != size) size = 0.0f;
  uint32_t v = sat_u32(minf_(size, (float)max_size));
  uint32_t hi = std::max(max_size, 2u);
  return std::min(std::max(v, 2u), hi);
}

}  // namespace dl

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
                    std::vector<uint32_t>& vals, std::vector<uint32_t>& bucket) {
    keys.clear();
    vals.clear();
    // jobs are in ascending bit order within each range, and the positive range precedes the
    // negative one, so concatenation is sorted by bit pattern.
    for (size_t j = 0; j < k.size(); ++j) {
      keys.insert(keys.end(), k[j].begin(), k[j].end());
      vals.insert(vals.end(), v[j].begin(), v[j].end());
    }
    bucket.assign(65537, 0);
    size_t i = 0;
    for (uint32_t b = 0; b <= 65536; ++b) {
      while (i < keys.size() && (keys[i] >> 16) < b) ++i;
      bucket[b] = (uint32_t)i;
    }
  };
  gather(sk, sv, out.sin_keys, out.sin_vals, out.sin_bucket);
  gather(ck, cv, out.cos_keys, out.cos_vals, out.cos_bucket);
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
  float cost(size_t first, size_t count, int ax, float split) const {
    Aabb3 l = empty_box(), r = empty_box();
    uint32_t lc = 0, rc = 0;
    for (size_t k = first; k < first + count; ++k) {
      uint32_t li = idx[k];
      if (axis(pos[li], ax) < split) { lc++; l = unite(l, bounds(li)); }
      else { rc++; r = unite(r, bounds(li)); }
    }
    if (lc == 0 || rc == 0) return 3.40282347e+38f;
    return (float)lc * half_area_x2(l) + (float)rc * half_area_x2(r);
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
    for (int a = 0; a < 3; ++a)
      for (uint32_t k = 0; k < count; ++k) {
        float cand = axis(pos[idx[first + k]], a);
        float c = cost(first, count, a, cand);
        if (c < best_cost) { best_cost = c; found = true; best_axis = a; best_split = cand; }
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
namespace {
using namespace dl;

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

uint32_t lightmap_size(float area, float density, uint32_t max_size) {
  float size = ceilf(sqrtf(density * density * area));
  if (size != size) size = 0.0f;
  uint32_t v = sat_u32(minf_(size, (float)max_size));
  uint32_t hi = std::max(max_size, 2u);
  return std::min(std::max(v, 2u), hi);
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);          \
      return e_ == hipErrorOutOfMemory ? (int)DL_E_OOM : (int)DL_E_DEVICE;   \
    }                                                                        \
  } while (0)

template <class T>
hipError_t upload(dl_ctx* ctx, DBuf<T>& b, const std::vector<T>& h) {
  hipError_t e = b.alloc(h.size());
  if (e != hipSuccess || h.empty()) return e;
  return hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
}
template <class T>
hipError_t download(dl_ctx* ctx, std::vector<T>& h, const T* p, size_t n) {
  h.resize(n);
  if (n == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->stream);
}
template <class T>
hipError_t zeroed(dl_ctx* ctx, DBuf<T>& b, size_t n) {
  hipError_t e = b.alloc(n);
  if (e != hipSuccess || n == 0) return e;
  return hipMemsetAsync(b.p, 0, n * sizeof(T), ctx->stream);
}

}  // namespace

int dl_bake_impl(dl_ctx* ctx, const dl_bake_input* input, const dl_bake_options* options, dl_progress* progress,
                 dl_baked_scene** out) {
  using namespace dl;
  auto t_start = std::chrono::steady_clock::now();
  ctx->timings = dl_timings{};
  if (!input || !options || !out) { ctx->err = "null argument"; return DL_E_ARGUMENT; }
  *out = nullptr;
  const dl_bake_options& opt = *options;
  auto cancelled = [&]() { return progress && progress->cancel; };
  auto stage = [&](uint32_t s, uint64_t total) {
    if (progress) { progress->stage = s; progress->done = 0; progress->total = total; }
  };
  stage(0, 0);
  int st = check_options(ctx, opt);
  if (st) return st;
  if (input->mesh_count && !input->meshes) { ctx->err = "null mesh array"; return DL_E_ARGUMENT; }
  if (input->light_count && !input->lights) { ctx->err = "null light array"; return DL_E_ARGUMENT; }
  for (uint32_t i = 0; i < input->light_count; ++i)
    if (input->lights[i].shape < 0 || input->lights[i].shape > 4) { ctx->err = "unknown light shape"; return DL_E_ARGUMENT; }
  std::vector<const dl_mesh*> meshes;
  for (uint32_t i = 0; i < input->mesh_count; ++i) meshes.push_back(&input->meshes[i]);
  std::stable_sort(meshes.begin(), meshes.end(), [](const dl_mesh* a, const dl_mesh* b) { return a->id < b->id; });
  for (size_t i = 1; i < meshes.size(); ++i)
    if (meshes[i - 1]->id == meshes[i]->id) { ctx->err = "two meshes share BakeId " + hex_id(meshes[i]->id); return DL_E_DUPLICATE_ID; }
  for (auto* m : meshes)
    if ((st = check_mesh(ctx, *m))) return st;
  std::vector<const dl_light*> lights_sorted;
  for (uint32_t i = 0; i < input->light_count; ++i) lights_sorted.push_back(&input->lights[i]);
  std::stable_sort(lights_sorted.begin(), lights_sorted.end(), [](const dl_light* a, const dl_light* b) { return a->id < b->id; });
  for (size_t i = 1; i < lights_sorted.size(); ++i)
    if (lights_sorted[i - 1]->id == lights_sorted[i]->id) { ctx->err = "two lights share BakeId " + hex_id(lights_sorted[i]->id); return DL_E_DUPLICATE_ID; }
  for (auto* l : lights_sorted)
    if ((st = check_light(ctx, *l))) return st;

  // ---- lights in slot order: BVH order, then suns, then skies (baker.rs:35-50) -------------
  std::vector<const dl_light*> local, suns, skies;
  for (auto* l : lights_sorted) {
    if (l->shape == DL_SHAPE_SUN) suns.push_back(l);
    else if (l->shape == DL_SHAPE_SKY) skies.push_back(l);
    else local.push_back(l);
  }
  std::vector<f3> lpos;
  std::vector<float> lrad;
  for (auto* l : local) { lpos.push_back(mk3(l->position[0], l->position[1], l->position[2])); lrad.push_back(l->radius); }
  LightBvhOut lbvh;
  build_light_bvh(lpos, lrad, lbvh);
  std::vector<const dl_light*> lights;
  for (uint32_t i : lbvh.order) lights.push_back(local[i]);
  for (auto* l : suns) lights.push_back(l);
  for (auto* l : skies) lights.push_back(l);
  const uint32_t nl = (uint32_t)lights.size(), nlocal = (uint32_t)local.size();
  std::vector<DevLight> hlights;
  bool bounce_in_scene = false;
  for (auto* l : lights) {
    hlights.push_back(make_dev_light(*l, opt));
    bounce_in_scene |= hlights.back().bakes_bounce != 0;
  }
  if (bounce_in_scene) {
    ctx->err = "bounce lighting (SingleBounce lights) is not supported by this build; bake with skip_bounce";
    return DL_E_ARGUMENT;
  }
  for (auto* m : meshes)
    if (m->alpha && m->occludes)
      for (auto* l : lights)
        if (l->transparency != DL_TRANSPARENCY_DISABLED) {
          ctx->err = "alpha-tested shadows are not supported by this build";
          return DL_E_ARGUMENT;
        }

  // ---- occluders (occluder.rs:80-115) --------------------------------------------------------
  std::vector<float> occ9;
  for (auto* m : meshes) {
    if (!m->occludes) continue;
    for (uint32_t t = 0; t < m->index_count / 3; ++t) {
      f3 v[3];
      for (int k = 0; k < 3; ++k) {
        const float* p = m->positions + 3 * (size_t)m->indices[t * 3 + k];
        v[k] = mk3(p[0], p[1], p[2]);
      }
      f3 n = cross(sub(v[1], v[0]), sub(v[2], v[0]));
      float l2 = len_sq(n);
      if (l2 > 0.0f && finite_(n.x) && finite_(n.y) && finite_(n.z))
        for (int k = 0; k < 3; ++k) { occ9.push_back(v[k].x); occ9.push_back(v[k].y); occ9.push_back(v[k].z); }
    }
  }
  Bvh4 bvh;
  build_bvh4(occ9.data(), (uint32_t)(occ9.size() / 9), ctx->host_threads, bvh);
  ctx->timings.bvh_nodes = bvh.nodes.size();
  ctx->timings.occluder_triangles = occ9.size() / 9;

  // ---- prepared meshes and receiver triangles (mesh_builder.rs:21-80) -----------------------
  const uint32_t nm = (uint32_t)meshes.size();
  std::vector<uint32_t> res(nm), first_tri(nm + 1, 0);
  std::vector<uint64_t> mid(nm), owned_off(nm + 1, 0);
  uint64_t ntri = 0;
  for (uint32_t i = 0; i < nm; ++i) {
    const dl_mesh* m = meshes[i];
    float area = -0.0f;
    for (uint32_t t = 0; t < m->index_count / 3; ++t) {
      const float* a = m->positions + 3 * (size_t)m->indices[t * 3];
      const float* b = m->positions + 3 * (size_t)m->indices[t * 3 + 1];
      const float* c = m->positions + 3 * (size_t)m->indices[t * 3 + 2];
      area = area + tri_area(mk3(a[0], a[1], a[2]), mk3(b[0], b[1], b[2]), mk3(c[0], c[1], c[2]));
    }
    res[i] = lightmap_size(area, opt.pixel_density, opt.max_lightmap_size);
    if (res[i] > 65535) { ctx->err = "mesh " + hex_id(m->id) + ": shadow map side above 65535 texels"; return DL_E_ARGUMENT; }
    mid[i] = m->id;
    first_tri[i] = (uint32_t)ntri;
    ntri += m->index_count / 3;
    owned_off[i + 1] = owned_off[i] + ((uint64_t)res[i] * res[i] + 31) / 32;
  }
  first_tri[nm] = (uint32_t)ntri;
  if (ntri >= (1ull << 31)) { ctx->err = "too many triangles"; return DL_E_ARGUMENT; }
  const uint32_t nt = (uint32_t)ntri;
  std::vector<float> hv((size_t)nt * 9), hn((size_t)nt * 9), huv((size_t)nt * 6);
  std::vector<uint32_t> hmesh(nt), hlocal(nt);
  for (uint32_t i = 0; i < nm; ++i) {
    const dl_mesh* m = meshes[i];
    bool flat = !m->normals || m->normal_count == 0;
    for (uint32_t t = 0; t < m->index_count / 3; ++t) {
      uint32_t g = first_tri[i] + t;
      hmesh[g] = i;
      hlocal[g] = t;
      f3 v[3];
      for (int k = 0; k < 3; ++k) {
        uint32_t c = m->indices[t * 3 + k];
        const float* p = m->positions + 3 * (size_t)c;
        v[k] = mk3(p[0], p[1], p[2]);
        hv[(size_t)g * 9 + 3 * k] = p[0];
        hv[(size_t)g * 9 + 3 * k + 1] = p[1];
        hv[(size_t)g * 9 + 3 * k + 2] = p[2];
        float u = m->uv1[2 * (size_t)c], w = m->uv1[2 * (size_t)c + 1];
        if (!(finite_(u) && finite_(w))) { u = 0.0f; w = 0.0f; }
        huv[(size_t)g * 6 + 2 * k] = u;
        huv[(size_t)g * 6 + 2 * k + 1] = w;
        if (!flat) {
          const float* nn = m->normals + 3 * (size_t)c;
          hn[(size_t)g * 9 + 3 * k] = nn[0];
          hn[(size_t)g * 9 + 3 * k + 1] = nn[1];
          hn[(size_t)g * 9 + 3 * k + 2] = nn[2];
        }
      }
      if (flat) {
        f3 f = normalize_or_zero(cross(sub(v[1], v[0]), sub(v[2], v[0])));
        for (int k = 0; k < 3; ++k) {
          hn[(size_t)g * 9 + 3 * k] = f.x;
          hn[(size_t)g * 9 + 3 * k + 1] = f.y;
          hn[(size_t)g * 9 + 3 * k + 2] = f.z;
        }
      }
    }
  }
  if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

  // ---- device state -----------------------------------------------------------------------
  auto t_up = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  DBuf<Node4> d_nodes;
  DBuf<float> d_occ;
  DBuf<uint32_t> d_overflow;
  HIP_TRY(upload(ctx, d_nodes, bvh.nodes));
  HIP_TRY(upload(ctx, d_occ, bvh.tris));
  HIP_TRY(zeroed(ctx, d_overflow, 1));
  DevScene sc;
  sc.nodes = d_nodes.p;
  sc.tris = reinterpret_cast<const float4*>(d_occ.p);
  sc.ntris = (uint32_t)(occ9.size() / 9);
  sc.stack = 3 * std::max(1u, bvh.depth) + 1;
  sc.overflow = d_overflow.p;
  sc.counters = nullptr;
  DBuf<DevLight> d_lights;
  HIP_TRY(upload(ctx, d_lights, hlights));
  DBuf<float> d_v, d_n, d_uv, d_face;
  DBuf<uint32_t> d_mesh, d_local, d_crop;
  HIP_TRY(upload(ctx, d_v, hv));
  HIP_TRY(upload(ctx, d_n, hn));
  HIP_TRY(upload(ctx, d_uv, huv));
  HIP_TRY(upload(ctx, d_mesh, hmesh));
  HIP_TRY(upload(ctx, d_local, hlocal));
  HIP_TRY(d_face.alloc((size_t)nt * 3));
  HIP_TRY(d_crop.alloc((size_t)nt * 4));
  DevTris tr;
  tr.n = nt;
  tr.v = d_v.p;
  tr.nrm = d_n.p;
  tr.uv = d_uv.p;
  tr.mesh = d_mesh.p;
  tr.local = d_local.p;
  tr.face = d_face.p;
  tr.crop = d_crop.p;
  DBuf<uint32_t> d_res, d_owned, d_visited;
  DBuf<uint64_t> d_mid, d_owned_off;
  DBuf<unsigned long long> d_mrays, d_msamples;
  HIP_TRY(upload(ctx, d_res, res));
  HIP_TRY(upload(ctx, d_mid, mid));
  HIP_TRY(upload(ctx, d_owned_off, owned_off));
  HIP_TRY(zeroed(ctx, d_owned, owned_off[nm] + 1));
  HIP_TRY(zeroed(ctx, d_visited, owned_off[nm] + 1));
  HIP_TRY(zeroed(ctx, d_mrays, nm + 1));
  HIP_TRY(zeroed(ctx, d_msamples, nm + 1));
  DevMeshes ms;
  ms.n = nm;
  ms.res = d_res.p;
  ms.id = d_mid.p;
  ms.owned_off = d_owned_off.p;
  ms.visited_off = d_owned_off.p;
  ms.owned = d_owned.p;
  ms.visited = d_visited.p;
  ms.rays = d_mrays.p;
  ms.samples = d_msamples.p;
  HIP_TRY(hipStreamSynchronize(s));
  ctx->timings.upload_ms = ms_since(t_up);

  // ---- association and rectangles -----------------------------------------------------------
  auto t_tex = std::chrono::steady_clock::now();
  stage(2, nm);
  DBuf<uint64_t> d_lc, d_tc, d_tcc, d_ent_start, d_trace_start, d_cand_start;
  HIP_TRY(zeroed(ctx, d_lc, (size_t)nt + 1));
  HIP_TRY(zeroed(ctx, d_tc, (size_t)nt + 1));
  HIP_TRY(zeroed(ctx, d_tcc, (size_t)nt + 1));
  HIP_TRY(d_ent_start.alloc((size_t)nt + 1));
  HIP_TRY(d_trace_start.alloc((size_t)nt + 1));
  HIP_TRY(d_cand_start.alloc((size_t)nt + 1));
  HIP_TRY(launch_setup(s, tr, ms, d_lights.p, nl, nlocal, d_lc.p, d_tc.p, d_tcc.p));
  HIP_TRY(scan_u64(s, d_lc.p, d_ent_start.p, nt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
  HIP_TRY(scan_u64(s, d_tc.p, d_trace_start.p, nt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
  HIP_TRY(scan_u64(s, d_tcc.p, d_cand_start.p, nt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
  std::vector<uint64_t> h_ent_start, h_trace_start, h_cand_start;
  HIP_TRY(download(ctx, h_ent_start, d_ent_start.p, (size_t)nt + 1));
  HIP_TRY(download(ctx, h_trace_start, d_trace_start.p, (size_t)nt + 1));
  HIP_TRY(download(ctx, h_cand_start, d_cand_start.p, (size_t)nt + 1));
  const uint64_t nent = h_ent_start[nt];
  DBuf<uint32_t> d_ent_light, d_ent_tri, d_ent_region;
  HIP_TRY(d_ent_light.alloc(nent + 1));
  HIP_TRY(d_ent_tri.alloc(nent + 1));
  HIP_TRY(launch_assoc_fill(s, tr, ms, d_lights.p, nl, nlocal, d_ent_start.p, d_ent_light.p, d_ent_tri.p));
  std::vector<uint32_t> h_ent_light, h_crop;
  HIP_TRY(download(ctx, h_ent_light, d_ent_light.p, nent));
  HIP_TRY(download(ctx, h_crop, d_crop.p, (size_t)nt * 4));

  // ---- light regions per mesh (mesh_bake.rs:175-192) ----------------------------------------
  std::vector<uint32_t> reg_rect, reg_bits, reg_mesh, h_ent_region(nent);
  std::vector<uint64_t> reg_off;
  std::vector<uint32_t> mesh_reg_first(nm + 1, 0);
  {
    std::vector<int32_t> slot_of(nl, -1);
    std::vector<uint32_t> rect(4 * (size_t)nl);
    std::vector<uint32_t> touched;
    uint64_t pool = 0;
    for (uint32_t i = 0; i < nm; ++i) {
      mesh_reg_first[i] = (uint32_t)reg_bits.size();
      touched.clear();
      for (uint32_t t = first_tri[i]; t < first_tri[i + 1]; ++t)
        for (uint64_t e = h_ent_start[t]; e < h_ent_start[t + 1]; ++e) {
          uint32_t l = h_ent_light[e];
          const uint32_t* c = &h_crop[4 * (size_t)t];
          if (slot_of[l] < 0) {
            slot_of[l] = 0;
            touched.push_back(l);
            std::memcpy(&rect[4 * (size_t)l], c, 16);
          } else {
            uint32_t* r = &rect[4 * (size_t)l];
            r[0] = std::min(r[0], c[0]); r[1] = std::min(r[1], c[1]);
            r[2] = std::max(r[2], c[2]); r[3] = std::max(r[3], c[3]);
          }
        }
      std::sort(touched.begin(), touched.end());
      for (uint32_t l : touched) {
        slot_of[l] = (int32_t)reg_bits.size();
        const uint32_t* r = &rect[4 * (size_t)l];
        reg_rect.insert(reg_rect.end(), r, r + 4);
        uint32_t bits = hlights[l].bits;
        reg_bits.push_back(bits);
        reg_mesh.push_back(i);
        reg_off.push_back(pool);
        uint64_t texels = (uint64_t)(r[2] - r[0] + 1) * (r[3] - r[1] + 1);
        pool += bits == 1 ? (texels + 31) / 32 : (texels + 3) / 4;
      }
      for (uint32_t t = first_tri[i]; t < first_tri[i + 1]; ++t)
        for (uint64_t e = h_ent_start[t]; e < h_ent_start[t + 1]; ++e) h_ent_region[e] = (uint32_t)slot_of[h_ent_light[e]];
      for (uint32_t l : touched) slot_of[l] = -1;
    }
    mesh_reg_first[nm] = (uint32_t)reg_bits.size();
    reg_off.push_back(pool);
  }
  const uint32_t nreg = (uint32_t)reg_bits.size();
  DBuf<uint32_t> d_reg_rect, d_reg_bits, d_reg_mesh, d_pool;
  DBuf<uint64_t> d_reg_off;
  HIP_TRY(upload(ctx, d_reg_rect, reg_rect));
  HIP_TRY(upload(ctx, d_reg_bits, reg_bits));
  HIP_TRY(upload(ctx, d_reg_mesh, reg_mesh));
  HIP_TRY(upload(ctx, d_reg_off, reg_off));
  HIP_TRY(upload(ctx, d_ent_region, h_ent_region));
  HIP_TRY(zeroed(ctx, d_pool, reg_off[nreg] + 1));
  DevRegions rg;
  rg.rect = d_reg_rect.p;
  rg.bits = d_reg_bits.p;
  rg.off = d_reg_off.p;
  rg.mesh = d_reg_mesh.p;
  rg.pool = d_pool.p;
  DevEntries ent;
  ent.n = nent;
  ent.light = d_ent_light.p;
  ent.region = d_ent_region.p;
  ent.tri = d_ent_tri.p;
  ent.start = d_ent_start.p;

  // ---- texels: owned, traced, burial, compaction (mesh_texels.rs:41-108) ---------------------
  const uint64_t ntrace = h_trace_start[nt], ncand = h_cand_start[nt];
  HIP_TRY(launch_owned(s, tr, ms, d_trace_start.p, ntrace));
  DBuf<unsigned long long> d_cnt;
  HIP_TRY(zeroed(ctx, d_cnt, (size_t)nt + 1));
  DevTexels none{};
  HIP_TRY(launch_texels(s, tr, ms, d_tc.p, d_cand_start.p, ncand, 0, d_cnt.p, nullptr, none));
  DBuf<uint64_t> d_tex0_start;
  HIP_TRY(d_tex0_start.alloc((size_t)nt + 1));
  HIP_TRY(scan_u64(s, reinterpret_cast<const uint64_t*>(d_cnt.p), d_tex0_start.p, nt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
  uint64_t ntex0 = 0;
  HIP_TRY(hipMemcpyAsync(&ntex0, d_tex0_start.p + nt, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  DBuf<uint32_t> d_t0xy, d_t0tri;
  DBuf<float4> d_t0pos, d_t0nrm;
  HIP_TRY(d_t0xy.alloc(ntex0 + 1));
  HIP_TRY(d_t0tri.alloc(ntex0 + 1));
  HIP_TRY(d_t0pos.alloc(ntex0 + 1));
  HIP_TRY(d_t0nrm.alloc(ntex0 + 1));
  DevTexels tx0{d_t0xy.p, d_t0pos.p, d_t0nrm.p, d_t0tri.p};
  HIP_TRY(hipMemsetAsync(d_cnt.p, 0, ((size_t)nt + 1) * 8, s));
  HIP_TRY(launch_texels(s, tr, ms, d_tc.p, d_cand_start.p, ncand, 1, d_cnt.p, d_tex0_start.p, tx0));
  DBuf<uint8_t> d_buried;
  HIP_TRY(d_buried.alloc(ntex0 + 1));
  HIP_TRY(hipMemsetAsync(d_cnt.p, 0, ((size_t)nt + 1) * 8, s));
  HIP_TRY(launch_burial(s, sc, tr, ms, tx0, ntex0, opt.seed, ctx->trig, d_buried.p, d_cnt.p, ctx->burial_block));
  DBuf<uint64_t> d_tex_start;
  HIP_TRY(d_tex_start.alloc((size_t)nt + 1));
  HIP_TRY(scan_u64(s, reinterpret_cast<const uint64_t*>(d_cnt.p), d_tex_start.p, nt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
  std::vector<uint64_t> h_tex_start;
  HIP_TRY(download(ctx, h_tex_start, d_tex_start.p, (size_t)nt + 1));
  const uint64_t ntex = h_tex_start[nt];
  DBuf<uint32_t> d_txy, d_ttri;
  DBuf<float4> d_tpos, d_tnrm;
  HIP_TRY(d_txy.alloc(ntex + 1));
  HIP_TRY(d_ttri.alloc(ntex + 1));
  HIP_TRY(d_tpos.alloc(ntex + 1));
  HIP_TRY(d_tnrm.alloc(ntex + 1));
  DevTexels tx{d_txy.p, d_tpos.p, d_tnrm.p, d_ttri.p};
  HIP_TRY(hipMemsetAsync(d_cnt.p, 0, ((size_t)nt + 1) * 8, s));
  HIP_TRY(launch_compact(s, tx0, ntex0, d_buried.p, d_tex_start.p, d_cnt.p, tx));
  HIP_TRY(hipStreamSynchronize(s));
  d_t0xy.release();
  d_t0tri.release();
  d_t0pos.release();
  d_t0nrm.release();
  d_buried.release();
  ctx->timings.texels_ms = ms_since(t_tex);
  if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

  // ---- shadow rays (mesh_bake.rs:203-267) ----------------------------------------------------
  auto t_trace = std::chrono::steady_clock::now();
  std::vector<uint64_t> h_units(nt + 1, 0);
  for (uint32_t t = 0; t < nt; ++t) {
    uint64_t texels = h_tex_start[t + 1] - h_tex_start[t];
    uint64_t k = h_ent_start[t + 1] - h_ent_start[t];
    h_units[t] = k * ((texels + 63) / 64);
  }
  std::vector<uint64_t> h_unit_start(nt + 1, 0);
  for (uint32_t t = 0; t < nt; ++t) h_unit_start[t + 1] = h_unit_start[t] + h_units[t];
  DBuf<uint64_t> d_unit_start;
  HIP_TRY(upload(ctx, d_unit_start, h_unit_start));
  TraceParams tp;
  tp.seed = opt.seed;
  tp.area_samples = opt.area_light_samples;
  tp.area_min = opt.area_light_min_samples;
  tp.sky_samples = opt.sky_samples;
  hipEvent_t ev0, ev1;
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  HIP_TRY(hipEventRecord(ev0, s));
  DBuf<unsigned long long> d_counters;
  DevScene sct = sc;
  if (ctx->counting) {
    HIP_TRY(zeroed(ctx, d_counters, 2));
    sct.counters = d_counters.p;
  }
  HIP_TRY(launch_trace(s, sct, tr, ms, d_lights.p, tx, d_tex_start.p, ent, rg, d_unit_start.p, h_unit_start[nt], tp, ctx->trig,
                       ctx->trace_block));
  HIP_TRY(hipEventRecord(ev1, s));
  HIP_TRY(hipStreamSynchronize(s));
  float kms = 0.0f;
  (void)hipEventElapsedTime(&kms, ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  ctx->timings.trace_kernel_ms = kms;
  if (ctx->counting) {
    unsigned long long c[2] = {0, 0};
    HIP_TRY(hipMemcpy(c, d_counters.p, 16, hipMemcpyDeviceToHost));
    ctx->timings.trace_node_visits = c[0];
    ctx->timings.trace_tri_tests = c[1];
  }
  ctx->timings.trace_launches = 1;
  ctx->timings.trace_ms = ms_since(t_trace);
  if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

  // ---- optimize, seams, crop, serialize (shadows.rs, seams.rs, triangle_words.rs) ------------
  auto t_post = std::chrono::steady_clock::now();
  DBuf<uint8_t> d_keep;
  HIP_TRY(d_keep.alloc(nent + 1));
  HIP_TRY(launch_optimize(s, tr, ms, ent, rg, nullptr, d_keep.p));
  std::vector<uint8_t> h_keep;
  HIP_TRY(download(ctx, h_keep, d_keep.p, nent));
  const uint32_t stride = bounce_in_scene ? 3 : 2;
  std::vector<uint8_t> used(nreg, 0);
  std::vector<uint64_t> mesh_removed(nm, 0), mesh_base(nm + 1, 0);
  std::vector<uint32_t> ld_off(nt, 0), kept_shadow_off;
  std::vector<uint64_t> tri_kept_start(nt + 1, 0), kept_entry, kept_word_start, kept_out;
  {
    uint64_t nk = 0;
    for (uint64_t e = 0; e < nent; ++e) nk += h_keep[e];
    kept_entry.reserve(nk);
    kept_word_start.reserve(nk + 1);
    kept_out.reserve(nk);
    kept_shadow_off.reserve(nk);
    uint64_t flat = 0;
    for (uint32_t i = 0; i < nm; ++i) {
      uint32_t T = first_tri[i + 1] - first_tri[i];
      uint64_t w = 4ull * T;
      for (uint32_t t = first_tri[i]; t < first_tri[i + 1]; ++t) {
        ld_off[t] = (uint32_t)w;
        uint64_t kept = 0;
        for (uint64_t e = h_ent_start[t]; e < h_ent_start[t + 1]; ++e) kept += h_keep[e];
        mesh_removed[i] += (h_ent_start[t + 1] - h_ent_start[t]) - kept;
        w += 1 + stride * kept;
      }
      for (uint32_t t = first_tri[i]; t < first_tri[i + 1]; ++t) {
        tri_kept_start[t] = kept_entry.size();
        const uint32_t* c = &h_crop[4 * (size_t)t];
        uint64_t texels = (uint64_t)(c[2] - c[0] + 1 + 4) * (c[3] - c[1] + 1 + 4);
        for (uint64_t e = h_ent_start[t]; e < h_ent_start[t + 1]; ++e) {
          if (!h_keep[e]) continue;
          uint32_t r = h_ent_region[e];
          used[r] = 1;
          uint32_t per = 32 / reg_bits[r];
          uint64_t words = (texels + per - 1) / per;
          kept_entry.push_back(e);
          kept_word_start.push_back(flat);
          kept_out.push_back(mesh_base[i] + w);
          kept_shadow_off.push_back((uint32_t)w);
          flat += words;
          w += words;
        }
      }
      mesh_base[i + 1] = mesh_base[i] + w;
    }
    tri_kept_start[nt] = kept_entry.size();
    kept_word_start.push_back(flat);
  }
  std::vector<uint32_t> used_list;
  std::vector<uint64_t> used_area(1, 0);
  for (uint32_t r = 0; r < nreg; ++r)
    if (used[r]) {
      used_list.push_back(r);
      const uint32_t* q = &reg_rect[4 * (size_t)r];
      used_area.push_back(used_area.back() + (uint64_t)(q[2] - q[0] + 1) * (q[3] - q[1] + 1));
    }
  DBuf<uint32_t> d_used;
  DBuf<uint64_t> d_used_area;
  HIP_TRY(upload(ctx, d_used, used_list));
  HIP_TRY(upload(ctx, d_used_area, used_area));
  HIP_TRY(launch_pad_seams(s, ms, rg, d_used.p, d_used_area.p, (uint32_t)used_list.size(), used_area.back()));
  const uint64_t total_words = mesh_base[nm];
  const uint64_t nkept = kept_entry.size();
  DBuf<uint32_t> d_out, d_ld_off, d_kept_shadow_off, d_first;
  DBuf<uint64_t> d_mesh_base, d_tri_kept_start, d_kept_entry, d_kept_word_start, d_kept_out;
  HIP_TRY(d_out.alloc(total_words + 1));
  HIP_TRY(upload(ctx, d_mesh_base, mesh_base));
  HIP_TRY(upload(ctx, d_ld_off, ld_off));
  HIP_TRY(upload(ctx, d_tri_kept_start, tri_kept_start));
  HIP_TRY(upload(ctx, d_kept_entry, kept_entry));
  HIP_TRY(upload(ctx, d_kept_word_start, kept_word_start));
  HIP_TRY(upload(ctx, d_kept_out, kept_out));
  HIP_TRY(upload(ctx, d_kept_shadow_off, kept_shadow_off));
  HIP_TRY(upload(ctx, d_first, first_tri));
  HIP_TRY(launch_serialize(s, tr, d_mesh_base.p, d_ld_off.p, d_tri_kept_start.p, d_kept_entry.p, d_kept_shadow_off.p, nullptr,
                           d_ent_light.p, stride, d_first.p, d_out.p));
  HIP_TRY(launch_crop(s, tr, ent, rg, d_kept_entry.p, d_kept_word_start.p, nkept, kept_word_start.back(), d_kept_out.p, d_out.p));
  HIP_TRY(hipStreamSynchronize(s));
  ctx->timings.post_ms = ms_since(t_post);

  // ---- distance cubes (distance_cubes.rs:29-72) -----------------------------------------------
  auto t_maps = std::chrono::steady_clock::now();
  stage(3, (uint64_t)nlocal * 6);
  std::unique_ptr<dl_baked_scene> scene(new dl_baked_scene());
  if (opt.dynamic_geometry != DL_DYNAMIC_LIGHTING_ONLY && nlocal > 0) {
    DBuf<uint32_t> d_cubes;
    HIP_TRY(d_cubes.alloc((size_t)nlocal * 6 * 2048));
    HIP_TRY(launch_distance_cubes(s, sc, d_lights.p, nlocal, d_cubes.p));
    HIP_TRY(download(ctx, scene->distance_cubes, d_cubes.p, (size_t)nlocal * 6 * 2048));
  }
  ctx->timings.maps_ms = ms_since(t_maps);

  // ---- read back ----------------------------------------------------------------------------
  auto t_down = std::chrono::steady_clock::now();
  std::vector<uint32_t> words;
  HIP_TRY(download(ctx, words, d_out.p, total_words));
  std::vector<unsigned long long> h_rays, h_samples;
  HIP_TRY(download(ctx, h_rays, d_mrays.p, nm));
  HIP_TRY(download(ctx, h_samples, d_msamples.p, nm));
  uint32_t overflow = 0;
  HIP_TRY(hipMemcpy(&overflow, d_overflow.p, 4, hipMemcpyDeviceToHost));
  if (overflow) { ctx->err = "BVH traversal stack overflow"; return DL_E_DEVICE; }
  ctx->timings.download_ms = ms_since(t_down);

  scene->settings.bounce = bounce_in_scene;
  scene->settings.colored_bounce = bounce_in_scene && opt.colored_bounce;
  scene->settings.dynamic_geometry = opt.dynamic_geometry;
  scene->settings.pixel_density = opt.pixel_density;
  scene->settings.max_lightmap_size = opt.max_lightmap_size;
  scene->settings.distance_cube_resolution = 64;
  for (size_t i = 0; i < lights.size(); ++i) {
    const dl_light* l = lights[i];
    dl_baked_light b{};
    b.id = l->id;
    std::memcpy(b.origin, l->position, 12);
    b.radius = l->radius;
    b.shape = l->shape;
    std::memcpy(b.rotation, l->rotation, 16);
    std::memcpy(b.size, l->size, 8);
    b.angular_diameter = l->angular_diameter;
    b.shadow_bits_per_texel = hlights[i].bits;
    b.bounce_compression = hlights[i].bakes_bounce ? (l->bounce_compression >= 0 ? l->bounce_compression : opt.bounce_compression) : -1;
    scene->lights.push_back(b);
  }
  scene->bvh = lbvh.nodes;
  scene->stats.meshes = nm;
  scene->stats.triangles = ntri;
  for (uint32_t i = 0; i < nm; ++i) {
    dl_baked_mesh_out mo;
    mo.id = meshes[i]->id;
    mo.content_hash = meshes[i]->content_hash;
    mo.resolution = res[i];
    mo.occludes = meshes[i]->occludes;
    mo.words.assign(words.begin() + (long)mesh_base[i], words.begin() + (long)mesh_base[i + 1]);
    scene->stats.triangle_words += mo.words.size();
    scene->meshes.push_back(std::move(mo));
    scene->stats.rays += h_rays[i];
    scene->stats.samples += h_samples[i];
    scene->stats.lights_removed += mesh_removed[i];
  }
  for (uint32_t t = 0; t < nt; ++t) scene->stats.texels += h_tex_start[t + 1] - h_tex_start[t];
  ctx->timings.trace_rays = scene->stats.rays;
  stage(4, 0);
  ctx->timings.total_ms = ms_since(t_start);
  *out = scene.release();
  return DL_OK;
}

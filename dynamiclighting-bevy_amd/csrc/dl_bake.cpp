// Orchestration of one bake (baker.rs:27-203).
//
// Scene-level passes run once: validation and ordering on the host, the occluder BVH (host
// build, uploaded once), photon cubes of every bounce light, and light association of every
// receiver triangle.  Receiver meshes are then processed in batches sized to an HBM budget:
//
//   owned texels -> traced texels -> burial -> compaction -> shadow rays -> bounce gather,
//   dilate, blur -> optimize -> seam padding -> crop/dilate/pack + serialization
//
// Host work between passes is O(triangles + light entries) integer bookkeeping (regions,
// windows and word offsets); every ray and texel is processed on the device.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "dl_host.h"

using namespace dl;

namespace {

struct Batch {
  uint32_t m0, m1;  // meshes [m0, m1)
};

// Gaussian weights of the bounce blur (bounce_map.rs:140-150): 7 taps, sigma 5, glibc expf.
BlurWeights blur_weights() {
  BlurWeights b;
  const float sigma2 = 5.0f * 5.0f;
  float sum = 0.0f;
  for (int i = -3; i <= 3; ++i) {
    b.w[i + 3] = (1.0f / sqrtf((2.0f * kPi) * sigma2)) * expf(-(float)(i * i) / (2.0f * sigma2));
    sum += b.w[i + 3];
  }
  float s2 = 0.0f;
  for (int i = 0; i < 7; ++i) {
    b.w[i] = b.w[i] / sum;
    s2 += b.w[i];
  }
  b.sum = s2;
  return b;
}

uint32_t bounce_bits(int32_t c) { return c == DL_BOUNCE_BITS8 ? 8u : c == DL_BOUNCE_BITS6 ? 6u : c == DL_BOUNCE_BITS5 ? 5u : 4u; }

// Burial of a batch's ntex texels (mesh_texels.rs:57-87) as ray streams, in chunks of texels:
// rays 0-3 of every texel of the chunk through the closest-hit stream, then rays 4-7 of the
// texels those leave undecided.  Writes buried[i] and counts the kept texels per triangle.
// With `counters` (6 words, counting on), counters[4..5] receive the node fetches and triangle
// tests of the burial rays; *rays counts the rays cast and *ms the device time.
// With `esc`, the rays that hit nothing are appended there (Capture::escaping's layout).
hipError_t burial_streamed(dl_ctx* ctx, DevScene sc, const DevTris& tr, const DevMeshes& ms, const DevTexels& tx, uint64_t ntex,
                           uint64_t seed, uint8_t* buried, unsigned long long* kept, unsigned long long* counters,
                           uint64_t* rays_cast, double* ms_spent, const Capture* esc) {
  if (ntex == 0) return hipSuccess;
  if (ntex >= (1ull << 32)) return hipErrorInvalidValue;  // texels are listed by 32-bit index
  hipStream_t s = ctx->stream;
  float* esc_out = esc && esc->escaping ? reinterpret_cast<float*>(esc->escaping->p) : nullptr;
  unsigned long long* esc_count = esc_out ? reinterpret_cast<unsigned long long*>(esc->escaping->p + 8 * esc->escaping_cap) : nullptr;
  if (ctx->burial_lanes && !esc_out) {  // one lane per texel instead (counters: node fetches and triangle tests only)
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess) (void)hipEventRecord(ev[0], s);
    sc.counters = counters ? counters + 4 : nullptr;
    hipError_t e = launch_burial_lanes(s, sc, tr, ms, tx, ntex, seed, ctx->trig, buried, kept, ctx->burial_lanes);
    float t = 0.0f;
    if (e == hipSuccess && ev[1] && hipEventRecord(ev[1], s) == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess &&
        hipEventElapsedTime(&t, ev[0], ev[1]) == hipSuccess)
      *ms_spent += t;
    for (hipEvent_t x : ev)
      if (x) (void)hipEventDestroy(x);
    return e;
  }
  const uint64_t chunk = std::min<uint64_t>(ntex, std::max<uint64_t>(1ull << 20, ctx->gather.chunk_rays / 8));
  GatherLaunch gl = ctx->gather;
  gl.order = 0;  // slots go out in storage order
  if (ctx->burial_block) gl.block = ctx->burial_block;
  // Burial rays (unbounded closest-hit) vary widely in length: a wave's finished lanes take new
  // rays once 8 of them are idle rather than when all are (0.66 s instead of 0.77 s of burial per
  // bake of the 200k-triangle interior).
  gl.refill = std::min<uint32_t>(gl.refill, 8u);
  DBuf<float4> rays, hits;
  DBuf<uint64_t> total;
  DBuf<uint8_t> back;
  DBuf<uint32_t> list;
  DBuf<unsigned long long> ctr;  // queue of pass 1, queue of pass 2, undecided count
  hipError_t e;
  if ((e = rays.alloc(8 * chunk)) || (e = hits.alloc(4 * chunk)) || (e = total.alloc(1)) || (e = back.alloc(ntex)) ||
      (e = list.alloc(chunk)) || (e = ctr.alloc(3)))
    return e;
  unsigned int* q1 = reinterpret_cast<unsigned int*>(ctr.p);
  unsigned int* q2 = reinterpret_cast<unsigned int*>(ctr.p + 1);
  sc.counters = counters ? counters + 4 : nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess) (void)hipEventRecord(ev[0], s);
  for (uint64_t c0 = 0; c0 < ntex; c0 += chunk) {
    const uint64_t n = std::min(chunk, ntex - c0);
    if ((e = hipMemsetAsync(ctr.p, 0, 3 * sizeof(unsigned long long), s)) ||
        (e = launch_burial_rays(s, tr, ms, tx, c0, nullptr, nullptr, n, 0, 4, seed, ctx->trig, rays.p, hits.p, total.p)) ||
        (e = launch_closest_stream(s, sc, rays.p, total.p, 4 * n, n, hits.p, q1, gl)) ||
        (esc_out && (e = launch_burial_misses(s, rays.p, hits.p, total.p, 4 * n, esc_out, esc->escaping_cap, esc_count))) ||
        (e = launch_burial_decide(s, sc, tx, c0, nullptr, nullptr, n, 0, 4, rays.p, hits.p, back.p, buried, kept, list.p, ctr.p + 2)) ||
        (e = launch_burial_rays(s, tr, ms, tx, 0, list.p, ctr.p + 2, n, 4, 4, seed, ctx->trig, rays.p, hits.p, total.p)) ||
        (e = launch_closest_stream(s, sc, rays.p, total.p, 4 * n, n, hits.p, q2, gl)) ||
        (esc_out && (e = launch_burial_misses(s, rays.p, hits.p, total.p, 4 * n, esc_out, esc->escaping_cap, esc_count))) ||
        (e = launch_burial_decide(s, sc, tx, 0, list.p, ctr.p + 2, n, 4, 4, rays.p, hits.p, back.p, buried, kept, nullptr, nullptr)))
      break;
    if (counters) {
      unsigned long long undecided = 0;
      if ((e = hipMemcpyAsync(&undecided, ctr.p + 2, 8, hipMemcpyDeviceToHost, s)) || (e = hipStreamSynchronize(s))) break;
      *rays_cast += 4 * (n + undecided);
    }
  }
  if (ev[1]) {
    float t = 0.0f;
    if (e == hipSuccess && hipEventRecord(ev[1], s) == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess &&
        hipEventElapsedTime(&t, ev[0], ev[1]) == hipSuccess)
      *ms_spent += t;
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  return e;
}

}  // namespace

int dl_bake_impl(dl_ctx* ctx, const dl_bake_input* input, const dl_bake_options* options, dl_progress* progress,
                 dl_baked_scene** out, Resident* resident, Capture* cap) {
  auto t_start = std::chrono::steady_clock::now();
  ctx->cache.stream = ctx->stream;
  const CacheScope cache_scope(&ctx->cache);  // before every device buffer of the bake
  ctx->timings = dl_timings{};
  ctx->captured.clear();
  ctx->captured_values.clear();
  if (!input || !options || !out) { ctx->err = "null argument"; return DL_E_ARGUMENT; }
  *out = nullptr;
  const dl_bake_options& opt = *options;
  auto cancelled = [&]() { return progress && progress->cancel; };
  auto stage = [&](uint32_t s, uint64_t total) {
    if (progress) { progress->stage = s; progress->done = 0; progress->total = total; }
  };
  auto advance = [&](uint64_t n) {
    if (progress) progress->done = progress->done + n;
  };
  stage(0, 0);
  // resident geometry: `use_res` bakes the resident scene, `fill_res` fills it from this bake
  const bool use_res = resident && resident->ready, fill_res = resident && !resident->ready;

  // ---- validation and ordering (baker.rs:31-50, 205-233) -------------------------------------
  int st = check_options(ctx, opt);
  if (st) return st;
  if (!use_res && input->mesh_count && !input->meshes) { ctx->err = "null mesh array"; return DL_E_ARGUMENT; }
  if (input->light_count && !input->lights) { ctx->err = "null light array"; return DL_E_ARGUMENT; }
  for (uint32_t i = 0; i < input->light_count; ++i)
    if (input->lights[i].shape < 0 || input->lights[i].shape > 4) { ctx->err = "unknown light shape"; return DL_E_ARGUMENT; }
  std::vector<const dl_mesh*> meshes;
  if (!use_res)
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
  std::vector<int32_t> comp_of_slot(nl + 1, 0), cube_of_slot(nl + 1, -1);
  std::vector<uint32_t> cube_slot;
  bool bounce_in_scene = false, dithered_bounce = false;
  for (uint32_t i = 0; i < nl; ++i) {
    const dl_light* l = lights[i];
    hlights.push_back(make_dev_light(*l, opt));
    bool bb = hlights.back().bakes_bounce != 0;
    bounce_in_scene |= bb;
    comp_of_slot[i] = l->bounce_compression >= 0 ? l->bounce_compression : opt.bounce_compression;
    if (bb) {
      dithered_bounce |= comp_of_slot[i] != DL_BOUNCE_BITS8;
      if (l->shape != DL_SHAPE_SUN) {
        cube_of_slot[i] = (int32_t)cube_slot.size();
        cube_slot.push_back(i);
      }
    }
  }
  bool alpha_lights = false;
  for (auto* l : lights) alpha_lights |= l->transparency != DL_TRANSPARENCY_DISABLED;

  // ---- occluders (occluder.rs:80-115) with their normals and albedo --------------------------
  std::vector<float> occ9_local, occ_nrm, occ_alb;
  std::vector<float>& occ9 = use_res ? resident->occ9 : occ9_local;
  std::vector<uint32_t> occ_flags;
  // alpha-tested occluders (occluder.rs:147-166): texture per occluder, corner UV0, opaque bits
  std::vector<uint32_t> occ_tex, tex_info, tex_bits;
  std::vector<float> occ_uv0;
  bool mesh_alpha = use_res && resident->has_alpha;
  for (auto* m : meshes) mesh_alpha |= m->occludes && m->alpha && m->uv0 && m->alpha->width > 0 && m->alpha->height > 0;
  // alpha data is used when a light tests alpha; a resident scene keeps it for later lights
  const bool any_alpha = alpha_lights && mesh_alpha;
  const bool build_alpha = any_alpha || (fill_res && mesh_alpha);
  // alpha textures first (per mesh, in order), then the occluder records in mesh and triangle
  // order: blocks of triangles count their valid triangles, then fill at the scanned offsets
  std::vector<uint32_t> mesh_tex(meshes.size(), 0xffffffffu);
  for (size_t mi = 0; mi < meshes.size(); ++mi) {
    const dl_mesh* m = meshes[mi];
    if (!m->occludes) continue;
    if (build_alpha && m->alpha && m->uv0 && m->alpha->width > 0 && m->alpha->height > 0) {
      const dl_alpha_texture* a = m->alpha;
      mesh_tex[mi] = (uint32_t)(tex_info.size() / 4);
      const uint64_t texels = (uint64_t)a->width * a->height;
      tex_info.push_back(a->width);
      tex_info.push_back(a->height);
      tex_info.push_back((uint32_t)tex_bits.size());
      tex_info.push_back(0);
      const size_t first = tex_bits.size();
      tex_bits.resize(first + (size_t)((texels + 31) / 32), 0u);
      for (uint64_t i = 0; i < texels; ++i)
        if ((float)a->rgba8[i * 4 + 3] / 255.0f >= a->cutoff) tex_bits[first + i / 32] |= 1u << (i % 32);
    }
  }
  struct OccBlock {
    uint32_t mesh, t0, t1;
    uint64_t at;  // valid triangles of the block, then its first occluder
  };
  constexpr uint32_t kOccBlock = 1u << 14;
  std::vector<OccBlock> oblocks;
  for (size_t mi = 0; mi < meshes.size(); ++mi) {
    if (!meshes[mi]->occludes) continue;
    const uint32_t n = meshes[mi]->index_count / 3;
    for (uint32_t t0 = 0; t0 < n; t0 += kOccBlock) oblocks.push_back({(uint32_t)mi, t0, std::min(n, t0 + kOccBlock), 0});
  }
  auto occ_corners = [](const dl_mesh* m, uint32_t t, f3* v) {
    for (int k = 0; k < 3; ++k) {
      const float* p = m->positions + 3 * (size_t)m->indices[t * 3 + k];
      v[k] = mk3(p[0], p[1], p[2]);
    }
    const f3 n = cross(sub(v[1], v[0]), sub(v[2], v[0]));
    return len_sq(n) > 0.0f && finite_(n.x) && finite_(n.y) && finite_(n.z);
  };
  run_parallel(oblocks.size(), ctx->host_threads, [&](uint64_t k) {
    OccBlock& b = oblocks[k];
    const dl_mesh* m = meshes[b.mesh];
    f3 v[3];
    for (uint32_t t = b.t0; t < b.t1; ++t) b.at += occ_corners(m, t, v) ? 1u : 0u;
  });
  uint64_t nvalid = 0;
  for (OccBlock& b : oblocks) {
    const uint64_t c = b.at;
    b.at = nvalid;
    nvalid += c;
  }
  if (!use_res) {
    occ9.resize(9 * nvalid);
    occ_nrm.resize(9 * nvalid);
    occ_flags.resize(nvalid);
    occ_alb.resize(3 * nvalid);
    if (build_alpha) {
      occ_tex.resize(nvalid);
      occ_uv0.resize(6 * nvalid);
    }
  }
  run_parallel(oblocks.size(), ctx->host_threads, [&](uint64_t k) {
    const OccBlock& b = oblocks[k];
    const dl_mesh* m = meshes[b.mesh];
    const uint32_t tex = mesh_tex[b.mesh];
    const bool has_n = m->normals && m->normal_count;
    uint64_t o = b.at;
    f3 v[3];
    for (uint32_t t = b.t0; t < b.t1; ++t) {
      if (!occ_corners(m, t, v)) continue;
      for (int k3 = 0; k3 < 3; ++k3) {
        occ9[9 * o + 3 * k3] = v[k3].x;
        occ9[9 * o + 3 * k3 + 1] = v[k3].y;
        occ9[9 * o + 3 * k3 + 2] = v[k3].z;
        for (int a = 0; a < 3; ++a)
          occ_nrm[9 * o + 3 * k3 + a] = has_n ? m->normals[3 * (size_t)m->indices[t * 3 + k3] + a] : 0.0f;
      }
      occ_flags[o] = has_n ? 1u : 0u;
      for (int a = 0; a < 3; ++a) occ_alb[3 * o + a] = m->albedo[a];
      if (build_alpha) {
        occ_tex[o] = tex;
        for (int k3 = 0; k3 < 3; ++k3)
          for (int a = 0; a < 2; ++a)
            occ_uv0[6 * o + 2 * k3 + a] = tex != 0xffffffffu ? m->uv0[2 * (size_t)m->indices[t * 3 + k3] + a] : 0.0f;
      }
      o++;
    }
  });
  const uint32_t nocc = use_res ? resident->nocc : (uint32_t)occ_flags.size();
  Bvh bvh;
  const auto t_bvh = std::chrono::steady_clock::now();
  if (!use_res) build_bvh(occ9.data(), nocc, ctx->host_threads, bvh);
  ctx->timings.bvh_ms = ms_since(t_bvh);
  ctx->timings.bvh_nodes = use_res ? resident->nodes : bvh.nodes.size();
  ctx->timings.occluder_triangles = nocc;

  // ---- prepared meshes and receiver triangles (mesh_builder.rs:21-80) -----------------------
  std::vector<MeshMeta> meta_local;
  std::vector<MeshMeta>& meta = use_res ? resident->meshes : meta_local;
  for (auto* m : meshes) meta.push_back(MeshMeta{m->id, m->content_hash, m->occludes});
  const uint32_t nm = (uint32_t)meta.size();
  std::vector<uint32_t> res_local(use_res ? 0 : nm), first_local(use_res ? 0 : nm + 1, 0);
  std::vector<uint64_t> mid_local(use_res ? 0 : nm), owned_local(use_res ? 0 : nm + 1, 0);
  std::vector<uint32_t>& res = use_res ? resident->res : res_local;
  std::vector<uint32_t>& first_tri = use_res ? resident->first_tri : first_local;
  std::vector<uint64_t>& mid = use_res ? resident->mid : mid_local;
  std::vector<uint64_t>& owned_off = use_res ? resident->owned_off : owned_local;
  uint64_t ntri = use_res ? resident->ntri : 0;
  for (uint32_t i = 0; i < (use_res ? 0u : nm); ++i) {
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
  if (!use_res) first_tri[nm] = (uint32_t)ntri;
  if (ntri >= (1ull << 31)) { ctx->err = "too many triangles"; return DL_E_ARGUMENT; }
  const uint32_t nt = (uint32_t)ntri;
  const size_t nth = use_res ? 0 : nt;  // the host copies of the receiver triangles
  std::vector<float> hv(nth * 9), hn(nth * 9), huv(nth * 6);
  std::vector<uint32_t> hmesh(nth), hlocal(nth);
  struct RecvBlock {
    uint32_t mesh, t0, t1;
  };
  std::vector<RecvBlock> rblocks;
  for (uint32_t i = 0; i < (use_res ? 0u : nm); ++i) {
    const uint32_t n = meshes[i]->index_count / 3;
    for (uint32_t t0 = 0; t0 < n; t0 += kOccBlock) rblocks.push_back({i, t0, std::min(n, t0 + kOccBlock)});
  }
  run_parallel(rblocks.size(), ctx->host_threads, [&](uint64_t bk) {
    const uint32_t i = rblocks[bk].mesh;
    const dl_mesh* m = meshes[i];
    bool flat = !m->normals || m->normal_count == 0;
    for (uint32_t t = rblocks[bk].t0; t < rblocks[bk].t1; ++t) {
      uint32_t g = first_tri[i] + t;
      hmesh[g] = i;
      hlocal[g] = t;
      f3 v[3];
      for (int k = 0; k < 3; ++k) {
        uint32_t c = m->indices[t * 3 + k];
        const float* p = m->positions + 3 * (size_t)c;
        v[k] = mk3(p[0], p[1], p[2]);
        for (int a = 0; a < 3; ++a) hv[(size_t)g * 9 + 3 * k + a] = p[a];
        float u = m->uv1[2 * (size_t)c], w = m->uv1[2 * (size_t)c + 1];
        if (!(finite_(u) && finite_(w))) { u = 0.0f; w = 0.0f; }
        huv[(size_t)g * 6 + 2 * k] = u;
        huv[(size_t)g * 6 + 2 * k + 1] = w;
        if (!flat)
          for (int a = 0; a < 3; ++a) hn[(size_t)g * 9 + 3 * k + a] = m->normals[3 * (size_t)c + a];
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
  });
  if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

  // ---- scene-level device state --------------------------------------------------------------
  // Geometry buffers live in the resident scene when there is one (filled here on its first
  // bake), otherwise in this bake; per-bake buffers are always this bake's.
  ctx->timings.prepare_ms = ms_since(t_start);
  auto t_up = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  Resident local_geo;
  Resident& geo = resident ? *resident : local_geo;
  DBuf<uint32_t> d_overflow;
  if (!use_res) {
    HIP_TRY(upload(ctx, geo.d_nodes, bvh.nodes));
    HIP_TRY(upload(ctx, geo.d_occ, bvh.tris));
    HIP_TRY(upload(ctx, geo.d_occ_nrm, occ_nrm));
    HIP_TRY(upload(ctx, geo.d_occ_alb, occ_alb));
    HIP_TRY(upload(ctx, geo.d_occ_flags, occ_flags));
    if (build_alpha) {
      HIP_TRY(upload(ctx, geo.d_occ_tex, occ_tex));
      HIP_TRY(upload(ctx, geo.d_occ_uv0, occ_uv0));
      HIP_TRY(upload(ctx, geo.d_tex_info, tex_info));
      HIP_TRY(upload(ctx, geo.d_tex_bits, tex_bits));
    }
    HIP_TRY(upload(ctx, geo.d_v, hv));
    HIP_TRY(upload(ctx, geo.d_n, hn));
    HIP_TRY(upload(ctx, geo.d_uv, huv));
    HIP_TRY(upload(ctx, geo.d_mesh, hmesh));
    HIP_TRY(upload(ctx, geo.d_local, hlocal));
    HIP_TRY(upload(ctx, geo.d_res, res));
    HIP_TRY(upload(ctx, geo.d_mid, mid));
    HIP_TRY(upload(ctx, geo.d_owned_off, owned_off));
    geo.nocc = nocc;
    geo.depth = bvh.depth;
    geo.nodes = bvh.nodes.size();
  }
  HIP_TRY(zeroed(ctx, d_overflow, 1));
  DevScene sc;
  sc.nodes = geo.d_nodes.p;
  sc.tris = reinterpret_cast<const float4*>(geo.d_occ.p);
  sc.ntris = nocc;
  sc.stack = bvh_stack_entries(geo.depth);
  sc.stack16 = geo.nodes < 65536 ? 1u : 0u;
  sc.overflow = d_overflow.p;
  sc.counters = nullptr;
  sc.occ_tex = nullptr;
  sc.occ_uv0 = nullptr;
  sc.tex = nullptr;
  sc.tex_bits = nullptr;
  if (any_alpha) {
    sc.occ_tex = geo.d_occ_tex.p;
    sc.occ_uv0 = geo.d_occ_uv0.p;
    sc.tex = geo.d_tex_info.p;
    sc.tex_bits = geo.d_tex_bits.p;
  }
  DevOcc occ{nullptr, nullptr, nullptr};
  if (bounce_in_scene) occ = DevOcc{geo.d_occ_nrm.p, geo.d_occ_alb.p, geo.d_occ_flags.p};
  DBuf<DevLight> d_lights;
  HIP_TRY(upload(ctx, d_lights, hlights));
  DBuf<int32_t> d_comp, d_cube_of;
  HIP_TRY(upload(ctx, d_comp, comp_of_slot));
  HIP_TRY(upload(ctx, d_cube_of, cube_of_slot));
  DBuf<float> d_face;
  DBuf<uint32_t> d_crop;
  HIP_TRY(d_face.alloc((size_t)nt * 3 + 3));
  HIP_TRY(d_crop.alloc((size_t)nt * 4 + 4));
  DevTris trall;
  trall.n = nt;
  trall.v = geo.d_v.p;
  trall.nrm = geo.d_n.p;
  trall.uv = geo.d_uv.p;
  trall.mesh = geo.d_mesh.p;
  trall.local = geo.d_local.p;
  trall.face = d_face.p;
  trall.crop = d_crop.p;
  DBuf<uint32_t> d_owned, d_visited;
  DBuf<unsigned long long> d_mrays, d_msamples, d_mbrays;
  if (!use_res) HIP_TRY(zeroed(ctx, d_owned, owned_off[nm] + 1));  // owned texels: only to find the texels
  HIP_TRY(zeroed(ctx, d_visited, owned_off[nm] + 1));
  HIP_TRY(zeroed(ctx, d_mrays, nm + 1));
  HIP_TRY(zeroed(ctx, d_msamples, nm + 1));
  HIP_TRY(zeroed(ctx, d_mbrays, nm + 1));
  DevMeshes ms;
  ms.n = nm;
  ms.res = geo.d_res.p;
  ms.id = geo.d_mid.p;
  ms.owned_off = geo.d_owned_off.p;
  ms.visited_off = geo.d_owned_off.p;
  ms.owned = d_owned.p;
  ms.visited = d_visited.p;
  ms.rays = d_mrays.p;
  ms.samples = d_msamples.p;
  HIP_TRY(hipStreamSynchronize(s));
  ctx->timings.upload_ms = ms_since(t_up);

  // ---- photon cubes (bounce.rs:126-151) ------------------------------------------------------
  auto t_bounce0 = std::chrono::steady_clock::now();
  stage(1, cube_slot.size());
  const uint32_t cres = opt.photon_cube_resolution;
  DBuf<uint32_t> d_cube_slot, d_photons;
  DBuf<unsigned long long> d_tile_live;
  DBuf<float4> d_tile_box;
  const uint64_t cube_texels = 6ull * cres * cres * cube_slot.size();
  const uint64_t cube_tiles = 6ull * photon_tiles_per_face(cres) * photon_tiles_per_face(cres) * cube_slot.size();
  // every texel of every cube, and their tiles
  auto build_cubes = [&]() -> int {
    HIP_TRY(d_photons.alloc(cube_texels * 4));  // one 16-byte record per photon (12 bytes used)
    HIP_TRY(launch_photons(s, sc, occ, d_lights.p, d_cube_slot.p, (uint32_t)cube_slot.size(), cres, d_photons.p, 128));
    if (ctx->photon_tiles) {
      HIP_TRY(d_tile_live.alloc(2 * cube_tiles));  // two summary words per tile
      HIP_TRY(d_tile_box.alloc(3 * cube_tiles));
      HIP_TRY(launch_photon_tiles(s, d_lights.p, d_cube_slot.p, (uint32_t)cube_slot.size(), cres, d_photons.p, d_tile_live.p,
                                  d_tile_box.p));
    }
    ctx->timings.photon_texels = cube_texels;
    return DL_OK;
  };
  // Lazy cubes wait for the gather's job lists (one mesh batch, the point-light gather with the
  // split setup), which decide between tracing the texels they can read and every texel; their
  // marks number texels 64 * tile + bit in 32 bits.
  bool cubes_pending = ctx->lazy_photons && ctx->gather_mode == 1 && ctx->gather_split && ctx->photon_tiles &&
                       !cube_slot.empty() && cube_tiles < (1ull << 26);
  const uint64_t cube_bytes = cube_texels * 16 + cube_tiles * 64;
  if (!cube_slot.empty()) {
    HIP_TRY(upload(ctx, d_cube_slot, cube_slot));
    if (!cubes_pending)
      if (int rc = build_cubes()) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    advance(cube_slot.size());
  }
  double bounce_ms = ms_since(t_bounce0);
  if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

  // ---- association and rectangles of every receiver triangle (mesh_bake.rs:141-173) ------------
  auto t_tex = std::chrono::steady_clock::now();
  stage(2, nm);
  DBuf<uint64_t> d_lc, d_tc, d_tcc, d_ent_start;
  HIP_TRY(zeroed(ctx, d_lc, (size_t)nt + 1));
  HIP_TRY(zeroed(ctx, d_tc, (size_t)nt + 1));
  HIP_TRY(zeroed(ctx, d_tcc, (size_t)nt + 1));
  HIP_TRY(d_ent_start.alloc((size_t)nt + 1));
  HIP_TRY(launch_setup(s, trall, ms, d_lights.p, nl, nlocal, d_lc.p, d_tc.p, d_tcc.p));
  HIP_TRY(scan_u64(s, d_lc.p, d_ent_start.p, nt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
  std::vector<uint64_t> h_ent_start, h_tc, h_tcc;
  HIP_TRY(download(ctx, h_ent_start, d_ent_start.p, (size_t)nt + 1));
  HIP_TRY(download(ctx, h_tc, d_tc.p, (size_t)nt));
  HIP_TRY(download(ctx, h_tcc, d_tcc.p, (size_t)nt));
  const uint64_t nent = h_ent_start[nt];
  DBuf<uint32_t> d_ent_light, d_ent_tri;
  HIP_TRY(d_ent_light.alloc(nent + 1));
  HIP_TRY(d_ent_tri.alloc(nent + 1));
  HIP_TRY(launch_assoc_fill(s, trall, ms, d_lights.p, nl, nlocal, d_ent_start.p, d_ent_light.p, d_ent_tri.p));
  std::vector<uint32_t> h_ent_light, h_crop;
  HIP_TRY(download(ctx, h_ent_light, d_ent_light.p, nent));
  HIP_TRY(download(ctx, h_crop, d_crop.p, (size_t)nt * 4));
  d_ent_tri.release();
  if (cap) {
    cap->pairs.clear();
    cap->cubes.clear();
    cap->crops = h_crop;
    cap->texels.assign(nm, 0);
    cap->escaping_count = 0;
    if (cap->escaping) HIP_TRY(hipMemsetAsync(cap->escaping->p + 8 * cap->escaping_cap, 0, 8, s));
  }

  // ---- regions (mesh_bake.rs:175-192) and bounce windows (bounce_map.rs:30-43) -----------------
  std::vector<uint32_t> reg_rect, reg_bits, reg_mesh, reg_light, h_ent_region(nent);
  std::vector<uint64_t> reg_words;
  std::vector<uint32_t> mesh_reg_first(nm + 1, 0);
  std::vector<int64_t> win_of_reg;
  std::vector<uint32_t> win_rect, win_res;
  std::vector<uint64_t> win_area;
  std::vector<uint32_t> mesh_win_first(nm + 1, 0);
  {
    std::vector<int32_t> slot_of(nl, -1);
    std::vector<uint32_t> rect(4 * (size_t)nl + 4);
    std::vector<uint32_t> touched;
    for (uint32_t i = 0; i < nm; ++i) {
      mesh_reg_first[i] = (uint32_t)reg_bits.size();
      mesh_win_first[i] = (uint32_t)win_area.size();
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
        reg_light.push_back(l);
        uint64_t texels = (uint64_t)(r[2] - r[0] + 1) * (r[3] - r[1] + 1);
        reg_words.push_back(bits == 1 ? (texels + 31) / 32 : (texels + 3) / 4);
        if (hlights[l].bakes_bounce) {
          win_of_reg.push_back((int64_t)win_area.size());
          uint32_t last = res[i] - 1;
          uint32_t w[4] = {r[0] >= 4 ? r[0] - 4 : 0, r[1] >= 4 ? r[1] - 4 : 0, std::min(r[2] + 4, last), std::min(r[3] + 4, last)};
          win_rect.insert(win_rect.end(), w, w + 4);
          win_res.push_back(res[i]);
          win_area.push_back((uint64_t)(w[2] - w[0] + 1) * (w[3] - w[1] + 1));
        } else {
          win_of_reg.push_back(-1);
        }
      }
      for (uint32_t t = first_tri[i]; t < first_tri[i + 1]; ++t)
        for (uint64_t e = h_ent_start[t]; e < h_ent_start[t + 1]; ++e) h_ent_region[e] = (uint32_t)slot_of[h_ent_light[e]];
      for (uint32_t l : touched) slot_of[l] = -1;
    }
    mesh_reg_first[nm] = (uint32_t)reg_bits.size();
    mesh_win_first[nm] = (uint32_t)win_area.size();
  }

  // ---- batches of meshes under the HBM budget ---------------------------------------------------
  // Default: half of the device memory still available after the scene-level buffers (free in
  // the driver, or held unused by the memory pool or this context's cache), over the number of
  // bakes running side by side.
  uint64_t budget = 48ull << 30;
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
      uint64_t avail = free_b + ctx->cache.cached;
      hipMemPool_t pool = nullptr;
      uint64_t reserved = 0, used = 0;
      if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess && pool &&
          hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) == hipSuccess &&
          hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess && reserved > used)
        avail += reserved - used;
      if (cubes_pending) avail -= std::min(avail, cube_bytes);  // the cubes are built later, from this memory
      if (avail > (2ull << 30)) budget = avail / 2 / std::max(1u, ctx->share);
    }
  }
  if (const char* env = std::getenv("DL_BATCH_BYTES")) budget = std::max<uint64_t>(1ull << 20, strtoull(env, nullptr, 10));
  std::vector<Batch> batches;
  {
    uint64_t acc = 0;
    uint32_t start = 0;
    for (uint32_t i = 0; i < nm; ++i) {
      uint64_t est = 0;
      for (uint32_t t = first_tri[i]; t < first_tri[i + 1]; ++t) {
        est += h_tcc[t] * 80;  // staged and final texel records
        const uint32_t* c = &h_crop[4 * (size_t)t];
        uint64_t padded = (uint64_t)(c[2] - c[0] + 5) * (c[3] - c[1] + 5);
        est += (h_ent_start[t + 1] - h_ent_start[t]) * (48 + padded * (bounce_in_scene ? 4 : 1) / 4);
        if (bounce_in_scene) est += (h_ent_start[t + 1] - h_ent_start[t]) * h_tcc[t] * 8;  // gather jobs
      }
      for (uint32_t r = mesh_reg_first[i]; r < mesh_reg_first[i + 1]; ++r) est += reg_words[r] * 4;
      for (uint32_t w = mesh_win_first[i]; w < mesh_win_first[i + 1]; ++w) est += win_area[w] * 24;
      if (i > start && acc + est > budget) {
        batches.push_back({start, i});
        start = i;
        acc = 0;
      }
      acc += est;
    }
    if (nm > start) batches.push_back({start, nm});
  }
  if (cubes_pending && batches.size() > 1) {  // the gathers of several batches read the same cubes
    cubes_pending = false;
    if (int rc = build_cubes()) return rc;
  }
  ctx->timings.texels_ms = ms_since(t_tex);

  std::unique_ptr<dl_baked_scene> scene(new dl_baked_scene());
  scene->meshes.resize(nm);
  DBuf<unsigned long long> d_counters;
  if (ctx->counting) HIP_TRY(zeroed(ctx, d_counters, 7));
  uint64_t burial_rays = 0;
  double burial_ms = 0.0;
  float bounce_kernel_ms = 0.0f;
  const BlurWeights blur = blur_weights();
  const uint32_t stride = bounce_in_scene ? 3 : 2;
  float trace_kernel_ms = 0.0f;
  hipEvent_t ev0, ev1;
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  struct EventGuard {
    hipEvent_t a, b;
    ~EventGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } guard{ev0, ev1};
  double trace_ms = 0, post_ms = 0, texels_ms = 0;

  for (const Batch& B : batches) {
    const uint32_t t0 = first_tri[B.m0], t1 = first_tri[B.m1], nbt = t1 - t0;
    const uint64_t e0 = h_ent_start[t0], e1 = h_ent_start[t1], nbe = e1 - e0;
    const uint32_t r0 = mesh_reg_first[B.m0], r1 = mesh_reg_first[B.m1], nbr = r1 - r0;
    const uint32_t w0 = mesh_win_first[B.m0], w1 = mesh_win_first[B.m1], nbw = w1 - w0;
    auto tb = std::chrono::steady_clock::now();
    DevTris tr = trall;
    tr.n = nbt;
    tr.v = trall.v + (size_t)t0 * 9;
    tr.nrm = trall.nrm + (size_t)t0 * 9;
    tr.uv = trall.uv + (size_t)t0 * 6;
    tr.mesh = trall.mesh + t0;
    tr.local = trall.local + t0;
    tr.face = trall.face + (size_t)t0 * 3;
    tr.crop = trall.crop + (size_t)t0 * 4;

    // entries and regions of the batch, rebased
    std::vector<uint64_t> b_ent_start(nbt + 1), b_trace_start(nbt + 1, 0), b_cand_start(nbt + 1, 0);
    for (uint32_t t = 0; t <= nbt; ++t) b_ent_start[t] = h_ent_start[t0 + t] - e0;
    for (uint32_t t = 0; t < nbt; ++t) {
      b_trace_start[t + 1] = b_trace_start[t] + h_tc[t0 + t];
      b_cand_start[t + 1] = b_cand_start[t] + h_tcc[t0 + t];
    }
    std::vector<uint32_t> b_ent_light(h_ent_light.begin() + (long)e0, h_ent_light.begin() + (long)e1), b_ent_tri(nbe), b_ent_region(nbe);
    std::vector<uint32_t> b_ent_window(nbe, 0xffffffffu);
    for (uint32_t t = 0; t < nbt; ++t)
      for (uint64_t e = b_ent_start[t]; e < b_ent_start[t + 1]; ++e) {
        b_ent_tri[e] = t;
        uint32_t r = h_ent_region[e0 + e];
        b_ent_region[e] = r - r0;
        if (win_of_reg[r] >= 0) b_ent_window[e] = (uint32_t)(win_of_reg[r] - w0);
      }
    std::vector<uint32_t> b_reg_rect(reg_rect.begin() + 4 * (size_t)r0, reg_rect.begin() + 4 * (size_t)r1);
    std::vector<uint32_t> b_reg_bits(reg_bits.begin() + r0, reg_bits.begin() + r1), b_reg_mesh(reg_mesh.begin() + r0, reg_mesh.begin() + r1);
    std::vector<uint64_t> b_reg_off(nbr + 1, 0);
    for (uint32_t r = 0; r < nbr; ++r) b_reg_off[r + 1] = b_reg_off[r] + reg_words[r0 + r];
    DBuf<uint64_t> d_bes, d_bts, d_bcs, d_reg_off;
    DBuf<uint32_t> d_bel, d_bet, d_ber, d_reg_rect, d_reg_bits, d_reg_mesh, d_pool;
    HIP_TRY(upload(ctx, d_bes, b_ent_start));
    HIP_TRY(upload(ctx, d_bts, b_trace_start));
    HIP_TRY(upload(ctx, d_bcs, b_cand_start));
    HIP_TRY(upload(ctx, d_bel, b_ent_light));
    HIP_TRY(upload(ctx, d_bet, b_ent_tri));
    HIP_TRY(upload(ctx, d_ber, b_ent_region));
    HIP_TRY(upload(ctx, d_reg_rect, b_reg_rect));
    HIP_TRY(upload(ctx, d_reg_bits, b_reg_bits));
    HIP_TRY(upload(ctx, d_reg_mesh, b_reg_mesh));
    HIP_TRY(upload(ctx, d_reg_off, b_reg_off));
    HIP_TRY(zeroed(ctx, d_pool, b_reg_off[nbr] + 1));
    DevRegions rg{d_reg_rect.p, d_reg_bits.p, d_reg_off.p, d_reg_mesh.p, d_pool.p};
    DevEntries ent{nbe, d_bel.p, d_ber.p, d_bet.p, d_bes.p};
    // capture: each region's own visited bits and counters
    std::vector<uint64_t> b_vis_off(nbr + 1, 0);
    DBuf<uint64_t> d_vis_off;
    DBuf<uint32_t> d_rvis;
    DBuf<unsigned long long> d_rcount, d_wcount;
    if (cap) {
      for (uint32_t r = 0; r < nbr; ++r) {
        const uint32_t* q = &b_reg_rect[4 * (size_t)r];
        b_vis_off[r + 1] = b_vis_off[r] + ((uint64_t)(q[2] - q[0] + 1) * (q[3] - q[1] + 1) + 31) / 32;
      }
      HIP_TRY(upload(ctx, d_vis_off, b_vis_off));
      HIP_TRY(zeroed(ctx, d_rvis, b_vis_off[nbr] + 1));
      HIP_TRY(zeroed(ctx, d_rcount, 2 * (size_t)nbr + 2));
      HIP_TRY(zeroed(ctx, d_wcount, (size_t)nbw + 1));
      rg.vis = d_rvis.p;
      rg.vis_off = d_vis_off.p;
      rg.rays = d_rcount.p;
      rg.samples = d_rcount.p + nbr + 1;
    }

    // ---- texels: owned, traced, burial, compaction (mesh_texels.rs:41-108) -------------------
    // A resident scene holds every triangle's texels already: the batch's are a contiguous run.
    std::vector<uint64_t> h_tex_start;
    DBuf<uint64_t> d_tex_start;
    DBuf<uint32_t> d_txy, d_ttri;
    DBuf<float4> d_tpos, d_tnrm;
    DevTexels tx{};
    uint64_t ntex = 0;
    if (use_res) {
      const uint64_t base = resident->tex_start[t0];
      h_tex_start.resize((size_t)nbt + 1);
      for (uint32_t t = 0; t <= nbt; ++t) h_tex_start[t] = resident->tex_start[t0 + t] - base;
      ntex = h_tex_start[nbt];
      HIP_TRY(upload(ctx, d_tex_start, h_tex_start));
      HIP_TRY(d_ttri.alloc(ntex + 1));
      HIP_TRY(launch_add_u32(s, resident->t_tri.p + base, d_ttri.p, ntex, 0u - t0));  // scene-wide -> batch triangle
      tx = DevTexels{resident->t_xy.p + base, resident->t_pos.p + base, resident->t_nrm.p + base, d_ttri.p};
    } else {
      DBuf<unsigned long long> d_cnt;
      HIP_TRY(zeroed(ctx, d_cnt, (size_t)nbt + 1));
      HIP_TRY(launch_owned(s, tr, ms, d_bts.p, b_trace_start[nbt]));
      DevTexels none{};
      const uint64_t* tc_b = d_tc.p + t0;
      HIP_TRY(launch_texels(s, tr, ms, tc_b, d_bcs.p, b_cand_start[nbt], 0, d_cnt.p, nullptr, none));
      DBuf<uint64_t> d_tex0_start;
      HIP_TRY(d_tex0_start.alloc((size_t)nbt + 1));
      HIP_TRY(scan_u64(s, reinterpret_cast<const uint64_t*>(d_cnt.p), d_tex0_start.p, nbt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
      uint64_t ntex0 = 0;
      HIP_TRY(hipMemcpyAsync(&ntex0, d_tex0_start.p + nbt, 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      DBuf<uint32_t> d_t0xy, d_t0tri;
      DBuf<float4> d_t0pos, d_t0nrm;
      HIP_TRY(d_t0xy.alloc(ntex0 + 1));
      HIP_TRY(d_t0tri.alloc(ntex0 + 1));
      HIP_TRY(d_t0pos.alloc(ntex0 + 1));
      HIP_TRY(d_t0nrm.alloc(ntex0 + 1));
      DevTexels tx0{d_t0xy.p, d_t0pos.p, d_t0nrm.p, d_t0tri.p};
      HIP_TRY(hipMemsetAsync(d_cnt.p, 0, ((size_t)nbt + 1) * 8, s));
      HIP_TRY(launch_texels(s, tr, ms, tc_b, d_bcs.p, b_cand_start[nbt], 1, d_cnt.p, d_tex0_start.p, tx0));
      DBuf<uint8_t> d_buried;
      HIP_TRY(d_buried.alloc(ntex0 + 1));
      HIP_TRY(hipMemsetAsync(d_cnt.p, 0, ((size_t)nbt + 1) * 8, s));
      HIP_TRY(burial_streamed(ctx, sc, tr, ms, tx0, ntex0, opt.seed, d_buried.p, d_cnt.p, ctx->counting ? d_counters.p : nullptr,
                              &burial_rays, &burial_ms, cap));
      HIP_TRY(d_tex_start.alloc((size_t)nbt + 1));
      HIP_TRY(scan_u64(s, reinterpret_cast<const uint64_t*>(d_cnt.p), d_tex_start.p, nbt, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
      HIP_TRY(download(ctx, h_tex_start, d_tex_start.p, (size_t)nbt + 1));
      ntex = h_tex_start[nbt];
      HIP_TRY(d_txy.alloc(ntex + 1));
      HIP_TRY(d_ttri.alloc(ntex + 1));
      HIP_TRY(d_tpos.alloc(ntex + 1));
      HIP_TRY(d_tnrm.alloc(ntex + 1));
      tx = DevTexels{d_txy.p, d_tpos.p, d_tnrm.p, d_ttri.p};
      HIP_TRY(hipMemsetAsync(d_cnt.p, 0, ((size_t)nbt + 1) * 8, s));
      HIP_TRY(launch_compact(s, tx0, ntex0, d_buried.p, d_tex_start.p, d_cnt.p, tx));
      HIP_TRY(hipStreamSynchronize(s));
      d_t0xy.release();
      d_t0tri.release();
      d_t0pos.release();
      d_t0nrm.release();
      d_buried.release();
      if (fill_res) {  // keep this batch's texels for the resident scene
        auto piece = std::make_unique<Resident::Piece>();
        piece->xy.take(d_txy);
        piece->pos.take(d_tpos);
        piece->nrm.take(d_tnrm);
        HIP_TRY(piece->tri.alloc(ntex + 1));
        HIP_TRY(launch_add_u32(s, d_ttri.p, piece->tri.p, ntex, t0));  // batch -> scene-wide triangle
        piece->start = h_tex_start;
        piece->n = ntex;
        resident->pieces.push_back(std::move(piece));
      }
    }
    for (uint32_t t = 0; t < nbt; ++t) scene->stats.texels += h_tex_start[t + 1] - h_tex_start[t];
    if (cap)
      for (uint32_t i = B.m0; i < B.m1; ++i) cap->texels[i] = h_tex_start[first_tri[i + 1] - t0] - h_tex_start[first_tri[i] - t0];
    texels_ms += ms_since(tb);
    if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

    // ---- shadow rays (mesh_bake.rs:203-267) ----------------------------------------------------
    auto t_trace = std::chrono::steady_clock::now();
    std::vector<uint64_t> h_unit_start(nbt + 1, 0);
    for (uint32_t t = 0; t < nbt; ++t) {
      uint64_t texels = h_tex_start[t + 1] - h_tex_start[t];
      h_unit_start[t + 1] = h_unit_start[t] + (b_ent_start[t + 1] - b_ent_start[t]) * ((texels + 63) / 64);
    }
    DBuf<uint64_t> d_unit_start;
    HIP_TRY(upload(ctx, d_unit_start, h_unit_start));
    TraceParams tp;
    tp.seed = opt.seed;
    tp.area_samples = opt.area_light_samples;
    tp.area_min = opt.area_light_min_samples;
    tp.sky_samples = opt.sky_samples;
    DevScene sct = sc;
    sct.counters = ctx->counting ? d_counters.p : nullptr;
    // Lights with several rays per texel (area lights, soft suns, skies) are streamed: their
    // (texel, entry) jobs, light by light, go through the persistent any-hit kernel.
    std::vector<uint32_t> vis_me;
    std::vector<uint64_t> vis_js(1, 0);
    uint32_t vis_slots_a = 0;
    bool per_lane = true;  // entries left to k_trace (with streaming: point lights and hard suns)
    if (ctx->trace_stream) {
      std::vector<std::pair<uint32_t, uint32_t>> multi;  // (light slot, entry) in entry order
      per_lane = false;
      for (uint32_t t = 0; t < nbt; ++t) {
        if (h_tex_start[t + 1] == h_tex_start[t]) continue;
        for (uint64_t e = b_ent_start[t]; e < b_ent_start[t + 1]; ++e) {
          const DevLight& L = hlights[b_ent_light[e]];
          if (L.shape == 1 || L.shape == 2 || L.shape == 4 || (L.shape == 3 && L.srad > 0.0f))
            multi.push_back({b_ent_light[e], (uint32_t)e});
          else
            per_lane = true;
        }
      }
      std::stable_sort(multi.begin(), multi.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      for (const auto& le : multi) {
        const uint32_t t = b_ent_tri[le.second];
        vis_me.push_back(le.second);
        vis_js.push_back(vis_js.back() + (h_tex_start[t + 1] - h_tex_start[t]));
        const DevLight& L = hlights[le.first];
        const uint32_t probes = L.shape == 4 ? 0u : (L.shape == 2 ? 3u : 5u);
        const uint32_t slots = L.shape == 4 ? tp.sky_samples : probes + std::min(tp.area_min, tp.area_samples);
        vis_slots_a = std::max(vis_slots_a, slots);
      }
      if (vis_js.back() > 0) tp.stream_multi = 1;
    }
    HIP_TRY(hipEventRecord(ev0, s));
    // with streaming, the per-lane kernel only meets point lights and hard suns
    if (per_lane)
      HIP_TRY(launch_trace(s, sct, tr, ms, d_lights.p, tx, d_tex_start.p, ent, rg, d_unit_start.p, h_unit_start[nbt], tp,
                           ctx->trig, ctx->trace_block, !ctx->trace_stream));
    if (tp.stream_multi) {
      const uint64_t njobs = vis_js.back();
      const uint32_t nme = (uint32_t)vis_me.size();
      const uint32_t slots_b = tp.area_samples > tp.area_min ? tp.area_samples - tp.area_min : 0u;
      const uint64_t slots = std::max<uint64_t>(1, std::max(vis_slots_a, slots_b));
      const uint64_t cn = std::min<uint64_t>(njobs, std::max<uint64_t>(1, ctx->gather.chunk_rays / slots));
      DBuf<uint64_t> d_vjs, d_vcnt, d_vrs, d_vneed, d_vbscan;
      DBuf<uint32_t> d_vme, d_vblist;
      DBuf<unsigned int> d_vq;
      DBuf<float4> d_vrays;
      DBuf<float> d_vcontrib, d_vwts;
      DBuf<VisState> d_vstate;
      HIP_TRY(upload(ctx, d_vjs, vis_js));
      HIP_TRY(upload(ctx, d_vme, vis_me));
      HIP_TRY(d_vcnt.alloc(cn + 1));
      HIP_TRY(d_vrs.alloc(cn + 1));
      HIP_TRY(d_vneed.alloc(cn + 1));
      HIP_TRY(d_vbscan.alloc(cn + 1));
      HIP_TRY(d_vblist.alloc(cn + 1));
      HIP_TRY(d_vq.alloc(1));
      HIP_TRY(d_vrays.alloc(2 * cn * slots));
      HIP_TRY(d_vcontrib.alloc(3 * cn * slots));
      HIP_TRY(d_vwts.alloc(cn * slots));
      HIP_TRY(d_vstate.alloc(cn + 1));
      DBuf<uint64_t> d_lcount, d_loff;  // traced slots of a phase, job by job
      DBuf<uint32_t> d_llist;
      HIP_TRY(d_lcount.alloc(cn + 1));
      HIP_TRY(d_loff.alloc(cn + 1));
      HIP_TRY(d_llist.alloc(cn * slots));
      GatherLaunch vgl = ctx->gather;
      vgl.order = 1;
      for (uint64_t q0 = 0; q0 < njobs; q0 += cn) {
        const uint64_t n = std::min(cn, njobs - q0);
        HIP_TRY(launch_vis_count(s, ent, d_lights.p, d_vjs.p, d_vme.p, nme, d_tex_start.p, q0, n, tp, d_vcnt.p));
        HIP_TRY(scan_u64(s, d_vcnt.p, d_vrs.p, n, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
        HIP_TRY(launch_vis_rays_a(s, tr, ms, d_lights.p, tx, ent, d_vjs.p, d_vme.p, nme, d_tex_start.p, q0, n, tp, ctx->trig,
                                  d_vrays.p, d_vcontrib.p, d_vwts.p, d_vstate.p));
        HIP_TRY(launch_ray_list(s, d_vrays.p, n, vis_slots_a, d_vrs.p, 1, d_lcount.p, d_loff.p, d_llist.p, &ctx->scan_tmp,
                                &ctx->scan_tmp_bytes));
        HIP_TRY(hipMemsetAsync(d_vq.p, 0, sizeof(unsigned int), s));
        HIP_TRY(launch_any_stream(s, sct, d_vrays.p, d_loff.p + n, n * vis_slots_a, d_vrs.p, n, d_vcontrib.p, d_vq.p, vgl,
                                  d_llist.p));
        HIP_TRY(launch_vis_mid(s, ent, d_lights.p, d_vjs.p, d_vme.p, nme, d_tex_start.p, q0, n, tp, d_vrays.p, d_vcontrib.p,
                               d_vwts.p, d_vstate.p, d_vneed.p));
        HIP_TRY(scan_u64(s, d_vneed.p, d_vbscan.p, n, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
        uint64_t nb = 0;
        HIP_TRY(hipMemcpyAsync(&nb, d_vbscan.p + n, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (nb > n) { ctx->err = "streamed visibility: bad phase B count"; return DL_E_DEVICE; }
        if (nb) {
          HIP_TRY(launch_vis_list(s, d_vneed.p, d_vbscan.p, n, d_vblist.p));
          HIP_TRY(launch_vis_rays_b(s, tr, ms, d_lights.p, tx, ent, d_vjs.p, d_vme.p, nme, d_tex_start.p, q0, d_vblist.p, nb, tp,
                                    ctx->trig, d_vrays.p, d_vcontrib.p, d_vwts.p));
          HIP_TRY(launch_ray_list(s, d_vrays.p, nb, slots_b, nullptr, 1, d_lcount.p, d_loff.p, d_llist.p, &ctx->scan_tmp,
                                  &ctx->scan_tmp_bytes));
          HIP_TRY(hipMemsetAsync(d_vq.p, 0, sizeof(unsigned int), s));
          HIP_TRY(launch_any_stream(s, sct, d_vrays.p, d_loff.p + nb, nb * slots_b, nullptr, nb, d_vcontrib.p, d_vq.p, vgl,
                                    d_llist.p));
        }
        HIP_TRY(launch_vis_finish(s, tr, ms, d_lights.p, tx, ent, rg, d_vjs.p, d_vme.p, nme, d_tex_start.p, q0, n, tp, d_vstate.p,
                                  d_vbscan.p, nb, d_vrays.p, d_vcontrib.p, d_vwts.p));
      }
    }
    HIP_TRY(hipEventRecord(ev1, s));
    // The host side of the bounce stage's setup, while the shadow-ray kernels run: bounce entries
    // grouped by light slot (triangle order within a light; the gather's results do not depend on
    // the order, and consecutive work then shares one light's photon cube and the BVH nodes around
    // it, which keeps the working set in L2/MALL), their 64-texel units, and the windows' offsets
    // and filter tiles.
    std::vector<uint64_t> win_off, be, be_units(1, 0), win_tiles;
    if (bounce_in_scene && nbw > 0) {
      win_off.assign(nbw + 1, 0);
      for (uint32_t w = 0; w < nbw; ++w) win_off[w + 1] = win_off[w] + win_area[w0 + w];
      std::vector<uint64_t> per_light(nl + 1, 0);
      for (uint32_t t = 0; t < nbt; ++t)
        for (uint64_t e = b_ent_start[t]; e < b_ent_start[t + 1]; ++e)
          if (hlights[b_ent_light[e]].bakes_bounce) per_light[b_ent_light[e] + 1]++;
      for (uint32_t l = 0; l < nl; ++l) per_light[l + 1] += per_light[l];
      be.resize(per_light[nl]);
      for (uint32_t t = 0; t < nbt; ++t)
        for (uint64_t e = b_ent_start[t]; e < b_ent_start[t + 1]; ++e)
          if (hlights[b_ent_light[e]].bakes_bounce) be[per_light[b_ent_light[e]]++] = e;
      be_units.resize(be.size() + 1);
      for (size_t i = 0; i < be.size(); ++i) {
        const uint32_t t = b_ent_tri[be[i]];
        be_units[i + 1] = be_units[i] + (h_tex_start[t + 1] - h_tex_start[t] + 63) / 64;
      }
      win_tiles.assign(nbw + 1, 0);  // square tiles of the filter per window
      for (uint32_t w = 0; w < nbw; ++w) {
        const uint32_t* q = &win_rect[4 * (size_t)(w0 + w)];
        win_tiles[w + 1] = win_tiles[w] + (uint64_t)((q[2] - q[0] + kFilterTile) / kFilterTile) *
                                              ((q[3] - q[1] + kFilterTile) / kFilterTile);
      }
    }
    HIP_TRY(hipStreamSynchronize(s));
    float kms = 0.0f;
    (void)hipEventElapsedTime(&kms, ev0, ev1);
    trace_kernel_ms += kms;
    ctx->timings.trace_launches++;
    trace_ms += ms_since(t_trace);
    if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

    // ---- bounce gather, dilate, blur (bounce.rs:155-217) -----------------------------------------
    auto t_b = std::chrono::steady_clock::now();
    DBuf<uint8_t> d_has;
    DBuf<float> d_map_a, d_map_b;
    DBuf<uint32_t> d_win_rect, d_win_res, d_ent_window;
    DBuf<uint64_t> d_win_off, d_win_start, d_be, d_be_units;  // d_win_start: filter tiles per window
    DBuf<uint8_t> d_tile_lit, d_tile_nb;
    DevBounce bb{};
    if (bounce_in_scene && nbw > 0) {
      std::vector<uint32_t> b_win_rect(win_rect.begin() + 4 * (size_t)w0, win_rect.begin() + 4 * (size_t)w1);
      std::vector<uint32_t> b_win_res(win_res.begin() + w0, win_res.begin() + w1);
      bool sun_bounce = false;
      for (uint64_t e : be) sun_bounce |= hlights[b_ent_light[e]].shape == 3;
      HIP_TRY(upload(ctx, d_win_rect, b_win_rect));
      HIP_TRY(upload(ctx, d_win_res, b_win_res));
      HIP_TRY(upload(ctx, d_win_off, win_off));
      HIP_TRY(upload(ctx, d_win_start, win_tiles));
      HIP_TRY(zeroed(ctx, d_tile_lit, win_tiles[nbw] + 1));
      HIP_TRY(d_tile_nb.alloc(win_tiles[nbw] + 1));  // written by the filter, tile by tile
      HIP_TRY(upload(ctx, d_ent_window, b_ent_window));
      HIP_TRY(upload(ctx, d_be, be));
      HIP_TRY(upload(ctx, d_be_units, be_units));
      HIP_TRY(zeroed(ctx, d_map_a, win_off[nbw] * 3 + 3));
      HIP_TRY(zeroed(ctx, d_map_b, win_off[nbw] * 3 + 3));
      bb.n = be.size();
      bb.entry = d_be.p;
      bb.unit_start = d_be_units.p;
      bb.win_rect = d_win_rect.p;
      bb.win_off = d_win_off.p;
      bb.ent_window = d_ent_window.p;
      bb.cube_of_slot = d_cube_of.p;
      bb.map_a = d_map_a.p;
      bb.map_b = d_map_b.p;
      bb.rays = d_mbrays.p;
      bb.tile_start = d_win_start.p;
      bb.tile_lit = d_tile_lit.p;
      bb.tile_nb = d_tile_nb.p;
      bb.wrays = cap ? d_wcount.p : nullptr;
      BounceParams bp;
      bp.seed = opt.seed;
      bp.colored = opt.colored_bounce ? 1u : 0u;
      bp.cube_res = cres;
      bp.photons = d_photons.p;
      bp.sun_only = 0;
      bp.photon_reads = ctx->counting ? d_counters.p + 6 : nullptr;
      if (d_tile_live.p) {
        bp.tiles_per_face = photon_tiles_per_face(cres);
        bp.tile_live = d_tile_live.p;
        bp.tile_box = d_tile_box.p;
      }
      HIP_TRY(launch_bounce_owner(s, tr, d_lights.p, tx, d_tex_start.p, ent, bb, be_units.back()));
      DevScene scb = sc;
      scb.counters = ctx->counting ? d_counters.p + 2 : nullptr;
      DBuf<unsigned long long> d_jobc;
      DBuf<uint2> d_tjobs, d_cjobs, d_sjobs;
      uint64_t ntjobs = 0, nsjobs = 0;
      if (ctx->gather_mode == 1) {
        // job lists of the streamed point-light gather (texel and entry indices are u32)
        if (ntex >= (1ull << 32) || nbe >= (1ull << 32)) {
          ctx->err = "a mesh batch holds 2^32 or more texels or light entries";
          return DL_E_DEVICE;
        }
        HIP_TRY(zeroed(ctx, d_jobc, 4));
        const uint64_t nb1 = bb.n + 1;
        DBuf<uint64_t> d_jcnt, d_joff;
        HIP_TRY(d_jcnt.alloc(3 * nb1));
        HIP_TRY(d_joff.alloc(3 * nb1));
        HIP_TRY(launch_gather_jobs(s, tr, d_lights.p, tx, d_tex_start.p, ent, bb, 0, d_jcnt.p, nullptr, nullptr, nullptr, nullptr));
        for (int l = 0; l < 3; ++l)
          HIP_TRY(scan_u64(s, d_jcnt.p + l * nb1, d_joff.p + l * nb1, bb.n, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
        std::vector<uint64_t> jc(3);
        for (int l = 0; l < 3; ++l)
          HIP_TRY(hipMemcpyAsync(&jc[l], d_joff.p + l * nb1 + bb.n, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        ntjobs = jc[0];
        nsjobs = jc[2];
        HIP_TRY(d_tjobs.alloc(std::max<uint64_t>(1, jc[0])));
        HIP_TRY(d_cjobs.alloc(std::max<uint64_t>(1, jc[1])));
        HIP_TRY(d_sjobs.alloc(std::max<uint64_t>(1, jc[2])));
        HIP_TRY(launch_gather_jobs(s, tr, d_lights.p, tx, d_tex_start.p, ent, bb, 1, nullptr, d_joff.p, d_tjobs.p, d_cjobs.p,
                                   d_sjobs.p));
        if (cubes_pending) {
          // The cubes, now that the jobs are known: when they set up fewer samples than twice the
          // cubes' texels, only the texels their samples can fall in (marked from the samples'
          // fast directions, exact where those are unsure), else every texel.
          cubes_pending = false;
          uint32_t most = 2;
          for (uint64_t e : be) most = std::max(most, hlights[b_ent_light[e]].bounce_samples);
          if ((jc[0] + jc[1]) * (uint64_t)most > 2 * cube_texels) {
            if (int rc = build_cubes()) return rc;
          } else {
            bp.tiles_per_face = photon_tiles_per_face(cres);
            DBuf<uint8_t> d_need;
            DBuf<unsigned long long> d_need_mask;
            DBuf<uint64_t> d_ncnt, d_noff;
            DBuf<uint32_t> d_nlist;
            HIP_TRY(zeroed(ctx, d_need, 64 * cube_tiles));
            HIP_TRY(launch_photon_mark(s, tr, ms, d_lights.p, tx, ent, bb, d_tjobs.p, jc[0], bp, ctx->trig, d_need.p));
            HIP_TRY(launch_photon_mark(s, tr, ms, d_lights.p, tx, ent, bb, d_cjobs.p, jc[1], bp, ctx->trig, d_need.p));
            HIP_TRY(d_need_mask.alloc(cube_tiles));
            HIP_TRY(d_ncnt.alloc(cube_tiles + 1));
            HIP_TRY(d_noff.alloc(cube_tiles + 1));
            HIP_TRY(launch_need_count(s, d_need.p, cube_tiles, d_need_mask.p, d_ncnt.p));
            HIP_TRY(scan_u64(s, d_ncnt.p, d_noff.p, cube_tiles, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
            uint64_t needed = 0;
            HIP_TRY(hipMemcpyAsync(&needed, d_noff.p + cube_tiles, 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            HIP_TRY(d_nlist.alloc(std::max<uint64_t>(1, needed)));
            HIP_TRY(launch_need_list(s, d_need_mask.p, d_noff.p, cube_tiles, d_nlist.p));
            HIP_TRY(d_photons.alloc(cube_texels * 4));
            HIP_TRY(launch_photons_listed(s, sc, occ, d_lights.p, d_cube_slot.p, cres, d_nlist.p, needed, d_photons.p, 128));
            HIP_TRY(d_tile_live.alloc(2 * cube_tiles));  // two summary words per tile
            HIP_TRY(d_tile_box.alloc(3 * cube_tiles));
            HIP_TRY(launch_photon_tiles(s, d_lights.p, d_cube_slot.p, (uint32_t)cube_slot.size(), cres, d_photons.p, d_tile_live.p,
                                        d_tile_box.p, d_need_mask.p));
            HIP_TRY(hipStreamSynchronize(s));  // before the marks and lists go
            ctx->timings.photon_texels = needed;
          }
          bp.photons = d_photons.p;
          bp.tiles_per_face = photon_tiles_per_face(cres);
          bp.tile_live = d_tile_live.p;
          bp.tile_box = d_tile_box.p;
        }
        HIP_TRY(launch_gather_count(s, tr, ms, d_lights.p, tx, ent, bb, d_cjobs.p, jc[1], bp, ctx->trig));
        bp.sun_only = 1;
      }
      HIP_TRY(hipEventRecord(ev0, s));
      if (ctx->gather_mode == 1) {
        // chunks of traced jobs: rays and contributions, any-hit visibility, ordered sums
        uint32_t max_samples = 2;
        for (uint64_t e : be) max_samples = std::max(max_samples, hlights[b_ent_light[e]].bounce_samples);
        const uint64_t per_chunk = std::max<uint64_t>(1, ctx->gather.chunk_rays / max_samples);
        const uint64_t cj = std::min<uint64_t>(per_chunk, std::max<uint64_t>(1, ntjobs));
        const uint32_t mask_words = (max_samples + 31) / 32;
        DBuf<uint64_t> d_nr, d_rs, d_tcount, d_toff;
        DBuf<float4> d_rays;
        DBuf<float> d_contrib;
        DBuf<uint32_t> d_tmask, d_list, d_omask;
        DBuf<uint64_t> d_ocount, d_ooff;
        // the split sample setup needs the photon tiles (phase A decides samples by them)
        const bool split = ctx->gather_split && bp.tile_live != nullptr;
        if (split) {
          HIP_TRY(d_omask.alloc(cj * mask_words));
          HIP_TRY(d_ocount.alloc(cj + 1));
          HIP_TRY(d_ooff.alloc(cj + 1));
        }
        HIP_TRY(d_nr.alloc(cj + 1));
        HIP_TRY(d_rs.alloc(cj + 1));
        HIP_TRY(d_tcount.alloc(cj + 1));
        HIP_TRY(d_toff.alloc(cj + 1));
        HIP_TRY(d_tmask.alloc(cj * mask_words));
        HIP_TRY(d_list.alloc(cj * max_samples));
        HIP_TRY(d_rays.alloc(2 * cj * max_samples));
        HIP_TRY(d_contrib.alloc(3 * cj * max_samples));
        std::vector<hipEvent_t> tev, gev;  // spans of the any-hit and sample-setup launches
        struct EvList {
          std::vector<hipEvent_t>& v;
          ~EvList() { for (hipEvent_t e : v) (void)hipEventDestroy(e); }
        } tev_guard{tev}, gev_guard{gev};
        for (uint64_t j0 = 0; j0 < ntjobs; j0 += cj) {
          const uint64_t n = std::min(cj, ntjobs - j0);
          const uint2* jobs = d_tjobs.p + j0;
          HIP_TRY(launch_gather_nrays(s, ent, d_lights.p, jobs, n, d_nr.p));
          HIP_TRY(scan_u64(s, d_nr.p, d_rs.p, n, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
          hipEvent_t g0, g1;
          HIP_TRY(hipEventCreate(&g0));
          gev.push_back(g0);
          HIP_TRY(hipEventCreate(&g1));
          gev.push_back(g1);
          HIP_TRY(hipEventRecord(g0, s));
          if (split) {
            // phase A decides most samples; the open ones (listed from their bits) get the exact chain
            HIP_TRY(launch_gather_split(s, tr, ms, d_lights.p, tx, ent, bb, jobs, n, d_rs.p, bp, d_rays.p, ctx->gather.order,
                                        d_omask.p, d_tmask.p, mask_words, d_ocount.p, d_tcount.p));
            HIP_TRY(scan_u64(s, d_ocount.p, d_ooff.p, n, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
            HIP_TRY(launch_gather_list(s, d_omask.p, mask_words, d_ooff.p, d_rs.p, n, ctx->gather.order, d_list.p));
            HIP_TRY(launch_gather_open(s, tr, ms, d_lights.p, tx, ent, bb, jobs, n, d_rs.p, bp, ctx->trig, d_rays.p, d_contrib.p,
                                       ctx->gather.order, d_list.p, d_ooff.p + n, n * max_samples, d_tmask.p, mask_words,
                                       d_tcount.p));
          } else {
            HIP_TRY(launch_gather_rays(s, tr, ms, d_lights.p, tx, ent, bb, jobs, n, d_rs.p, bp, ctx->trig, d_rays.p, d_contrib.p,
                                       ctx->gather.order, d_tmask.p, mask_words, d_tcount.p));
          }
          HIP_TRY(hipEventRecord(g1, s));
          // the traced rays only, job by job
          HIP_TRY(scan_u64(s, d_tcount.p, d_toff.p, n, &ctx->scan_tmp, &ctx->scan_tmp_bytes));
          if (ctx->counting) {  // the chunk's samples, traced and open samples, for the instrumented figures
            uint64_t three[3] = {0, 0, 0};
            HIP_TRY(hipMemcpyAsync(&three[0], d_rs.p + n, 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(&three[1], d_toff.p + n, 8, hipMemcpyDeviceToHost, s));
            if (split) HIP_TRY(hipMemcpyAsync(&three[2], d_ooff.p + n, 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            ctx->timings.gather_jobs += n;
            ctx->timings.gather_samples += three[0];
            ctx->timings.gather_traced += three[1];
            ctx->timings.gather_open += three[2];
          }
          HIP_TRY(launch_gather_list(s, d_tmask.p, mask_words, d_toff.p, d_rs.p, n, ctx->gather.order, d_list.p));
          HIP_TRY(hipMemsetAsync(d_jobc.p + 2, 0, sizeof(unsigned long long), s));
          hipEvent_t a, b2;
          HIP_TRY(hipEventCreate(&a));
          tev.push_back(a);
          HIP_TRY(hipEventCreate(&b2));
          tev.push_back(b2);
          HIP_TRY(hipEventRecord(a, s));
          HIP_TRY(launch_any_stream(s, scb, d_rays.p, d_toff.p + n, n * max_samples, d_rs.p, n, d_contrib.p,
                                    reinterpret_cast<unsigned int*>(d_jobc.p + 2), ctx->gather, d_list.p));
          HIP_TRY(hipEventRecord(b2, s));
          HIP_TRY(launch_gather_sum(s, ent, d_lights.p, tx, bb, jobs, n, d_rs.p, d_contrib.p, ctx->gather.order, d_tmask.p,
                                    mask_words));
        }
        if (sun_bounce && nsjobs) {
          // sun bounce in chunks of jobs: cosine rays, closest hits, sun rays, their visibility, sums
          uint32_t sun_slots = 2;
          for (uint64_t e : be)
            if (hlights[b_ent_light[e]].shape == 3) sun_slots = std::max(sun_slots, hlights[b_ent_light[e]].bounce_samples);
          const uint64_t sj = std::min<uint64_t>(nsjobs, std::max<uint64_t>(1, ctx->gather.chunk_rays / 2 / sun_slots));
          DBuf<float4> d_srays, d_shits;
          DBuf<float> d_scontrib;
          DBuf<uint64_t> d_stotal, d_scount, d_soff;
          DBuf<uint32_t> d_slist;
          HIP_TRY(d_scount.alloc(sj + 1));
          HIP_TRY(d_soff.alloc(sj + 1));
          HIP_TRY(d_slist.alloc(sj * sun_slots));
          HIP_TRY(d_srays.alloc(2 * sj * sun_slots));
          HIP_TRY(d_shits.alloc(sj * sun_slots));
          HIP_TRY(d_scontrib.alloc(3 * sj * sun_slots));
          HIP_TRY(d_stotal.alloc(1));
          for (uint64_t j0 = 0; j0 < nsjobs; j0 += sj) {
            const uint64_t n = std::min(sj, nsjobs - j0);
            const uint2* jobs = d_sjobs.p + j0;
            HIP_TRY(launch_sun_rays(s, tr, ms, d_lights.p, tx, ent, jobs, n, sun_slots, ctx->gather.order, bp, ctx->trig, d_srays.p,
                                    d_shits.p, d_stotal.p));
            HIP_TRY(hipMemsetAsync(d_jobc.p + 2, 0, 2 * sizeof(unsigned long long), s));
            HIP_TRY(launch_closest_stream(s, scb, d_srays.p, d_stotal.p, n * sun_slots, n, d_shits.p,
                                          reinterpret_cast<unsigned int*>(d_jobc.p + 2), ctx->gather));
            HIP_TRY(launch_sun_mid(s, scb, occ, d_lights.p, ent, tx, bb, jobs, n, sun_slots, ctx->gather.order, bp.colored,
                                   d_srays.p, d_shits.p, d_scontrib.p));
            HIP_TRY(launch_ray_list(s, d_srays.p, n, sun_slots, nullptr, ctx->gather.order, d_scount.p, d_soff.p, d_slist.p,
                                    &ctx->scan_tmp, &ctx->scan_tmp_bytes));
            HIP_TRY(launch_any_stream(s, scb, d_srays.p, d_soff.p + n, n * sun_slots, nullptr, n, d_scontrib.p,
                                      reinterpret_cast<unsigned int*>(d_jobc.p + 3), ctx->gather, d_slist.p));
            HIP_TRY(launch_sun_sum(s, tr, ent, d_lights.p, tx, bb, jobs, n, sun_slots, ctx->gather.order, d_srays.p,
                                   d_scontrib.p));
          }
        }
        HIP_TRY(hipStreamSynchronize(s));
        for (size_t k = 0; k + 1 < tev.size(); k += 2) {
          float ms_k = 0.0f;
          (void)hipEventElapsedTime(&ms_k, tev[k], tev[k + 1]);
          ctx->timings.gather_trace_ms += ms_k;
          ctx->timings.gather_trace_launches++;
        }
        for (size_t k = 0; k + 1 < gev.size(); k += 2) {
          float ms_k = 0.0f;
          (void)hipEventElapsedTime(&ms_k, gev[k], gev[k + 1]);
          ctx->timings.gather_rays_ms += ms_k;
          ctx->timings.gather_rays_launches++;
        }
      } else {
        HIP_TRY(launch_bounce_gather(s, scb, occ, tr, ms, d_lights.p, tx, d_tex_start.p, ent, bb, be_units.back(), bp, ctx->trig,
                                     ctx->trace_block));
      }
      HIP_TRY(hipEventRecord(ev1, s));
      HIP_TRY(hipEventSynchronize(ev1));
      float gms = 0.0f;
      (void)hipEventElapsedTime(&gms, ev0, ev1);
      bounce_kernel_ms += gms;
      ctx->timings.bounce_launches++;
      HIP_TRY(launch_bounce_filter(s, bb, d_win_res.p, d_win_start.p, nbw, win_tiles[nbw], blur));
      HIP_TRY(d_has.alloc(nbe + 1));
      HIP_TRY(launch_bounce_has(s, tr, ent, bb, d_has.p));
      HIP_TRY(hipStreamSynchronize(s));
    }
    bounce_ms += ms_since(t_b);
    if (cancelled()) { ctx->err = "the bake was cancelled"; return DL_E_CANCELLED; }

    // ---- optimize, seams, crop, pack, serialize (shadows.rs, seams.rs, triangle_words.rs) ------
    auto t_post = std::chrono::steady_clock::now();
    DBuf<uint8_t> d_keep;
    HIP_TRY(d_keep.alloc(nbe + 1));
    HIP_TRY(launch_optimize(s, tr, ms, ent, rg, d_has.p, d_keep.p));
    std::vector<uint8_t> h_keep, h_has;
    HIP_TRY(download(ctx, h_keep, d_keep.p, nbe));
    if (d_has.p) HIP_TRY(download(ctx, h_has, d_has.p, nbe));
    std::vector<uint8_t> used(nbr, 0);
    std::vector<uint64_t> mesh_base(nm + 1, 0);
    std::vector<uint32_t> ld_off(nbt, 0), kept_shadow_off, kept_bounce_off, first_adj(nm, 0);
    std::vector<uint64_t> tri_kept_start(nbt + 1, 0), kept_entry, kept_word_start, kept_out;
    std::vector<uint64_t> kb_entry, kb_word_start, kb_out, kd_entry, kd_word_start, kd_out;
    uint64_t flat = 0, kb_flat = 0, kd_flat = 0, bw = 0;
    for (uint32_t i = B.m0; i < B.m1; ++i) {
      first_adj[i] = first_tri[i] - t0;
      uint32_t tm0 = first_tri[i] - t0, tm1 = first_tri[i + 1] - t0;
      uint64_t w = 4ull * (tm1 - tm0);
      uint64_t removed = 0;
      for (uint32_t t = tm0; t < tm1; ++t) {
        ld_off[t] = (uint32_t)w;
        uint64_t kept = 0;
        for (uint64_t e = b_ent_start[t]; e < b_ent_start[t + 1]; ++e) kept += h_keep[e];
        removed += (b_ent_start[t + 1] - b_ent_start[t]) - kept;
        w += 1 + stride * kept;
      }
      scene->stats.lights_removed += removed;
      for (uint32_t t = tm0; t < tm1; ++t) {
        tri_kept_start[t] = kept_entry.size();
        const uint32_t* c = &h_crop[4 * (size_t)(t0 + t)];
        uint64_t texels = (uint64_t)(c[2] - c[0] + 5) * (c[3] - c[1] + 5);
        for (uint64_t e = b_ent_start[t]; e < b_ent_start[t + 1]; ++e) {
          if (!h_keep[e]) continue;
          uint32_t r = b_ent_region[e];
          used[r] = 1;
          uint64_t words = (texels + (32 / b_reg_bits[r]) - 1) / (32 / b_reg_bits[r]);
          kept_entry.push_back(e);
          kept_word_start.push_back(flat);
          kept_out.push_back(bw + w);
          kept_shadow_off.push_back((uint32_t)w);
          flat += words;
          w += words;
          uint32_t boff = 0;
          if (!h_has.empty() && h_has[e]) {
            int32_t comp = comp_of_slot[b_ent_light[e]];
            uint32_t bits = bounce_bits(comp);
            uint64_t bwords = opt.colored_bounce ? (texels * 3 * bits + 31) / 32 : (texels + (32 / bits) - 1) / (32 / bits);
            boff = (uint32_t)w;
            if (comp == DL_BOUNCE_BITS8) {
              kb_entry.push_back(e);
              kb_word_start.push_back(kb_flat);
              kb_out.push_back(bw + w);
              kb_flat += bwords;
            } else {
              kd_entry.push_back(e);
              kd_word_start.push_back(kd_flat);
              kd_out.push_back(bw + w);
              kd_flat += bwords;
            }
            w += bwords;
          }
          kept_bounce_off.push_back(boff);
        }
      }
      mesh_base[i] = bw;
      bw += w;
      mesh_base[i + 1] = bw;
    }
    tri_kept_start[nbt] = kept_entry.size();
    kept_word_start.push_back(flat);
    kb_word_start.push_back(kb_flat);
    kd_word_start.push_back(kd_flat);
    std::vector<uint32_t> used_list;
    std::vector<uint64_t> used_area(1, 0);
    for (uint32_t r = 0; r < nbr; ++r)
      if (used[r]) {
        used_list.push_back(r);
        const uint32_t* q = &b_reg_rect[4 * (size_t)r];
        used_area.push_back(used_area.back() + (uint64_t)(q[2] - q[0] + 1) * (q[3] - q[1] + 1));
      }
    std::shared_ptr<DevBlock> cap_block;
    std::vector<unsigned long long> h_rcount, h_wcount;
    if (cap) {  // the regions' levels before seam padding, and their visited bits
      const uint64_t pool_words = b_reg_off[nbr], vis_words = b_vis_off[nbr];
      cap_block = std::make_shared<DevBlock>();
      HIP_TRY(cap_block->alloc(pool_words + vis_words, s));
      if (pool_words) HIP_TRY(hipMemcpyAsync(cap_block->p, d_pool.p, pool_words * 4, hipMemcpyDeviceToDevice, s));
      if (vis_words) HIP_TRY(hipMemcpyAsync(cap_block->p + pool_words, d_rvis.p, vis_words * 4, hipMemcpyDeviceToDevice, s));
      HIP_TRY(download(ctx, h_rcount, d_rcount.p, 2 * (size_t)nbr + 2));
      HIP_TRY(download(ctx, h_wcount, d_wcount.p, (size_t)nbw + 1));
    }
    DBuf<uint32_t> d_used;
    DBuf<uint64_t> d_used_area;
    HIP_TRY(upload(ctx, d_used, used_list));
    HIP_TRY(upload(ctx, d_used_area, used_area));
    HIP_TRY(launch_pad_seams(s, ms, rg, d_used.p, d_used_area.p, (uint32_t)used_list.size(), used_area.back()));
    const uint64_t nkept = kept_entry.size();
    DBuf<uint32_t> d_out, d_ld_off, d_kso, d_kbo, d_first;
    DBuf<uint64_t> d_mesh_base, d_tks, d_ke, d_kws, d_kout, d_kbe, d_kbws, d_kbout, d_kde, d_kdws, d_kdout;
    HIP_TRY(d_out.alloc(bw + 1));
    HIP_TRY(upload(ctx, d_mesh_base, mesh_base));
    HIP_TRY(upload(ctx, d_ld_off, ld_off));
    HIP_TRY(upload(ctx, d_tks, tri_kept_start));
    HIP_TRY(upload(ctx, d_ke, kept_entry));
    HIP_TRY(upload(ctx, d_kws, kept_word_start));
    HIP_TRY(upload(ctx, d_kout, kept_out));
    HIP_TRY(upload(ctx, d_kso, kept_shadow_off));
    HIP_TRY(upload(ctx, d_kbo, kept_bounce_off));
    HIP_TRY(upload(ctx, d_first, first_adj));
    HIP_TRY(launch_serialize(s, tr, d_mesh_base.p, d_ld_off.p, d_tks.p, d_ke.p, d_kso.p, d_kbo.p, d_bel.p, stride, d_first.p,
                             d_out.p));
    HIP_TRY(launch_crop(s, tr, ent, rg, d_ke.p, d_kws.p, nkept, flat, d_kout.p, d_out.p));
    if (!kb_entry.empty()) {
      HIP_TRY(upload(ctx, d_kbe, kb_entry));
      HIP_TRY(upload(ctx, d_kbws, kb_word_start));
      HIP_TRY(upload(ctx, d_kbout, kb_out));
      HIP_TRY(launch_bounce_pack(s, tr, ms, d_lights.p, ent, bb, d_kbe.p, d_kbws.p, kb_entry.size(), kb_flat, d_kbout.p, d_comp.p,
                                 opt.colored_bounce ? 1u : 0u, opt.seed, 0, d_out.p));
    }
    if (!kd_entry.empty()) {
      HIP_TRY(upload(ctx, d_kde, kd_entry));
      HIP_TRY(upload(ctx, d_kdws, kd_word_start));
      HIP_TRY(upload(ctx, d_kdout, kd_out));
      HIP_TRY(launch_bounce_pack(s, tr, ms, d_lights.p, ent, bb, d_kde.p, d_kdws.p, kd_entry.size(), kd_flat, d_kdout.p, d_comp.p,
                                 opt.colored_bounce ? 1u : 0u, opt.seed, 1, d_out.p));
    }
    if (ctx->capture_bounce && !(kb_entry.empty() && kd_entry.empty())) {
      // test hook: the f32 crops the packing above read, entry by entry
      std::vector<uint64_t> cap_entry(kb_entry), vstart(1, 0);
      cap_entry.insert(cap_entry.end(), kd_entry.begin(), kd_entry.end());
      for (uint64_t e : cap_entry) {
        const uint32_t* c = &h_crop[4 * (size_t)(t0 + b_ent_tri[e])];
        vstart.push_back(vstart.back() + 3ull * (c[2] - c[0] + 5) * (c[3] - c[1] + 5));
      }
      DBuf<uint64_t> d_ce, d_vs;
      DBuf<float> d_vals;
      HIP_TRY(upload(ctx, d_ce, cap_entry));
      HIP_TRY(upload(ctx, d_vs, vstart));
      HIP_TRY(d_vals.alloc(vstart.back() + 1));
      HIP_TRY(launch_bounce_dump(s, tr, ent, bb, d_ce.p, cap_entry.size(), d_vs.p, d_vals.p));
      std::vector<float> vals;
      HIP_TRY(download(ctx, vals, d_vals.p, vstart.back()));
      for (size_t k = 0; k < cap_entry.size(); ++k) {
        const uint64_t e = cap_entry[k];
        const uint32_t tg = t0 + b_ent_tri[e];
        const uint32_t mi = (uint32_t)(std::upper_bound(first_tri.begin(), first_tri.begin() + nm + 1, tg) - first_tri.begin()) - 1;
        const uint32_t* c = &h_crop[4 * (size_t)tg];
        dl_ctx::CapturedCrop cc{meta[mi].id, lights[b_ent_light[e]]->id, tg - first_tri[mi], c[2] - c[0] + 5, c[3] - c[1] + 5,
                                ctx->captured_values.size()};
        ctx->captured.push_back(cc);
        ctx->captured_values.insert(ctx->captured_values.end(), vals.begin() + (long)vstart[k], vals.begin() + (long)vstart[k + 1]);
      }
    }
    if (cap) {
      // the batch's words stay on the device: the kept entries' bounce payloads are read from them
      auto wblock = std::make_shared<DevBlock>();
      HIP_TRY(wblock->alloc(bw, s));
      if (bw) HIP_TRY(hipMemcpyAsync(wblock->p, d_out.p, bw * 4, hipMemcpyDeviceToDevice, s));
      std::vector<uint64_t> boff(nbe, 0), bwords(nbe, 0);  // bounce payload of each kept entry
      for (size_t k = 0; k < kb_entry.size(); ++k) {
        boff[kb_entry[k]] = kb_out[k];
        bwords[kb_entry[k]] = kb_word_start[k + 1] - kb_word_start[k];
      }
      for (size_t k = 0; k < kd_entry.size(); ++k) {
        boff[kd_entry[k]] = kd_out[k];
        bwords[kd_entry[k]] = kd_word_start[k + 1] - kd_word_start[k];
      }
      const size_t first = cap->pairs.size();
      for (uint32_t r = 0; r < nbr; ++r) {
        CapturedPair c;
        c.mesh = b_reg_mesh[r];
        c.light_id = lights[reg_light[r0 + r]]->id;
        std::memcpy(c.rect, &b_reg_rect[4 * (size_t)r], 16);
        c.bits = b_reg_bits[r];
        c.grid_off = b_reg_off[r];
        c.grid_words = b_reg_off[r + 1] - b_reg_off[r];
        c.vis_off = b_reg_off[nbr] + b_vis_off[r];
        c.vis_words = b_vis_off[r + 1] - b_vis_off[r];
        c.rays = h_rcount[r];
        c.samples = h_rcount[nbr + 1 + r];
        c.bounce_rays = win_of_reg[r0 + r] >= 0 ? h_wcount[(size_t)(win_of_reg[r0 + r] - w0)] : 0;
        c.entries = 0;
        c.block = cap_block;
        c.words = wblock;
        cap->pairs.push_back(std::move(c));
      }
      for (uint32_t t = 0; t < nbt; ++t)
        for (uint64_t e = b_ent_start[t]; e < b_ent_start[t + 1]; ++e) {
          CapturedPair& c = cap->pairs[first + b_ent_region[e]];
          c.entries++;
          if (!h_keep[e]) continue;
          c.kept.push_back(t0 + t - first_tri[c.mesh]);
          c.bounce_off.push_back(boff[e]);
          c.bounce_words.push_back(bwords[e]);
        }
    }
    // the batch's words in one buffer, each mesh holding its slice
    auto words = std::make_shared<std::vector<uint32_t, NoInit<uint32_t>>>(bw);
    HIP_TRY(copy_to_host(ctx, words->data(), d_out.p, bw * sizeof(uint32_t)));
    for (uint32_t i = B.m0; i < B.m1; ++i) {
      dl_baked_mesh_out& mo = scene->meshes[i];
      mo.id = meta[i].id;
      mo.content_hash = meta[i].content_hash;
      mo.resolution = res[i];
      mo.occludes = meta[i].occludes;
      mo.words.share(words, words->data() + mesh_base[i], mesh_base[i + 1] - mesh_base[i]);
    }
    post_ms += ms_since(t_post);
    advance(B.m1 - B.m0);
    (void)dithered_bounce;
  }
  if (fill_res) {  // join the batches' texels into the resident scene's, in triangle order
    uint64_t total = 0;
    for (auto& piece : resident->pieces) total += piece->n;
    HIP_TRY(resident->t_xy.alloc(total + 1));
    HIP_TRY(resident->t_tri.alloc(total + 1));
    HIP_TRY(resident->t_pos.alloc(total + 1));
    HIP_TRY(resident->t_nrm.alloc(total + 1));
    resident->tex_start.clear();
    resident->tex_start.reserve((size_t)nt + 1);
    uint64_t off = 0;
    for (auto& piece : resident->pieces) {
      const uint64_t n = piece->n;
      if (n) {
        HIP_TRY(hipMemcpyAsync(resident->t_xy.p + off, piece->xy.p, n * 4, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(resident->t_tri.p + off, piece->tri.p, n * 4, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(resident->t_pos.p + off, piece->pos.p, n * 16, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(resident->t_nrm.p + off, piece->nrm.p, n * 16, hipMemcpyDeviceToDevice, s));
      }
      for (size_t t = 0; t + 1 < piece->start.size(); ++t) resident->tex_start.push_back(off + piece->start[t]);
      off += n;
    }
    resident->tex_start.push_back(off);
    HIP_TRY(hipStreamSynchronize(s));
    resident->pieces.clear();
    if (resident->tex_start.size() != (size_t)nt + 1) { ctx->err = "resident texels do not cover the scene"; return DL_E_DEVICE; }
  }
  ctx->timings.texels_ms += texels_ms;
  ctx->timings.trace_ms = trace_ms;
  ctx->timings.trace_kernel_ms = trace_kernel_ms;
  ctx->timings.bounce_ms = bounce_ms;
  ctx->timings.post_ms = post_ms;
  ctx->timings.bounce_kernel_ms = bounce_kernel_ms;
  ctx->timings.burial_ms = burial_ms;
  if (ctx->counting) {
    unsigned long long c[7] = {0, 0, 0, 0, 0, 0, 0};
    HIP_TRY(copy_to_host(ctx, c, d_counters.p, 56));
    ctx->timings.gather_photon_reads = c[6];
    ctx->timings.burial_rays = burial_rays;
    ctx->timings.burial_node_visits = c[4];
    ctx->timings.burial_tri_tests = c[5];
    ctx->timings.trace_node_visits = c[0];
    ctx->timings.trace_tri_tests = c[1];
    ctx->timings.bounce_node_visits = c[2];
    ctx->timings.bounce_tri_tests = c[3];
  }

  // ---- distance cubes (distance_cubes.rs:29-72) -----------------------------------------------
  auto t_maps = std::chrono::steady_clock::now();
  stage(3, (uint64_t)nlocal * 6);
  if (opt.dynamic_geometry != DL_DYNAMIC_LIGHTING_ONLY && nlocal > 0) {
    DBuf<uint32_t> d_cubes;
    HIP_TRY(d_cubes.alloc((size_t)nlocal * 6 * 2048));
    HIP_TRY(launch_distance_cubes(s, sc, d_lights.p, nlocal, d_cubes.p));
    HIP_TRY(download(ctx, scene->distance_cubes, d_cubes.p, (size_t)nlocal * 6 * 2048));
    if (cap)
      for (uint32_t i = 0; i < nlocal; ++i)
        cap->cubes.push_back({lights[i]->id, std::vector<uint32_t>(scene->distance_cubes.begin() + (long)i * 6 * 2048,
                                                                   scene->distance_cubes.begin() + (long)(i + 1) * 6 * 2048)});
  }
  // ---- sun maps (sun_map.rs:25-75) and sky probes (sky_probes.rs:146-191), baker.rs:104-152 --
  if (opt.dynamic_geometry != DL_DYNAMIC_LIGHTING_ONLY) {
    bool sky = false;
    uint32_t sky_alpha = 0;
    for (uint32_t l = nlocal; l < nl; ++l) {
      if (hlights[l].shape == DL_SHAPE_SKY) {
        sky = true;
        sky_alpha |= hlights[l].alpha;
      }
      if (hlights[l].shape != DL_SHAPE_SUN) continue;
      dl_sun_map_out m{};
      SunMapParams sp{};
      sp.alpha = hlights[l].alpha;
      if (!sun_map_setup(occ9.data(), (size_t)nocc * 3, hlights[l].toward, opt.sun_map_resolution, &sp, m.rows)) {
        m.resolution = 1;
        m.texels.push_back(pack_sun_depth(1.0f, 1.0f));
      } else {
        m.resolution = opt.sun_map_resolution;
        const uint64_t words = ((uint64_t)m.resolution * m.resolution + 1) / 2;
        DBuf<uint32_t> d_map;
        HIP_TRY(d_map.alloc(words));
        HIP_TRY(launch_sun_map(s, sc, sp, d_map.p, words));
        HIP_TRY(download(ctx, m.texels, d_map.p, words));
      }
      scene->sun_maps.push_back(std::move(m));
    }
    if (sky && nocc > 0) {
      f3 mn = mk3(occ9[0], occ9[1], occ9[2]), mx = mn;
      for (size_t i = 0; i < (size_t)nocc * 3; ++i) {
        const f3 v = mk3(occ9[3 * i], occ9[3 * i + 1], occ9[3 * i + 2]);
        mn = vmin(mn, v);
        mx = vmax(mx, v);
      }
      SkyProbeParams pp{};
      f3 origin;
      probe_grid(mn, mx, opt.sky_probe_spacing, &origin, &pp.spacing, pp.counts);
      pp.origin[0] = origin.x;
      pp.origin[1] = origin.y;
      pp.origin[2] = origin.z;
      pp.samples = opt.sky_samples;
      pp.alpha = sky_alpha;
      pp.seed = opt.seed;
      const size_t n = (size_t)pp.counts[0] * pp.counts[1] * pp.counts[2];
      DBuf<float> d_open;
      DBuf<uint32_t> d_back;
      HIP_TRY(d_open.alloc(n * 6));
      HIP_TRY(d_back.alloc(n * 6));
      HIP_TRY(launch_sky_probes(s, sc, ctx->trig, pp, d_open.p, d_back.p));
      std::vector<float> open;
      std::vector<uint32_t> back;
      HIP_TRY(download(ctx, open, d_open.p, n * 6));
      HIP_TRY(download(ctx, back, d_back.p, n * 6));
      // a probe is inside geometry when more than a quarter of its rays reach a back face
      std::vector<uint8_t> valid(n);
      const float rays = (float)(std::max(opt.sky_samples, 1u) * 6u);
      for (size_t i = 0; i < n; ++i) {
        uint32_t b = 0;
        for (int a = 0; a < 6; ++a) b += back[i * 6 + a];
        valid[i] = (float)b <= 0.25f * rays ? 1 : 0;
      }
      fill_inside(pp.counts, open, valid);
      scene->probes = true;
      std::memcpy(scene->probe_origin, pp.origin, sizeof pp.origin);
      scene->probe_spacing = pp.spacing;
      std::memcpy(scene->probe_counts, pp.counts, sizeof pp.counts);
      scene->probe_words.resize(n * 2);
      const float ones[6] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
      for (size_t i = 0; i < n; ++i) pack_probe(valid[i] ? &open[i * 6] : ones, &scene->probe_words[i * 2]);
    }
  }
  ctx->timings.maps_ms = ms_since(t_maps);

  // ---- stats, settings, lights --------------------------------------------------------------
  auto t_down = std::chrono::steady_clock::now();
  std::vector<unsigned long long> h_rays, h_samples, h_brays;
  HIP_TRY(download(ctx, h_rays, d_mrays.p, nm));
  HIP_TRY(download(ctx, h_samples, d_msamples.p, nm));
  HIP_TRY(download(ctx, h_brays, d_mbrays.p, nm));
  uint32_t overflow = 0;
  HIP_TRY(copy_to_host(ctx, &overflow, d_overflow.p, 4));
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
    // BakedLight keeps only what its LightShape holds: an area light's rotation and size, a sun's
    // rotation and angular diameter, nothing for points and skies
    const float identity[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    std::memcpy(b.rotation, identity, 16);
    if (l->shape == DL_SHAPE_RECTANGLE || l->shape == DL_SHAPE_TUBE) {
      std::memcpy(b.rotation, l->rotation, 16);
      std::memcpy(b.size, l->size, 8);
    } else if (l->shape == DL_SHAPE_SUN) {
      std::memcpy(b.rotation, l->rotation, 16);
      b.angular_diameter = l->angular_diameter;
    }
    b.shadow_bits_per_texel = hlights[i].bits;
    b.bounce_compression = hlights[i].bakes_bounce ? comp_of_slot[i] : -1;
    scene->lights.push_back(b);
  }
  scene->bvh = lbvh.nodes;
  scene->stats.meshes = nm;
  scene->stats.triangles = ntri;
  for (uint32_t i = 0; i < nm; ++i) {
    scene->stats.triangle_words += scene->meshes[i].words.size();
    scene->stats.rays += h_rays[i];
    scene->stats.samples += h_samples[i];
    scene->stats.bounce_rays += h_brays[i];
  }
  ctx->timings.trace_rays = scene->stats.rays;
  if (fill_res) {  // the geometry of this bake becomes the resident scene
    resident->meshes = std::move(meta_local);
    resident->res = std::move(res_local);
    resident->first_tri = std::move(first_local);
    resident->mid = std::move(mid_local);
    resident->owned_off = std::move(owned_local);
    resident->ntri = ntri;
    resident->occ9 = std::move(occ9_local);
    resident->has_alpha = mesh_alpha;
    resident->ready = true;
  }
  if (cap && cap->escaping) {
    unsigned long long n = 0;
    HIP_TRY(copy_to_host(ctx, &n, cap->escaping->p + 8 * cap->escaping_cap, 8));
    cap->escaping_count = n;
  }
  stage(4, 0);
  ctx->timings.total_ms = ms_since(t_start);
  *out = scene.release();
  return DL_OK;
}

// The C ABI of libdynlight_hip.so (include/dynlight.h, include/dynlight_testing.h).  Nothing
// unwinds across it: every entry point catches, stores the message and returns a status.
#include <cstring>
#include <new>
#include <thread>

#include "../../include/dynlight_testing.h"
#include "dl_host.h"

namespace {
thread_local std::string g_null_ctx_err;

int set_err(dl_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  else g_null_ctx_err = msg;
  return code;
}

int upload_trig(dl_ctx* ctx) {
  using namespace dl;
  const TrigHost& t = ctx->trig_host;
  // four arrays, each padded to an even word count (8-byte alignment of the run pairs)
  auto words = [](const std::vector<uint32_t>& v) { return (v.size() + 2) & ~(size_t)1; };
  size_t total = words(t.sin_runs) + words(t.cos_runs) + words(t.sin_vals) + words(t.cos_vals);
  hipError_t e = ctx->trig_buf.alloc(total);
  if (e != hipSuccess) return set_err(ctx, DL_E_OOM, "allocating the sin/cos tables failed");
  uint32_t* base = ctx->trig_buf.p;
  size_t off = 0;
  hipError_t copy = hipSuccess;
  auto put = [&](const std::vector<uint32_t>& v) -> const uint32_t* {
    const uint32_t* p = base + off;
    if (!v.empty() && copy == hipSuccess) copy = hipMemcpy(base + off, v.data(), v.size() * 4, hipMemcpyHostToDevice);
    off += words(v);
    return p;
  };
  ctx->trig.sin.runs = reinterpret_cast<const uint2*>(put(t.sin_runs));
  ctx->trig.cos.runs = reinterpret_cast<const uint2*>(put(t.cos_runs));
  ctx->trig.sin.vals = put(t.sin_vals);
  ctx->trig.cos.vals = put(t.cos_vals);
  ctx->trig.band = t.band;
  if (copy != hipSuccess || off > total) return set_err(ctx, DL_E_DEVICE, "uploading the sin/cos tables failed");
  return DL_OK;
}
}  // namespace

namespace {
const dl::TrigHost& host_trig() {
  static std::once_flag once;
  static dl::TrigHost host;
  std::call_once(once, [] { dl::build_trig_tables(host, std::max(1u, std::thread::hardware_concurrency())); });
  return host;
}
}  // namespace

extern "C" {

int dl_ctx_create(int device, dl_ctx** out) {
  if (!out) return set_err(nullptr, DL_E_ARGUMENT, "null output pointer");
  *out = nullptr;
  try {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      return set_err(nullptr, DL_E_DEVICE, "no HIP device is available");
    if (device < 0 || device >= count) return set_err(nullptr, DL_E_ARGUMENT, "device index out of range");
    dl_ctx* ctx = new dl_ctx();
    ctx->device = device;
    ctx->host_threads = std::max(1u, std::thread::hardware_concurrency());
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return set_err(nullptr, DL_E_DEVICE, "creating the HIP stream failed");
    }
    dl::build_trig_tables(ctx->trig_host, ctx->host_threads);
    int st = upload_trig(ctx);
    if (st) {
      g_null_ctx_err = ctx->err;
      dl_ctx_destroy(ctx);
      return st;
    }
    *out = ctx;
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return set_err(nullptr, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_err(nullptr, DL_E_DEVICE, "unexpected failure");
  }
}

void dl_ctx_destroy(dl_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  ctx->trig_buf.release();
  if (ctx->scan_tmp) (void)hipFree(ctx->scan_tmp);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* dl_last_error(const dl_ctx* ctx) { return ctx ? ctx->err.c_str() : g_null_ctx_err.c_str(); }

void dl_default_options(dl_bake_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->pixel_density = 128.0f;
  o->max_lightmap_size = 2048;
  o->bounce_compression = DL_BOUNCE_BITS8;
  o->dynamic_geometry = DL_DYNAMIC_ANGULAR;
  o->skip_bounce = 0;
  o->photon_cube_resolution = 512;
  o->area_light_samples = 64;
  o->area_light_min_samples = 16;
  o->sun_map_resolution = 1024;
  o->sky_samples = 64;
  o->sky_probe_spacing = 1.0f;
  o->colored_bounce = 0;
  o->seed = 0;
}

int dl_bake(dl_ctx* ctx, const dl_bake_input* input, const dl_bake_options* options, dl_progress* progress,
            dl_baked_scene** out) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "hipSetDevice failed");
    return dl_bake_impl(ctx, input, options, progress, out);
  } catch (const std::bad_alloc&) {
    return set_err(ctx, DL_E_OOM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_err(ctx, DL_E_DEVICE, e.what());
  } catch (...) {
    return set_err(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

void dl_baked_scene_free(dl_baked_scene* s) { delete s; }
void dl_scene_settings(const dl_baked_scene* s, dl_baked_settings* out) { *out = s->settings; }
void dl_scene_stats(const dl_baked_scene* s, dl_bake_stats* out) { *out = s->stats; }
uint32_t dl_scene_light_count(const dl_baked_scene* s) { return (uint32_t)s->lights.size(); }
int dl_scene_light(const dl_baked_scene* s, uint32_t i, dl_baked_light* out) {
  if (!s || i >= s->lights.size()) return DL_E_ARGUMENT;
  *out = s->lights[i];
  return DL_OK;
}
uint32_t dl_scene_light_bvh(const dl_baked_scene* s, const dl_light_bvh_node** nodes) {
  *nodes = s->bvh.data();
  return (uint32_t)s->bvh.size();
}
uint32_t dl_scene_mesh_count(const dl_baked_scene* s) { return (uint32_t)s->meshes.size(); }
int dl_scene_mesh(const dl_baked_scene* s, uint32_t i, dl_baked_mesh_view* out) {
  if (!s || i >= s->meshes.size()) return DL_E_ARGUMENT;
  const auto& m = s->meshes[i];
  out->id = m.id;
  out->content_hash = m.content_hash;
  out->resolution = m.resolution;
  out->occludes = m.occludes;
  out->triangles = m.words.data();
  out->triangle_word_count = m.words.size();
  return DL_OK;
}
uint64_t dl_scene_distance_cubes(const dl_baked_scene* s, const uint32_t** words) {
  *words = s->distance_cubes.data();
  return s->distance_cubes.size();
}
uint32_t dl_scene_sun_map_count(const dl_baked_scene* s) { return s ? (uint32_t)s->sun_maps.size() : 0u; }
int dl_scene_sun_map(const dl_baked_scene* s, uint32_t i, dl_sun_map_view* out) {
  if (!s || !out || i >= s->sun_maps.size()) return DL_E_ARGUMENT;
  const dl_sun_map_out& m = s->sun_maps[i];
  std::memcpy(out->rows, m.rows, sizeof m.rows);
  out->resolution = m.resolution;
  out->texels = m.texels.data();
  out->texel_word_count = m.texels.size();
  return DL_OK;
}
void dl_scene_sky_probes(const dl_baked_scene* s, dl_sky_probes_view* out) {
  if (!out) return;
  std::memset(out, 0, sizeof *out);
  if (!s || !s->probes) return;
  out->present = 1;
  std::memcpy(out->origin, s->probe_origin, sizeof s->probe_origin);
  out->spacing = s->probe_spacing;
  std::memcpy(out->counts, s->probe_counts, sizeof s->probe_counts);
  out->words = s->probe_words.data();
  out->word_count = s->probe_words.size();
}

int dl_last_timings(const dl_ctx* ctx, dl_timings* out) {
  if (!ctx || !out) return DL_E_ARGUMENT;
  *out = ctx->timings;
  return DL_OK;
}

int dl_test_ray_queries(dl_ctx* ctx, const float* tris9, uint32_t tri_count, const float* rays8, uint64_t ray_count,
                        int closest, float* t_out, uint32_t* hit_or_index) {
  using namespace dl;
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  Bvh4 bvh;
  build_bvh4(tris9, tri_count, ctx->host_threads, bvh);
  DBuf<CNode> n;
  DBuf<float> t, r, tout;
  DBuf<uint32_t> idx, ovf;
  auto fail = [&](hipError_t e) { return set_err(ctx, e == hipErrorOutOfMemory ? DL_E_OOM : DL_E_DEVICE, hipGetErrorString(e)); };
  hipError_t e;
  if ((e = n.alloc(bvh.nodes.size())) || (e = t.alloc(bvh.tris.size())) || (e = r.alloc(ray_count * 8 + 1)) ||
      (e = tout.alloc(ray_count + 1)) || (e = idx.alloc(ray_count + 1)) || (e = ovf.alloc(1)))
    return fail(e);
  if (!bvh.nodes.empty()) (void)hipMemcpy(n.p, bvh.nodes.data(), bvh.nodes.size() * sizeof(CNode), hipMemcpyHostToDevice);
  if (!bvh.tris.empty()) (void)hipMemcpy(t.p, bvh.tris.data(), bvh.tris.size() * 4, hipMemcpyHostToDevice);
  if (ray_count) (void)hipMemcpy(r.p, rays8, ray_count * 32, hipMemcpyHostToDevice);
  (void)hipMemset(ovf.p, 0, 4);
  DevScene sc{};
  sc.nodes = n.p;
  sc.tris = reinterpret_cast<const float4*>(t.p);
  sc.ntris = tri_count;
  sc.stack = 3 * std::max(1u, bvh.depth) + 1;
  sc.stack16 = bvh.nodes.size() < 65536 ? 1u : 0u;
  sc.overflow = ovf.p;
  sc.counters = nullptr;
  if ((e = launch_ray_queries(ctx->stream, sc, r.p, ray_count, closest, tout.p, idx.p))) return fail(e);
  if ((e = hipStreamSynchronize(ctx->stream))) return fail(e);
  if (ray_count) {
    if (closest && t_out) (void)hipMemcpy(t_out, tout.p, ray_count * 4, hipMemcpyDeviceToHost);
    (void)hipMemcpy(hit_or_index, idx.p, ray_count * 4, hipMemcpyDeviceToHost);
  }
  uint32_t o = 0;
  (void)hipMemcpy(&o, ovf.p, 4, hipMemcpyDeviceToHost);
  if (o) return set_err(ctx, DL_E_DEVICE, "BVH traversal stack overflow");
  return DL_OK;
}

int dl_test_trig(dl_ctx* ctx, const float* x, uint64_t n, float* sin_out, float* cos_out) {
  using namespace dl;
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  DBuf<float> dx, ds, dc;
  hipError_t e;
  if ((e = dx.alloc(n + 1)) || (e = ds.alloc(n + 1)) || (e = dc.alloc(n + 1)))
    return set_err(ctx, DL_E_OOM, hipGetErrorString(e));
  if (n) (void)hipMemcpy(dx.p, x, n * 4, hipMemcpyHostToDevice);
  if ((e = launch_trig_eval(ctx->stream, ctx->trig, dx.p, n, ds.p, dc.p)) || (e = hipStreamSynchronize(ctx->stream)))
    return set_err(ctx, DL_E_DEVICE, hipGetErrorString(e));
  if (n) {
    (void)hipMemcpy(sin_out, ds.p, n * 4, hipMemcpyDeviceToHost);
    (void)hipMemcpy(cos_out, dc.p, n * 4, hipMemcpyDeviceToHost);
  }
  return DL_OK;
}


int dl_test_trig_exhaustive(uint32_t lo, uint32_t hi, uint64_t* mismatches) {
  using namespace dl;
  const TrigHost& host = host_trig();
  TrigTables t;
  t.sin = {reinterpret_cast<const uint2*>(host.sin_runs.data()), host.sin_vals.data()};
  t.cos = {reinterpret_cast<const uint2*>(host.cos_runs.data()), host.cos_vals.data()};
  t.band = host.band;
  const unsigned n = std::max(1u, std::thread::hardware_concurrency());
  std::vector<uint64_t> bad(n, 0);
  std::vector<std::thread> ts;
  for (unsigned k = 0; k < n; ++k)
    ts.emplace_back([&, k] {
      for (uint64_t b = (uint64_t)lo + k; b <= hi; b += n) {
        const float x = u2f((uint32_t)b);
        if (f2u(lsinf(t, x)) != f2u(sinf(x))) bad[k]++;
        if (f2u(lcosf(t, x)) != f2u(cosf(x))) bad[k]++;
      }
    });
  for (auto& th : ts) th.join();
  *mismatches = 0;
  for (uint64_t v : bad) *mismatches += v;
  return DL_OK;
}

uint64_t dl_test_trig_keys(int which, uint32_t* out, uint64_t cap) {
  const dl::TrigHost& host = host_trig();
  const std::vector<uint32_t>& keys = which ? host.cos_keys : host.sin_keys;
  if (out) std::memcpy(out, keys.data(), std::min<uint64_t>(cap, keys.size()) * 4);
  return keys.size();
}

int dl_test_trig_host(const float* x, uint64_t n, float* sin_out, float* cos_out) {
  using namespace dl;
  const TrigHost& host = host_trig();
  TrigTables t;
  t.sin = {reinterpret_cast<const uint2*>(host.sin_runs.data()), host.sin_vals.data()};
  t.cos = {reinterpret_cast<const uint2*>(host.cos_runs.data()), host.cos_vals.data()};
  t.band = host.band;
  for (uint64_t i = 0; i < n; ++i) {
    sin_out[i] = lsinf(t, x[i]);
    cos_out[i] = lcosf(t, x[i]);
  }
  return DL_OK;
}

void dl_test_trig_table_sizes(const dl_ctx* ctx, uint64_t* s, uint64_t* c) {
  *s = ctx ? ctx->trig_host.sin_keys.size() : 0;
  *c = ctx ? ctx->trig_host.cos_keys.size() : 0;
}

int dl_set_counting(dl_ctx* ctx, int enable) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->counting = enable != 0;
  return DL_OK;
}

int dl_set_launch(dl_ctx* ctx, uint32_t trace_block, uint32_t burial_block) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  auto ok = [](uint32_t b) { return b == 0 || (b % 64 == 0 && b >= 64 && b <= 1024); };
  if (!ok(trace_block) || !ok(burial_block)) return set_err(ctx, DL_E_ARGUMENT, "block sizes must be multiples of 64 in [64, 1024]");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (trace_block) ctx->trace_block = trace_block;
  if (burial_block) ctx->burial_block = burial_block;
  return DL_OK;
}

int dl_set_gather(dl_ctx* ctx, uint32_t mode, uint32_t block, uint32_t chunk, uint64_t chunk_rays, uint32_t blocks_per_cu,
                  uint32_t min_waves, uint32_t refill) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  if (mode > 1) return set_err(ctx, DL_E_ARGUMENT, "gather mode must be 0 or 1");
  if (block && (block % 64 != 0 || block > 1024)) return set_err(ctx, DL_E_ARGUMENT, "block must be a multiple of 64 up to 1024");
  if (chunk_rays && chunk_rays < 4096) return set_err(ctx, DL_E_ARGUMENT, "chunks need at least 4096 rays");
  if (blocks_per_cu > 64) return set_err(ctx, DL_E_ARGUMENT, "at most 64 blocks per CU");
  if (chunk > (1u << 20)) return set_err(ctx, DL_E_ARGUMENT, "at most 2^20 rays per queue run");
  if (min_waves != 0 && min_waves != 8) return set_err(ctx, DL_E_ARGUMENT, "min_waves must be 0 or 8");
  if (refill > 64) return set_err(ctx, DL_E_ARGUMENT, "refill must be at most 64");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->gather_mode = mode;
  if (block) ctx->gather.block = block;
  if (chunk) ctx->gather.chunk = chunk;
  if (chunk_rays) ctx->gather.chunk_rays = chunk_rays;
  ctx->gather.blocks_per_cu = blocks_per_cu;
  ctx->gather.min_waves = min_waves;
  if (refill) ctx->gather.refill = refill;
  return DL_OK;
}

}  // extern "C"

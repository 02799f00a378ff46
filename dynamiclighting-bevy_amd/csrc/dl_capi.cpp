// The C ABI of libdynlight_hip.so (include/dynlight.h, include/dynlight_testing.h).  Nothing
// unwinds across it: every entry point catches, stores the message and returns a status.
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
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
  size_t total = words(t.sin_runs) + words(t.cos_runs) + words(t.sin_vals) + words(t.cos_vals) + words(t.sin_any) +
                 words(t.cos_any);
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
  ctx->trig.sin.any = put(t.sin_any);
  ctx->trig.cos.any = put(t.cos_any);
  ctx->trig.tol = t.tol;
  if (copy != hipSuccess || off > total) return set_err(ctx, DL_E_DEVICE, "uploading the sin/cos tables failed");
  return DL_OK;
}
}  // namespace

namespace {
// CPUs this process may run on (its affinity mask), which can be fewer than the machine has
unsigned usable_cpus() {
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
  return std::max(1u, std::thread::hardware_concurrency());
}

// A rough relative cost of baking one job, to order a batch: texels estimated from the mesh
// areas and the density, times the rays a texel casts (one per light, 16 per area light, 32 more
// per bounce light), plus the photon rays of the bounce lights.
double job_cost(const dl_bake_input& in, const dl_bake_options& o) {
  double area = 0.0;
  for (uint32_t m = 0; m < in.mesh_count; ++m) {
    const dl_mesh& mesh = in.meshes[m];
    if (!mesh.positions || !mesh.indices) continue;
    for (uint32_t t = 0; t + 2 < mesh.index_count; t += 3) {
      const uint32_t a = mesh.indices[t], b = mesh.indices[t + 1], c = mesh.indices[t + 2];
      if (a >= mesh.vertex_count || b >= mesh.vertex_count || c >= mesh.vertex_count) continue;
      const float* p = mesh.positions;
      const double e1[3] = {p[3 * b] - (double)p[3 * a], p[3 * b + 1] - (double)p[3 * a + 1], p[3 * b + 2] - (double)p[3 * a + 2]};
      const double e2[3] = {p[3 * c] - (double)p[3 * a], p[3 * c + 1] - (double)p[3 * a + 1], p[3 * c + 2] - (double)p[3 * a + 2]};
      const double x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2], z = e1[0] * e2[1] - e1[1] * e2[0];
      const double a2 = std::sqrt(x * x + y * y + z * z) * 0.5;
      if (std::isfinite(a2)) area += a2;
    }
  }
  double rays = 1.0, photons = 0.0;
  for (uint32_t l = 0; l < in.light_count; ++l) {
    const dl_light& L = in.lights[l];
    rays += (L.shape == DL_SHAPE_RECTANGLE || L.shape == DL_SHAPE_TUBE) ? 16.0 : 1.0;
    if (L.illumination == DL_ILLUMINATION_SINGLE_BOUNCE) {
      rays += 32.0;
      photons += 6.0 * o.photon_cube_resolution * (double)o.photon_cube_resolution;
    }
  }
  const double density = std::isfinite(o.pixel_density) ? o.pixel_density : 0.0;
  return area * density * density * rays + photons;
}

// Returns the device pool's unused memory beyond 4 GiB to the driver (after every stream that
// used it is idle).
void trim_pool(int device) {
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) (void)hipMemPoolTrimTo(pool, 4ull << 30);
}

// Worker contexts 0..streams-1 of ctx, each with its own stream and the parent's sin/cos tables,
// set up to share the device `streams` ways.  Returns false when a stream cannot be created.
bool prepare_workers(dl_ctx* ctx, uint32_t streams) {
  const unsigned hw = usable_cpus();
  while (ctx->workers.size() < streams) {
    dl_ctx* w = new dl_ctx();
    w->device = ctx->device;
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
      delete w;
      return false;
    }
    w->trig = ctx->trig;  // read-only tables owned by ctx (the worker's trig_buf stays empty)
    ctx->workers.push_back(w);
  }
  for (uint32_t k = 0; k < streams; ++k) {
    dl_ctx* w = ctx->workers[k];
    w->host_threads = std::max(1u, hw / streams);
    w->share = streams;
    w->cache.cap = std::max<size_t>(4ull << 30, (32ull << 30) / streams);
    w->trace_block = ctx->trace_block;
    w->burial_block = ctx->burial_block;
    w->burial_lanes = ctx->burial_lanes;
    w->gather = ctx->gather;
    // A worker's persistent ray-stream grid fills half of the device: a grid of 1/streams left
    // CUs idle whenever fewer streams had one in flight (the 512-job batch took 4.35 s on 8
    // streams that way, 3.91 s with half-device grids; whole-device grids 3.98 s).
    w->gather.share = std::min<uint32_t>(streams, 2);
    // smaller ray chunks, so that each worker's gather buffers stay within its cache
    w->gather.chunk_rays = std::max<uint64_t>(1ull << 22, ctx->gather.chunk_rays / streams);
    w->gather_mode = ctx->gather_mode;
    w->trace_stream = ctx->trace_stream;
    w->photon_tiles = ctx->photon_tiles;
    w->gather_split = ctx->gather_split;
    w->lazy_photons = ctx->lazy_photons;
    w->counting = false;
  }
  return true;
}

const dl::TrigHost& host_trig() {
  static std::once_flag once;
  static dl::TrigHost host;
  std::call_once(once, [] { dl::build_trig_tables(host, usable_cpus()); });
  return host;
}
}  // namespace

void dl_set_thread_error(const std::string& msg) { g_null_ctx_err = msg; }

namespace dl {
int set_error(dl_ctx* ctx, int code, const std::string& msg) { return set_err(ctx, code, msg); }
bool make_workers(dl_ctx* ctx, uint32_t streams) { return prepare_workers(ctx, streams); }
double batch_job_cost(const dl_bake_input& in, const dl_bake_options& o) { return job_cost(in, o); }
}  // namespace dl

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
    ctx->host_threads = usable_cpus();
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return set_err(nullptr, DL_E_DEVICE, "creating the HIP stream failed");
    }
    // Bakes allocate from the device's default pool.  It keeps freed memory instead of returning
    // it to the driver at every synchronization (releasing can wait for the whole device, which
    // stalls bakes running side by side), and it reuses memory freed on another stream only once
    // that free has completed, so no stream is made to wait for another.  Contexts trim the pool
    // after a batch and when they are destroyed.
    // The context's buffer cache keeps up to a quarter of the device's memory: a bake's buffers
    // (about 40 GB for a scene of 10^8 texels with bounce) then come back from it whole at the next
    // bake, and no bake waits for the pool to map (and clear) fresh memory.  Worker contexts keep
    // 32 GB between them.
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) ctx->cache.cap = std::max<size_t>(ctx->cache.cap, total_b / 4);
    if (const char* env = std::getenv("DL_CACHE_BYTES")) ctx->cache.cap = strtoull(env, nullptr, 10);
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      int no = 0;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &no);
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
  for (dl_ctx* w : ctx->workers) dl_ctx_destroy(w);
  ctx->workers.clear();
  (void)hipSetDevice(ctx->device);
  ctx->cache.stream = ctx->stream;
  ctx->cache.clear();
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  trim_pool(ctx->device);
  if (ctx->staging.host) (void)hipHostFree(ctx->staging.host);
  ctx->staging = dl::Staging{};
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

int dl_bake_batch(dl_ctx* ctx, const dl_bake_input* inputs, const dl_bake_options* options, uint32_t count, uint32_t streams,
                  dl_baked_scene** out, int* status) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  if (count && (!inputs || !options || !out || !status)) return set_err(ctx, DL_E_ARGUMENT, "null argument");
  if (streams == 0 || streams > 16) return set_err(ctx, DL_E_ARGUMENT, "streams must be 1 to 16");
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "hipSetDevice failed");
    for (uint32_t i = 0; i < count; ++i) {
      out[i] = nullptr;
      status[i] = DL_E_CANCELLED;
    }
    if (!prepare_workers(ctx, streams)) return set_err(ctx, DL_E_DEVICE, "creating a worker stream failed");
    // the host hands out jobs largest first (by a rough cost estimate), so that the last jobs
    // running alone are small ones; each worker bakes on its own stream
    std::vector<uint32_t> order(count);
    {
      std::vector<double> cost(count);
      for (uint32_t i = 0; i < count; ++i) cost[i] = job_cost(inputs[i], options[i]);
      for (uint32_t i = 0; i < count; ++i) order[i] = i;
      std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::string> errs(count);
    ctx->batch_timings.assign(count, dl_timings{});
    ctx->batch_worker.assign(count, 0);
    ctx->batch_span.assign(2 * (size_t)count, 0.0);
    const auto t_batch = std::chrono::steady_clock::now();
    auto run = [&](dl_ctx* w, uint32_t wi) {
      (void)hipSetDevice(w->device);
      for (;;) {
        const uint32_t n = next.fetch_add(1);
        if (n >= count) break;
        const uint32_t i = order[n];
        ctx->batch_span[2 * (size_t)i] = dl::ms_since(t_batch);
        ctx->batch_worker[i] = wi;
        try {
          status[i] = dl_bake_impl(w, &inputs[i], &options[i], nullptr, &out[i]);
        } catch (const std::bad_alloc&) {
          w->err = "host allocation failed";
          status[i] = DL_E_OOM;
        } catch (...) {
          w->err = "unexpected failure";
          status[i] = DL_E_DEVICE;
        }
        if (status[i] != DL_OK) errs[i] = "job " + std::to_string(i) + ": " + w->err;
        ctx->batch_timings[i] = w->timings;
        ctx->batch_span[2 * (size_t)i + 1] = dl::ms_since(t_batch);
      }
    };
    std::vector<std::thread> ts;
    for (uint32_t k = 1; k < streams; ++k) ts.emplace_back(run, ctx->workers[k], k);
    run(ctx->workers[0], 0);
    for (auto& t : ts) t.join();
    trim_pool(ctx->device);
    for (uint32_t i = 0; i < count; ++i)
      if (status[i] != DL_OK) return set_err(ctx, status[i], errs[i]);
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return set_err(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_err(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

// ---- rebake service ----------------------------------------------------------------------
int dl_service_create(dl_ctx* ctx, const dl_bake_input* scene, const dl_bake_options* options, dl_service** out,
                      dl_baked_scene** initial) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  if (!scene || !options || !out) return set_err(ctx, DL_E_ARGUMENT, "null argument");
  *out = nullptr;
  if (initial) *initial = nullptr;
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "hipSetDevice failed");
    std::unique_ptr<dl_service> svc(new dl_service());
    svc->ctx = ctx;
    svc->options = *options;
    dl_baked_scene* first = nullptr;
    dl::Capture cap;
    cap.escaping = std::make_shared<dl::DevBlock>();  // for the isolation checks of rooms added later
    cap.escaping_cap = dl::kEscapingCap;
    if (cap.escaping->alloc(8 * cap.escaping_cap + 2, ctx->stream) != hipSuccess)
      return set_err(ctx, DL_E_OOM, "allocating the burial ray record failed");
    int st = dl_bake_impl(ctx, scene, options, nullptr, &first, &svc->res, &cap);
    if (st != DL_OK) return st;
    std::unique_ptr<dl_baked_scene> hold(first);
    if ((st = dl::service_init_live(svc.get(), scene, cap, *first)) != DL_OK) return st;
    // The service's later tracing runs on worker contexts (with buffers of other sizes) and its
    // context assembles: its cache keeps 32 GB of the initial bake's buffers, the rest goes back
    // to the device's pool for the workers.
    ctx->cache.stream = ctx->stream;
    ctx->cache.cap = std::min<size_t>(ctx->cache.cap, 32ull << 30);
    ctx->cache.trim(ctx->cache.cap);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "releasing the initial bake's buffers failed");
    if (initial) *initial = hold.release();
    *out = svc.release();
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return set_err(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_err(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

int dl_service_bake(dl_service* svc, const dl_light* lights, uint32_t light_count, dl_baked_scene** out) {
  if (!svc) return set_err(nullptr, DL_E_ARGUMENT, "null service");
  dl_ctx* ctx = svc->ctx;
  if (!out || (light_count && !lights)) return set_err(ctx, DL_E_ARGUMENT, "null argument");
  *out = nullptr;
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "hipSetDevice failed");
    if (svc->parts.size() > 1) return set_err(ctx, DL_E_ARGUMENT, "dl_service_bake bakes services without added rooms");
    const dl_bake_input in{nullptr, 0, lights, light_count};
    return dl_bake_impl(ctx, &in, &svc->options, nullptr, out, &svc->res);
  } catch (const std::bad_alloc&) {
    return set_err(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_err(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

int dl_service_replay(dl_service* svc, const dl_service_request* requests, uint32_t count, uint32_t workers,
                      double* latency_ms, double* service_ms, int* status, dl_baked_scene** out) {
  if (!svc) return set_err(nullptr, DL_E_ARGUMENT, "null service");
  dl_ctx* ctx = svc->ctx;
  if (count && (!requests || !latency_ms || !status)) return set_err(ctx, DL_E_ARGUMENT, "null argument");
  if (workers == 0 || workers > 16) return set_err(ctx, DL_E_ARGUMENT, "workers must be 1 to 16");
  for (uint32_t i = 0; i < count; ++i) {
    if (requests[i].light_count && !requests[i].lights) return set_err(ctx, DL_E_ARGUMENT, "null light array in a request");
    if (!(requests[i].arrival_s >= 0.0) || (i && requests[i].arrival_s < requests[i - 1].arrival_s))
      return set_err(ctx, DL_E_ARGUMENT, "arrival times must be non-negative and in order");
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "hipSetDevice failed");
    if (svc->parts.size() > 1) return set_err(ctx, DL_E_ARGUMENT, "dl_service_replay serves services without added rooms");
    if (!prepare_workers(ctx, workers)) return set_err(ctx, DL_E_DEVICE, "creating a worker stream failed");
    for (uint32_t i = 0; i < count; ++i) {
      status[i] = DL_E_CANCELLED;
      latency_ms[i] = 0.0;
      if (service_ms) service_ms[i] = 0.0;
      if (out) out[i] = nullptr;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::string> errs(count);
    const auto t0 = std::chrono::steady_clock::now();
    auto run = [&](dl_ctx* w) {
      (void)hipSetDevice(w->device);
      for (;;) {
        const uint32_t i = next.fetch_add(1);  // requests are admitted in arrival order
        if (i >= count) break;
        const auto arrival = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                      std::chrono::duration<double>(requests[i].arrival_s));
        std::this_thread::sleep_until(arrival);
        const auto started = std::chrono::steady_clock::now();
        const dl_bake_input in{nullptr, 0, requests[i].lights, requests[i].light_count};
        dl_baked_scene* scene = nullptr;
        try {
          status[i] = dl_bake_impl(w, &in, &svc->options, nullptr, &scene, &svc->res);
        } catch (const std::bad_alloc&) {
          w->err = "host allocation failed";
          status[i] = DL_E_OOM;
        } catch (...) {
          w->err = "unexpected failure";
          status[i] = DL_E_DEVICE;
        }
        const auto done = std::chrono::steady_clock::now();
        latency_ms[i] = std::chrono::duration<double, std::milli>(done - arrival).count();
        if (service_ms) service_ms[i] = std::chrono::duration<double, std::milli>(done - started).count();
        if (status[i] != DL_OK) errs[i] = "request " + std::to_string(i) + ": " + w->err;
        if (out) out[i] = scene;
        else delete scene;
      }
    };
    std::vector<std::thread> ts;
    for (uint32_t k = 1; k < workers; ++k) ts.emplace_back(run, ctx->workers[k]);
    run(ctx->workers[0]);
    for (auto& t : ts) t.join();
    for (uint32_t i = 0; i < count; ++i)
      if (status[i] != DL_OK) return set_err(ctx, status[i], errs[i]);
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return set_err(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_err(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

int dl_service_ray_queries_noalloc(dl_service* svc, const float* rays8, uint64_t ray_count, int closest, float* t_out,
                                   uint32_t* hit_or_index, uint64_t* hits) {
  using namespace dl;
  if (!svc) return set_err(nullptr, DL_E_ARGUMENT, "null service");
  dl_ctx* ctx = svc->ctx;
  if (ray_count > DL_NOALLOC_MAX_RAYS)
    return set_err(ctx, DL_E_ARGUMENT, "at most " + std::to_string((unsigned long long)DL_NOALLOC_MAX_RAYS) + " rays per call");
  if (ray_count && (!rays8 || !hit_or_index || (closest && !t_out))) return set_err(ctx, DL_E_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "hipSetDevice failed");
  if (svc->parts.size() > 1) return set_err(ctx, DL_E_ARGUMENT, "ray queries see services without added rooms");
  const Resident& r = svc->res;
  DevScene sc{};
  sc.nodes = r.d_nodes.p;
  sc.tris = reinterpret_cast<const float4*>(r.d_occ.p);
  sc.ntris = r.nocc;
  sc.stack = bvh_stack_entries(r.depth);
  sc.stack16 = r.nodes < 65536 ? 1u : 0u;
  uint32_t overflow = 0;
  const hipError_t e = launch_ray_queries_noalloc(ctx->stream, ctx->device, sc, rays8, ray_count, closest ? 1 : 0, t_out,
                                                  hit_or_index, hits, &overflow);
  if (e != hipSuccess) return set_err(ctx, DL_E_DEVICE, hipGetErrorString(e));
  if (overflow) return set_err(ctx, DL_E_DEVICE, "BVH traversal stack overflow");
  return DL_OK;
}

void dl_service_destroy(dl_service* svc) {
  if (!svc) return;
  dl_ctx* ctx = svc->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();  // no worker still reads the resident buffers
  delete svc;
}

uint32_t dl_test_batch_trace(dl_ctx* ctx, dl_timings* timings, uint32_t* worker, double* span, uint32_t cap) {
  if (!ctx) return 0;
  std::lock_guard<std::mutex> lock(ctx->mu);
  const uint32_t n = (uint32_t)std::min<size_t>(cap, ctx->batch_worker.size());
  for (uint32_t i = 0; i < n; ++i) {
    if (timings) timings[i] = ctx->batch_timings[i];
    if (worker) worker[i] = ctx->batch_worker[i];
    if (span) {
      span[2 * (size_t)i] = ctx->batch_span[2 * (size_t)i];
      span[2 * (size_t)i + 1] = ctx->batch_span[2 * (size_t)i + 1];
    }
  }
  return (uint32_t)ctx->batch_worker.size();
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
uint32_t dl_scene_meshes(const dl_baked_scene* s, uint32_t first, uint32_t count, dl_baked_mesh_view* out) {
  if (!s || !out || first >= s->meshes.size()) return 0;
  const uint32_t n = (uint32_t)std::min<uint64_t>(count, s->meshes.size() - first);
  for (uint32_t k = 0; k < n; ++k) (void)dl_scene_mesh(s, first + k, out + k);
  return n;
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

int dl_test_bvh_structure(const float* tris9, uint32_t tri_count, uint32_t threads, uint64_t out[4]) {
  using namespace dl;
  if ((!tris9 && tri_count) || !out) return set_err(nullptr, DL_E_ARGUMENT, "null argument");
  Bvh bvh;
  build_bvh(tris9, tri_count, threads, bvh);
  const size_t nn = bvh.nodes.size(), nr = bvh.order.size();
  auto bad = [](const char* what) { return set_err(nullptr, DL_E_ARGUMENT, std::string("BVH structure: ") + what); };
  if (tri_count == 0) {
    if (nn || nr) return bad("nodes without triangles");
  } else {
    std::vector<uint8_t> node_seen(nn, 0), ref_seen(nr, 0);
    std::vector<uint32_t> depth(nn, 0), stack{0};
    node_seen[0] = 1;
    depth[0] = 1;
    uint32_t deepest = 0;
    while (!stack.empty()) {
      const uint32_t n = stack.back();
      stack.pop_back();
      deepest = std::max(deepest, depth[n]);
      for (int k = 0; k < kBvhWidth; ++k) {
        const int32_t c = bvh.nodes[n].child[k];
        if (c > 0) {
          if ((size_t)c >= nn || node_seen[c]++) return bad("an inner child out of range or reached twice");
          depth[c] = depth[n] + 1;
          stack.push_back((uint32_t)c);
        } else if (c < 0) {
          const uint32_t e = (uint32_t)(-1 - c), first = e >> 2, count = (e & 3u) + 1u;
          for (uint32_t i = first; i < first + count; ++i)
            if (i >= nr || ref_seen[i]++) return bad("a reference out of range or in two leaves");
        }
      }
    }
    for (size_t i = 0; i < nn; ++i)
      if (!node_seen[i]) return bad("a node not reached from the root");
    for (size_t i = 0; i < nr; ++i)
      if (!ref_seen[i]) return bad("a reference in no leaf");
    if (deepest != bvh.depth) return bad("depth differs from the longest path");
    for (size_t k = 0; k < nr; ++k) {
      const uint32_t p = bvh.order[k];
      if (p >= tri_count) return bad("a reference of no triangle");
      const float* t = tris9 + (size_t)p * 9;
      const float* o = bvh.tris.data() + k * 12;
      uint32_t q;
      std::memcpy(&q, &o[3], 4);
      if (q != p || o[0] != t[0] || o[1] != t[1] || o[2] != t[2] || o[4] != t[3] - t[0] || o[5] != t[4] - t[1] ||
          o[6] != t[5] - t[2] || o[8] != t[6] - t[0] || o[9] != t[7] - t[1] || o[10] != t[8] - t[2])
        return bad("a leaf record differs from its triangle");
    }
  }
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the reference order
  for (size_t k = 0; k < nr; ++k) h = (h ^ bvh.order[k]) * 1099511628211ull;
  out[0] = nn;
  out[1] = nr;
  out[2] = bvh.depth;
  out[3] = h;
  return DL_OK;
}

int dl_test_ray_queries(dl_ctx* ctx, const float* tris9, uint32_t tri_count, const float* rays8, uint64_t ray_count,
                        int closest, float* t_out, uint32_t* hit_or_index) {
  using namespace dl;
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  Bvh bvh;
  build_bvh(tris9, tri_count, ctx->host_threads, bvh);
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
  sc.stack = bvh_stack_entries(bvh.depth);
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
  t.sin = {reinterpret_cast<const uint2*>(host.sin_runs.data()), host.sin_vals.data(), host.sin_any.data()};
  t.cos = {reinterpret_cast<const uint2*>(host.cos_runs.data()), host.cos_vals.data(), host.cos_any.data()};
  t.tol = host.tol;
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
  t.sin = {reinterpret_cast<const uint2*>(host.sin_runs.data()), host.sin_vals.data(), host.sin_any.data()};
  t.cos = {reinterpret_cast<const uint2*>(host.cos_runs.data()), host.cos_vals.data(), host.cos_any.data()};
  t.tol = host.tol;
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

int dl_set_bounce_capture(dl_ctx* ctx, int enable) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->capture_bounce = enable != 0;
  ctx->captured.clear();
  ctx->captured_values.clear();
  return DL_OK;
}

uint64_t dl_test_bounce_capture(dl_ctx* ctx, uint64_t index, uint64_t meta[5], const float** values) {
  if (!ctx) return 0;
  std::lock_guard<std::mutex> lock(ctx->mu);
  auto& c = ctx->captured;
  std::sort(c.begin(), c.end(), [](const dl_ctx::CapturedCrop& a, const dl_ctx::CapturedCrop& b) {
    if (a.mesh_id != b.mesh_id) return a.mesh_id < b.mesh_id;
    if (a.triangle != b.triangle) return a.triangle < b.triangle;
    return a.light_id < b.light_id;
  });
  if (index < c.size()) {
    if (meta) {
      meta[0] = c[index].mesh_id;
      meta[1] = c[index].triangle;
      meta[2] = c[index].light_id;
      meta[3] = c[index].width;
      meta[4] = c[index].height;
    }
    if (values) *values = ctx->captured_values.data() + c[index].offset;
  }
  return c.size();
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

int dl_set_burial(dl_ctx* ctx, uint32_t lanes_block) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  if (lanes_block != 0 && (lanes_block % 64 != 0 || lanes_block > 1024))
    return set_err(ctx, DL_E_ARGUMENT, "block sizes must be multiples of 64 in [64, 1024]");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->burial_lanes = lanes_block;
  return DL_OK;
}

int dl_set_trace_stream(dl_ctx* ctx, int enable) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->trace_stream = enable ? 1u : 0u;
  return DL_OK;
}

int dl_set_photon_tiles(dl_ctx* ctx, int enable) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->photon_tiles = enable ? 1u : 0u;
  return DL_OK;
}

int dl_set_gather_split(dl_ctx* ctx, int enable) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->gather_split = enable ? 1u : 0u;
  return DL_OK;
}

uint32_t dl_test_light_bvh(const float* positions, const float* radii, uint32_t n, dl_light_bvh_node* nodes, uint32_t cap,
                           uint32_t* order) {
  if (n && (!positions || !radii || !order)) return 0;
  for (uint64_t i = 0; i < 3ull * n; ++i)
    if (!std::isfinite(positions[i])) return 0;  // (the bake validates its lights before the build)
  try {
    std::vector<dl::f3> pos(n);
    std::vector<float> rad(radii, radii + n);
    for (uint32_t i = 0; i < n; ++i) pos[i] = dl::mk3(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]);
    dl::LightBvhOut out;
    dl::build_light_bvh(pos, rad, out);
    if (out.nodes.size() > cap || (!out.nodes.empty() && !nodes)) return 0;
    std::copy(out.nodes.begin(), out.nodes.end(), nodes);
    std::copy(out.order.begin(), out.order.end(), order);
    return (uint32_t)out.nodes.size();
  } catch (...) {
    return 0;
  }
}

int dl_set_lazy_photons(dl_ctx* ctx, int enable) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->lazy_photons = enable ? 1u : 0u;
  return DL_OK;
}

int dl_test_gather_directions(dl_ctx* ctx, const float* dir3, const float* z, const float* azimuth, const float* spread, uint64_t n,
                              float* exact3, float* fast3, uint32_t* fast_ok) {
  using namespace dl;
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  if (n && (!dir3 || !z || !azimuth || !spread || !exact3 || !fast3 || !fast_ok)) return set_err(ctx, DL_E_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, DL_E_DEVICE, "hipSetDevice failed");
    if (n == 0) return DL_OK;
    const CacheScope scope(&ctx->cache);
    DBuf<float> a, zz, aa, ss, e, f;
    DBuf<uint32_t> ok;
    hipStream_t s = ctx->stream;
    if (a.alloc(3 * n) || zz.alloc(n) || aa.alloc(n) || ss.alloc(n) || e.alloc(3 * n) || f.alloc(3 * n) || ok.alloc(n) ||
        copy_to_device(ctx, a.p, dir3, 12 * n) || copy_to_device(ctx, zz.p, z, 4 * n) ||
        copy_to_device(ctx, aa.p, azimuth, 4 * n) || copy_to_device(ctx, ss.p, spread, 4 * n) ||
        launch_test_directions(s, ctx->trig, a.p, zz.p, aa.p, ss.p, n, e.p, f.p, ok.p) ||
        copy_to_host(ctx, exact3, e.p, 12 * n) || copy_to_host(ctx, fast3, f.p, 12 * n) || copy_to_host(ctx, fast_ok, ok.p, 4 * n))
      return set_err(ctx, DL_E_DEVICE, "device pass failed");
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return set_err(ctx, DL_E_OOM, "host allocation failed");
  }
}

int dl_set_gather(dl_ctx* ctx, uint32_t mode, uint32_t block, uint32_t chunk, uint64_t chunk_rays, uint32_t blocks_per_cu,
                  uint32_t min_waves, uint32_t refill, uint32_t order) {
  if (!ctx) return set_err(nullptr, DL_E_ARGUMENT, "null context");
  if (mode > 1) return set_err(ctx, DL_E_ARGUMENT, "gather mode must be 0 or 1");
  if (block && (block % 64 != 0 || block > 1024)) return set_err(ctx, DL_E_ARGUMENT, "block must be a multiple of 64 up to 1024");
  if (chunk_rays && chunk_rays < 4096) return set_err(ctx, DL_E_ARGUMENT, "chunks need at least 4096 rays");
  if (chunk_rays > (1ull << 31)) return set_err(ctx, DL_E_ARGUMENT, "at most 2^31 rays per chunk (ray slots are 32-bit)");
  if (blocks_per_cu > 64) return set_err(ctx, DL_E_ARGUMENT, "at most 64 blocks per CU");
  if (chunk > (1u << 20)) return set_err(ctx, DL_E_ARGUMENT, "at most 2^20 rays per queue run");
  if (min_waves != 0 && min_waves != 8) return set_err(ctx, DL_E_ARGUMENT, "min_waves must be 0 or 8");
  if (refill > 64) return set_err(ctx, DL_E_ARGUMENT, "refill must be at most 64");
  if (order > 2) return set_err(ctx, DL_E_ARGUMENT, "order must be 0 (keep), 1 (job-major) or 2 (sample-major)");
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->gather_mode = mode;
  if (order) ctx->gather.order = order - 1;
  if (block) ctx->gather.block = block;
  if (chunk) ctx->gather.chunk = chunk;
  if (chunk_rays) ctx->gather.chunk_rays = chunk_rays;
  ctx->gather.blocks_per_cu = blocks_per_cu;
  ctx->gather.min_waves = min_waves;
  if (refill) ctx->gather.refill = refill;
  return DL_OK;
}

}  // extern "C"

// Host-side pieces of the bake shared between the orchestration and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/dynlight.h"
#include "dl_kernels.h"

namespace dl {

// Device memory of one context's bakes.  Sizes are rounded up to classes of eight steps per
// power of two (at most 12.5% more; 256 bytes at least), and a freed buffer is kept, up to `cap`
// bytes in all, for a later allocation of its class: every use is on the context's one stream,
// so stream order alone makes the reuse safe.  New memory is a stream-ordered allocation from
// the device's pool.  hipMalloc/hipFree, and the pool when it has to grow, can block the host for
// a long time, which stalls the stream that waits for this host thread and every bake running
// beside it.
struct DevCache {
  hipStream_t stream = nullptr;
  std::unordered_map<size_t, std::vector<void*>> lists;
  size_t cached = 0, cap = 32ull << 30;
  static size_t class_of(size_t bytes) {
    if (bytes <= 256) return 256;
    const int k = 63 - __builtin_clzll((unsigned long long)(bytes - 1));  // 2^k < bytes <= 2^(k+1)
    const size_t step = (size_t)1 << (k - 3);
    return ((size_t)1 << k) + ((bytes - ((size_t)1 << k) + step - 1) / step) * step;
  }
  hipError_t get(size_t bytes, void** p, size_t* cls) {
    *cls = class_of(bytes);
    auto it = lists.find(*cls);
    if (it != lists.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      cached -= *cls;
      return hipSuccess;
    }
    hipError_t e = hipMallocAsync(p, *cls, stream);
    if (e == hipErrorOutOfMemory && cached) {  // the cache's buffers go back to the pool, and the request is tried again
      (void)hipGetLastError();
      clear();
      e = hipMallocAsync(p, *cls, stream);
    }
    return e;
  }
  void put(void* p, size_t cls) {
    if (cached + cls > cap) {
      (void)hipFreeAsync(p, stream);
      return;
    }
    lists[cls].push_back(p);
    cached += cls;
  }
  // Returns cached buffers to the pool, the largest first, until at most `keep` bytes are cached.
  void trim(size_t keep) {
    while (cached > keep) {
      auto big = lists.end();
      for (auto it = lists.begin(); it != lists.end(); ++it)
        if (!it->second.empty() && (big == lists.end() || it->first > big->first)) big = it;
      if (big == lists.end()) break;
      (void)hipFreeAsync(big->second.back(), stream);
      big->second.pop_back();
      cached -= big->first;
    }
  }
  void clear() {
    for (auto& l : lists)
      for (void* q : l.second) (void)hipFreeAsync(q, stream);
    lists.clear();
    cached = 0;
  }
};

// While a CacheScope is alive, DBuf allocations of this thread come from its cache.
inline thread_local DevCache* tls_cache = nullptr;
struct CacheScope {
  DevCache* prev;
  explicit CacheScope(DevCache* c) : prev(tls_cache) { tls_cache = c; }
  ~CacheScope() { tls_cache = prev; }
  CacheScope(const CacheScope&) = delete;
  CacheScope& operator=(const CacheScope&) = delete;
};

// Device allocation owned by the caller scope.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DevCache* cache = nullptr;  // non-null: taken from (and returned to) this cache
  size_t cls = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      if (cache) cache->put(p, cls);
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    cache = nullptr;
  }
  void take(DBuf& o) {  // takes over o's allocation
    if (&o == this) return;
    release();
    p = o.p;
    n = o.n;
    cache = o.cache;
    cls = o.cls;
    o.p = nullptr;
    o.n = 0;
    o.cache = nullptr;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    cache = tls_cache;
    if (cache) {
      hipError_t e = cache->get(count * sizeof(T), (void**)&p, &cls);
      if (e != hipSuccess) {
        p = nullptr;
        cache = nullptr;
      }
      return e;
    }
    return hipMalloc((void**)&p, count * sizeof(T));
  }
};

// Pinned host memory through which a context's small copies go: copies from pageable memory are
// staged by the runtime and cost far more per call.
struct Staging {
  char* host = nullptr;
  size_t cap = 0, off = 0;
};
constexpr size_t kStagingBytes = 64ull << 20;  // staging buffer of a context
constexpr size_t kStagedMax = 8ull << 20;      // larger copies use the pageable path

// The exception tables of f32 sin/cos (see dl_trig.h), built from the host's libm.
struct TrigHost {
  std::vector<uint32_t> sin_keys, sin_vals, cos_keys, cos_vals;  // exceptions in ascending order
  std::vector<uint32_t> sin_runs, cos_runs;  // see TrigTable::runs (bitmap, rank) word pairs
  std::vector<uint32_t> sin_any, cos_any;    // see TrigTable::any
  uint32_t tol = 0;  // see TrigTables::tol
};
void build_trig_tables(TrigHost& out, unsigned threads);

struct LightBvhOut {
  std::vector<dl_light_bvh_node> nodes;
  std::vector<uint32_t> order;
};
// build_light_bvh (light_bvh.rs:59-189)
void build_light_bvh(const std::vector<f3>& pos, const std::vector<float>& radii, LightBvhOut& out);

}  // namespace dl

struct dl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  dl::TrigHost trig_host;
  dl::DBuf<uint32_t> trig_buf;
  dl::TrigTables trig{};
  dl_timings timings{};
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  uint32_t trace_block = 128;
  uint32_t burial_block = 0;  // threads per block of the burial ray stream (0: the gather's)
  uint32_t burial_lanes = 0;  // nonzero: the burial test with one lane per texel, this many threads per block
  dl::GatherLaunch gather{};
  uint32_t gather_mode = 1;  // 0: one wave per (triangle, light, 64 texels); 1: streamed jobs
  uint32_t trace_stream = 1;  // 1: area lights, soft suns and skies through the streamed any-hit path
  uint32_t photon_tiles = 1;  // 1: the point-light gather reads photons only where the photon tiles leave a sample open
  uint32_t lazy_photons = 1;  // 1: bakes with few gather samples per cube texel trace only the texels the samples can read
  uint32_t gather_split = 1;  // 1: the gather's sample setup in a fast pass and an exact pass for the samples it leaves
  unsigned host_threads = 1;
  bool counting = false;
  // capture mode (test hook): later bakes keep each kept bounce entry's pre-pack f32 crop
  bool capture_bounce = false;
  struct CapturedCrop {
    uint64_t mesh_id, light_id;
    uint32_t triangle, width, height;
    uint64_t offset;  // into captured_values (floats)
  };
  std::vector<CapturedCrop> captured;
  std::vector<float> captured_values;
  // batched bakes: worker contexts with their own streams (sharing this context's sin/cos
  // tables); a worker plans for 1/share of the device: of half the free memory and of the
  // resident blocks of persistent kernels
  std::vector<dl_ctx*> workers;
  uint32_t share = 1;
  dl::DevCache cache;
  dl::Staging staging;
  // per job of the last batch: its timings, worker, and start/end in ms from the batch start
  std::vector<dl_timings> batch_timings;
  std::vector<uint32_t> batch_worker;
  std::vector<double> batch_span;
};

// A run of words that shares ownership of the buffer holding it: a bake downloads the triangle
// words of a batch of meshes into one buffer and each mesh keeps its slice of it.
struct dl_word_slice {
  std::shared_ptr<const uint32_t> first;  // aliases the owner of the whole buffer
  size_t n = 0;
  const uint32_t* data() const { return first.get(); }
  size_t size() const { return n; }
  bool empty() const { return n == 0; }
  const uint32_t& operator[](size_t i) const { return first.get()[i]; }
  const uint32_t* begin() const { return first.get(); }
  const uint32_t* end() const { return first.get() + n; }
  // a slice of [p, p + count) kept alive by `owner`
  template <class Owner>
  void share(const std::shared_ptr<Owner>& owner, const uint32_t* p, size_t count) {
    first = std::shared_ptr<const uint32_t>(owner, p);
    n = count;
  }
  void own(std::vector<uint32_t>&& v) {
    auto h = std::make_shared<std::vector<uint32_t>>(std::move(v));
    share(h, h->data(), h->size());
  }
};

struct dl_baked_mesh_out {
  uint64_t id, content_hash;
  uint32_t resolution, occludes;
  dl_word_slice words;
};

struct dl_sun_map_out {
  float rows[12];
  uint32_t resolution;
  std::vector<uint32_t> texels;
};

struct dl_baked_scene {
  dl_baked_settings settings{};
  dl_bake_stats stats{};
  std::vector<dl_baked_light> lights;
  std::vector<dl_light_bvh_node> bvh;
  std::vector<dl_baked_mesh_out> meshes;
  std::vector<uint32_t> distance_cubes;
  std::vector<dl_sun_map_out> sun_maps;
  bool probes = false;
  float probe_origin[3] = {0, 0, 0};
  float probe_spacing = 0.0f;
  uint32_t probe_counts[3] = {0, 0, 0};
  std::vector<uint32_t> probe_words;
  bool triangles_taken = false;  // the meshes' words were moved into buffers (take_scene_buffers)
};

namespace dl {
// The geometry side of a bake kept on the device between bakes of one scene with different lights
// (a rebake service): receiver meshes in BakeId order, the occluder BVH and triangle records, the
// receiver triangles, and every triangle's unburied texels.  None of these depend on the lights.
struct MeshMeta {
  uint64_t id, content_hash;
  uint32_t occludes;
};
struct Resident {
  bool ready = false;  // filled by a first bake, then read by later bakes (by several at once)
  std::vector<MeshMeta> meshes;
  std::vector<uint32_t> res, first_tri;
  std::vector<uint64_t> mid, owned_off;
  uint64_t ntri = 0;
  std::vector<float> occ9;  // occluder corners (sun map frames, sky probe bounds)
  uint32_t nocc = 0, depth = 0;
  uint64_t nodes = 0;
  bool has_alpha = false;  // some occluder has an alpha texture (used when a light tests alpha)
  DBuf<CNode> d_nodes;
  DBuf<float> d_occ, d_occ_nrm, d_occ_alb, d_occ_uv0;
  DBuf<uint32_t> d_occ_flags, d_occ_tex, d_tex_info, d_tex_bits;
  DBuf<float> d_v, d_n, d_uv;
  DBuf<uint32_t> d_mesh, d_local, d_res;
  DBuf<uint64_t> d_mid, d_owned_off;
  // unburied texels of every triangle in triangle order; tri holds the scene-wide triangle index
  DBuf<uint32_t> t_xy, t_tri;
  DBuf<float4> t_pos, t_nrm;
  std::vector<uint64_t> tex_start;  // ntri + 1
  // while the first bake runs: the texels of each of its mesh batches
  struct Piece {
    DBuf<uint32_t> xy, tri;
    DBuf<float4> pos, nrm;
    std::vector<uint64_t> start;  // per batch triangle, relative
    uint64_t n = 0;
  };
  std::vector<std::unique_ptr<Piece>> pieces;
};
}  // namespace dl

namespace dl {
// Device memory outside any context's cache, for data that outlives a bake (a rebake service's
// stores): allocated from the device pool on `stream`, freed on the stream it was made on.
struct DevBlock {
  uint32_t* p = nullptr;
  size_t words = 0;
  hipStream_t stream = nullptr;
  DevBlock() = default;
  DevBlock(const DevBlock&) = delete;
  DevBlock& operator=(const DevBlock&) = delete;
  ~DevBlock() {
    if (p) (void)hipFreeAsync(p, stream);
  }
  hipError_t alloc(size_t n, hipStream_t s) {
    stream = s;
    words = n;
    return hipMallocAsync((void**)&p, std::max<size_t>(n, 1) * 4, s);
  }
};

// Per-(mesh, light) results of a bake, kept for a rebake service.  A "pair" is one region of a
// mesh batch: a mesh and a light that reaches some of its triangles.
struct CapturedPair {
  uint32_t mesh;          // mesh index of the bake (sorted order)
  uint64_t light_id;
  uint32_t rect[4], bits;
  uint64_t grid_off, grid_words;  // pre-padding level words, in `block`
  uint64_t vis_off, vis_words;    // region-local visited bits, in `block`
  uint64_t rays, samples, bounce_rays;
  uint64_t entries;               // (triangle, light) entries of the pair; the kept ones follow
  std::vector<uint32_t> kept;     // mesh-local triangles whose entry survived optimize, ascending
  std::vector<uint64_t> bounce_off, bounce_words;  // per kept triangle: its bounce payload in `words`
  std::shared_ptr<DevBlock> block, words;
};
// Burial rays that hit nothing kept per part of a rebake service (8 floats each): up to 2^27 for
// the scene a service is created with (4 GiB; the 200k-triangle interior sends 119 M of them out),
// up to 2^22 for a room that joins it (128 MiB).
constexpr uint64_t kEscapingCap = 1ull << 27;
constexpr uint64_t kRoomEscapingCap = 1ull << 22;
struct Capture {
  std::vector<CapturedPair> pairs;
  std::vector<std::pair<uint64_t, std::vector<uint32_t>>> cubes;  // distance cube words per local light id
  std::vector<uint32_t> crops;    // 4 per receiver triangle of the bake (scene order)
  std::vector<uint64_t> texels;   // per mesh
  // Burial rays that hit nothing (8 floats each: origin, direction, +inf, 0), when `escaping` is
  // set before the bake: it holds 8 * escaping_cap floats, then the 64-bit count of such rays
  // (all of them, also those past the cap), which the bake leaves in escaping_count.
  std::shared_ptr<DevBlock> escaping;
  uint64_t escaping_cap = 0, escaping_count = 0;
};
}  // namespace dl

namespace dl {
// One (mesh, light) of a service's live scene: the light's region of the mesh (levels before seam
// padding and its own visited bits, in the service's level store or, while new, in a capture
// block), its counters, and its kept triangles with their payloads in the live words.
struct LivePair {
  uint64_t light_id = 0;
  uint32_t rect[4] = {0, 0, 0, 0}, bits = 1;
  uint64_t grid_off = 0, grid_words = 0, vis_off = 0, vis_words = 0;
  uint64_t rays = 0, samples = 0, bounce_rays = 0, entries = 0;
  std::vector<uint32_t> kept;          // mesh-local triangles, ascending
  std::vector<uint64_t> shadow_off;    // per kept triangle: its shadow payload, from its mesh's first word
  // per kept triangle: its bounce payload (0 words: none), from its mesh's first word, or in
  // src_words while that is set (the words of the bake that captured the pair)
  std::vector<uint64_t> bounce_off, bounce_words;
  std::shared_ptr<DevBlock> levels;     // holds the pair's unpadded levels and visited bits
  std::shared_ptr<DevBlock> src_words;
};
// A pair's visited bits to add to or take from its mesh's visit counts (the block holding them
// lives until the counts are updated).
struct VisDelta {
  uint32_t mesh = 0;
  uint32_t rect[4] = {0, 0, 0, 0};
  std::shared_ptr<DevBlock> block;
  uint64_t vis_off = 0, vis_words = 0;
};
// A part of a service's live scene: its first scene, or a room added later.  Its meshes are
// live meshes mesh0 .. mesh0 + res->meshes.size() - 1; the bounds hold all its vertices.
struct ScenePart {
  Resident* res = nullptr;
  uint32_t mesh0 = 0;
  float bmin[3] = {0.0f, 0.0f, 0.0f}, bmax[3] = {0.0f, 0.0f, 0.0f};
};
struct LiveMesh {
  std::vector<LivePair> pairs;         // sorted by light id
  uint64_t w_off = 0, w_len = 0;       // the mesh's words in the live words
  uint64_t texels = 0;
};
}  // namespace dl

// A rebake service: one context, the bake options, the resident scene, and its live scene (the
// resident meshes with the current light set; every edit is applied to it).
struct dl_service {
  dl_ctx* ctx = nullptr;
  dl_bake_options options{};
  dl::Resident res;
  std::vector<dl_light> lights;                 // the live light set, sorted by id
  std::vector<dl::LiveMesh> mesh;               // resident mesh order
  std::vector<uint32_t> crops;                  // 4 per receiver triangle
  std::unique_ptr<dl::DevBlock> W;              // live words
  // per texel of every mesh's grid: how many of its pairs visited it (16 bits, two per word),
  // from word vc_off[mesh]; and the changes of the pairs merged since the last assembly
  std::unique_ptr<dl::DevBlock> vc;
  std::vector<uint64_t> vc_off;
  std::vector<dl::VisDelta> vis_add, vis_sub;
  std::unique_ptr<dl::DevBlock> d_crops;        // the receiver triangles' crops on the device
  // the live meshes of every part in mesh order (the first scene's, then each added room's):
  // their meta, grid resolutions and first triangles (a scene-wide numbering, mesh count + 1)
  std::vector<dl::MeshMeta> mesh_meta;
  std::vector<uint32_t> mesh_res, first_tri;
  uint64_t ntri = 0;
  std::vector<dl::ScenePart> parts;              // part 0 is `res`
  std::vector<std::unique_ptr<dl::Resident>> rooms;  // the residents of the added rooms
  // burial rays of the parts that hit nothing in their own part (8 floats each, see
  // Capture::escaping), one block and count per part; esc_open when a part had more than its
  // capture kept
  std::vector<std::pair<std::shared_ptr<dl::DevBlock>, uint64_t>> esc;
  uint64_t esc_total = 0;  // escaping rays of the scene the service was created with (also past the cap)
  bool esc_open = false;
  std::unordered_map<uint64_t, std::vector<uint32_t>> cubes;  // distance cube words per local light id
  std::unordered_map<uint64_t, uint32_t> slot;  // light id -> slot in the live words
  std::unordered_map<uint64_t, std::vector<uint32_t>> light_meshes;  // light id -> meshes holding its pairs
  std::unique_ptr<dl_baked_scene> base;         // settings, sun maps and sky probes of the first bake
  std::vector<dl_baked_light> baked;            // the live scene's lights in slot order
  std::vector<dl_light_bvh_node> bvh;
  dl_bake_stats stats{};
  uint32_t stride = 2;
  uint64_t version = 0;
};

namespace dl {
// The C ABI's error bookkeeping and worker contexts, for the service's entry points.
int set_error(dl_ctx* ctx, int code, const std::string& msg);
bool make_workers(dl_ctx* ctx, uint32_t streams);
// A bake job's rough cost (texels times rays per texel, plus photons), by which dl_bake_batch
// hands jobs out largest first.
double batch_job_cost(const dl_bake_input& in, const dl_bake_options& o);
// Builds the live scene of a service from the capture of its first bake.
int service_init_live(dl_service* svc, const dl_bake_input* scene, Capture& cap, const dl_baked_scene& first);
}  // namespace dl

// Sets the error that dl_last_error(NULL) reports on this thread.
void dl_set_thread_error(const std::string& msg);

// The bake.  With `res` not ready, a normal bake that also fills it; with `res` ready, the meshes
// of `input` are ignored and the resident ones are baked with the lights of `input`.

int dl_bake_impl(dl_ctx* ctx, const dl_bake_input* input, const dl_bake_options* options, dl_progress* progress,
                 dl_baked_scene** out, dl::Resident* res = nullptr, dl::Capture* cap = nullptr);

namespace dl {

// Validation with the reference's messages (input.rs:233-373); errors go to ctx->err.
std::string hex_id(uint64_t id);
int check_options(dl_ctx* ctx, const dl_bake_options& o);
int check_mesh(dl_ctx* ctx, const dl_mesh& m);
int check_light(dl_ctx* ctx, const dl_light& l);
DevLight make_dev_light(const dl_light& l, const dl_bake_options& o);
// lightmap_size (geometry.rs:66-70)
uint32_t lightmap_size(float area, float density, uint32_t max_size);

// Sun map frame and bounds over the occluder corners (sun_map.rs:100-151).  Returns false when
// there is no static geometry (the map is then one texel nothing is behind).
bool sun_map_setup(const float* corners, size_t ncorners, const float toward[3], uint32_t resolution, SunMapParams* p,
                   float rows[12]);
uint32_t pack_sun_depth(float lower, float upper);
// ProbeGrid::over (sky_probes.rs:46-62)
void probe_grid(f3 mn, f3 mx, float spacing, f3* origin, float* spacing_out, uint32_t counts[3]);
// fill_inside (sky_probes.rs:110-144): open holds 6 values per probe, valid one flag per probe.
void fill_inside(const uint32_t counts[3], std::vector<float>& open, std::vector<uint8_t>& valid);
// SkyProbes::pack (format/types.rs:105-112)
void pack_probe(const float* open6, uint32_t* words2);

// fn(k) for k in [0, n), handed out one at a time to up to `threads` threads (this one included,
// at most 16).
template <class F>
void run_parallel(uint64_t n, unsigned threads, const F& fn) {
  const unsigned th = (unsigned)std::min<uint64_t>(std::min(std::max(1u, threads), 16u), n);
  std::atomic<uint64_t> next{0};
  auto work = [&]() {
    for (uint64_t k; (k = next.fetch_add(1, std::memory_order_relaxed)) < n;) fn(k);
  };
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < th; ++t) ts.emplace_back(work);
  work();
  for (auto& t : ts) t.join();
}

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Host-to-device and device-to-host copies on the context's stream, through its pinned staging
// buffer when they are small.  copy_to_device returns once the source may be reused;
// copy_to_host returns once the data has arrived.
hipError_t copy_to_device(dl_ctx* ctx, void* dst, const void* src, size_t bytes);
hipError_t copy_to_host(dl_ctx* ctx, void* dst, const void* src, size_t bytes);

template <class T, class A>
hipError_t upload(dl_ctx* ctx, DBuf<T>& b, const std::vector<T, A>& h) {
  hipError_t e = b.alloc(h.size());
  if (e != hipSuccess || h.empty()) return e;
  return copy_to_device(ctx, b.p, h.data(), h.size() * sizeof(T));
}
template <class T>
hipError_t download(dl_ctx* ctx, std::vector<T>& h, const T* p, size_t n) {
  h.resize(n);
  if (n == 0) return hipSuccess;
  return copy_to_host(ctx, h.data(), p, n * sizeof(T));
}
template <class T>
hipError_t zeroed(dl_ctx* ctx, DBuf<T>& b, size_t n) {
  hipError_t e = b.alloc(n);
  if (e != hipSuccess || n == 0) return e;
  return hipMemsetAsync(b.p, 0, n * sizeof(T), ctx->stream);
}

}  // namespace dl

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);          \
      return e_ == hipErrorOutOfMemory ? (int)DL_E_OOM : (int)DL_E_DEVICE;   \
    }                                                                        \
  } while (0)

// Host-side pieces of the bake shared between the orchestration and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/dynlight.h"
#include "dl_kernels.h"

namespace dl {

// Device allocation owned by the caller scope.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)&p, count * sizeof(T));
  }
};

// The exception tables of f32 sin/cos (see dl_trig.h), built from the host's libm.
struct TrigHost {
  std::vector<uint32_t> sin_keys, sin_vals, cos_keys, cos_vals;
  std::vector<uint32_t> sin_bucket, cos_bucket;
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
  uint32_t trace_block = 256;
  uint32_t burial_block = 128;
  unsigned host_threads = 1;
  bool counting = false;
};

struct dl_baked_mesh_out {
  uint64_t id, content_hash;
  uint32_t resolution, occludes;
  std::vector<uint32_t> words;
};

struct dl_baked_scene {
  dl_baked_settings settings{};
  dl_bake_stats stats{};
  std::vector<dl_baked_light> lights;
  std::vector<dl_light_bvh_node> bvh;
  std::vector<dl_baked_mesh_out> meshes;
  std::vector<uint32_t> distance_cubes;
};

int dl_bake_impl(dl_ctx* ctx, const dl_bake_input* input, const dl_bake_options* options, dl_progress* progress,
                 dl_baked_scene** out);

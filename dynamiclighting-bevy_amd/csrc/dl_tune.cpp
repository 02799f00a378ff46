// Autotuner of the batched bake's launch shapes (dl_autotune_batch).
//
// Candidate settings are timed on a sample of the batch's own jobs: the jobs sorted by their
// estimated cost, every (count / sample)-th of them, so small and large jobs are both in it.
// Coordinate descent: each parameter in turn takes each of its candidate values with the others
// at their best so far, timed as the best of two dl_bake_batch passes over the sample.  Results
// are exact whatever the settings (they change launch shapes only), so tuning is free to try any.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dynlight.h"
#include "../../include/dynlight_testing.h"
#include "dl_host.h"

namespace {

constexpr uint32_t kSample = 32;

uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  return h * 0xff51afd7ed558ccdull;
}
uint64_t mix_f(uint64_t h, double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return mix(h, b);
}

// What the best settings depend on: the jobs' sizes and kinds, the options that shape the work,
// the device and the hardware queues HIP spreads streams over.
std::string signature(const dl_bake_input* in, const dl_bake_options* op, uint32_t count, int device) {
  uint64_t h = mix(0, count);
  for (uint32_t j = 0; j < count; ++j) {
    uint64_t tris = 0;
    for (uint32_t m = 0; m < in[j].mesh_count; ++m) tris += in[j].meshes[m].index_count / 3;
    h = mix(h, tris);
    h = mix(h, in[j].mesh_count);
    for (uint32_t l = 0; l < in[j].light_count; ++l)
      h = mix(h, ((uint64_t)in[j].lights[l].shape << 8) | (uint64_t)in[j].lights[l].illumination);
    h = mix_f(h, op[j].pixel_density);
    h = mix(h, op[j].photon_cube_resolution);
    h = mix(h, ((uint64_t)op[j].area_light_samples << 32) | op[j].sky_samples);
    h = mix(h, op[j].skip_bounce);
  }
  hipDeviceProp_t prop{};
  std::string dev = "device";
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) dev = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
  const char* q = std::getenv("GPU_MAX_HW_QUEUES");
  std::ostringstream s;
  s << std::hex << h << std::dec << ":" << dev << ":q" << (q ? q : "default");
  return s.str();
}

bool read_cache(const char* path, const std::string& sig, dl_tuned_launch* out) {
  if (!path) return false;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream s(line);
    std::string key;
    dl_tuned_launch t{};
    if (!(s >> key >> t.streams >> t.trace_block >> t.stream_block >> t.packet >> t.burial_block >> t.burial_lanes >>
          t.trace_stream >> t.ms))
      continue;
    if (key != sig) continue;
    *out = t;
    out->from_cache = 1;
    out->trials = 0;
    return true;
  }
  return false;
}

void write_cache(const char* path, const std::string& sig, const dl_tuned_launch& t) {
  if (!path) return;
  std::vector<std::string> keep;
  {
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line))
      if (line.compare(0, sig.size() + 1, sig + " ") != 0 && !line.empty()) keep.push_back(line);
  }
  std::ostringstream s;
  s << sig << " " << t.streams << " " << t.trace_block << " " << t.stream_block << " " << t.packet << " " << t.burial_block << " "
    << t.burial_lanes << " " << t.trace_stream << " " << t.ms;
  keep.push_back(s.str());
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream f(tmp, std::ios::trunc);
    for (const std::string& l : keep) f << l << "\n";
    if (!f) return;
  }
  std::rename(tmp.c_str(), path);
}

int apply(dl_ctx* ctx, const dl_tuned_launch& t) {
  int st = dl_set_launch(ctx, t.trace_block, t.burial_block);
  if (st == DL_OK) st = dl_set_burial(ctx, t.burial_lanes);
  if (st == DL_OK) st = dl_set_trace_stream(ctx, t.trace_stream ? 1 : 0);
  if (st != DL_OK) return st;
  // block and packet of the ray-stream kernel; every other gather setting as it is
  return dl_set_gather(ctx, ctx->gather_mode, t.stream_block, t.packet, 0, ctx->gather.blocks_per_cu, ctx->gather.min_waves, 0, 0);
}

}  // namespace

extern "C" int dl_autotune_batch(dl_ctx* ctx, const dl_bake_input* inputs, const dl_bake_options* options, uint32_t count,
                                 uint32_t max_streams, const char* cache_path, dl_tuned_launch* out) {
  if (!ctx) return dl::set_error(nullptr, DL_E_ARGUMENT, "null context");
  if (!out || (count && (!inputs || !options))) return dl::set_error(ctx, DL_E_ARGUMENT, "null argument");
  if (count == 0) return dl::set_error(ctx, DL_E_ARGUMENT, "no jobs to tune on");
  if (max_streams == 0 || max_streams > 16) return dl::set_error(ctx, DL_E_ARGUMENT, "max_streams must be 1 to 16");
  try {
    const std::string sig = signature(inputs, options, count, ctx->device);
    dl_tuned_launch best{};
    if (read_cache(cache_path, sig, &best) && best.streams >= 1 && best.streams <= max_streams) {
      const int st = apply(ctx, best);
      if (st != DL_OK) return st;
      *out = best;
      return DL_OK;
    }
    // the sample: jobs in cost order, evenly spaced
    std::vector<uint32_t> order(count);
    std::vector<double> cost(count);
    for (uint32_t i = 0; i < count; ++i) {
      order[i] = i;
      cost[i] = dl::batch_job_cost(inputs[i], options[i]);
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    const uint32_t n = std::min(count, kSample);
    std::vector<dl_bake_input> sin(n);
    std::vector<dl_bake_options> sop(n);
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t i = order[(uint64_t)k * count / n];
      sin[k] = inputs[i];
      sop[k] = options[i];
    }
    std::vector<dl_baked_scene*> res(n, nullptr);
    std::vector<int> status(n, 0);
    uint32_t trials = 0;
    auto time_it = [&](const dl_tuned_launch& t, double* ms) -> int {
      int st = apply(ctx, t);
      if (st != DL_OK) return st;
      double bestms = 1e300;
      for (int rep = 0; rep < 2; ++rep) {
        const auto t0 = std::chrono::steady_clock::now();
        st = dl_bake_batch(ctx, sin.data(), sop.data(), n, t.streams, res.data(), status.data());
        const double d = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (dl_baked_scene*& r : res) {
          dl_baked_scene_free(r);
          r = nullptr;
        }
        if (st != DL_OK) return st;
        bestms = std::min(bestms, d);
      }
      ++trials;
      *ms = bestms;
      return DL_OK;
    };
    best.streams = std::min<uint32_t>(8, max_streams);
    best.trace_block = ctx->trace_block;
    best.stream_block = ctx->gather.block;
    best.packet = ctx->gather.chunk;
    best.burial_block = ctx->burial_block ? ctx->burial_block : ctx->gather.block;
    best.burial_lanes = ctx->burial_lanes;
    best.trace_stream = ctx->trace_stream;
    {  // warm-up: allocations, code objects and worker streams
      double ms;
      int st = time_it(best, &ms);
      if (st != DL_OK) return st;
      st = time_it(best, &best.ms);
      if (st != DL_OK) return st;
    }
    std::vector<uint32_t> streams;
    for (uint32_t s : {2u, 4u, 6u, 8u, 12u, 16u})
      if (s <= max_streams) streams.push_back(s);
    if (streams.empty()) streams.push_back(max_streams);
    struct Param {
      uint32_t dl_tuned_launch::*field;
      std::vector<uint32_t> values;
    };
    const Param params[] = {
        {&dl_tuned_launch::streams, streams},
        {&dl_tuned_launch::trace_block, {64, 128, 256}},
        {&dl_tuned_launch::stream_block, {128, 256, 512}},
        {&dl_tuned_launch::packet, {64, 128, 256}},
        {&dl_tuned_launch::burial_block, {64, 128, 256}},
        {&dl_tuned_launch::burial_lanes, {0, 64, 128}},
        {&dl_tuned_launch::trace_stream, {1, 0}},
    };
    for (const Param& p : params)
      for (uint32_t v : p.values) {
        if (best.*(p.field) == v) continue;
        dl_tuned_launch t = best;
        t.*(p.field) = v;
        int st = time_it(t, &t.ms);
        if (st != DL_OK) return st;
        if (t.ms < best.ms) best = t;
      }
    const int st = apply(ctx, best);
    if (st != DL_OK) return st;
    best.trials = trials;
    best.from_cache = 0;
    write_cache(cache_path, sig, best);
    *out = best;
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return dl::set_error(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return dl::set_error(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

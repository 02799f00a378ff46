// Launch wrappers for the bake kernels (dl_kernels.hip).  Everything here is plain data: device
// pointers and sizes.  All wrappers enqueue on `stream` and return the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dl_bvh.h"
#include "dl_math.h"
#include "dl_trig.h"

namespace dl {

// One baked light in device form (per BVH slot); host-derived per-light quantities included.
struct DevLight {
  float pos[3];
  float radius;
  float bmin[3];  // sphere bounds Aabb::from_sphere (light_bvh.rs:74-78)
  float bmax[3];
  int32_t shape;
  uint32_t bits;
  uint32_t alpha;
  uint32_t bakes_bounce;
  uint32_t distant;
  float reach;    // shape_reach (light_shape.rs:22-28)
  float ax[3], ay[3], af[3];  // area light axes x, y, forward (light/area.rs:46-57)
  float hx, hy;
  uint32_t rect;
  float toward[3];  // sun direction toward the sun
  float srad, stan, ssin;
  uint32_t bounce_samples;
  float bounce_intensity;
  uint64_t id;
};

// Occluder scene on the device.
struct DevScene {
  const CNode* nodes;
  const float4* tris;
  uint32_t ntris;
  uint32_t stack;  // LDS stack entries per lane
  uint32_t stack16;  // 1: 16-bit stack entries (fewer than 65536 nodes), half the LDS per wave
  uint32_t* overflow;
  unsigned long long* counters;  // null, or [node fetches, triangle tests] for instrumented runs
  // Alpha-tested occluders (null when the scene has none): per occluder its texture or ~0u, the
  // UV0 of its corners (6 floats), per texture {width, height, first word of its bits, 0}, and
  // one bit per texel (alpha / 255 >= cutoff).
  const uint32_t* occ_tex;
  const float* occ_uv0;
  const uint32_t* tex;
  const uint32_t* tex_bits;
};

// Receiver triangles of the meshes being baked (SoA, one entry per triangle).
struct DevTris {
  uint32_t n;
  const float* v;       // 9 per triangle
  const float* nrm;     // 9 per triangle (vertex normals or the face normal three times)
  const float* uv;      // 6 per triangle (sanitized UV1)
  const uint32_t* mesh; // mesh index
  const uint32_t* local;// triangle index within its mesh
  float* face;          // 3 per triangle (written by setup)
  uint32_t* crop;       // 4 per triangle (written by setup)
};

struct DevMeshes {
  uint32_t n;
  const uint32_t* res;
  const uint64_t* id;
  const uint64_t* owned_off;    // word offset of the owned bit grid
  const uint64_t* visited_off;  // word offset of the visited bit grid
  uint32_t* owned;              // word pool
  uint32_t* visited;            // word pool
  unsigned long long* rays;     // per mesh counters
  unsigned long long* samples;
};

struct DevTexels {
  uint32_t* xy;     // x | y << 16
  float4* pos;
  float4* nrm;
  uint32_t* tri;
};

struct DevRegions {
  const uint32_t* rect;   // 4 per region
  const uint32_t* bits;
  const uint64_t* off;    // word offset in pool
  const uint32_t* mesh;
  uint32_t* pool;
  // Capture of per-(mesh, light) data (null when not wanted): a region-local visited bit grid
  // (1 bit per texel of the rect, LSB first, at vis[vis_off[r]]) and the region's rays and samples.
  uint32_t* vis = nullptr;
  const uint64_t* vis_off = nullptr;
  unsigned long long* rays = nullptr;
  unsigned long long* samples = nullptr;
};

struct DevEntries {
  uint64_t n;
  const uint32_t* light;
  const uint32_t* region;
  const uint32_t* tri;
  const uint64_t* start;  // per triangle, n_tris + 1
};

hipError_t launch_setup(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, uint32_t nlights,
                        uint32_t local_lights, uint64_t* light_count, uint64_t* trace_count, uint64_t* touched_count);
hipError_t launch_assoc_fill(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, uint32_t nlights,
                             uint32_t local_lights, const uint64_t* ent_start, uint32_t* ent_light, uint32_t* ent_tri);
hipError_t launch_owned(hipStream_t s, DevTris tris, DevMeshes meshes, const uint64_t* trace_start, uint64_t total);
hipError_t launch_texels(hipStream_t s, DevTris tris, DevMeshes meshes, const uint64_t* trace_count,
                         const uint64_t* cand_start, uint64_t total, int fill, unsigned long long* counts,
                         const uint64_t* tex_start, DevTexels out);
// *out += a fingerprint of n words (order-free sum of splitmix64(index << 32 | word)).
hipError_t launch_words_hash(hipStream_t s, const uint32_t* words, uint64_t n, unsigned long long* out);
// out[i] = i for i < n.
hipError_t launch_iota_u64(hipStream_t s, uint64_t* out, uint64_t n);
hipError_t launch_add_u32(hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t delta);
// Renumbers in place the light slots (through `table`, old slot -> new) of n meshes' words, mesh j
// at base[j] with nt[j] triangles and `stride` words per light entry.
hipError_t launch_remap_slots(hipStream_t s, uint32_t* words, const uint64_t* base, const uint32_t* nt, uint32_t n, uint32_t stride,
                              const uint32_t* table);
// The burial test with one lane per texel (block threads per block).
hipError_t launch_burial_lanes(hipStream_t s, DevScene scene, DevTris tris, DevMeshes meshes, DevTexels tx, uint64_t ntex,
                               uint64_t seed, TrigTables trig, uint8_t* buried, unsigned long long* kept, uint32_t block);
// The burial test streamed (burial rays, the closest-hit stream, the tally): rays k0..k0+nk-1
// (k0 + nk <= 8) of max_n texels, tx index first + j or list[j] (then *count of them).
hipError_t launch_burial_rays(hipStream_t s, DevTris tris, DevMeshes meshes, DevTexels tx, uint64_t first, const uint32_t* list,
                              const unsigned long long* count, uint64_t max_n, uint32_t k0, uint32_t nk, uint64_t seed,
                              TrigTables trig, float4* rays, float4* hits, uint64_t* total);
// The rays of a burial pass (slots below *total, at most max_rays) that hit nothing, appended to
// `out` as 8 floats each while fewer than `cap` are there; *count counts them all.
hipError_t launch_burial_misses(hipStream_t s, const float4* rays, const float4* hits, const uint64_t* total, uint64_t max_rays,
                                float* out, uint64_t cap, unsigned long long* count);
// *count += the nonzero flags among n.
hipError_t launch_count_nonzero(hipStream_t s, const uint32_t* flags, uint64_t n, unsigned long long* count);
hipError_t launch_burial_decide(hipStream_t s, DevScene scene, DevTexels tx, uint64_t first, const uint32_t* list,
                                const unsigned long long* count, uint64_t max_n, uint32_t k0, uint32_t nk, const float4* rays,
                                const float4* hits, uint8_t* back_count, uint8_t* buried, unsigned long long* kept, uint32_t* next,
                                unsigned long long* next_n);
hipError_t launch_compact(hipStream_t s, DevTexels in, uint64_t ntex, const uint8_t* buried, const uint64_t* start,
                          unsigned long long* cursor, DevTexels out);
struct TraceParams {
  uint64_t seed;
  uint32_t area_samples, area_min, sky_samples;
  uint32_t stream_multi = 0;  // 1: k_trace leaves area lights, soft suns and skies to the k_vis_* kernels
};
// Streamed multi-ray visibility (area lights, soft suns, skies).  A job is one (texel, entry) of
// such a light: job q of the batch is texel q - js[k] of entry me[k] (js: nme + 1 prefix counts of
// the entries' texels).  A chunk holds jobs q0 .. q0 + n - 1.  The running share of a job:
struct VisState {
  float total, lit;
  uint32_t nr, flags;  // bit 0 any ray lit, 1 any ray blocked, 2 texel in reach, 3 needs phase B
};
hipError_t launch_vis_count(hipStream_t s, DevEntries ent, const DevLight* lights, const uint64_t* js, const uint32_t* me,
                            uint32_t nme, const uint64_t* tex_start, uint64_t q0, uint64_t n, TraceParams p, uint64_t* out);
hipError_t launch_vis_rays_a(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                             const uint64_t* js, const uint32_t* me, uint32_t nme, const uint64_t* tex_start, uint64_t q0,
                             uint64_t n, TraceParams p, TrigTables trig, float4* rays, float* contrib, float* wts,
                             VisState* state);
hipError_t launch_vis_mid(hipStream_t s, DevEntries ent, const DevLight* lights, const uint64_t* js, const uint32_t* me,
                          uint32_t nme, const uint64_t* tex_start, uint64_t q0, uint64_t n, TraceParams p, const float4* rays,
                          const float* contrib, const float* wts, VisState* state, uint64_t* need);
hipError_t launch_vis_list(hipStream_t s, const uint64_t* need, const uint64_t* scan, uint64_t n, uint32_t* blist);
hipError_t launch_vis_rays_b(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                             const uint64_t* js, const uint32_t* me, uint32_t nme, const uint64_t* tex_start, uint64_t q0,
                             const uint32_t* blist, uint64_t nb, TraceParams p, TrigTables trig, float4* rays, float* contrib,
                             float* wts);
hipError_t launch_vis_finish(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                             DevRegions rg, const uint64_t* js, const uint32_t* me, uint32_t nme, const uint64_t* tex_start,
                             uint64_t q0, uint64_t n, TraceParams p, const VisState* state, const uint64_t* bscan, uint64_t nb,
                             const float4* rays, const float* contrib, const float* wts);
hipError_t launch_trace(hipStream_t s, DevScene scene, DevTris tris, DevMeshes meshes, const DevLight* lights,
                        DevTexels tx, const uint64_t* tex_start, DevEntries ent, DevRegions reg,
                        const uint64_t* unit_start, uint64_t units, TraceParams p, TrigTables trig, uint32_t block,
                        bool multi = true);
hipError_t launch_optimize(hipStream_t s, DevTris tris, DevMeshes meshes, DevEntries ent, DevRegions reg,
                           const uint8_t* has_bounce, uint8_t* keep);
hipError_t launch_pad_seams(hipStream_t s, DevMeshes meshes, DevRegions reg, const uint32_t* used_regions,
                            const uint64_t* area_start, uint32_t nused, uint64_t total);
// Crop + dilate + pack of kept entries into the output words (shadows.rs:148-160, seams.rs:43-95).
hipError_t launch_crop(hipStream_t s, DevTris tris, DevEntries ent, DevRegions reg, const uint64_t* kept_entry,
                       const uint64_t* kept_word_start, uint64_t nkept, uint64_t total_words,
                       const uint64_t* kept_out, uint32_t* out);
hipError_t launch_serialize(hipStream_t s, DevTris tris, const uint64_t* mesh_base, const uint32_t* ld_off,
                            const uint64_t* tri_kept_start, const uint64_t* kept_entry, const uint32_t* kept_shadow_off,
                            const uint32_t* kept_bounce_off, const uint32_t* ent_light, uint32_t stride,
                            const uint32_t* tri_first_in_mesh, uint32_t* out);
hipError_t launch_distance_cubes(hipStream_t s, DevScene scene, const DevLight* lights, uint32_t nlocal, uint32_t* out);

// ---- bounce (bounce.rs, sun_bounce.rs, bounce_map.rs, bounce_pack.rs) -------------------------
// Per occluder: vertex normals (9 floats, used when flags bit 0), albedo (3 floats).
struct DevOcc {
  const float* nrm;
  const float* albedo;
  const uint32_t* flags;
};
struct BounceParams {
  uint64_t seed;
  uint32_t colored;
  uint32_t cube_res;
  const uint32_t* photons;  // 4 words per photon (distance, normal, albedo, 0), 6 * res^2 per cube
  uint32_t sun_only;        // k_bounce_gather: point lights are gathered by the streamed kernel
  // Photon tiles (launch_photon_tiles): the tiles step by kPhotonTile texels over a cube face
  // (tiles_per_face^2 per face), and each summarizes the kTileSpan^2 texels from its corner (its
  // own 8x8 and the first row and column of its neighbours'): tile_live holds two words per tile,
  // one bit per texel whose photon has a lit surface (bit dy * kTileSpan + dx), tile_box three
  // float4 (shared normal | uniform flag, box minimum | coordinate scale, box maximum).  Any 2x2
  // block of texels lies in the summary of the tile holding its first texel.  Null tile_live: the
  // gather fetches every sample's photon.
  uint32_t tiles_per_face = 0;
  const unsigned long long* tile_live = nullptr;
  const float4* tile_box = nullptr;
  unsigned long long* photon_reads = nullptr;  // counting: photon records read by launch_gather_rays
};
static constexpr uint32_t kPhotonTile = 8;  // photon tile step (its own texels: 8x8)
static constexpr uint32_t kTileSpan = kPhotonTile + 1;  // texels a tile summarizes along each axis
inline uint32_t photon_tiles_per_face(uint32_t res) { return (res + kPhotonTile - 1) / kPhotonTile; }
// Bounce work: entries (triangle-light pairs) whose light bakes bounce, each with its window.
struct DevBounce {
  uint64_t n;                 // bounce entries
  const uint64_t* entry;      // entry index of each bounce entry
  const uint64_t* unit_start; // n + 1: 64-texel chunks per bounce entry
  const uint32_t* win_rect;   // 4 per window
  const uint64_t* win_off;    // float3 offset of each window
  const uint32_t* ent_window; // window of each entry (all entries; ~0u when none)
  const int32_t* cube_of_slot;// cube index per light slot, -1 when none
  float* map_a;               // 3 floats per window texel
  float* map_b;
  unsigned long long* rays;   // per mesh
  const uint64_t* tile_start; // filter tiles (kFilterTile squares) before each window, windows + 1
  uint8_t* tile_lit;          // per filter tile, zeroed: 1 once a gathered value in it has a positive component
  uint8_t* tile_nb;           // per filter tile, written by the filter: 1 when it wrote the tile's values in
                              // map_b, 0 when they are all zero (no mark among its 3x3 tiles; left unwritten)
  unsigned long long* wrays = nullptr;  // per window: bounce rays of its (mesh, light) (capture; may be null)
};
hipError_t launch_photons(hipStream_t s, DevScene scene, DevOcc occ, const DevLight* lights, const uint32_t* cube_slot,
                          uint32_t ncubes, uint32_t res, uint32_t* photons, uint32_t block);
// Tile summaries of built photon cubes (BounceParams::tile_live / tile_box): 6 * tpf^2 tiles per
// cube with tpf = photon_tiles_per_face(res); live holds one u64 per tile, box three float4.
// need_mask (lazy cubes, else null): per tile the texels that hold a photon; the others count as
// unlit.
hipError_t launch_photon_tiles(hipStream_t s, const DevLight* lights, const uint32_t* cube_slot, uint32_t ncubes, uint32_t res,
                               const uint32_t* photons, unsigned long long* live, float4* box,
                               const unsigned long long* need_mask = nullptr);
// Lazy photon cubes, for bakes whose gathers read few of their cubes' texels.  need: one byte per
// cube texel in tile order (64 * tile + bit), zeroed; launch_photon_mark sets the bytes of the
// texels a list of gather jobs can read (p.tiles_per_face set); launch_need_count turns them into
// per-tile masks and counts (tiles + 1 entries, for a scan), launch_need_list lists the texels
// (64 * tile + bit) at the scanned offsets, and launch_photons_listed traces those texels only.
hipError_t launch_photon_mark(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* jobs, uint64_t n, BounceParams p, TrigTables trig, uint8_t* need);
hipError_t launch_need_count(hipStream_t s, const uint8_t* need, uint64_t tiles, unsigned long long* need_mask, uint64_t* cnt);
hipError_t launch_need_list(hipStream_t s, const unsigned long long* need_mask, const uint64_t* off, uint64_t tiles, uint32_t* list);
hipError_t launch_photons_listed(hipStream_t s, DevScene scene, DevOcc occ, const DevLight* lights, const uint32_t* cube_slot,
                                 uint32_t res, const uint32_t* list, uint64_t total, uint32_t* photons, uint32_t block);
hipError_t launch_bounce_owner(hipStream_t s, DevTris tris, const DevLight* lights, DevTexels tx, const uint64_t* tex_start,
                               DevEntries ent, DevBounce b, uint64_t units);
hipError_t launch_bounce_gather(hipStream_t s, DevScene scene, DevOcc occ, DevTris tris, DevMeshes meshes,
                                const DevLight* lights, DevTexels tx, const uint64_t* tex_start, DevEntries ent, DevBounce b,
                                uint64_t units, BounceParams p, TrigTables trig, uint32_t block);
// Split point-light gather.  Jobs are (texel, entry) pairs within the light's radius: pass 0 of
// launch_gather_jobs writes per bounce entry its traced and count-only job counts, and its sun
// job count (every texel of a sun bounce entry), 3 (n + 1) words; pass 1 fills the three lists at
// the scanned offsets.  Traced jobs then run in chunks: ray offsets
// (launch_gather_nrays + scan), rays and contributions (launch_gather_rays), any-hit visibility
// (launch_any_stream, persistent) and the per-job sums (launch_gather_sum).
struct GatherLaunch {
  uint32_t block = 256;         // threads per block of the any-hit kernel
  uint32_t chunk = 128;         // rays a wave takes from the queue at once
  uint32_t blocks_per_cu = 0;   // 0: as many as fit
  uint32_t min_waves = 0;       // 8: registers capped for 8 waves per SIMD; otherwise uncapped
  uint32_t refill = 64;         // idle lanes before finished lanes take new rays (1-64)
  uint64_t chunk_rays = 1ull << 28;  // rays per chunk of jobs (buffer size)
  uint32_t share = 1;           // the persistent grid fills 1/share of the device (bakes side by side)
  uint32_t order = 1;           // ray slots: 0 job-major, 1 sample-major
};
hipError_t launch_gather_jobs(hipStream_t s, DevTris tris, const DevLight* lights, DevTexels tx, const uint64_t* tex_start,
                              DevEntries ent, DevBounce b, int pass, uint64_t* cnt, const uint64_t* off, uint2* trace_jobs,
                              uint2* count_jobs, uint2* sun_jobs);
hipError_t launch_gather_count(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                               DevBounce b, const uint2* jobs, uint64_t njobs, BounceParams p, TrigTables trig);
hipError_t launch_gather_nrays(hipStream_t s, DevEntries ent, const DevLight* lights, const uint2* jobs, uint64_t n, uint64_t* out);
hipError_t launch_gather_rays(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* jobs, uint64_t n, const uint64_t* rs, BounceParams p, TrigTables trig,
                              float4* rays, float* contrib, uint32_t order, uint32_t* tmask, uint32_t words, uint64_t* tcount);
// The split sample setup (instead of launch_gather_rays, with the photon tiles): phase A decides
// most samples with fast arithmetic and marks the others open (omask, ocount per job, n + 1
// entries; tmask and tcount zeroed); after a scan of ocount and launch_gather_list over omask,
// phase B evaluates the `*count` open slots of `list` exactly and fills tmask and tcount.
hipError_t launch_gather_split(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                               DevBounce b, const uint2* jobs, uint64_t n, const uint64_t* rs, BounceParams p, float4* rays,
                               uint32_t order, uint32_t* omask, uint32_t* tmask, uint32_t words, uint64_t* ocount,
                               uint64_t* tcount);
hipError_t launch_gather_open(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* jobs, uint64_t n, const uint64_t* rs, BounceParams p, TrigTables trig,
                              float4* rays, float* contrib, uint32_t order, const uint32_t* list, const uint64_t* count,
                              uint64_t max_count, uint32_t* tmask, uint32_t words, uint64_t* tcount);
// Exact and fast gather sample directions (3 floats each per input) and whether the fast one
// applies, for n inputs (a3: 3 floats each).
hipError_t launch_test_directions(hipStream_t s, TrigTables trig, const float* a3, const float* z, const float* ang,
                                  const float* sp, uint64_t n, float* exact3, float* fast3, uint32_t* fast_ok);
// The slots of the traced rays of a chunk (a job's in sample order, jobs in order) from the
// per-job bits `tmask` (words per job) and the scanned traced counts `toff`.
hipError_t launch_gather_list(hipStream_t s, const uint32_t* tmask, uint32_t words, const uint64_t* toff, const uint64_t* rs,
                              uint64_t n, uint32_t order, uint32_t* list);
// The traced slots (flag 2) of n jobs of `slots` ray slots each (the first rs[i + 1] - rs[i] of
// job i when rs is given) stored at ray_slot(order, ...), job by job: count (n + 1), off (n + 1;
// off[n] is the total) and list (up to n * slots).
hipError_t launch_ray_list(hipStream_t s, const float4* rays, uint64_t n, uint32_t slots, const uint64_t* rs, uint32_t order,
                           uint64_t* count, uint64_t* off, uint32_t* list, void** tmp, size_t* tmp_bytes);
// max_rays: the chunk's ray slots (jobs x the most samples of any of its lights); rs, njobs: the
// jobs' ray offsets and count, which sample-major order needs to skip the slots past a job's samples.
// list (optional): the slots to trace, *total of them, in this order.
hipError_t launch_any_stream(hipStream_t s, DevScene scene, const float4* rays, const uint64_t* total, uint64_t max_rays,
                             const uint64_t* rs, uint64_t njobs, float* contrib, unsigned int* queue, const GatherLaunch& gl,
                             const uint32_t* list = nullptr);
hipError_t launch_gather_sum(hipStream_t s, DevEntries ent, const DevLight* lights, DevTexels tx, DevBounce b, const uint2* jobs,
                             uint64_t n, const uint64_t* rs, const float* contrib, uint32_t order, const uint32_t* tmask,
                             uint32_t words);
// Closest-hit form of launch_any_stream: every traced ray that hits writes (t, u, v, leaf position)
// to its slot of `hits`; slots of rays that miss keep what they held.  rs is implied uniform.
hipError_t launch_closest_stream(hipStream_t s, DevScene scene, const float4* rays, const uint64_t* total, uint64_t max_rays,
                                 uint64_t njobs, float4* hits, unsigned int* queue, const GatherLaunch& gl);
// Streamed sun bounce over a chunk of n sun jobs with `slots` ray slots each (the most bounce
// samples of the chunk's suns): cosine rays and cleared hits (launch_sun_rays, which also stores
// n * slots to *total), closest hits (launch_closest_stream), sun rays and contributions
// (launch_sun_mid), their visibility (launch_any_stream), ordered sums and ray counts
// (launch_sun_sum).
hipError_t launch_sun_rays(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                           const uint2* jobs, uint64_t n, uint32_t slots, uint32_t order, BounceParams p, TrigTables trig,
                           float4* rays, float4* hits, uint64_t* total);
hipError_t launch_sun_mid(hipStream_t s, DevScene scene, DevOcc occ, const DevLight* lights, DevEntries ent, DevTexels tx,
                          DevBounce b, const uint2* jobs, uint64_t n, uint32_t slots, uint32_t order, uint32_t colored,
                          float4* rays, const float4* hits, float* contrib);
hipError_t launch_sun_sum(hipStream_t s, DevTris tris, DevEntries ent, const DevLight* lights, DevTexels tx, DevBounce b,
                          const uint2* jobs, uint64_t n, uint32_t slots, uint32_t order, const float4* rays, const float* contrib);
// Sun map of one sun (sun_map.rs:25-75): frame axes, bounds minimum, texel size and depth.
struct SunMapParams {
  float side[3], up[3], along[3];
  float min[3];
  float texel, depth;
  uint32_t res;
  uint32_t alpha;
};
hipError_t launch_sun_map(hipStream_t s, DevScene scene, const SunMapParams& p, uint32_t* out, uint64_t words);
// Sky probe grid (sky_probes.rs:46-106): per (probe, axis) openness and back-face hit count.
struct SkyProbeParams {
  float origin[3];
  float spacing;
  uint32_t counts[3];
  uint32_t samples;
  uint32_t alpha;
  uint64_t seed;
};
hipError_t launch_sky_probes(hipStream_t s, DevScene scene, TrigTables trig, const SkyProbeParams& p, float* open, uint32_t* back);
struct BlurWeights {
  float w[7];
  float sum;
};
// Dilate and blur of every window from map_a into map_b; tile_start (nwin + 1): the first tile of
// each window (ceil(width / kFilterTile) x ceil(height / kFilterTile) square tiles, row-major).
constexpr int kFilterTile = 16;
hipError_t launch_bounce_filter(hipStream_t s, DevBounce b, const uint32_t* win_mesh_res, const uint64_t* tile_start,
                                uint32_t nwin, uint64_t ntiles, BlurWeights weights);
hipError_t launch_bounce_has(hipStream_t s, DevTris tris, DevEntries ent, DevBounce b, uint8_t* has_bounce);
// ORs every region's region-local visited bits into its mesh's visited grid.
// Visit counts of a rebake service's meshes: adds or subtracts n pairs' visited bits (pair p:
// its bits at device address vis[p], rect, mesh resolution, counts from word count_off[p]; start:
// word offsets, total words); then the visited bit grids of n meshes from their counts.
hipError_t launch_vis_count(hipStream_t s, const uint64_t* vis, const uint32_t* rect, const uint32_t* res, const uint64_t* count_off,
                            uint32_t n, const uint64_t* start, uint64_t total, uint32_t* counts, int sign);
hipError_t launch_visited_from_counts(hipStream_t s, const uint32_t* res, const uint64_t* count_off, uint32_t n, const uint64_t* start,
                                      uint64_t total, const uint32_t* counts, uint32_t* visited);
hipError_t launch_region_visited_or(hipStream_t s, DevMeshes meshes, DevRegions rg, uint32_t nregions, const uint64_t* vis_words_start,
                                    uint64_t total_words);
// Copies n word ranges: dst[r.y .. r.y + r.z) = src[r.x .. r.x + r.z) for r = ranges[i] (x, y, z as
// u64 triples), from a scan of the lengths (range_start, n + 1 entries).
hipError_t launch_copy_ranges(hipStream_t s, const uint32_t* src, uint32_t* dst, const uint64_t* ranges,
                              const uint64_t* range_start, uint64_t n, uint64_t total_words);
// Ray queries through a fixed per-GPU device scratch (queue head, hit count, overflow flag): no
// device memory is allocated.  Rays, t_out and idx_out are device memory; *hits and *overflow are
// read back when the queries are done.
hipError_t launch_ray_queries_noalloc(hipStream_t s, int device, DevScene scene, const float* rays, uint64_t n, int closest,
                                      float* t_out, uint32_t* idx_out, uint64_t* hits, uint32_t* overflow);
// Copies the f32 RGB crop (with 2 texels of zero padding) each kept bounce entry's packing reads
// to out + value_start[j] (3 floats per texel).
hipError_t launch_bounce_dump(hipStream_t s, DevTris tris, DevEntries ent, DevBounce b, const uint64_t* kept_entry, uint64_t nkept,
                              const uint64_t* value_start, float* out);
// Packs the bounce of kept entries into `out` (bounce_pack.rs:18-86): per-word when no dither is
// drawn, one thread per entry otherwise.
hipError_t launch_bounce_pack(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevEntries ent,
                              DevBounce b, const uint64_t* kept_entry, const uint64_t* word_start, uint64_t nkept,
                              uint64_t total_words, const uint64_t* kept_out, const int32_t* comp_of_slot,
                              uint32_t colored, uint64_t seed, int dithered, uint32_t* out);

// Batched closest-hit / any-hit queries (tests and maps): rays as 8 floats (o, d, max, pad).
hipError_t launch_ray_queries(hipStream_t s, DevScene scene, const float* rays, uint64_t n, int closest,
                              float* t_out, uint32_t* idx_out);
// f32 sin/cos through the exception tables, for the exhaustive host-libm check.
hipError_t launch_trig_eval(hipStream_t s, TrigTables trig, const float* x, uint64_t n, float* sin_out, float* cos_out);

// Exclusive prefix sum over n+1 entries of `in` (in[n] may be anything; out[n] = total of in[0..n)).
hipError_t scan_u64(hipStream_t s, const uint64_t* in, uint64_t* out, uint64_t n, void** tmp, size_t* tmp_bytes);

}  // namespace dl

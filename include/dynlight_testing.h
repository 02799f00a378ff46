/*
 * dynlight_testing.h -- diagnostic entry points of libdynlight_hip.so.
 *
 * No counterpart in the reference: these expose single device passes so tests can check them
 * in isolation (ray queries against a scan of every occluder, f32 sin/cos against the host's
 * libm).  They run on the context's device and stream and allocate their own scratch.
 */
#ifndef DYNLIGHT_TESTING_H
#define DYNLIGHT_TESTING_H

#include "dynlight.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Builds the occluder BVH over `tri_count` triangles (9 floats each) and answers `ray_count`
 * rays given as 8 floats (origin, direction, max distance, unused).  closest = 0: hit_or_index[i]
 * is 1 for a hit; closest = 1: t_out[i] is the nearest distance (-1 on a miss) and
 * hit_or_index[i] the occluder index of the nearest hit (lowest index on ties). */
int dl_test_ray_queries(dl_ctx* ctx, const float* tris9, uint32_t tri_count, const float* rays8,
                        uint64_t ray_count, int closest, float* t_out, uint32_t* hit_or_index);

/* Host only (no device): builds the occluder BVH over `tri_count` triangles (9 floats each) on
 * `threads` threads and checks its structure: every node is reached once from the root, the
 * leaves' reference ranges partition the references, and each leaf record holds its triangle
 * (v0, occluder index, edges).  out: node count, reference count, depth, and a hash of the
 * references' triangle order (equal for every thread count).  DL_E_ARGUMENT when a check fails. */
int dl_test_bvh_structure(const float* tris9, uint32_t tri_count, uint32_t threads, uint64_t out[4]);

/* Evaluates the device f32 sin and cos at `n` inputs. */
int dl_test_trig(dl_ctx* ctx, const float* x, uint64_t n, float* sin_out, float* cos_out);

/* The same f32 sin/cos evaluated on the host from the tables this process builds (no device
 * needed): the check that the tables reproduce the host's libm. */
int dl_test_trig_host(const float* x, uint64_t n, float* sin_out, float* cos_out);

/* Number of exception entries of the sin and cos tables built for this context. */
void dl_test_trig_table_sizes(const dl_ctx* ctx, uint64_t* sin_entries, uint64_t* cos_entries);

/* The host evaluation of the table path (no device) compared with the host libm for every float
 * with bit pattern in [lo, hi]: the number of sin and cos results that differ. */
int dl_test_trig_exhaustive(uint32_t lo, uint32_t hi, uint64_t* mismatches);

/* The exception inputs of sin (which = 0) or cos (1): copies up to cap keys, returns the count. */
uint64_t dl_test_trig_keys(int which, uint32_t* out, uint64_t cap);

/* Counting mode: when enabled, later bakes on this context count the BVH nodes fetched and the
 * triangles tested by the shadow-ray kernel (dl_timings.trace_node_visits / trace_tri_tests). */
int dl_set_counting(dl_ctx* ctx, int enable);

/* Bounce capture: when enabled, each later bake on this context keeps, for every kept (triangle,
 * light) entry with bounce, the f32 RGB values its bounce packing reads: the filtered bounce map
 * cropped to the triangle with 2 texels of zero padding (bounce_map.rs:123-137), before the square
 * root and quantization of bounce_pack.rs.  Enabling or disabling clears what was kept. */
int dl_set_bounce_capture(dl_ctx* ctx, int enable);
/* The last bake's captured entries, sorted by (mesh id, triangle, light id): returns their count;
 * for index < count fills meta = {mesh id, triangle (mesh-local), light id, width, height} and
 * *values = width * height * 3 floats (row-major, RGB), valid until the context's next bake. */
uint64_t dl_test_bounce_capture(dl_ctx* ctx, uint64_t index, uint64_t meta[5], const float** values);

/* Launch shape of the traversal kernels (0 keeps the current value): threads per block of the
 * shadow-ray kernel and of the burial test's closest-hit ray stream. */
int dl_set_launch(dl_ctx* ctx, uint32_t trace_block, uint32_t burial_block);
/* The burial test's form: 0 (default) streams its rays through the closest-hit ray stream;
 * otherwise one lane per texel casts its rays in order, lanes_block threads per block. */
int dl_set_burial(dl_ctx* ctx, uint32_t lanes_block);
/* Per job of the context's last dl_bake_batch (up to cap jobs): its dl_timings, the worker that
 * baked it, and its start and end in ms from the start of the batch (two doubles per job).  Any
 * output may be null.  Returns the job count of that batch. */
uint32_t dl_test_batch_trace(dl_ctx* ctx, dl_timings* timings, uint32_t* worker, double* span, uint32_t cap);

/* 1 (default): area lights, soft suns and skies are traced as streamed jobs through the persistent
 * any-hit kernel; 0: each lane of the shadow-ray kernel casts its texel's rays in turn. */
int dl_set_trace_stream(dl_ctx* ctx, int enable);

/* 1 (default): the point-light gather summarizes each photon cube in 8x8-texel tiles (which
 * texels hold a lit photon, their shared normal, a box around the photon points a sample can
 * reconstruct) and reads a sample's photon only when its tile cannot show that the sample casts
 * no ray or brings no light; 0: every sample reads its photon.  Results are identical. */
int dl_set_photon_tiles(dl_ctx* ctx, int enable);

/* 1 (default): with the photon tiles, the gather's sample setup runs in two passes: every sample
 * once in fast f32 arithmetic (hardware sin, cos, reciprocal and square roots), which decides the
 * samples whose cube texel is certain and whose tile settles them, and the exact chain for the
 * others; 0: the exact chain for every sample.  Results are identical. */
int dl_set_gather_split(dl_ctx* ctx, int enable);

/* 1 (default): a bake whose gathers set up fewer samples than twice its photon cubes' texels (one
 * mesh batch, split setup) traces only the cube texels its gather samples can fall in, marked from
 * the samples' fast directions before the cubes are built; 0: every texel of every cube.  Results
 * are identical. */
int dl_set_lazy_photons(dl_ctx* ctx, int enable);

/* The sample directions of the gather for n inputs (dir3: the job's direction to its texel, 3
 * floats; z and azimuth: on_unit_sphere's draws; spread: the slerp factor): exact3 as the bake
 * computes them (bounce.rs:262-270), fast3 as the split setup's fast pass does, fast_ok[i] = 1
 * when the fast pass applies to input i.  For the test of the fast pass's error bound. */
int dl_test_gather_directions(dl_ctx* ctx, const float* dir3, const float* z, const float* azimuth, const float* spread,
                              uint64_t n, float* exact3, float* fast3, uint32_t* fast_ok);

/* The light BVH the bake builds on the host (LightBvh::build, light_bvh.rs:59-189) over n point
 * lights (positions: 3 floats each, radii): writes up to `cap` nodes and the lights' slot order
 * (n entries) and returns the node count (0 when cap is too small or a position is not
 * finite).  Needs no device. */
uint32_t dl_test_light_bvh(const float* positions, const float* radii, uint32_t n, dl_light_bvh_node* nodes, uint32_t cap,
                           uint32_t* order);

/* Point-light bounce gather: mode 0 = one wave per (triangle, light, 64 texels) tracing its
 * samples in turn; mode 1 = split gather (rays, persistent any-hit traversal, ordered sums;
 * default).  For mode 1: threads per block and rays per queue run of the any-hit kernel, rays per
 * chunk of jobs (the ray buffer's size), blocks per CU (0 = as many as fit) and the occupancy the
 * any-hit kernel's registers are capped for (8 waves per SIMD, or 0 for no cap), and how many of a
 * wave's lanes must be idle before they take new rays (1-64), and the order of the ray slots
 * (1 job-major: a job's samples together; 2 sample-major: one sample of every job together, the
 * default).  Zero keeps the current value for block, chunk, chunk_rays, refill and order. */
int dl_set_gather(dl_ctx* ctx, uint32_t mode, uint32_t block, uint32_t chunk, uint64_t chunk_rays, uint32_t blocks_per_cu,
                  uint32_t min_waves, uint32_t refill, uint32_t order);

#ifdef __cplusplus
}
#endif
#endif

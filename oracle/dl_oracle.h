/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement of the reference crate's bake (`dynlight`, src/bake/ files), kept in the
 * reference's f32 operation order (built with -ffp-contract=off, glibc sinf/cosf/tanf/expf as
 * Rust std calls them).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the CPU baseline: the product library never links it.
 *
 * Parity status of each function is stated beside it in dl_oracle.cpp; functions pinned by the
 * reference's own known-answer tests are exercised by tests/test_oracle_kats.py.
 */
#ifndef DL_ORACLE_H
#define DL_ORACLE_H

#include "../include/dynlight.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dlo_scene dlo_scene;

/* Same contract as dl_bake; `threads` = worker threads (0 = all cores), `brute_force` = 1 to
 * scan every occluder per ray instead of using the oracle's BVH. */
int dlo_bake(const dl_bake_input* input, const dl_bake_options* options, uint32_t threads,
             uint32_t brute_force, dlo_scene** out);
/* The bench's CPU-baseline sample: bakes only the meshes with receivers[i] != 0 (indexed like
 * input->meshes) while every mesh still occludes.  Not a reference entry point. */
int dlo_bake_receivers(const dl_bake_input* input, const dl_bake_options* options, uint32_t threads,
                       const uint8_t* receivers, dlo_scene** out);
const char* dlo_last_error(void);
/* Wall times (ms) of the last bake on this thread: photon cubes, meshes, distance cubes. */
void dlo_last_timings(double* out3);
void dlo_scene_free(dlo_scene* scene);
void dlo_scene_settings(const dlo_scene* scene, dl_baked_settings* out);
void dlo_scene_stats(const dlo_scene* scene, dl_bake_stats* out);
uint32_t dlo_scene_light_count(const dlo_scene* scene);
int dlo_scene_light(const dlo_scene* scene, uint32_t index, dl_baked_light* out);
uint32_t dlo_scene_light_bvh(const dlo_scene* scene, const dl_light_bvh_node** nodes);
uint32_t dlo_scene_mesh_count(const dlo_scene* scene);
int dlo_scene_mesh(const dlo_scene* scene, uint32_t index, dl_baked_mesh_view* out);
uint64_t dlo_scene_distance_cubes(const dlo_scene* scene, const uint32_t** words);
uint32_t dlo_scene_sun_map_count(const dlo_scene* scene);
int dlo_scene_sun_map(const dlo_scene* scene, uint32_t index, dl_sun_map_view* out);
void dlo_scene_sky_probes(const dlo_scene* scene, dl_sky_probes_view* out);

/* ---- unit-level entry points for the reference's known-answer tests ---------------------- */
uint64_t dlo_splitmix64(uint64_t v);
uint64_t dlo_seed(const uint64_t* parts, uint32_t n);
/* PCG32 stream: fills `out` with n next_u32 values of EffectRng::new(seed, stream). */
void dlo_pcg32(uint64_t seed, uint64_t stream, uint32_t n, uint32_t* out);
/* random::stream(parts) then n next_f32 values. */
void dlo_stream_f32(const uint64_t* parts, uint32_t nparts, uint32_t n, float* out);
uint32_t dlo_lightmap_size(float area, float density, uint32_t max_size);
/* rects: kind 0 crop, 1 trace, 2 optimize -> out[4] = min_x, min_y, max_x, max_y */
void dlo_rect(uint32_t kind, const float* uvs6, uint32_t size, uint32_t* out4);
void dlo_closest_point(const float* p, const float* a, const float* b, const float* c, float* out);
int dlo_uv_barycentrics(const float* t6, const float* p2, float* out3);
/* Centered texels of a triangle: returns count, fills up to cap records of
 * (x, y, px, py, pz, nx, ny, nz) as floats (x,y stored as float values). */
uint32_t dlo_triangle_texels(const float* v9, const float* n9, const float* uv6, const float* face3,
                             uint32_t size, float* out, uint32_t cap);
/* Touched texels with an `owned` bit grid of size*size bits (LSB-first words). */
uint32_t dlo_touched_texels(const float* v9, const float* n9, const float* uv6, const float* face3,
                            uint32_t size, const uint32_t* owned, float* out, uint32_t cap);
uint8_t dlo_level(float fraction, uint32_t bits);
/* spread(index, offset, rectangle) -> out2 */
void dlo_spread(uint32_t index, const float* offset2, uint32_t rectangle, float* out2);
void dlo_concentric_disk(const float* square2, float* out2);
/* is_buried with a scripted back_face sequence (answers[i] for ray i); returns buried flag,
 * writes the number of rays asked and their directions (up to 8). */
int dlo_is_buried(const float* normal3, float turn, const uint8_t* answers, uint32_t* rays,
                  float* dirs24);
/* adaptive area-light share with a scripted occlusion rule; see dl_oracle.cpp. */
int dlo_area_visibility(const dl_light* light, float reach_override, const float* pos3,
                        const float* normal3, uint32_t samples, uint32_t minimum,
                        const uint64_t* key, uint32_t nkey, int32_t blocker_kind,
                        const float* blocker, float* fraction, uint32_t* rays);
int dlo_sun_visibility(const dl_light* light, const float* normal3, uint32_t samples,
                       uint32_t minimum, const uint64_t* key, uint32_t nkey, int32_t blocker_kind,
                       const float* blocker, float* fraction, uint32_t* rays);
float dlo_sky_visibility(const float* normal3, uint32_t samples, const uint64_t* key,
                         uint32_t nkey, int32_t blocker_kind, const float* blocker);
/* Light BVH over spheres; returns node count, fills nodes (cap) and order (n). */
uint32_t dlo_light_bvh(const float* positions, const float* radii, uint32_t n,
                       dl_light_bvh_node* nodes, uint32_t cap, uint32_t* order);
/* pad_seams / dilate / crop KATs on level grids (one byte per texel). */
void dlo_pad_seams(uint32_t width, uint32_t height, uint8_t* levels, const uint8_t* visited);
void dlo_dilate(uint32_t width, uint32_t height, uint8_t* levels);
/* Pack levels into words at `bits` per texel (ShadowGrid::into_words). Returns word count. */
uint32_t dlo_pack_levels(const uint8_t* levels, uint32_t count, uint32_t bits, uint32_t* out);
/* serialize_triangles: records given as flat arrays, see dl_oracle.cpp. */
uint32_t dlo_serialize(uint32_t triangles, const uint32_t* crops, const uint32_t* light_counts,
                       const uint32_t* lights, const uint32_t* shadow_lens,
                       const uint32_t* shadow_words, const uint32_t* bounce_lens,
                       const uint32_t* bounce_words, uint32_t bounce_layout, uint32_t* out,
                       uint32_t cap);
/* pack_bounce / pack_bounce_rgb with rng stream `key`. */
uint32_t dlo_pack_bounce(const float* values, uint32_t count, uint32_t rgb, int32_t compression,
                         const uint64_t* key, uint32_t nkey, uint32_t* out, uint32_t cap);
uint32_t dlo_f16_pair(float lower, float upper);
/* pack_sun_depth (sun_map.rs:77-80) */
uint32_t dlo_sun_depth_pair(float lower, float upper);
/* ProbeGrid::over (sky_probes.rs:46-62) */
void dlo_probe_grid(const float* min3, const float* max3, float spacing, float* origin3, float* spacing_out, uint32_t* counts3);
/* fill_inside (sky_probes.rs:110-144): open is 6 floats per probe, valid 1 byte per probe */
void dlo_fill_inside(const uint32_t* counts3, float* open, uint8_t* valid);
/* SkyProbes::pack (format/types.rs:105-112) */
void dlo_pack_probe(const float* open6, uint32_t* words2);
/* Ray queries on an occluder soup (brute force or BVH). tris: n*9 floats. */
int dlo_ray_any(const float* tris, uint32_t n, const float* origin, const float* dir, float max,
                uint32_t brute);
int dlo_ray_closest(const float* tris, uint32_t n, const float* origin, const float* dir,
                    float max, uint32_t brute, float* t, uint32_t* index);
/* Batched ray queries against one BVH build (for brute-force equality tests). */
void dlo_rays_batch(const float* tris, uint32_t n, const float* rays, uint32_t count,
                    uint32_t closest, uint32_t brute, float* t_out, uint32_t* idx_out);
/* glibc sinf/cosf as the reference's f32::sin/cos reach them. */
float dlo_sinf(float x);
float dlo_cosf(float x);

/* ---- bounce side (shared/cube.rs, bounce_map.rs, bounce.rs, random.rs) -------------------- */
void dlo_cube_face_uv(const float* d3, uint32_t* face, float* uv2);              /* direction_to_face_uv */
void dlo_cube_direction(uint32_t face, const float* uv2, float* out3);            /* face_uv_to_direction */
void dlo_cube_texel_direction(uint32_t face, uint32_t x, uint32_t y, uint32_t res, float* out3);
void dlo_cube_texel(const float* d3, uint32_t res, uint32_t* out3);               /* direction_to_texel */
uint32_t dlo_cube_texel_index(uint32_t face, uint32_t x, uint32_t y, uint32_t res);
void dlo_gaussian_weights(float* w7, float* sum);
/* BounceMap of side `size` windowed on crops4 with n texels (x, y, r, g, b); ops bit 0 dilates,
 * bit 1 blurs; out receives get(x, y) for every texel of the map (size*size*3 floats). */
void dlo_bounce_map(uint32_t size, const uint32_t* crops4, const float* texels5, uint32_t n, uint32_t ops, float* out);
/* BounceMap::crop of crop4 into out ((w+4)*(h+4)*3 floats); returns 1 when lit. */
int dlo_bounce_crop(uint32_t size, const uint32_t* crops4, const float* texels5, uint32_t n, const uint32_t* crop4, float* out);
void dlo_slerp(const float* a3, const float* b3, float s, float* out3);
float dlo_acos_approx(float v);
void dlo_on_unit_sphere(const uint64_t* key, uint32_t nkey, uint32_t n, float* out3n);
/* Texel window of later bakes (rect4 NULL ends it): in mesh `mesh_id` only the triangles whose
 * crop rectangle meets rect4 = (min_x, min_y, max_x, max_y) are baked; the others keep neither
 * texels nor lights.  Words of triangles whose crop lies 8 or more texels inside the rectangle
 * are then those of the whole bake (padding, ring dilation and the bounce filter reach at most
 * 4 texels). */
void dlo_set_window(uint64_t mesh_id, const uint32_t* rect4);
/* Capture of later bakes: mode 1 keeps each kept bounce entry's f32 crop (the values the bounce
 * packing reads) and the photon cubes; mode 2 the FP64 crops of the same entries; 0 stops.  The
 * setter clears what was captured. */
void dlo_set_bounce_capture(uint32_t mode);
/* Sorts the captured crops by (mesh id, triangle, light id) and returns their count. */
uint64_t dlo_bounce_capture_count(void);
/* meta5 = mesh id, triangle (mesh-local), light id, width, height; rgb: width*height*3 doubles. */
int dlo_bounce_capture_get(uint64_t i, uint64_t* meta5, double* rgb, uint64_t cap);
/* The captured photon cube of light `light_id`: 6*res*res distances, 3 normal bytes and 3 albedo
 * bytes per photon. */
int dlo_photon_cube(uint64_t light_id, uint32_t* res, const float** dist, const int8_t** nrm, const uint8_t** alb);

#ifdef __cplusplus
}
#endif
#endif

/*
 * dynlight.h -- C ABI of the MI355X bake backend (libdynlight_hip.so).
 *
 * A drop-in for the reference crate's pure bake entry point and its plain-data boundary:
 *
 *   pub fn bake(input: &BakeInput, options: &BakeOptions, progress: &BakeProgress)
 *       -> Result<BakedScene, BakeError>                 (src/bake/baker.rs:27-31)
 *
 * Every struct here is plain C: pointers and sizes, borrowed for the duration of one call.
 * Results are owned by the library and released with the matching *_free function.  Nothing
 * unwinds across this boundary; failures return a negative dl_status and leave a message that
 * dl_last_error() returns.
 */
#ifndef DYNLIGHT_H
#define DYNLIGHT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: BakeError (src/bake/input.rs:198-225) plus device classes ------------- */
typedef enum dl_status {
    DL_OK = 0,
    DL_E_CANCELLED = -1,       /* BakeError::Cancelled                                        */
    DL_E_DUPLICATE_ID = -2,    /* BakeError::DuplicateId { kind, id }                         */
    DL_E_INVALID_MESH = -3,    /* BakeError::InvalidMesh { id, reason }                       */
    DL_E_INVALID_LIGHT = -4,   /* BakeError::InvalidLight { id, reason }                      */
    DL_E_INVALID_OPTIONS = -5, /* BakeError::InvalidOptions(reason)                           */
    DL_E_DEVICE = -16,         /* a HIP runtime call failed                                   */
    DL_E_OOM = -17,            /* device or host allocation failed                            */
    DL_E_ARGUMENT = -18,       /* null pointer or out-of-range argument at the ABI itself     */
    DL_E_FORMAT = -19,         /* FormatError of decode (src/bake/format/mod.rs:53-77)        */
    DL_E_TRIANGLES_TAKEN = -20 /* TrianglesTaken (src/bake/format/mod.rs:84-95)               */
} dl_status;

/* ---- enums mirroring src/light/kinds.rs:97-158 and src/shared/settings.rs:229-237 -------- */
enum { DL_ILLUMINATION_DIRECT = 0, DL_ILLUMINATION_SINGLE_BOUNCE = 1 };
enum { DL_TRANSPARENCY_DISABLED = 0, DL_TRANSPARENCY_ENABLED = 1, DL_TRANSPARENCY_ENABLED_MAX = 2 };
/* BounceCompression; DL_BOUNCE_DEFAULT means "None: use BakeOptions::bounce_compression". */
enum { DL_BOUNCE_BITS8 = 0, DL_BOUNCE_BITS6 = 1, DL_BOUNCE_BITS5 = 2, DL_BOUNCE_BITS4 = 3,
       DL_BOUNCE_DEFAULT = -1 };
/* LightShape (src/light/shape.rs:19-103): Point, Area(Rectangle|Tube), Sun, Sky. */
enum { DL_SHAPE_POINT = 0, DL_SHAPE_RECTANGLE = 1, DL_SHAPE_TUBE = 2, DL_SHAPE_SUN = 3,
       DL_SHAPE_SKY = 4 };
/* DynamicGeometryMode: LightingOnly bakes no distance cubes, sun maps or sky probes. */
enum { DL_DYNAMIC_LIGHTING_ONLY = 0, DL_DYNAMIC_ANGULAR = 1, DL_DYNAMIC_DISTANCE_CUBES = 2 };

/* AlphaTexture (src/bake/input.rs:66-97). */
typedef struct dl_alpha_texture {
    uint32_t width, height;
    const uint8_t* rgba8; /* width*height RGBA8 texels, rows top to bottom */
    uint64_t rgba8_len;   /* bytes at rgba8; must be width*height*4 */
    float cutoff;         /* Henry uses 0.5 */
} dl_alpha_texture;

/* BakeMesh (src/bake/input.rs:24-45), world space. */
typedef struct dl_mesh {
    uint64_t id;                    /* BakeId                                               */
    uint64_t content_hash;          /* stored for stale detection                           */
    uint32_t vertex_count;
    const float* positions;         /* vertex_count * 3                                     */
    const float* normals;           /* vertex_count * 3, or NULL for flat normals           */
    uint32_t index_count;           /* multiple of 3                                        */
    const uint32_t* indices;
    const float* uv1;               /* vertex_count * 2                                     */
    const float* uv0;               /* vertex_count * 2 or NULL                             */
    const dl_alpha_texture* alpha;  /* NULL or an alpha texture (needs uv0)                 */
    uint32_t occludes;              /* 0 = every ray passes through the mesh                */
    float albedo[3];                /* mean linear reflectance, (1,1,1) by default          */
    uint32_t uv1_count;             /* must equal vertex_count (validation mirrors input.rs) */
    uint32_t uv0_count;
    uint32_t normal_count;          /* 0 for flat normals                                   */
} dl_mesh;

/* BakeLight (src/bake/input.rs:101-132) with its LightShape. */
typedef struct dl_light {
    uint64_t id;
    float position[3];
    float radius;
    int32_t illumination;   /* DL_ILLUMINATION_*                          */
    int32_t transparency;   /* DL_TRANSPARENCY_*                          */
    uint32_t bounce_samples; /* BakeBounce::samples (>= 2 used)            */
    float bounce_intensity;
    int32_t bounce_compression; /* DL_BOUNCE_* or DL_BOUNCE_DEFAULT        */
    int32_t shape;           /* DL_SHAPE_*                                */
    float rotation[4];       /* x, y, z, w: area light / sun rotation     */
    float size[2];           /* area light width, height (tube: height 0) */
    float angular_diameter;  /* sun, degrees                              */
} dl_light;

typedef struct dl_bake_input {
    const dl_mesh* meshes;
    uint32_t mesh_count;
    const dl_light* lights;
    uint32_t light_count;
} dl_bake_input;

/* BakeOptions (src/bake/input.rs:136-164). */
typedef struct dl_bake_options {
    float pixel_density;
    uint32_t max_lightmap_size;
    int32_t bounce_compression;     /* DL_BOUNCE_BITS* */
    int32_t dynamic_geometry;       /* DL_DYNAMIC_*    */
    uint32_t skip_bounce;
    uint32_t photon_cube_resolution;
    uint32_t area_light_samples;
    uint32_t area_light_min_samples;
    uint32_t sun_map_resolution;
    uint32_t sky_samples;
    float sky_probe_spacing;
    uint32_t colored_bounce;
    uint64_t seed;
} dl_bake_options;

/* BakeProgress (src/bake/progress.rs:38-112): written by the library, read by any thread.
 * Setting `cancel` to non-zero makes the bake stop at the next check and return
 * DL_E_CANCELLED. */
typedef struct dl_progress {
    volatile uint32_t stage; /* 0 Preparing, 1 PhotonCubes, 2 Meshes, 3 DistanceCubes, 4 Done */
    volatile uint64_t done;
    volatile uint64_t total;
    volatile uint32_t cancel;
} dl_progress;

/* BakeStats (src/bake/format/types.rs:209-223). */
typedef struct dl_bake_stats {
    uint32_t meshes;
    uint64_t triangles;
    uint64_t texels;
    uint64_t rays;
    uint64_t lights_removed;
    uint64_t bounce_rays;
    uint64_t triangle_words;
    /* Not in the reference's stats: texel-light visibility samples (texel_shadow returned
     * Some), the unit of the throughput metric. */
    uint64_t samples;
} dl_bake_stats;

/* LightBvhNode (src/bake/light_bvh.rs:17-25), 32 bytes as the shaders read it. */
typedef struct dl_light_bvh_node {
    float aabb_min[3];
    uint32_t left_first;
    float aabb_max[3];
    uint32_t count;
} dl_light_bvh_node;

/* BakedLight (src/bake/format/types.rs). */
typedef struct dl_baked_light {
    uint64_t id;
    float origin[3];
    float radius;
    int32_t shape;
    float rotation[4];
    float size[2];
    float angular_diameter;
    uint32_t shadow_bits_per_texel;
    int32_t bounce_compression; /* DL_BOUNCE_* or -1 when the light bakes no bounce */
} dl_baked_light;

/* A view of BakedMesh (src/bake/format/types.rs): words stay owned by the scene. */
typedef struct dl_baked_mesh_view {
    uint64_t id;
    uint64_t content_hash;
    uint32_t resolution;
    uint32_t occludes;
    const uint32_t* triangles; /* Henry's dynamic_triangles words (triangle_words.rs:1-12) */
    uint64_t triangle_word_count;
} dl_baked_mesh_view;

/* SunMap / SkyProbes views. */
typedef struct dl_sun_map_view {
    float rows[12];            /* three Vec4 rows, world to map space */
    uint32_t resolution;
    const uint32_t* texels;    /* unorm16 pairs */
    uint64_t texel_word_count;
} dl_sun_map_view;

typedef struct dl_sky_probes_view {
    uint32_t present;
    float origin[3];
    float spacing;
    uint32_t counts[3];
    const uint32_t* words;
    uint64_t word_count;
} dl_sky_probes_view;

/* BakedSettings. */
typedef struct dl_baked_settings {
    uint32_t bounce;
    uint32_t colored_bounce;
    int32_t dynamic_geometry;
    float pixel_density;
    uint32_t max_lightmap_size;
    uint32_t distance_cube_resolution;
} dl_baked_settings;

typedef struct dl_ctx dl_ctx;
typedef struct dl_baked_scene dl_baked_scene;

/* ---- context ------------------------------------------------------------------------------
 * One context owns one HIP device, its streams and the resident lookup tables. Calls on one
 * context are serialized internally; contexts are independent. */
int dl_ctx_create(int device, dl_ctx** out);
void dl_ctx_destroy(dl_ctx* ctx);
/* The message of the last failed call on this context (thread-local when ctx is NULL). */
const char* dl_last_error(const dl_ctx* ctx);
/* BakeOptions::default() (src/bake/input.rs:166-171 with BakeSettings::default(),
 * src/bake/settings.rs:67-81). */
void dl_default_options(dl_bake_options* out);

/* ---- the bake (replaces bake(), src/bake/baker.rs:27-203) --------------------------------- */
int dl_bake(dl_ctx* ctx, const dl_bake_input* input, const dl_bake_options* options,
            dl_progress* progress, dl_baked_scene** out);
void dl_baked_scene_free(dl_baked_scene* scene);

/* ---- the .dynlight file (src/bake/format/mod.rs:1-21, 101-301; format/stream.rs:34-71) ----------
 * dl_encode writes the scene's file bytes (encode); the same scene always gives the same bytes.
 * The buffer is the caller's to release with dl_bytes_free.  dl_decode parses a file held in
 * memory (decode), checking magic, version, length and checksum before the payload;
 * dl_decode_stream parses one as it is read (decode_stream): `read` fills up to cap bytes and
 * returns how many (0 at the end, negative on failure).  Both return DL_E_FORMAT with the
 * reference's FormatError message in dl_last_error(NULL) and its variant in
 * dl_format_error_kind().  A decoded scene's stats have no sample count (the file holds none). */
enum { DL_FORMAT_NOT_DYNLIGHT = 1, DL_FORMAT_UNSUPPORTED_VERSION = 2, DL_FORMAT_TRUNCATED = 3,
       DL_FORMAT_DECOMPRESS = 4, DL_FORMAT_CHECKSUM = 5, DL_FORMAT_INVALID = 6 };
typedef int64_t (*dl_read_fn)(void* user, uint8_t* buf, uint64_t cap);
int dl_encode(const dl_baked_scene* scene, uint8_t** bytes, uint64_t* len);
void dl_bytes_free(uint8_t* bytes);
int dl_decode(const uint8_t* bytes, uint64_t len, dl_baked_scene** out);
int dl_decode_stream(dl_read_fn read, void* user, dl_baked_scene** out);
int dl_format_error_kind(void);

/* ---- BakedBuffers (src/baked_lighting/layout.rs:1-215, apply.rs:27-98) ----------------------
 * The `triangles` and `aux` word buffers the runtime's shaders read, and the receiver lookup
 * (BakeId -> Receiver), built from a baked scene as scene_buffers does.  Receivers are listed in
 * the scene's mesh order: receiver i has tag i + 1 (MeshTag 0 means "not baked"). */
typedef struct dl_receiver {
    uint64_t id;
    uint64_t content_hash;
    uint32_t tag;
    uint32_t triangle_count;
    uint32_t occludes;
    uint32_t reserved;
} dl_receiver;
typedef struct dl_baked_buffers dl_baked_buffers;
int dl_scene_buffers(const dl_baked_scene* scene, dl_baked_buffers** out);
/* take_scene_buffers (apply.rs:45-61): the same buffers, releasing each mesh's triangle words as
 * they are copied so the scene and the buffers are never both whole in memory (the loader path,
 * baked_lighting/loader.rs:47-68).  Afterwards the scene is marked taken: dl_encode and
 * dl_scene_buffers refuse it with DL_E_TRIANGLES_TAKEN, as encode and scene_buffers do. */
int dl_take_scene_buffers(dl_baked_scene* scene, dl_baked_buffers** out);
/* BakedScene::triangles_taken */
int dl_scene_triangles_taken(const dl_baked_scene* scene);
uint64_t dl_buffers_triangles(const dl_baked_buffers* buffers, const uint32_t** words);
uint64_t dl_buffers_aux(const dl_baked_buffers* buffers, const uint32_t** words);
uint32_t dl_buffers_receiver_count(const dl_baked_buffers* buffers);
int dl_buffers_receiver(const dl_baked_buffers* buffers, uint32_t index, dl_receiver* out);
void dl_baked_buffers_free(dl_baked_buffers* buffers);

/* ---- the ECS caller's conversions (src/bake/snapshot.rs:83-158, 229-272; identity.rs:64-101)
 * For callers outside Bevy that assemble a dl_bake_input the way SceneSnapshot::take does. */
/* mesh_content_hash (src/shared/mesh_hash.rs:16-33): positions (3 floats) and UV (2 floats) of
 * every triangle corner in draw order by their bit patterns; indices past a buffer hash as
 * missing. */
uint64_t dl_mesh_content_hash(const float* positions, uint32_t vertex_count, const uint32_t* indices, uint64_t index_count,
                              const float* uv, uint32_t uv_count);
/* BakeId::from_path (src/shared/bake_id.rs:15-23): FNV-1a 64 over length-prefixed segments. */
uint64_t dl_bake_id_from_path(const char* const* segments, const uint64_t* lengths, uint32_t count);
/* A mesh or light awaiting its id (identity.rs:17-22): an explicit BakeId, or its Name path;
 * tiebreak_hash and position order candidates whose derived ids collide (meshes pass their
 * content hash, lights 0). */
typedef struct dl_id_candidate {
    uint32_t has_explicit;
    uint32_t segment_count;
    uint64_t explicit_id;
    const char* const* segments;
    const uint64_t* segment_lengths;
    uint64_t tiebreak_hash;
    float position[3];
    uint32_t reserved;
} dl_id_candidate;
/* assign_ids (identity.rs:64-101): explicit ids are kept; derived ids that collide keep the plain
 * id for the first in tiebreak order and get a "#k" segment for the k-th after it. */
int dl_assign_ids(const dl_id_candidate* candidates, uint32_t count, uint64_t* out);
/* MeshData::to_world (snapshot.rs:229-272): `affine` is 12 floats, the columns of the linear
 * part then the translation.  Positions are transformed as points; normals (may be NULL) by the
 * inverse transpose and normalized, or not written when the determinant is not normal (the
 * reference then bakes flat normals); a negative determinant swaps corners 1 and 2 of every
 * triangle so faces keep their front; *content_hash (may be NULL) hashes the local data.
 * Returns 1 when normals were written, 0 when not, or a negative error. */
int dl_mesh_to_world(const float* affine, const float* positions, uint32_t vertex_count, const float* normals,
                     const uint32_t* indices, uint64_t index_count, const float* uv1, uint32_t uv_count, float* positions_out,
                     float* normals_out, uint32_t* indices_out, uint64_t* content_hash);

/* ---- rebake service (no counterpart in the reference, which re-runs bake() on every change) ----
 * A service keeps one scene's geometry resident on the device: receiver meshes, the occluder BVH
 * and every triangle's unburied texels (none of which depend on the lights).  A rebake then bakes
 * a set of lights (moved or added ones) against it, re-tracing only the (triangle, light) pairs of
 * those lights: the result equals dl_bake of the scene's meshes with exactly those lights and the
 * service's options.  Destroy services before their context. */
typedef struct dl_service dl_service;
typedef struct dl_service_request {
    const dl_light* lights;  /* the lights to rebake                                   */
    uint32_t light_count;
    double arrival_s;        /* arrival time in seconds after the replay starts (sorted) */
} dl_service_request;

/* Bakes `scene` (meshes and lights) once and keeps its geometry resident; *initial (if not NULL)
 * receives that bake's result. */
int dl_service_create(dl_ctx* ctx, const dl_bake_input* scene, const dl_bake_options* options, dl_service** out,
                      dl_baked_scene** initial);
/* Bakes `lights` against the resident scene. */
int dl_service_bake(dl_service* svc, const dl_light* lights, uint32_t light_count, dl_baked_scene** out);
/* Serves requests as they arrive: each is admitted, in arrival order, by the first of `workers`
 * (1-16) worker streams that is free, so requests that arrive while others run overlap on the
 * device.  latency_ms[i] runs from the request's arrival to its result being readable on the host,
 * service_ms[i] (may be NULL) from its admission; out (may be NULL) receives the results, which
 * equal dl_service_bake of each request alone. */
int dl_service_replay(dl_service* svc, const dl_service_request* requests, uint32_t count, uint32_t workers,
                      double* latency_ms, double* service_ms, int* status, dl_baked_scene** out);
void dl_service_destroy(dl_service* svc);

/* The live scene of a service: the resident meshes baked with its current light set (the lights
 * it was created with, then every edit applied).  An edit adds lights, replaces lights by id (a
 * moved or changed light) and removes lights; only local lights (point, rectangle, tube) are
 * edited.  Applying it traces only the changed lights' (triangle, light) pairs against the
 * resident geometry, pads and crops again the meshes whose pairs changed (seam padding reads the
 * mesh-wide visited grid every light sets, mesh_bake.rs:246, seams.rs:15-39), and writes every
 * mesh's words with the new light slots (baker.rs:35-50); afterwards the live scene equals dl_bake
 * of the resident meshes with the edited light set (words, lights, light BVH, distance cubes,
 * stats).  No counterpart in the reference, which re-runs bake() for every change. */
typedef struct dl_service_edit {
    const dl_light* lights;   /* lights to add, or to replace (same id)                        */
    uint32_t light_count;
    const uint64_t* remove;   /* ids of lights to remove                                       */
    uint32_t remove_count;
    double arrival_s;         /* for dl_service_replay_edits: arrival in seconds (sorted)      */
    /* Non-NULL (and no lights or removals): the edit adds this room, a scene of its own (local
     * lights only, mesh ids above every live one) that joins the live scene as a part of its own.
     * The result is the bake of the whole scene: the room is refused (DL_E_ARGUMENT) unless no
     * light of one part reaches the other's bounds and no burial ray that leaves its own part
     * hits the other.  Light edits then trace each light against the part it reaches (a light
     * that reaches two parts is refused). */
    const dl_bake_input* room;
} dl_service_edit;
typedef struct dl_service_update {
    uint64_t version;         /* edits applied to the live scene so far                        */
    uint32_t dirty_meshes;    /* meshes padded and cropped again                               */
    uint32_t reserved;
    uint64_t traced_pairs;    /* (mesh, light) pairs traced for the edit                       */
    double trace_ms, assemble_ms, total_ms;
    uint64_t traced_samples;  /* texel-light samples, shadow rays and bounce rays the edit's     */
    uint64_t traced_rays;     /* tracing cast (the dl_bake_stats of those pairs)               */
    uint64_t traced_bounce_rays;
    double queue_ms;          /* dl_service_replay_edits: arrival to the start of its tracing  */
} dl_service_update;
int dl_service_apply(dl_service* svc, const dl_service_edit* edit, dl_service_update* info);
/* Serves edits as they arrive, with continuous batching: each edit's changed lights are traced on
 * the first of `workers` (1-16) worker streams that is free, admitted in arrival order while the
 * device memory the edits being traced may hold fits `budget_bytes` (0: half the free device
 * memory), so edits that arrive while others run overlap on the device.  Edits join the live scene
 * in arrival order; those whose tracing has ended by the time the scene is free join together
 * (up to 16), in one assembly that gives the scene of applying them one by one.
 * latency_ms[i] runs from edit i's arrival to the live scene holding it.  checksum (may be NULL)
 * receives, for the last edit of each joint assembly, a fingerprint of the live words, stats and
 * light order (the same as after that edit applied alone), and 0 for the others.  info (may be
 * NULL) receives each edit's update record (dirty_meshes and assemble_ms are those of its joint
 * assembly; `reserved` is the number of edits in it). */
int dl_service_replay_edits(dl_service* svc, const dl_service_edit* edits, uint32_t count, uint32_t workers,
                            uint64_t budget_bytes, double* latency_ms, uint64_t* checksum, dl_service_update* info);
/* A copy of the live scene on the host (downloads its words). */
int dl_service_scene(dl_service* svc, dl_baked_scene** out);
/* The live scene's triangle words on the device, every mesh's words back to back in mesh order
 * (the `triangles` buffer of BakedBuffers, layout.rs:148-169); valid until the next edit. */
int dl_service_words(dl_service* svc, const uint32_t** device_words, uint64_t* count);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Answers `ray_count`
 * rays against the occluder BVH resident in `svc`, like dl_test_ray_queries against the same
 * triangles: rays8 holds 8 floats per ray (origin, direction, max distance, unused); closest = 0
 * sets hit_or_index[i] to 1 for a hit, closest = 1 sets t_out[i] to the nearest distance (-1 on a
 * miss) and hit_or_index[i] to its occluder index (lowest on ties).  rays8, t_out and
 * hit_or_index are device memory of the service's GPU; *hits (may be NULL) receives the number
 * of rays that hit.  Allocates nothing: the kernels' queue head, hit count and overflow flag live
 * in a fixed 256-byte device scratch per GPU, kept until the process ends and shared by every
 * call on that GPU, so calls on one GPU from different contexts must not overlap (they race).
 * At most DL_NOALLOC_MAX_RAYS rays per call. */
#define DL_NOALLOC_MAX_RAYS (1ull << 30)
int dl_service_ray_queries_noalloc(dl_service* svc, const float* rays8, uint64_t ray_count, int closest, float* t_out,
                                   uint32_t* hit_or_index, uint64_t* hits);

/* Many independent bakes (no counterpart in the reference, which bakes one scene per call):
 * job i bakes inputs[i] with options[i] into out[i] (status[i] is its dl_bake result).  Jobs are
 * handed out in order to `streams` worker contexts (1-16), each baking on its own HIP stream, so
 * small jobs overlap on the device; each result equals dl_bake of the same job alone.  Returns
 * DL_OK, or the first failing job's code with its message in dl_last_error. */
int dl_bake_batch(dl_ctx* ctx, const dl_bake_input* inputs, const dl_bake_options* options, uint32_t count,
                  uint32_t streams, dl_baked_scene** out, int* status);

/* Launch settings of a batch, picked by dl_autotune_batch. */
typedef struct dl_tuned_launch {
    uint32_t streams;       /* worker streams to pass to dl_bake_batch                          */
    uint32_t trace_block;   /* threads per block of the shadow-ray kernel                       */
    uint32_t stream_block;  /* threads per block of the persistent ray-stream kernel            */
    uint32_t packet;        /* rays a wave of the ray-stream kernel takes from its queue at once */
    uint32_t burial_block;  /* threads per block of the burial test's ray stream                */
    uint32_t burial_lanes;  /* 0: burial rays streamed; else one lane per texel, this block size */
    uint32_t trace_stream;  /* 1: area, soft-sun and sky rays streamed; 0: per lane             */
    uint32_t trials;        /* settings timed (0 when read from the cache file)                 */
    double ms;              /* the sample's batch time with these settings                      */
    uint32_t from_cache;    /* 1: the settings came from cache_path                             */
} dl_tuned_launch;

/* Autotuner of dl_bake_batch's launch shapes (no counterpart in the reference): times
 * dl_bake_batch on a deterministic sample of the jobs (up to 32, spread over the jobs' cost order)
 * for candidate worker-stream counts (up to max_streams), shadow-ray and ray-stream block sizes,
 * ray packet widths, burial block sizes and the forms of the burial test and of multi-ray lights
 * (streamed or one lane per texel), one parameter at a time from the current settings, and keeps
 * the fastest.  The result is applied to ctx (block sizes and packet) and returned;
 * `streams` is for the caller's dl_bake_batch.  With cache_path, settings are looked up there by a
 * signature of the batch (job count, sizes and options), the device and GPU_MAX_HW_QUEUES, and
 * stored there after tuning.  The stack depth of the traversal kernels is not tuned: it is the
 * BVH's need (3 entries per level), below which traversal would fail. */
int dl_autotune_batch(dl_ctx* ctx, const dl_bake_input* inputs, const dl_bake_options* options, uint32_t count,
                      uint32_t max_streams, const char* cache_path, dl_tuned_launch* out);

/* ---- BakedScene accessors (src/bake/format/types.rs:13-37) --------------------------------- */
void dl_scene_settings(const dl_baked_scene* scene, dl_baked_settings* out);
void dl_scene_stats(const dl_baked_scene* scene, dl_bake_stats* out);
uint32_t dl_scene_light_count(const dl_baked_scene* scene);
int dl_scene_light(const dl_baked_scene* scene, uint32_t index, dl_baked_light* out);
uint32_t dl_scene_light_bvh(const dl_baked_scene* scene, const dl_light_bvh_node** nodes);
uint32_t dl_scene_mesh_count(const dl_baked_scene* scene);
int dl_scene_mesh(const dl_baked_scene* scene, uint32_t index, dl_baked_mesh_view* out);
/* Views of meshes first .. first + count - 1 (those that exist) in one call; returns how many
 * were written.  Same views as dl_scene_mesh, for callers that read every mesh. */
uint32_t dl_scene_meshes(const dl_baked_scene* scene, uint32_t first, uint32_t count, dl_baked_mesh_view* out);
uint64_t dl_scene_distance_cubes(const dl_baked_scene* scene, const uint32_t** words);
uint32_t dl_scene_sun_map_count(const dl_baked_scene* scene);
int dl_scene_sun_map(const dl_baked_scene* scene, uint32_t index, dl_sun_map_view* out);
void dl_scene_sky_probes(const dl_baked_scene* scene, dl_sky_probes_view* out);

/* ---- timing of the last dl_bake on a context (milliseconds, device events) ----------------- */
typedef struct dl_timings {
    double total_ms;      /* wall time of dl_bake                                          */
    double upload_ms;     /* host-to-device copies                                         */
    double texels_ms;     /* owned/touched texel enumeration and burial                    */
    double trace_ms;      /* shadow rays (the trace pass)                                  */
    double bounce_ms;     /* photon cubes and gather                                       */
    double post_ms;       /* optimize, seams, crop, dilate, serialize                      */
    double maps_ms;       /* distance cubes, sun maps, sky probes                          */
    double download_ms;   /* device-to-host copies                                         */
    double trace_kernel_ms; /* summed device time of the shadow-ray kernel launches        */
    uint64_t trace_launches;
    uint64_t trace_rays;  /* rays cast by the shadow-ray kernel                             */
    uint64_t bvh_nodes;   /* occluder BVH node count                                       */
    uint64_t occluder_triangles;
    uint64_t trace_node_visits; /* BVH nodes fetched by the shadow-ray kernel (counting on)  */
    uint64_t trace_tri_tests;   /* triangles tested by the shadow-ray kernel (counting on)  */
    double bounce_kernel_ms;    /* device time of the gather passes, one span per mesh batch */
    uint64_t bounce_node_visits;/* BVH nodes fetched by the gather's visibility rays (counting) */
    uint64_t bounce_tri_tests;  /* triangles tested by the gather's visibility rays (counting) */
    uint64_t bounce_launches;   /* mesh batches with bounce work                              */
    double gather_trace_ms;     /* summed device time of the any-hit gather kernel launches   */
    uint64_t gather_trace_launches;
    double prepare_ms;          /* host: validation, occluders, BVH build, receiver triangles */
    double bvh_ms;              /* host: the occluder BVH build alone                         */
    double gather_rays_ms;      /* summed device time of the gather's sample-setup launches     */
    uint64_t gather_rays_launches;
    uint64_t gather_jobs;       /* traced (texel, light) gather jobs (counting on)            */
    uint64_t gather_samples;    /* samples those jobs set up (counting on)                    */
    uint64_t gather_traced;     /* samples with a ray to trace (counting on)                  */
    double burial_ms;           /* device time of the burial test (rays, traversal, tally)     */
    uint64_t burial_rays;       /* closest-hit rays cast by the burial test (counting on)      */
    uint64_t burial_node_visits;/* BVH nodes fetched by those rays (counting on)               */
    uint64_t burial_tri_tests;  /* triangles tested by those rays (counting on)                */
    uint64_t gather_photon_reads; /* photon records the traced gather jobs read (counting on)  */
    uint64_t gather_open;       /* samples the split setup's fast pass left to the exact one (counting on) */
    uint64_t photon_texels;     /* photon cube texels traced (all of every cube, or the marked ones of lazy cubes) */
} dl_timings;
int dl_last_timings(const dl_ctx* ctx, dl_timings* out);

#ifdef __cplusplus
}
#endif
#endif /* DYNLIGHT_H */

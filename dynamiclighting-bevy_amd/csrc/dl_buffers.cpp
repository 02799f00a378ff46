// BakedBuffers (src/baked_lighting/layout.rs:1-215, apply.rs:27-98): the two word buffers the
// runtime's shaders read, built from a baked scene.
//
//   triangles  every baked mesh's triangle words, back to back (one 0 word when there are none)
//   aux        a 12-word header, a 4-word record per receiver [triangles_base, resolution,
//              triangle_count, 0], the light BVH nodes (8 words each), the distance cubes, a
//              16-word record per sun map (its three rows as f32 bits, resolution, first texel
//              word, 0, 0), the sun map texels, then the sky probe record (origin, spacing as f32
//              bits, counts, first probe word) and the probes.
//
// Receiver n (tag n + 1) is the scene's n-th mesh; a mesh's triangle count is its first word / 4
// (the light data offset of its first triangle).
#include <cstring>
#include <new>
#include <vector>

#include "dl_host.h"

struct dl_baked_buffers {
  std::vector<uint32_t> triangles, aux;
  std::vector<dl_receiver> receivers;
};

namespace {
// aux header word indices (layout.rs:30-52) and flag bits (layout.rs:55-63)
enum : size_t {
  kReceivers = 0, kReceiverCount = 1, kBvh = 2, kBvhNodes = 3, kCubes = 4, kCubeCount = 5, kFlags = 6,
  kBakedLights = 7, kSunMaps = 8, kSunMapCount = 9, kSkyProbes = 10, kHeaderWords = 12
};
constexpr uint32_t kBounceLayout = 1, kCubeBias = 2, kCubeAngular = 4;

uint32_t bits_of(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  return u;
}

// BakedScene::local_light_count (format/types.rs:48-59): the lights before the trailing suns and skies
uint32_t local_lights(const dl_baked_scene& s) {
  size_t n = s.lights.size();
  while (n > 0 && (s.lights[n - 1].shape == DL_SHAPE_SUN || s.lights[n - 1].shape == DL_SHAPE_SKY)) --n;
  return (uint32_t)n;
}
}  // namespace

namespace {
// scene_buffers / take_scene_buffers (apply.rs:27-61): with `take`, each mesh's triangle words
// are released as soon as they are copied, so the scene and the buffers are never both whole.
int build_buffers(dl_baked_scene* scene, bool take, dl_baked_buffers** out) {
  if (!scene || !out) {
    dl_set_thread_error("null argument");
    return DL_E_ARGUMENT;
  }
  *out = nullptr;
  if (scene->triangles_taken) {
    dl_set_thread_error(take ? "the scene's triangle words were already taken"
                             : "scene_buffers needs a full scene, but this one's triangle words were taken; use the "
                               "buffers taken from it");
    return DL_E_TRIANGLES_TAKEN;
  }
  try {
    dl_baked_scene& s = *scene;
    std::unique_ptr<dl_baked_buffers> b(new dl_baked_buffers());
    // BufferBuilder::push (layout.rs:148-169)
    size_t words = 0;
    for (const auto& m : s.meshes) words += m.words.size();
    b->triangles.reserve(words ? words : 1);
    std::vector<uint32_t> records;
    records.reserve(s.meshes.size() * 4);
    for (auto& m : s.meshes) {
      const uint32_t triangle_count = m.words.empty() ? 0u : m.words[0] / 4;
      dl_receiver r{};
      r.id = m.id;
      r.content_hash = m.content_hash;
      r.tag = (uint32_t)b->receivers.size() + 1;
      r.triangle_count = triangle_count;
      r.occludes = m.occludes ? 1u : 0u;
      b->receivers.push_back(r);
      records.insert(records.end(), {(uint32_t)b->triangles.size(), m.resolution, triangle_count, 0u});
      b->triangles.insert(b->triangles.end(), m.words.begin(), m.words.end());
      if (take) m.words = dl_word_slice{};
    }
    if (take) s.triangles_taken = true;
    // BufferBuilder::finish (layout.rs:171-214) with scene_input (apply.rs:73-97)
    if (b->triangles.empty()) b->triangles.push_back(0);
    std::vector<uint32_t>& aux = b->aux;
    aux.assign(kHeaderWords, 0u);
    aux[kReceivers] = (uint32_t)aux.size();
    aux[kReceiverCount] = (uint32_t)(records.size() / 4);
    aux.insert(aux.end(), records.begin(), records.end());
    aux[kBvh] = (uint32_t)aux.size();
    aux[kBvhNodes] = (uint32_t)s.bvh.size();
    const uint32_t* bvh = reinterpret_cast<const uint32_t*>(s.bvh.data());
    aux.insert(aux.end(), bvh, bvh + s.bvh.size() * 8);
    aux[kCubes] = (uint32_t)aux.size();
    aux[kCubeCount] = s.distance_cubes.empty() ? 0u : local_lights(s);
    aux.insert(aux.end(), s.distance_cubes.begin(), s.distance_cubes.end());
    uint32_t flags = s.settings.bounce ? kBounceLayout : 0u;
    if (s.settings.dynamic_geometry == DL_DYNAMIC_DISTANCE_CUBES) flags |= kCubeBias;
    if (s.settings.dynamic_geometry == DL_DYNAMIC_ANGULAR) flags |= kCubeAngular;
    aux[kFlags] = flags;
    aux[kBakedLights] = (uint32_t)s.lights.size();
    aux[kSunMaps] = (uint32_t)aux.size();
    aux[kSunMapCount] = (uint32_t)s.sun_maps.size();
    size_t texels = aux.size() + s.sun_maps.size() * 16;
    for (const auto& m : s.sun_maps) {
      for (float v : m.rows) aux.push_back(bits_of(v));
      aux.insert(aux.end(), {m.resolution, (uint32_t)texels, 0u, 0u});
      texels += m.texels.size();
    }
    for (const auto& m : s.sun_maps) aux.insert(aux.end(), m.texels.begin(), m.texels.end());
    if (s.probes) {
      aux[kSkyProbes] = (uint32_t)aux.size();
      for (float v : s.probe_origin) aux.push_back(bits_of(v));
      aux.push_back(bits_of(s.probe_spacing));
      for (uint32_t v : s.probe_counts) aux.push_back(v);
      aux.push_back((uint32_t)aux.size() + 1);
      aux.insert(aux.end(), s.probe_words.begin(), s.probe_words.end());
    }
    *out = b.release();
    return DL_OK;
  } catch (const std::bad_alloc&) {
    dl_set_thread_error("host allocation failed");
    return DL_E_OOM;
  }
}
}  // namespace

extern "C" {

int dl_scene_buffers(const dl_baked_scene* scene, dl_baked_buffers** out) {
  return build_buffers(const_cast<dl_baked_scene*>(scene), false, out);
}

int dl_take_scene_buffers(dl_baked_scene* scene, dl_baked_buffers** out) { return build_buffers(scene, true, out); }

int dl_scene_triangles_taken(const dl_baked_scene* scene) { return scene && scene->triangles_taken ? 1 : 0; }

uint64_t dl_buffers_triangles(const dl_baked_buffers* b, const uint32_t** words) {
  if (!b) return 0;
  if (words) *words = b->triangles.data();
  return b->triangles.size();
}

uint64_t dl_buffers_aux(const dl_baked_buffers* b, const uint32_t** words) {
  if (!b) return 0;
  if (words) *words = b->aux.data();
  return b->aux.size();
}

uint32_t dl_buffers_receiver_count(const dl_baked_buffers* b) { return b ? (uint32_t)b->receivers.size() : 0u; }

int dl_buffers_receiver(const dl_baked_buffers* b, uint32_t index, dl_receiver* out) {
  if (!b || !out || index >= b->receivers.size()) return DL_E_ARGUMENT;
  *out = b->receivers[index];
  return DL_OK;
}

void dl_baked_buffers_free(dl_baked_buffers* b) { delete b; }

}  // extern "C"

// The ECS caller's side of a bake, restated for callers outside Bevy: what SceneSnapshot::take
// (src/bake/snapshot.rs:83-158) does to each entity before it becomes a BakeMesh or BakeLight.
//
//   * dl_mesh_content_hash   mesh_content_hash (src/shared/mesh_hash.rs:16-33, WordHash 83-111)
//   * dl_bake_id_from_path   BakeId::from_path (src/shared/bake_id.rs:15-23, FNV-1a 64)
//   * dl_assign_ids          assign_ids (src/bake/identity.rs:64-101): explicit ids kept, derived
//                            ids that collide get a "#k" segment in tiebreak order
//   * dl_mesh_to_world       MeshData::to_world (src/bake/snapshot.rs:229-272): affine positions,
//                            inverse-transpose normals, winding swap for mirroring transforms,
//                            the content hash of the local data
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dl_host.h"

namespace {

// WordHash (mesh_hash.rs:83-111): a word-at-a-time multiplicative hash, SplitMix64 finalizer.
struct WordHash {
  uint64_t v = 0x6479'6e6c'6967'6874ull;  // "dynlight"
  void write(uint32_t w) {
    v = (((v << 5) | (v >> 59)) ^ (uint64_t)w) * 0x517c'c1b7'2722'0a95ull;
  }
  void write_len(uint64_t n) {
    write((uint32_t)n);
    write((uint32_t)(n >> 32));
  }
  uint64_t finish() const {
    uint64_t z = v;
    z = (z ^ (z >> 30)) * 0xbf58'476d'1ce4'e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d0'49bb'1331'11ebull;
    return z ^ (z >> 31);
  }
};

uint32_t bits_of(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// FNV-1a 64 over length-prefixed segments (bake_id.rs:15-23, 36-55).
uint64_t path_id(const std::vector<std::string>& segments) {
  uint64_t h = 0xcbf2'9ce4'8422'2325ull;
  auto put = [&](const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 0x0000'0100'0000'01b3ull;
    }
  };
  for (const std::string& s : segments) {
    const uint64_t len = s.size();
    uint8_t le[8];
    for (int i = 0; i < 8; ++i) le[i] = (uint8_t)(len >> (8 * i));
    put(le, 8);
    put(s.data(), s.size());
  }
  return h;
}

struct F3 {
  float x, y, z;
};
F3 f3(float x, float y, float z) { return {x, y, z}; }
F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
F3 cross(F3 a, F3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
// Mat3 * Vec3 with columns x, y, z: x * v.x + y * v.y + z * v.z
F3 mat_mul(const F3* cols, F3 v) { return add(add(mul(cols[0], v.x), mul(cols[1], v.y)), mul(cols[2], v.z)); }
// Vec3::normalize_or_zero: v * (1 / length) when that reciprocal is finite and positive
F3 normalize_or_zero(F3 v) {
  const float rcp = 1.0f / std::sqrt(dot(v, v));
  return (std::isfinite(rcp) && rcp > 0.0f) ? mul(v, rcp) : f3(0.0f, 0.0f, 0.0f);
}
// f32::is_normal: neither zero, subnormal, infinite nor NaN
bool is_normal(float v) { return std::isnormal(v); }

}  // namespace

extern "C" {

uint64_t dl_mesh_content_hash(const float* positions, uint32_t vertex_count, const uint32_t* indices, uint64_t index_count,
                              const float* uv, uint32_t uv_count) {
  constexpr uint32_t kMissing = 0xffffffffu;
  WordHash h;
  h.write_len(index_count);
  for (uint64_t i = 0; i < index_count; ++i) {
    const uint32_t v = indices[i];
    if (v < vertex_count)
      for (int a = 0; a < 3; ++a) h.write(bits_of(positions[3 * (size_t)v + a]));
    else
      h.write(kMissing);
    if (v < uv_count)
      for (int a = 0; a < 2; ++a) h.write(bits_of(uv[2 * (size_t)v + a]));
    else
      h.write(kMissing);
  }
  return h.finish();
}

uint64_t dl_bake_id_from_path(const char* const* segments, const uint64_t* lengths, uint32_t count) {
  std::vector<std::string> s;
  for (uint32_t i = 0; i < count; ++i) s.emplace_back(segments[i], segments[i] + lengths[i]);
  return path_id(s);
}

int dl_assign_ids(const dl_id_candidate* candidates, uint32_t count, uint64_t* out) {
  if (count && (!candidates || !out)) {
    dl_set_thread_error("null argument");
    return DL_E_ARGUMENT;
  }
  try {
    auto segments_of = [&](uint32_t i) {
      std::vector<std::string> s;
      const dl_id_candidate& c = candidates[i];
      for (uint32_t k = 0; k < c.segment_count; ++k) s.emplace_back(c.segments[k], c.segments[k] + c.segment_lengths[k]);
      return s;
    };
    std::map<uint64_t, std::vector<uint32_t>> groups;  // ordered by id, as a BTreeMap
    for (uint32_t i = 0; i < count; ++i) {
      out[i] = candidates[i].has_explicit ? candidates[i].explicit_id : path_id(segments_of(i));
      if (!candidates[i].has_explicit) groups[out[i]].push_back(i);
    }
    for (auto& g : groups) {
      std::vector<uint32_t>& group = g.second;
      if (group.size() < 2) continue;
      // tiebreak: content hash, then the position's bit patterns (identity.rs:21-22, 58-60)
      auto key = [&](uint32_t i) {
        const dl_id_candidate& c = candidates[i];
        return std::make_tuple(c.tiebreak_hash, bits_of(c.position[0]), bits_of(c.position[1]), bits_of(c.position[2]));
      };
      std::stable_sort(group.begin(), group.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
      for (size_t k = 1; k < group.size(); ++k) {
        std::vector<std::string> s = segments_of(group[k]);
        s.push_back("#" + std::to_string(k));
        out[group[k]] = path_id(s);
      }
    }
    return DL_OK;
  } catch (const std::bad_alloc&) {
    dl_set_thread_error("host allocation failed");
    return DL_E_OOM;
  }
}

int dl_mesh_to_world(const float* affine, const float* positions, uint32_t vertex_count, const float* normals,
                     const uint32_t* indices, uint64_t index_count, const float* uv1, uint32_t uv_count, float* positions_out,
                     float* normals_out, uint32_t* indices_out, uint64_t* content_hash) {
  if (!affine || (vertex_count && (!positions || !positions_out)) || (index_count && (!indices || !indices_out))) {
    dl_set_thread_error("null argument");
    return DL_E_ARGUMENT;
  }
  if (index_count % 3) {
    dl_set_thread_error("the index count is not a multiple of 3");
    return DL_E_ARGUMENT;
  }
  // columns of the linear part, then the translation (Affine3A: matrix3, translation)
  const F3 cols[3] = {f3(affine[0], affine[1], affine[2]), f3(affine[3], affine[4], affine[5]), f3(affine[6], affine[7], affine[8])};
  const F3 t = f3(affine[9], affine[10], affine[11]);
  // Mat3::determinant: z . (x cross y)
  const float det = dot(cols[2], cross(cols[0], cols[1]));
  for (uint32_t v = 0; v < vertex_count; ++v) {  // Affine3A::transform_point3
    const F3 p = add(mat_mul(cols, f3(positions[3 * (size_t)v], positions[3 * (size_t)v + 1], positions[3 * (size_t)v + 2])), t);
    positions_out[3 * (size_t)v] = p.x;
    positions_out[3 * (size_t)v + 1] = p.y;
    positions_out[3 * (size_t)v + 2] = p.z;
  }
  int wrote_normals = 0;
  if (normals && normals_out && is_normal(det)) {
    // Mat3::inverse: rows of the inverse are the cross products over the determinant; the normal
    // matrix is its transpose, whose columns are those cross products
    const float inv = 1.0f / det;
    const F3 n[3] = {mul(cross(cols[1], cols[2]), inv), mul(cross(cols[2], cols[0]), inv), mul(cross(cols[0], cols[1]), inv)};
    for (uint32_t v = 0; v < vertex_count; ++v) {
      const F3 q = normalize_or_zero(mat_mul(n, f3(normals[3 * (size_t)v], normals[3 * (size_t)v + 1], normals[3 * (size_t)v + 2])));
      normals_out[3 * (size_t)v] = q.x;
      normals_out[3 * (size_t)v + 1] = q.y;
      normals_out[3 * (size_t)v + 2] = q.z;
    }
    wrote_normals = 1;
  }
  for (uint64_t i = 0; i < index_count; i += 3) {  // mirroring transforms swap corners 1 and 2
    indices_out[i] = indices[i];
    indices_out[i + 1] = det < 0.0f ? indices[i + 2] : indices[i + 1];
    indices_out[i + 2] = det < 0.0f ? indices[i + 1] : indices[i + 2];
  }
  if (content_hash) *content_hash = dl_mesh_content_hash(positions, vertex_count, indices, index_count, uv1, uv_count);
  return wrote_normals;
}

}  // extern "C"

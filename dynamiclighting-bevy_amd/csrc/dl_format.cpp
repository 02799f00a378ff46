// The `.dynlight` file of a baked scene (src/bake/format/mod.rs:1-21): a 24-byte header (magic
// "DYNLIGHT", format version 6, payload length u64, CRC-32 of the payload) followed by the
// payload as one gzip stream.  The payload is little-endian with u64 counts; its sections, in
// order, are settings, lights (60 bytes each), light BVH words, meshes, distance cube words, sun
// maps, sky probes and stats.  Encoding is deterministic: the same scene gives the same bytes
// (gzip with mtime 0, OS byte 255 and zlib's level 6).  Decoding checks everything the
// reference's decoder checks and fails with the same error kinds and messages.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "dl_host.h"

namespace {

constexpr uint8_t kMagic[8] = {'D', 'Y', 'N', 'L', 'I', 'G', 'H', 'T'};
constexpr uint32_t kVersion = 6;            // FORMAT_VERSION (format/mod.rs:46)
constexpr size_t kHeaderBytes = 24;
constexpr uint32_t kMaxSunMapSide = 8192;   // MAX_SUN_MAP_RESOLUTION (format/mod.rs:49)
constexpr uint64_t kMaxSkyProbes = 65536;   // MAX_SKY_PROBES (sky_probes.rs:19)
constexpr size_t kLightBytes = 60;          // lights.rs:17
constexpr size_t kSunMapBytes = 60;         // maps.rs:14

thread_local int g_format_kind = 0;

// FormatError (format/mod.rs:53-77) as an exception inside this file.
struct FormatError {
  int kind;
  std::string message;
};
[[noreturn]] void fail(int kind, const std::string& detail = std::string()) {
  switch (kind) {
    case DL_FORMAT_NOT_DYNLIGHT: throw FormatError{kind, "not a .dynlight file (bad magic)"};
    case DL_FORMAT_TRUNCATED: throw FormatError{kind, ".dynlight file is truncated or has corrupt lengths"};
    case DL_FORMAT_DECOMPRESS: throw FormatError{kind, ".dynlight payload doesn't decompress: " + detail};
    case DL_FORMAT_CHECKSUM: throw FormatError{kind, ".dynlight payload checksum mismatch"};
    case DL_FORMAT_INVALID: throw FormatError{kind, ".dynlight file has an invalid " + detail};
    default: throw FormatError{kind, detail};
  }
}
[[noreturn]] void unsupported_version(uint32_t found) {
  throw FormatError{DL_FORMAT_UNSUPPORTED_VERSION, ".dynlight format version " + std::to_string(found) +
                                                       " is not supported (this build reads version " +
                                                       std::to_string(kVersion) + "); re-bake the scene"};
}

// ---- payload writer (codec.rs:7-36) ------------------------------------------------------------
struct Writer {
  std::vector<uint8_t> b;
  void u32(uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void u64(uint64_t v) {
    for (int i = 0; i < 8; ++i) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void f32(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u32(u);
  }
  void words(const uint32_t* w, size_t n) {
    u64(n);
    const size_t at = b.size();
    b.resize(at + 4 * n);
    for (size_t i = 0; i < n; ++i)
      for (int k = 0; k < 4; ++k) b[at + 4 * i + k] = (uint8_t)(w[i] >> (8 * k));
  }
};

// ---- byte sources and the payload reader (codec.rs:38-140) --------------------------------------
struct Source {
  virtual ~Source() = default;
  // Reads up to `cap` bytes; 0 at the end.  Fails with DL_FORMAT_DECOMPRESS on a source error.
  virtual size_t read(uint8_t* buf, size_t cap) = 0;
};

struct MemorySource : Source {
  const uint8_t* p;
  size_t n;
  MemorySource(const uint8_t* p_, size_t n_) : p(p_), n(n_) {}
  size_t read(uint8_t* buf, size_t cap) override {
    const size_t k = cap < n ? cap : n;
    if (k) std::memcpy(buf, p, k);
    p += k;
    n -= k;
    return k;
  }
};

struct CallbackSource : Source {
  dl_read_fn fn;
  void* user;
  CallbackSource(dl_read_fn f, void* u) : fn(f), user(u) {}
  size_t read(uint8_t* buf, size_t cap) override {
    const int64_t got = fn(user, buf, (uint64_t)cap);
    if (got < 0) fail(DL_FORMAT_DECOMPRESS, "the source failed to read");
    return (size_t)got;
  }
};

// Inflates a gzip stream from `src`, hashing and counting the bytes it hands out.
struct Inflater : Source {
  Source& src;
  z_stream zs{};
  std::vector<uint8_t> in;
  bool open = false, ended = false;
  uLong crc = crc32(0L, Z_NULL, 0);
  uint64_t count = 0;
  explicit Inflater(Source& s) : src(s), in(1 << 16) {
    if (inflateInit2(&zs, 15 + 16) != Z_OK) fail(DL_FORMAT_DECOMPRESS, "inflate could not start");
    open = true;
  }
  ~Inflater() override {
    if (open) inflateEnd(&zs);
  }
  size_t read(uint8_t* buf, size_t cap) override {
    if (ended || cap == 0) return 0;
    zs.next_out = buf;
    zs.avail_out = (uInt)cap;
    while (zs.avail_out == cap) {
      if (zs.avail_in == 0) {
        const size_t got = src.read(in.data(), in.size());
        if (got == 0) fail(DL_FORMAT_DECOMPRESS, "unexpected end of the gzip stream");
        zs.next_in = in.data();
        zs.avail_in = (uInt)got;
      }
      const int r = inflate(&zs, Z_NO_FLUSH);
      if (r == Z_STREAM_END) {
        ended = true;
        break;
      }
      if (r != Z_OK && r != Z_BUF_ERROR) fail(DL_FORMAT_DECOMPRESS, zs.msg ? zs.msg : "corrupt deflate stream");
    }
    const size_t n = cap - zs.avail_out;
    crc = crc32(crc, buf, (uInt)n);
    count += n;
    return n;
  }
};

struct Reader {
  Source& src;
  uint64_t remaining;
  Reader(Source& s, uint64_t len) : src(s), remaining(len) {}
  void fill(uint8_t* buf, size_t n) {
    if (n > remaining) fail(DL_FORMAT_TRUNCATED);
    size_t have = 0;
    while (have < n) {
      const size_t got = src.read(buf + have, n - have);
      if (got == 0) fail(DL_FORMAT_TRUNCATED);
      have += got;
    }
    remaining -= n;
  }
  uint32_t u32() {
    uint8_t b[4];
    fill(b, 4);
    return (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24;
  }
  uint64_t u64() {
    const uint64_t lo = u32();
    return lo | (uint64_t)u32() << 32;
  }
  float f32() {
    const uint32_t u = u32();
    float v;
    std::memcpy(&v, &u, 4);
    return v;
  }
  // A count of items of at least `item_bytes` each, no more than the rest of the payload holds.
  uint64_t count(uint64_t item_bytes) {
    const uint64_t n = u64();
    const uint64_t per = item_bytes ? item_bytes : 1;
    if (n > remaining / per) fail(DL_FORMAT_TRUNCATED);
    return n;
  }
  std::vector<uint32_t> words() {
    const uint64_t n = count(4);
    std::vector<uint32_t> w;
    w.reserve((size_t)std::min<uint64_t>(n, 1ull << 26));
    uint8_t chunk[4096 * 4];
    while (w.size() < n) {
      const size_t take = (size_t)std::min<uint64_t>(n - w.size(), 4096);
      fill(chunk, take * 4);
      for (size_t i = 0; i < take; ++i)
        w.push_back((uint32_t)chunk[4 * i] | (uint32_t)chunk[4 * i + 1] << 8 | (uint32_t)chunk[4 * i + 2] << 16 |
                    (uint32_t)chunk[4 * i + 3] << 24);
    }
    return w;
  }
};

// ---- enum mappings ------------------------------------------------------------------------------
uint32_t bounce_bits_of(int32_t c) {  // BounceCompression::bits_per_pixel, 0 for none
  switch (c) {
    case DL_BOUNCE_BITS8: return 8;
    case DL_BOUNCE_BITS6: return 6;
    case DL_BOUNCE_BITS5: return 5;
    case DL_BOUNCE_BITS4: return 4;
    default: return 0;
  }
}
uint32_t file_dynamic(int32_t d) {  // 0 LightingOnly, 1 DistanceCubes, 2 Angular (format/mod.rs:182-186)
  return d == DL_DYNAMIC_DISTANCE_CUBES ? 1u : d == DL_DYNAMIC_ANGULAR ? 2u : 0u;
}
int group_of(int32_t shape) { return shape == DL_SHAPE_SUN ? 1 : shape == DL_SHAPE_SKY ? 2 : 0; }
bool finite4(const float* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

// ---- encode_payload (format/mod.rs:173-212, lights.rs:27-53, maps.rs:16-38) ---------------------
std::vector<uint8_t> encode_payload(const dl_baked_scene& s) {
  Writer w;
  w.u32((s.settings.bounce ? 1u : 0u) | (s.settings.colored_bounce ? 2u : 0u));
  w.u32(file_dynamic(s.settings.dynamic_geometry));
  w.f32(s.settings.pixel_density);
  w.u32(s.settings.max_lightmap_size);
  w.u32(s.settings.distance_cube_resolution);

  w.u64(s.lights.size());
  for (const dl_baked_light& l : s.lights) {
    w.u64(l.id);
    for (int k = 0; k < 3; ++k) w.f32(l.origin[k]);
    w.f32(l.radius);
    w.u32(bounce_bits_of(l.bounce_compression));
    float rot[4] = {0.0f, 0.0f, 0.0f, 1.0f}, size[2] = {0.0f, 0.0f};
    if (l.shape == DL_SHAPE_RECTANGLE || l.shape == DL_SHAPE_TUBE) {
      std::memcpy(rot, l.rotation, 16);
      std::memcpy(size, l.size, 8);
    } else if (l.shape == DL_SHAPE_SUN) {
      std::memcpy(rot, l.rotation, 16);
      size[0] = l.angular_diameter;
    }
    w.u32((uint32_t)l.shape);
    for (float v : rot) w.f32(v);
    for (float v : size) w.f32(v);
    w.u32(l.shadow_bits_per_texel);
  }

  w.words(reinterpret_cast<const uint32_t*>(s.bvh.data()), s.bvh.size() * 8);

  w.u64(s.meshes.size());
  for (const dl_baked_mesh_out& m : s.meshes) {
    w.u64(m.id);
    w.u64(m.content_hash);
    w.u32(m.resolution);
    w.u32(m.occludes ? 1u : 0u);
    w.words(m.words.data(), m.words.size());
  }

  w.words(s.distance_cubes.data(), s.distance_cubes.size());

  w.u64(s.sun_maps.size());
  for (const dl_sun_map_out& m : s.sun_maps) {
    w.u32(m.resolution);
    for (float v : m.rows) w.f32(v);
    w.words(m.texels.data(), m.texels.size());
  }
  w.u32(s.probes ? 1u : 0u);
  if (s.probes) {
    for (float v : s.probe_origin) w.f32(v);
    w.f32(s.probe_spacing);
    for (uint32_t v : s.probe_counts) w.u32(v);
    w.words(s.probe_words.data(), s.probe_words.size());
  }

  w.u32(s.stats.meshes);
  for (uint64_t v : {s.stats.triangles, s.stats.texels, s.stats.rays, s.stats.lights_removed, s.stats.bounce_rays,
                     s.stats.triangle_words})
    w.u64(v);
  return std::move(w.b);
}

// ---- parse_payload (format/mod.rs:218-301, lights.rs:55-112, maps.rs:40-97) ---------------------
dl_baked_scene* parse_payload(Reader& r) {
  std::unique_ptr<dl_baked_scene> s(new dl_baked_scene());
  const uint32_t flags = r.u32();
  if (flags > 3) fail(DL_FORMAT_INVALID, "settings flag");
  const uint32_t dyn = r.u32();
  if (dyn > 2) fail(DL_FORMAT_INVALID, "dynamic geometry mode");
  s->settings.bounce = flags & 1u;
  s->settings.colored_bounce = (flags >> 1) & 1u;
  s->settings.dynamic_geometry = dyn == 0 ? DL_DYNAMIC_LIGHTING_ONLY : dyn == 1 ? DL_DYNAMIC_DISTANCE_CUBES : DL_DYNAMIC_ANGULAR;
  s->settings.pixel_density = r.f32();
  s->settings.max_lightmap_size = r.u32();
  s->settings.distance_cube_resolution = r.u32();

  const uint64_t nl = r.count(kLightBytes);
  uint32_t suns = 0, skies = 0;
  for (uint64_t i = 0; i < nl; ++i) {
    dl_baked_light l{};
    l.id = r.u64();
    for (int k = 0; k < 3; ++k) l.origin[k] = r.f32();
    l.radius = r.f32();
    switch (r.u32()) {
      case 0: l.bounce_compression = -1; break;
      case 8: l.bounce_compression = DL_BOUNCE_BITS8; break;
      case 6: l.bounce_compression = DL_BOUNCE_BITS6; break;
      case 5: l.bounce_compression = DL_BOUNCE_BITS5; break;
      case 4: l.bounce_compression = DL_BOUNCE_BITS4; break;
      default: fail(DL_FORMAT_INVALID, "bounce compression");
    }
    const uint32_t kind = r.u32();
    float rot[4], size[2];
    for (float& v : rot) v = r.f32();
    for (float& v : size) v = r.f32();
    if (!(finite4(rot, 4) && finite4(size, 2))) fail(DL_FORMAT_INVALID, "light shape");
    if (kind > 4) fail(DL_FORMAT_INVALID, "light shape");
    l.shape = (int32_t)kind;
    const float identity[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    std::memcpy(l.rotation, identity, 16);
    if (kind == DL_SHAPE_RECTANGLE || kind == DL_SHAPE_TUBE) {
      std::memcpy(l.rotation, rot, 16);
      std::memcpy(l.size, size, 8);
    } else if (kind == DL_SHAPE_SUN) {
      std::memcpy(l.rotation, rot, 16);
      l.angular_diameter = size[0];
    }
    if (!s->lights.empty() && group_of(s->lights.back().shape) > group_of(l.shape)) fail(DL_FORMAT_INVALID, "light order");
    l.shadow_bits_per_texel = r.u32();
    if (!(l.shadow_bits_per_texel == 1 || l.shadow_bits_per_texel == 2 || l.shadow_bits_per_texel == 4 ||
          l.shadow_bits_per_texel == 8))
      fail(DL_FORMAT_INVALID, "shadow bits per texel");
    suns += kind == DL_SHAPE_SUN;
    skies += kind == DL_SHAPE_SKY;
    s->lights.push_back(l);
  }

  const std::vector<uint32_t> bvh = r.words();
  if (bvh.size() % 8) fail(DL_FORMAT_INVALID, "light BVH");
  s->bvh.resize(bvh.size() / 8);
  if (!bvh.empty()) std::memcpy(s->bvh.data(), bvh.data(), bvh.size() * 4);

  const uint64_t nm = r.count(32);
  s->meshes.reserve((size_t)nm);
  for (uint64_t i = 0; i < nm; ++i) {
    dl_baked_mesh_out m;
    m.id = r.u64();
    m.content_hash = r.u64();
    m.resolution = r.u32();
    const uint32_t occ = r.u32();
    if (occ > 1) fail(DL_FORMAT_INVALID, "mesh shadow flag");
    m.occludes = occ;
    m.words.own(r.words());
    s->meshes.push_back(std::move(m));
  }

  s->distance_cubes = r.words();

  const uint64_t maps = r.count(kSunMapBytes);
  if (maps != 0 && maps != suns) fail(DL_FORMAT_INVALID, "sun map count");
  for (uint64_t i = 0; i < maps; ++i) {
    dl_sun_map_out m;
    m.resolution = r.u32();
    for (float& v : m.rows) v = r.f32();
    m.texels = r.words();
    const bool valid = m.resolution >= 1 && m.resolution <= kMaxSunMapSide && finite4(m.rows, 12) &&
                       m.texels.size() == ((uint64_t)m.resolution * m.resolution + 1) / 2;
    if (!valid) fail(DL_FORMAT_INVALID, "sun map");
    s->sun_maps.push_back(std::move(m));
  }
  const uint32_t has_probes = r.u32();
  if (has_probes == 1 && skies > 0) {
    for (float& v : s->probe_origin) v = r.f32();
    s->probe_spacing = r.f32();
    for (uint32_t& v : s->probe_counts) v = r.u32();
    s->probe_words = r.words();
    const double probes = (double)s->probe_counts[0] * s->probe_counts[1] * s->probe_counts[2];
    const bool valid = finite4(s->probe_origin, 3) && std::isfinite(s->probe_spacing) && s->probe_spacing > 0.0f &&
                       s->probe_counts[0] >= 1 && s->probe_counts[1] >= 1 && s->probe_counts[2] >= 1 &&
                       probes <= (double)kMaxSkyProbes && (double)s->probe_words.size() == probes * 2.0;
    if (!valid) fail(DL_FORMAT_INVALID, "sky probes");
    s->probes = true;
  } else if (has_probes != 0) {
    fail(DL_FORMAT_INVALID, "sky probes");
  }

  s->stats.meshes = r.u32();
  s->stats.triangles = r.u64();
  s->stats.texels = r.u64();
  s->stats.rays = r.u64();
  s->stats.lights_removed = r.u64();
  s->stats.bounce_rays = r.u64();
  s->stats.triangle_words = r.u64();
  // stats.samples is not in the file (the reference's BakeStats has no sample count)
  if (r.remaining != 0) fail(DL_FORMAT_INVALID, "trailing payload");
  return s.release();
}

struct Header {
  uint64_t length;
  uint32_t checksum;
};
// parse_header (format/mod.rs:128-150)
Header parse_header(const uint8_t* b, size_t n) {
  if (n < kHeaderBytes || std::memcmp(b, kMagic, 8) != 0) {
    if (n >= 8 && std::memcmp(b, kMagic, 8) != 0) fail(DL_FORMAT_NOT_DYNLIGHT);
    fail(DL_FORMAT_TRUNCATED);
  }
  MemorySource ms(b + 8, kHeaderBytes - 8);
  Reader r(ms, kHeaderBytes - 8);
  const uint32_t version = r.u32();
  if (version != kVersion) unsupported_version(version);
  Header h;
  h.length = r.u64();
  h.checksum = r.u32();
  return h;
}

// CRC-32 (IEEE), zlib's crc32, over any length
uint32_t crc_of(const uint8_t* p, size_t n) {
  uLong crc = crc32(0L, Z_NULL, 0);
  for (size_t off = 0; off < n;) {
    const uInt k = (uInt)std::min<size_t>(n - off, 1u << 30);
    crc = crc32(crc, p + off, k);
    off += k;
  }
  return (uint32_t)crc;
}

int format_error(const FormatError& e) {
  g_format_kind = e.kind;
  dl_set_thread_error(e.message);
  return DL_E_FORMAT;
}

}  // namespace

extern "C" {

int dl_encode(const dl_baked_scene* scene, uint8_t** bytes, uint64_t* len) {
  if (!scene || !bytes || !len) {
    dl_set_thread_error("null argument");
    return DL_E_ARGUMENT;
  }
  *bytes = nullptr;
  *len = 0;
  if (scene->triangles_taken) {  // TrianglesTaken (format/mod.rs:84-104)
    dl_set_thread_error("the scene's triangle words were moved into its GPU buffers, so it can't be written");
    return DL_E_TRIANGLES_TAKEN;
  }
  try {
    // encode (format/mod.rs:101-119): header, then the payload as one gzip member
    const std::vector<uint8_t> payload = encode_payload(*scene);
    Writer out;
    out.b.assign(kMagic, kMagic + 8);
    out.u32(kVersion);
    out.u64(payload.size());
    out.u32(crc_of(payload.data(), payload.size()));
    z_stream zs{};
    if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
      dl_set_thread_error("deflate could not start");
      return DL_E_OOM;
    }
    gz_header gh{};
    gh.time = 0;    // mtime 0
    gh.os = 255;    // "unknown", as GzBuilder::operating_system(255)
    deflateSetHeader(&zs, &gh);
    std::vector<uint8_t>& o = out.b;
    size_t done = o.size(), off = 0;
    o.resize(done + payload.size() / 4 + (1u << 20));
    int r = Z_OK;
    while (r != Z_STREAM_END) {
      const size_t chunk = std::min<size_t>(payload.size() - off, 1u << 30);
      const bool last = off + chunk == payload.size();
      zs.next_in = const_cast<Bytef*>(payload.data() + off);
      zs.avail_in = (uInt)chunk;
      for (;;) {
        if (o.size() - done < (1u << 20)) o.resize(o.size() + std::max<size_t>(o.size() / 2, 1u << 24));
        zs.next_out = o.data() + done;
        const size_t room = std::min<size_t>(o.size() - done, 1u << 30);
        zs.avail_out = (uInt)room;
        r = deflate(&zs, last ? Z_FINISH : Z_NO_FLUSH);
        done += room - zs.avail_out;
        if (r == Z_STREAM_ERROR) {
          deflateEnd(&zs);
          dl_set_thread_error("deflate failed");
          return DL_E_DEVICE;
        }
        if (last ? r == Z_STREAM_END : zs.avail_in == 0) break;
      }
      off += chunk;
    }
    deflateEnd(&zs);
    uint8_t* buf = static_cast<uint8_t*>(std::malloc(done ? done : 1));
    if (!buf) {
      dl_set_thread_error("host allocation failed");
      return DL_E_OOM;
    }
    std::memcpy(buf, o.data(), done);
    *bytes = buf;
    *len = done;
    return DL_OK;
  } catch (const std::bad_alloc&) {
    dl_set_thread_error("host allocation failed");
    return DL_E_OOM;
  }
}

void dl_bytes_free(uint8_t* bytes) { std::free(bytes); }

int dl_decode(const uint8_t* bytes, uint64_t len, dl_baked_scene** out) {
  if (!out || (len && !bytes)) {
    dl_set_thread_error("null argument");
    return DL_E_ARGUMENT;
  }
  *out = nullptr;
  g_format_kind = 0;
  try {
    // decode (format/mod.rs:152-169): length and checksum are checked before parsing
    const Header h = parse_header(bytes, (size_t)len);
    MemorySource packed(bytes + kHeaderBytes, (size_t)len - kHeaderBytes);
    std::vector<uint8_t> payload;
    {
      Inflater inf(packed);
      uint8_t chunk[1 << 16];
      for (;;) {
        // at most length + 1 bytes are taken from the stream
        const uint64_t want = std::min<uint64_t>(sizeof chunk, h.length + 1 - std::min<uint64_t>(payload.size(), h.length + 1));
        if (want == 0) break;
        const size_t got = inf.read(chunk, (size_t)want);
        if (got == 0) break;
        payload.insert(payload.end(), chunk, chunk + got);
      }
    }
    if (payload.size() != h.length) fail(DL_FORMAT_TRUNCATED);
    if (crc_of(payload.data(), payload.size()) != h.checksum) fail(DL_FORMAT_CHECKSUM);
    MemorySource ps(payload.data(), payload.size());
    Reader r(ps, payload.size());
    *out = parse_payload(r);
    return DL_OK;
  } catch (const FormatError& e) {
    return format_error(e);
  } catch (const std::bad_alloc&) {
    dl_set_thread_error("host allocation failed");
    return DL_E_OOM;
  }
}

int dl_decode_stream(dl_read_fn read, void* user, dl_baked_scene** out) {
  if (!out || !read) {
    dl_set_thread_error("null argument");
    return DL_E_ARGUMENT;
  }
  *out = nullptr;
  g_format_kind = 0;
  try {
    // decode_stream (format/stream.rs:34-71): parse while inflating, then check the rest
    CallbackSource src(read, user);
    uint8_t head[kHeaderBytes];
    size_t have = 0;
    while (have < kHeaderBytes) {
      const size_t got = src.read(head + have, kHeaderBytes - have);
      if (got == 0) break;
      have += got;
    }
    const Header h = parse_header(head, have);
    Inflater inf(src);
    Reader r(inf, h.length);
    dl_baked_scene* parsed = nullptr;
    FormatError parse_err{0, std::string()};
    try {
      parsed = parse_payload(r);
    } catch (const FormatError& e) {
      if (e.kind == DL_FORMAT_DECOMPRESS) throw;
      parse_err = e;
    }
    std::unique_ptr<dl_baked_scene> keep(parsed);
    uint8_t sink[1 << 16];
    try {
      if (parsed) {
        inf.read(sink, 1);  // one byte past the declared length must not exist
      } else {
        while (inf.read(sink, sizeof sink) != 0) {
        }
      }
    } catch (const FormatError&) {
      if (parsed) throw;
    }
    if (inf.count != h.length) fail(DL_FORMAT_TRUNCATED);
    if ((uint32_t)inf.crc != h.checksum) fail(DL_FORMAT_CHECKSUM);
    if (!parsed) throw parse_err;
    *out = keep.release();
    return DL_OK;
  } catch (const FormatError& e) {
    return format_error(e);
  } catch (const std::bad_alloc&) {
    dl_set_thread_error("host allocation failed");
    return DL_E_OOM;
  }
}

int dl_format_error_kind(void) { return g_format_kind; }

}  // extern "C"

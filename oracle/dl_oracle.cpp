// ORACLE -- test infrastructure only (see dl_oracle.h).
//
// A CPU restatement of the reference crate's bake, function by function, in the reference's f32
// operation order.  Each function cites the reference file:line it follows.  Built with
// -ffp-contract=off; f32::sin/cos/tan/exp go to glibc sinf/cosf/tanf/expf exactly as Rust std
// does on Linux.
//
// Parity status (what pins each part; see tests/test_oracle_kats.py):
//   * pinned by the reference's own known-answer tests: splitmix64/PCG32 (effects/rng.rs),
//     lightmap_size, rects, texel enumeration, level, bit/level packing, pad_seams, dilation,
//     crop, serialize_triangles, pack_bounce, light BVH, is_buried ray counts, adaptive
//     sampler ray counts, spread/Sobol, concentric disk, f16 packing, occluder queries;
//   * pinned by closed forms and brute force: Moller-Trumbore, closest point, the oracle BVH
//     (equal to a scan of every occluder), analytic shadow footprints;
//   * parity unpinned (no reference run possible without a Rust toolchain): glam's exact
//     any_orthonormal_pair / slerp / acos_approx formulas, and which libm glam's from_angle
//     reaches (restated here as glibc through std, per the reference's std build).
#include "dl_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <vector>

namespace o {

// ------------------------------------------------------------------------------------------
// glam-equivalent scalar vector math (glam 0.33, scalar Vec3/Vec2; SSE2 Vec3A/Quat give the same
// bits for the operations used here: dot is (x*x' + y*y') + z*z', cross by components).
// ------------------------------------------------------------------------------------------
struct V2 { float x, y; };
struct V3 { float x, y, z; };

static inline V2 v2(float x, float y) { return {x, y}; }
static inline V3 v3(float x, float y, float z) { return {x, y, z}; }
static inline V3 v3p(const float* p) { return {p[0], p[1], p[2]}; }
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline bool operator!=(V3 a, V3 b) { return !(a == b); }
static inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
static inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
static inline V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }
static inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
static inline float length_sq(V3 a) { return dot(a, a); }
static inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// f32::min / f32::max (IEEE minNum/maxNum: a NaN operand yields the other).
static inline float fmin_(float a, float b) { return std::fmin(a, b); }
static inline float fmax_(float a, float b) { return std::fmax(a, b); }
static inline V3 vmin(V3 a, V3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
static inline V3 vmax(V3 a, V3 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
static inline V2 vmin(V2 a, V2 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y)}; }
static inline V2 vmax(V2 a, V2 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y)}; }
static inline V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
static inline float max_element(V3 a) { return fmax_(a.x, fmax_(a.y, a.z)); }
static inline float min_element(V3 a) { return fmin_(a.x, fmin_(a.y, a.z)); }
static inline bool finite3(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
// normalize = v * (1 / length)
static inline V3 normalize(V3 a) { return a * (1.0f / length(a)); }
// normalize_or(fallback): rcp = 1/length; finite and > 0 -> v * rcp
static inline V3 normalize_or(V3 a, V3 fallback) {
  float rcp = 1.0f / length(a);
  if (std::isfinite(rcp) && rcp > 0.0f) return a * rcp;
  return fallback;
}
static inline V3 normalize_or_zero(V3 a) { return normalize_or(a, v3(0, 0, 0)); }
static inline bool try_normalize(V3 a, V3* out) {
  float rcp = 1.0f / length(a);
  if (std::isfinite(rcp) && rcp > 0.0f) { *out = a * rcp; return true; }
  return false;
}
// f32::signum
static inline float signum(float v) {
  if (std::isnan(v)) return v;
  return std::signbit(v) ? -1.0f : 1.0f;
}
// Vec3::any_orthonormal_pair (Duff et al., "Building an Orthonormal Basis, Revisited").
static inline void any_orthonormal_pair(V3 n, V3* a_out, V3* b_out) {
  float sign = signum(n.z);
  float a = -1.0f / (sign + n.z);
  float b = (n.x * n.y) * a;
  *a_out = v3(1.0f + ((sign * n.x) * n.x) * a, sign * b, (-sign) * n.x);
  *b_out = v3(b, sign + (n.y * n.y) * a, -n.y);
}
// Quat * Vec3: v*(w^2 - |b|^2) + b*(2 (v.b)) + (b x v)*(2 w)
static inline V3 quat_mul(const float* q, V3 v) {
  V3 b = v3(q[0], q[1], q[2]);
  float w = q[3];
  float b2 = dot(b, b);
  return (v * (w * w - b2) + b * (dot(v, b) * 2.0f)) + cross(b, v) * (w * 2.0f);
}

static const float TAU = 6.28318530717958647692f;
static const float PI_F = 3.14159265358979323846f;
static const float FRAC_PI_4 = 0.785398163397448309616f;
static const float NDOTL_TOLERANCE = -0.1f;  // mesh_bake.rs:25
static const float RAY_OFFSET = 0.001f;       // mesh_bake.rs:27

// Rust `as u32` / `as u8` / `as i8`: saturating, NaN -> 0, truncation toward zero.
static inline uint32_t as_u32(float v) {
  if (!(v > 0.0f)) return 0;  // NaN, negatives, zero
  if (v >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)v;
}
static inline uint8_t as_u8(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)v;
}
static inline int8_t as_i8(float v) {
  if (std::isnan(v)) return 0;
  if (v <= -128.0f) return -128;
  if (v >= 127.0f) return 127;
  return (int8_t)v;
}
// f32::clamp (NaN stays NaN)
static inline float clampf(float v, float lo, float hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}

// ------------------------------------------------------------------------------------------
// RNG: src/effects/rng.rs:10-71, src/bake/random.rs:9-28
// ------------------------------------------------------------------------------------------
static inline uint64_t splitmix64(uint64_t value) {
  value += 0x9e3779b97f4a7c15ull;
  value = (value ^ (value >> 30)) * 0xbf58476d1ce4e5b9ull;
  value = (value ^ (value >> 27)) * 0x94d049bb133111ebull;
  return value ^ (value >> 31);
}
struct Rng {
  uint64_t state, inc;
  uint32_t next_u32() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    uint32_t x = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (x >> rot) | (x << ((32 - rot) & 31));
  }
  float next_f32() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
};
static inline Rng rng_new(uint64_t seed, uint64_t stream) {
  Rng r{0, (stream << 1) | 1};
  r.next_u32();
  r.state += seed;
  r.next_u32();
  return r;
}
static inline uint64_t seed_of(const uint64_t* parts, size_t n) {
  uint64_t acc = 0x0b4ce5eed0000000ull;
  for (size_t i = 0; i < n; ++i) acc = splitmix64(acc ^ splitmix64(parts[i]));
  return acc;
}
static inline Rng stream(std::initializer_list<uint64_t> parts) {
  uint64_t s = seed_of(parts.begin(), parts.size());
  return rng_new(s, (s << 17) | (s >> 47));
}
static inline Rng stream_n(const uint64_t* parts, size_t n) {
  uint64_t s = seed_of(parts, n);
  return rng_new(s, (s << 17) | (s >> 47));
}
// random.rs:22-28
static inline V3 on_unit_sphere(Rng& rng) {
  float z = rng.next_f32() * 2.0f - 1.0f;
  float angle = rng.next_f32() * TAU;
  float r = std::sqrt(fmax_(1.0f - z * z, 0.0f));
  return v3(r * cosf(angle), r * sinf(angle), z);
}

// ------------------------------------------------------------------------------------------
// Rects and geometry: bits.rs:7-41, mesh_builder.rs:83-137, geometry.rs:8-140
// ------------------------------------------------------------------------------------------
struct Rect {
  uint32_t min_x, min_y, max_x, max_y;
  uint32_t width() const { return max_x - min_x + 1; }
  uint32_t height() const { return max_y - min_y + 1; }
  Rect unite(Rect o) const {
    return {std::min(min_x, o.min_x), std::min(min_y, o.min_y), std::max(max_x, o.max_x),
            std::max(max_y, o.max_y)};
  }
  bool contains(int64_t x, int64_t y) const {
    return x >= min_x && y >= min_y && x <= (int64_t)max_x && y <= (int64_t)max_y;
  }
};

static inline void uv_bounds(const V2* uvs, V2* mn, V2* mx) {
  *mn = vmin(vmin(uvs[0], uvs[1]), uvs[2]);
  *mx = vmax(vmax(uvs[0], uvs[1]), uvs[2]);
}
static inline uint32_t clamp_texel(float value, uint32_t size) {
  return as_u32(clampf(value, 0.0f, (float)(size - 1)));
}
// mesh_builder.rs:103-113
static Rect crop_rect(const V2* uvs, uint32_t size) {
  V2 mn, mx;
  uv_bounds(uvs, &mn, &mx);
  float s = (float)size;
  return {clamp_texel(std::floor(mn.x * s) - 2.0f, size), clamp_texel(std::floor(mn.y * s) - 2.0f, size),
          clamp_texel(std::ceil(mx.x * s) + 2.0f, size), clamp_texel(std::ceil(mx.y * s) + 2.0f, size)};
}
// mesh_builder.rs:127-137
static Rect grown_center_rect(const V2* uvs, uint32_t size, float grow) {
  V2 mn, mx;
  uv_bounds(uvs, &mn, &mx);
  float s = (float)size;
  return {clamp_texel(std::floor(mn.x * s - 0.5f) - grow, size),
          clamp_texel(std::floor(mn.y * s - 0.5f) - grow, size),
          clamp_texel(std::ceil(mx.x * s - 0.5f) + grow, size),
          clamp_texel(std::ceil(mx.y * s - 0.5f) + grow, size)};
}
static Rect trace_rect(const V2* uvs, uint32_t size) { return grown_center_rect(uvs, size, 0.0f); }
static Rect optimize_rect(const V2* uvs, uint32_t size) { return grown_center_rect(uvs, size, 1.0f); }

// geometry.rs:66-70
static uint32_t lightmap_size(float surface_area, float pixel_density, uint32_t max_size) {
  float size = std::ceil(std::sqrt(pixel_density * pixel_density * surface_area));
  if (std::isnan(size)) size = 0.0f;
  uint32_t v = as_u32(fmin_(size, (float)max_size));
  uint32_t hi = std::max(max_size, 2u);
  return std::min(std::max(v, 2u), hi);
}
// geometry.rs:59-61
static inline float triangle_area(V3 a, V3 b, V3 c) { return length(cross(b - a, c - a)) * 0.5f; }

// geometry.rs:73-108 (Ericson)
static V3 closest_point_on_triangle(V3 p, V3 a, V3 b, V3 c) {
  V3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  V3 bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + ab * (d1 / (d1 - d3));
  V3 cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + ac * (d2 / (d2 - d6));
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
    return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
  float denom = 1.0f / ((va + vb) + vc);
  return (a + ab * (vb * denom)) + ac * (vc * denom);
}
static inline bool sphere_intersects_triangle(V3 center, float radius, V3 a, V3 b, V3 c) {
  V3 closest = closest_point_on_triangle(center, a, b, c);
  return length_sq(closest - center) <= radius * radius;
}
// geometry.rs:117-134
static inline float signed_area(V2 a, V2 b, V2 c) {
  V2 a2 = a - c, b2 = b - c;
  return (a2.x * b2.y - a2.y * b2.x) * 0.5f;
}
static bool uv_barycentrics(const V2* t, V2 p, V3* out) {
  float area = signed_area(t[0], t[1], t[2]);
  if (area == 0.0f) return false;
  float a1 = signed_area(t[1], t[2], p) / area;
  float a2 = signed_area(t[2], t[0], p) / area;
  float a3 = signed_area(t[0], t[1], p) / area;
  if (a1 >= 0.0f && a2 >= 0.0f && a3 >= 0.0f) { *out = v3(a1, a2, a3); return true; }
  return false;
}
// geometry.rs:138-140
static inline bool approx_eq(V3 a, V3 b, float eps) { return max_element(vabs(a - b)) < eps; }

struct Aabb {
  V3 min, max;
  static Aabb empty() {
    return {v3(INFINITY, INFINITY, INFINITY), v3(-INFINITY, -INFINITY, -INFINITY)};
  }
  static Aabb sphere(V3 c, float r) { return {c - v3(r, r, r), c + v3(r, r, r)}; }
  Aabb grow(V3 p) const { return {vmin(min, p), vmax(max, p)}; }
  Aabb unite(const Aabb& o) const { return {vmin(min, o.min), vmax(max, o.max)}; }
  bool intersects(const Aabb& o) const {
    return min.x <= o.max.x && min.y <= o.max.y && min.z <= o.max.z && o.min.x <= max.x &&
           o.min.y <= max.y && o.min.z <= max.z;
  }
  float half_area_x2() const {
    V3 e = max - min;
    if (min_element(e) < 0.0f) return 0.0f;
    return 2.0f * ((e.x * e.y + e.y * e.z) + e.z * e.x);
  }
};

// ------------------------------------------------------------------------------------------
// Texels: texels.rs:24-141
// ------------------------------------------------------------------------------------------
struct SurfaceTexel {
  uint32_t x, y;
  V3 position, normal;
};
static inline V2 texel_center(uint32_t x, uint32_t y, uint32_t size) {
  return v2(((float)x + 0.5f) / (float)size, ((float)y + 0.5f) / (float)size);
}
static bool uv_to_surface(const V3* vs, const V3* ns, const V2* uvs, V3 face, V2 uv, V3* pos, V3* nrm) {
  V3 w;
  if (!uv_barycentrics(uvs, uv, &w)) return false;
  *pos = (vs[0] * w.x + vs[1] * w.y) + vs[2] * w.z;
  V3 n = (ns[0] * w.x + ns[1] * w.y) + ns[2] * w.z;
  if (!try_normalize(n, nrm)) *nrm = face;
  return true;
}
static void triangle_texels(const V3* vs, const V3* ns, const V2* uvs, V3 face, uint32_t size,
                            std::vector<SurfaceTexel>& out) {
  Rect r = trace_rect(uvs, size);
  for (uint32_t y = r.min_y; y <= r.max_y; ++y)
    for (uint32_t x = r.min_x; x <= r.max_x; ++x) {
      V3 p, n;
      if (uv_to_surface(vs, ns, uvs, face, texel_center(x, y, size), &p, &n)) out.push_back({x, y, p, n});
    }
}
static inline V3 lift(V2 uv) { return v3(uv.x, uv.y, 0.0f); }
// texels.rs:125-141
static bool nearest_weights(const V2* uvs, V2 point, V3* out) {
  V3 a = lift(uvs[0]), b = lift(uvs[1]), c = lift(uvs[2]);
  V3 n3 = closest_point_on_triangle(lift(point), a, b, c);
  V2 nearest = v2(n3.x, n3.y);
  float area = cross(b - a, c - a).z;
  if (area == 0.0f || !std::isfinite(area)) return false;
  auto weight = [&](V2 p, V2 q) { return cross(lift(p) - lift(nearest), lift(q) - lift(nearest)).z; };
  V3 w = v3(weight(uvs[1], uvs[2]), weight(uvs[2], uvs[0]), weight(uvs[0], uvs[1])) / area;
  w = vmax(w, v3(0, 0, 0));
  float sum = (w.x + w.y) + w.z;
  *out = w / fmax_(sum, 1.17549435e-38f);
  return true;
}
template <class Owned>
static void touched_texels(const V3* vs, const V3* ns, const V2* uvs, V3 face, uint32_t size,
                           Owned owned, std::vector<SurfaceTexel>& out) {
  float s = (float)size;
  V2 mn, mx;
  uv_bounds(uvs, &mn, &mx);
  V2 first = vmax(v2(std::floor(mn.x * s), std::floor(mn.y * s)), v2(0, 0));
  V2 last = vmin(v2(std::floor(mx.x * s), std::floor(mx.y * s)), v2(s - 1.0f, s - 1.0f));
  float half = 0.5f / s;
  int64_t y0 = as_u32(first.y), y1 = as_u32(last.y), x0 = as_u32(first.x), x1 = as_u32(last.x);
  for (int64_t y = y0; y <= y1; ++y)
    for (int64_t x = x0; x <= x1; ++x) {
      if (owned((uint32_t)x, (uint32_t)y)) continue;
      V2 center = texel_center((uint32_t)x, (uint32_t)y, size);
      V3 w;
      if (!nearest_weights(uvs, center, &w)) continue;
      V2 nearest = v2((uvs[0].x * w.x + uvs[1].x * w.y) + uvs[2].x * w.z,
                      (uvs[0].y * w.x + uvs[1].y * w.y) + uvs[2].y * w.z);
      V2 d = nearest - center;
      if (fmax_(std::fabs(d.x), std::fabs(d.y)) > half) continue;
      V3 p = (vs[0] * w.x + vs[1] * w.y) + vs[2] * w.z;
      V3 n = (ns[0] * w.x + ns[1] * w.y) + ns[2] * w.z, nn;
      if (!try_normalize(n, &nn)) nn = face;
      out.push_back({(uint32_t)x, (uint32_t)y, p, nn});
    }
}

// ------------------------------------------------------------------------------------------
// Low-discrepancy points: spread.rs:16-58
// ------------------------------------------------------------------------------------------
static inline uint32_t reverse_bits(uint32_t v) {
  v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
  v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
  v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
  v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
  return (v >> 16) | (v << 16);
}
static inline uint32_t sobol_second(uint32_t index) {
  uint32_t value = 0, direction = 1u << 31;
  while (index != 0) {
    if (index & 1) value ^= direction;
    index >>= 1;
    direction ^= direction >> 1;
  }
  return value;
}
static inline float unit_(uint32_t bits) { return (float)(bits >> 8) / 16777216.0f; }
static inline float fract_(float v) { return v - std::trunc(v); }
static inline V2 spread(uint32_t index, V2 offset, bool rectangle) {
  float across = unit_(reverse_bits(index));
  float down = rectangle ? unit_(sobol_second(index)) : 0.0f;
  return v2(fract_(across + offset.x), fract_(down + offset.y));
}
static inline V2 concentric_disk(V2 square) {
  V2 p = v2(square.x * 2.0f - 1.0f, square.y * 2.0f - 1.0f);
  if (p.x == 0.0f && p.y == 0.0f) return v2(0, 0);
  float radius, angle;
  if (std::fabs(p.x) > std::fabs(p.y)) {
    radius = p.x;
    angle = FRAC_PI_4 * (p.y / p.x);
  } else {
    radius = p.y;
    angle = 2.0f * FRAC_PI_4 - FRAC_PI_4 * (p.x / p.y);
  }
  // Vec2::from_angle(angle) = (cos, sin), std f32::sin_cos -> glibc
  float sn = sinf(angle), cs = cosf(angle);
  return v2(cs * radius, sn * radius);
}

// ------------------------------------------------------------------------------------------
// Lights: input.rs:101-132, light_shape.rs:13-65, light/area.rs:46-123, light/shape.rs:95-102
// ------------------------------------------------------------------------------------------
struct AreaShape {
  V3 center, x, y, forward;
  V2 half;
  bool rectangle;
  V3 nearest_point(V3 world) const {
    V3 local = world - center;
    float px = clampf(dot(local, x), -half.x, half.x);
    float py = clampf(dot(local, y), -half.y, half.y);
    return (center + x * px) + y * py;
  }
  V3 point_at(V2 uv) const {
    V2 off = v2((uv.x * 2.0f - 1.0f) * half.x, (uv.y * 2.0f - 1.0f) * half.y);
    return (center + x * off.x) + y * off.y;
  }
  void corners(V3* out) const {
    V3 hx = x * half.x, hy = y * half.y;
    out[0] = (center - hx) - hy;
    out[1] = (center + hx) - hy;
    out[2] = (center + hx) + hy;
    out[3] = (center - hx) + hy;
  }
};

struct Light {
  const dl_light* src;
  uint64_t id;
  V3 position;
  float radius;
  int shape;
  bool distant, sky, area, sun;
  AreaShape as;
  float reach;
  V3 sun_toward;
  float sun_radius, sun_tan, sun_sin;
  uint32_t bits;
  bool alpha_tested;
  bool bakes_bounce;
};

static Light make_light(const dl_light* l, const dl_bake_options& opt) {
  Light L{};
  L.src = l;
  L.id = l->id;
  L.position = v3p(l->position);
  L.radius = l->radius;
  L.shape = l->shape;
  L.sun = l->shape == DL_SHAPE_SUN;
  L.sky = l->shape == DL_SHAPE_SKY;
  L.area = l->shape == DL_SHAPE_RECTANGLE || l->shape == DL_SHAPE_TUBE;
  L.distant = L.sun || L.sky;
  if (L.area) {
    L.as.center = L.position;
    L.as.x = quat_mul(l->rotation, v3(1, 0, 0));
    L.as.y = quat_mul(l->rotation, v3(0, 1, 0));
    L.as.forward = quat_mul(l->rotation, v3(0, 0, -1));
    L.as.half = v2(l->size[0] * 0.5f, l->size[1] * 0.5f);
    L.as.rectangle = l->shape == DL_SHAPE_RECTANGLE;
    float hx = L.as.half.x, hy = L.as.half.y;
    L.reach = l->radius - std::sqrt(hx * hx + hy * hy);
  } else {
    L.reach = l->radius;
  }
  if (L.sun) {
    L.sun_toward = normalize(quat_mul(l->rotation, v3(0, 0, 1)));
    L.sun_radius = (l->angular_diameter * 0.5f) * (PI_F / 180.0f);
    L.sun_tan = tanf(L.sun_radius);
    L.sun_sin = sinf(L.sun_radius);
  }
  if (L.shape == DL_SHAPE_POINT) L.bits = 1;
  else if (L.area) L.bits = 4;
  else if (L.sun) L.bits = l->angular_diameter > 0.0f ? 4 : 1;
  else L.bits = 8;
  L.alpha_tested = l->transparency != DL_TRANSPARENCY_DISABLED;
  L.bakes_bounce = l->illumination == DL_ILLUMINATION_SINGLE_BOUNCE && !opt.skip_bounce && !L.sky;
  return L;
}

// light_shape.rs:47-65
static bool may_face(const Light& L, const V3* v, V3 normal) {
  V3 center = ((v[0] + v[1]) + v[2]) / 3.0f;
  auto faces = [&](V3 p) { return dot(normal, normalize_or_zero(p - center)) > NDOTL_TOLERANCE; };
  if (L.sky) return true;
  if (L.sun) return dot(normal, L.sun_toward) + L.sun_sin > NDOTL_TOLERANCE;
  if (!L.area) return faces(L.position);
  bool lit_side = !L.as.rectangle;
  if (!lit_side)
    for (int k = 0; k < 3; ++k)
      if (dot(L.as.forward, v[k] - L.as.center) > 0.0f) lit_side = true;
  if (!lit_side) return false;
  V3 c[4];
  L.as.corners(c);
  for (int k = 0; k < 4; ++k)
    if (faces(c[k])) return true;
  return false;
}

// ------------------------------------------------------------------------------------------
// Light BVH: light_bvh.rs:59-220
// ------------------------------------------------------------------------------------------
struct LightBvh {
  std::vector<dl_light_bvh_node> nodes;
  std::vector<uint32_t> order;
  std::vector<Aabb> bounds;
};
struct LightBvhBuilder {
  const V3* pos;
  const float* radii;
  std::vector<uint32_t> idx;
  std::vector<dl_light_bvh_node> nodes;
  Aabb bounds_of(uint32_t l) const { return Aabb::sphere(pos[l], radii[l]); }
  static Aabb node_aabb(const dl_light_bvh_node& n) {
    return {v3p(n.aabb_min), v3p(n.aabb_max)};
  }
  void update_bounds(size_t node) {
    uint32_t first = nodes[node].left_first, count = nodes[node].count;
    Aabb a = Aabb::empty();
    for (uint32_t k = first; k < first + count; ++k) a = a.unite(bounds_of(idx[k]));
    nodes[node].aabb_min[0] = a.min.x; nodes[node].aabb_min[1] = a.min.y; nodes[node].aabb_min[2] = a.min.z;
    nodes[node].aabb_max[0] = a.max.x; nodes[node].aabb_max[1] = a.max.y; nodes[node].aabb_max[2] = a.max.z;
  }
  static float axis_of(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
  float split_cost(size_t first, size_t count, int axis, float split) const {
    Aabb l = Aabb::empty(), r = Aabb::empty();
    uint32_t lc = 0, rc = 0;
    for (size_t k = first; k < first + count; ++k) {
      uint32_t light = idx[k];
      if (axis_of(pos[light], axis) < split) { lc++; l = l.unite(bounds_of(light)); }
      else { rc++; r = r.unite(bounds_of(light)); }
    }
    if (lc == 0 || rc == 0) return 3.40282347e+38f;
    return (float)lc * l.half_area_x2() + (float)rc * r.half_area_x2();
  }
  void subdivide(size_t node) {
    uint32_t left_first = nodes[node].left_first, count = nodes[node].count;
    if (count <= 2) return;
    size_t first = left_first;
    float parent_cost = (float)count * node_aabb(nodes[node]).half_area_x2();
    bool have = false;
    int best_axis = 0;
    float best_split = 0.0f, best_cost = 3.40282347e+38f;
    for (int axis = 0; axis < 3; ++axis)
      for (size_t k = 0; k < count; ++k) {
        float candidate = axis_of(pos[idx[first + k]], axis);
        float cost = split_cost(first, count, axis, candidate);
        if (cost < best_cost) { best_cost = cost; have = true; best_axis = axis; best_split = candidate; }
      }
    if (!have) return;
    if (best_cost >= parent_cost) return;
    int64_t i = (int64_t)first, j = (int64_t)(first + count) - 1;
    while (i <= j) {
      if (axis_of(pos[idx[(size_t)i]], best_axis) < best_split) i++;
      else { std::swap(idx[(size_t)i], idx[(size_t)j]); j--; }
    }
    size_t left_count = (size_t)i - first;
    if (left_count == 0 || left_count == count) return;
    size_t left = nodes.size();
    dl_light_bvh_node a{}, b{};
    a.left_first = (uint32_t)first; a.count = (uint32_t)left_count;
    b.left_first = (uint32_t)i; b.count = (uint32_t)(count - left_count);
    nodes.push_back(a);
    nodes.push_back(b);
    nodes[node].left_first = (uint32_t)left;
    nodes[node].count = 0;
    update_bounds(left);
    update_bounds(left + 1);
    subdivide(left);
    subdivide(left + 1);
  }
};
static LightBvh build_light_bvh(const std::vector<V3>& positions, const std::vector<float>& radii) {
  LightBvh out;
  if (positions.empty()) return out;
  LightBvhBuilder b{positions.data(), radii.data(), {}, {}};
  for (uint32_t i = 0; i < positions.size(); ++i) b.idx.push_back(i);
  dl_light_bvh_node root{};
  root.left_first = 0;
  root.count = (uint32_t)positions.size();
  b.nodes.push_back(root);
  b.update_bounds(0);
  b.subdivide(0);
  out.nodes = b.nodes;
  out.order = b.idx;
  for (uint32_t i : b.idx) out.bounds.push_back(Aabb::sphere(positions[i], radii[i]));
  return out;
}

// ------------------------------------------------------------------------------------------
// Occluders: occluder.rs:80-251.  Queries follow brute-force semantics over all occluders
// (any-hit: some triangle with 0 <= t <= max and t finite; closest: lowest (t, index)); the BVH
// below is an accelerator whose culling is conservative, checked against the scan in tests.
// ------------------------------------------------------------------------------------------
struct OccTri {
  V3 v[3];
  uint32_t mesh, triangle;
};

// occluder.rs:225-251 two-sided Moller-Trumbore
static inline bool intersect(const V3* v, V3 origin, V3 direction, float max, float* t_out, V3* bary) {
  V3 e1 = v[1] - v[0];
  V3 e2 = v[2] - v[0];
  V3 p = cross(direction, e2);
  float det = dot(e1, p);
  if (det == 0.0f || !std::isfinite(det)) return false;
  float inverse = 1.0f / det;
  V3 s = origin - v[0];
  float u = dot(s, p) * inverse;
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  V3 q = cross(s, e1);
  float vv = dot(direction, q) * inverse;
  if (vv < 0.0f || u + vv > 1.0f) return false;
  float t = dot(e2, q) * inverse;
  if (!(t >= 0.0f && t <= max)) return false;
  *t_out = t;
  if (bary) *bary = v3((1.0f - u) - vv, u, vv);
  return true;
}

struct BvhNode {
  float lo[3], hi[3];
  int32_t left;    // inner: child indices left, left+1; leaf: -1
  uint32_t first, count;
};

struct Scene;
struct Occluders {
  std::vector<OccTri> tris;
  std::vector<const dl_mesh*> meshes;  // sorted mesh order
  std::vector<BvhNode> nodes;
  std::vector<uint32_t> prim;          // BVH leaf order -> occluder index
  bool brute = false;

  void build_bvh();
  bool blocks(const OccTri& o, V3 w) const;
  float hit_distance(uint32_t index, V3 origin, V3 dir, float max, bool alpha) const {
    const OccTri& o = tris[index];
    float t;
    V3 w;
    if (intersect(o.v, origin, dir, max, &t, &w) && (!alpha || blocks(o, w))) return t;
    return INFINITY;
  }
  bool occluded(V3 origin, V3 dir, float max, bool alpha) const;
  bool closest_hit(V3 origin, V3 dir, float max, bool alpha, float* dist, uint32_t* occ, V3* bary) const;
  V3 face_normal(uint32_t occ) const {
    const OccTri& o = tris[occ];
    return normalize_or_zero(cross(o.v[1] - o.v[0], o.v[2] - o.v[0]));
  }
  V3 surface_normal(uint32_t occ, V3 w) const;
  V3 albedo(uint32_t occ) const {
    const dl_mesh* m = meshes[tris[occ].mesh];
    return v3p(m->albedo);
  }
};

// input.rs:79-97
static inline uint32_t wrap_coord(float coordinate, uint32_t size) {
  float r = std::fmod(coordinate, 1.0f);
  if (r < 0.0f) r = r + 1.0f;
  float texel = std::floor(r * (float)size);
  if (std::isfinite(texel)) return std::min(as_u32(texel), size - 1);
  return 0;
}
bool Occluders::blocks(const OccTri& o, V3 w) const {
  const dl_mesh* m = meshes[o.mesh];
  if (!m->alpha || !m->uv0) return true;
  const uint32_t* c = m->indices + o.triangle * 3;
  const float* uv0 = m->uv0;
  float ux = (uv0[c[0] * 2] * w.x + uv0[c[1] * 2] * w.y) + uv0[c[2] * 2] * w.z;
  float uy = (uv0[c[0] * 2 + 1] * w.x + uv0[c[1] * 2 + 1] * w.y) + uv0[c[2] * 2 + 1] * w.z;
  const dl_alpha_texture* a = m->alpha;
  if (a->width == 0 || a->height == 0) return true;
  uint32_t x = wrap_coord(ux, a->width), y = wrap_coord(uy, a->height);
  uint8_t alpha = a->rgba8[((size_t)y * a->width + x) * 4 + 3];
  return (float)alpha / 255.0f >= a->cutoff;
}
V3 Occluders::surface_normal(uint32_t occ, V3 w) const {
  const OccTri& o = tris[occ];
  V3 face = face_normal(occ);
  const dl_mesh* m = meshes[o.mesh];
  if (!m->normals || m->normal_count == 0) return face;
  const uint32_t* c = m->indices + o.triangle * 3;
  V3 n = (v3p(m->normals + c[0] * 3) * w.x + v3p(m->normals + c[1] * 3) * w.y) + v3p(m->normals + c[2] * 3) * w.z;
  V3 r;
  if (!try_normalize(n, &r)) return face;
  return r;
}

// Box padding: child boxes are grown by a relative 2^-16 of their coordinate magnitude plus
// extent, so f32 rounding in the slab test or in Moller-Trumbore cannot cull a real hit.
static void pad_box(float* lo, float* hi) {
  for (int a = 0; a < 3; ++a) {
    float pad = (std::fabs(lo[a]) + std::fabs(hi[a]) + (hi[a] - lo[a])) * (1.0f / 65536.0f) + 1e-30f;
    lo[a] -= pad;
    hi[a] += pad;
  }
}

void Occluders::build_bvh() {
  nodes.clear();
  prim.clear();
  if (tris.empty()) return;
  size_t n = tris.size();
  std::vector<float> cen(n * 3), blo(n * 3), bhi(n * 3);
  for (size_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      float v0 = (&tris[i].v[0].x)[a], v1 = (&tris[i].v[1].x)[a], v2 = (&tris[i].v[2].x)[a];
      blo[i * 3 + a] = std::min(v0, std::min(v1, v2));
      bhi[i * 3 + a] = std::max(v0, std::max(v1, v2));
      cen[i * 3 + a] = 0.5f * (blo[i * 3 + a] + bhi[i * 3 + a]);
    }
  }
  prim.resize(n);
  for (size_t i = 0; i < n; ++i) prim[i] = (uint32_t)i;
  struct Item { size_t node, first, count; };
  nodes.push_back({});
  std::vector<Item> work{{0, 0, n}};
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t k = it.first; k < it.first + it.count; ++k) {
      uint32_t p = prim[k];
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], blo[p * 3 + a]);
        hi[a] = std::max(hi[a], bhi[p * 3 + a]);
        clo[a] = std::min(clo[a], cen[p * 3 + a]);
        chi[a] = std::max(chi[a], cen[p * 3 + a]);
      }
    }
    pad_box(lo, hi);
    BvhNode& nd = nodes[it.node];
    for (int a = 0; a < 3; ++a) { nd.lo[a] = lo[a]; nd.hi[a] = hi[a]; }
    if (it.count <= 4) {
      nd.left = -1; nd.first = (uint32_t)it.first; nd.count = (uint32_t)it.count;
      continue;
    }
    // binned SAH over the longest centroid axis
    int axis = 0;
    float ext = chi[0] - clo[0];
    for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > ext) { ext = chi[a] - clo[a]; axis = a; }
    size_t mid;
    if (!(ext > 0.0f)) {
      mid = it.first + it.count / 2;
    } else {
      const int B = 16;
      size_t cnt[B] = {};
      float blo_[B][3], bhi_[B][3];
      for (int b = 0; b < B; ++b) for (int a = 0; a < 3; ++a) { blo_[b][a] = INFINITY; bhi_[b][a] = -INFINITY; }
      auto bin_of = [&](uint32_t p) {
        int b = (int)((cen[p * 3 + axis] - clo[axis]) / ext * B);
        return std::min(std::max(b, 0), B - 1);
      };
      for (size_t k = it.first; k < it.first + it.count; ++k) {
        uint32_t p = prim[k];
        int b = bin_of(p);
        cnt[b]++;
        for (int a = 0; a < 3; ++a) { blo_[b][a] = std::min(blo_[b][a], blo[p * 3 + a]); bhi_[b][a] = std::max(bhi_[b][a], bhi[p * 3 + a]); }
      }
      auto area = [](const float* l, const float* h) {
        float e0 = h[0] - l[0], e1 = h[1] - l[1], e2 = h[2] - l[2];
        if (e0 < 0 || e1 < 0 || e2 < 0) return 0.0f;
        return e0 * e1 + e1 * e2 + e2 * e0;
      };
      float best = INFINITY;
      int best_b = -1;
      for (int s = 1; s < B; ++s) {
        float l1[3] = {INFINITY, INFINITY, INFINITY}, h1[3] = {-INFINITY, -INFINITY, -INFINITY};
        float l2[3] = {INFINITY, INFINITY, INFINITY}, h2[3] = {-INFINITY, -INFINITY, -INFINITY};
        size_t c1 = 0, c2 = 0;
        for (int b = 0; b < B; ++b) {
          float* L = b < s ? l1 : l2; float* H = b < s ? h1 : h2;
          (b < s ? c1 : c2) += cnt[b];
          for (int a = 0; a < 3; ++a) { L[a] = std::min(L[a], blo_[b][a]); H[a] = std::max(H[a], bhi_[b][a]); }
        }
        if (c1 == 0 || c2 == 0) continue;
        float cost = (float)c1 * area(l1, h1) + (float)c2 * area(l2, h2);
        if (cost < best) { best = cost; best_b = s; }
      }
      if (best_b < 0) {
        mid = it.first + it.count / 2;
      } else {
        auto beg = prim.begin() + (long)it.first, end = beg + (long)it.count;
        auto m = std::stable_partition(beg, end, [&](uint32_t p) { return bin_of(p) < best_b; });
        mid = (size_t)(m - prim.begin());
        if (mid == it.first || mid == it.first + it.count) mid = it.first + it.count / 2;
      }
    }
    size_t left = nodes.size();
    nodes.push_back({});
    nodes.push_back({});
    nodes[it.node].left = (int32_t)left;
    nodes[it.node].first = 0;
    nodes[it.node].count = 0;
    work.push_back({left, it.first, mid - it.first});
    work.push_back({left + 1, mid, it.first + it.count - mid});
  }
}

struct RayPre {
  V3 o;
  float inv[3];
};
static inline RayPre ray_pre(V3 o, V3 d) {
  RayPre r{o, {0, 0, 0}};
  float dd[3] = {d.x, d.y, d.z};
  for (int a = 0; a < 3; ++a) {
    float c = dd[a];
    if (std::fabs(c) < 1e-20f) c = std::signbit(c) ? -1e-20f : 1e-20f;
    r.inv[a] = 1.0f / c;
  }
  return r;
}
// Conservative slab test: true when the padded box may hold a point of the ray with
// 0 <= t <= tmax.
static inline bool slab(const RayPre& r, const float* lo, const float* hi, float tmax) {
  float o[3] = {r.o.x, r.o.y, r.o.z};
  float tn = -INFINITY, tf = INFINITY;
  for (int a = 0; a < 3; ++a) {
    float t0 = (lo[a] - o[a]) * r.inv[a], t1 = (hi[a] - o[a]) * r.inv[a];
    tn = std::max(tn, std::min(t0, t1));
    tf = std::min(tf, std::max(t0, t1));
  }
  tn = tn - std::fabs(tn) * (1.0f / 65536.0f);
  tf = tf + std::fabs(tf) * (1.0f / 65536.0f);
  return tn <= tf && tf >= 0.0f && tn <= tmax * (1.0f + 1.0f / 65536.0f);
}

bool Occluders::occluded(V3 origin, V3 dir, float max, bool alpha) const {
  if (tris.empty()) return false;
  if (brute || nodes.empty()) {
    for (uint32_t i = 0; i < tris.size(); ++i) {
      float t = hit_distance(i, origin, dir, max, alpha);
      if (t != INFINITY) return true;
    }
    return false;
  }
  RayPre r = ray_pre(origin, dir);
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const BvhNode& nd = nodes[stack[--sp]];
    if (!slab(r, nd.lo, nd.hi, max)) continue;
    if (nd.left < 0) {
      for (uint32_t k = nd.first; k < nd.first + nd.count; ++k)
        if (hit_distance(prim[k], origin, dir, max, alpha) != INFINITY) return true;
    } else {
      stack[sp++] = (uint32_t)nd.left;
      stack[sp++] = (uint32_t)nd.left + 1;
    }
  }
  return false;
}

bool Occluders::closest_hit(V3 origin, V3 dir, float max, bool alpha, float* dist, uint32_t* occ, V3* bary) const {
  if (tris.empty()) return false;
  bool have = false;
  float best_t = INFINITY;
  uint32_t best_i = 0;
  auto consider = [&](uint32_t index) {
    float t = hit_distance(index, origin, dir, max, alpha);
    if (std::isfinite(t) && (!have || t < best_t || (t == best_t && index < best_i))) {
      have = true; best_t = t; best_i = index;
    }
  };
  if (brute || nodes.empty()) {
    for (uint32_t i = 0; i < tris.size(); ++i) consider(i);
  } else {
    RayPre r = ray_pre(origin, dir);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const BvhNode& nd = nodes[stack[--sp]];
      if (!slab(r, nd.lo, nd.hi, have ? best_t : max)) continue;
      if (nd.left < 0) {
        for (uint32_t k = nd.first; k < nd.first + nd.count; ++k) consider(prim[k]);
      } else {
        stack[sp++] = (uint32_t)nd.left;
        stack[sp++] = (uint32_t)nd.left + 1;
      }
    }
  }
  if (!have) return false;
  float t2;
  V3 w;
  if (!intersect(tris[best_i].v, origin, dir, INFINITY, &t2, &w)) return false;
  *dist = best_t;
  *occ = best_i;
  if (bary) *bary = w;
  return true;
}

// ------------------------------------------------------------------------------------------
// Thread pool: rayon-like data parallelism; nested calls run inline.
// ------------------------------------------------------------------------------------------
static unsigned g_threads = 1;
static thread_local bool tl_worker = false;
template <class F>
static void par_for(size_t n, F&& f) {
  if (tl_worker || g_threads <= 1 || n <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  auto work = [&]() {
    bool was = tl_worker;
    tl_worker = true;
    for (;;) {
      size_t i = next.fetch_add(1);
      if (i >= n) break;
      f(i);
    }
    tl_worker = was;
  };
  unsigned k = (unsigned)std::min<size_t>(g_threads, n);
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < k; ++t) ts.emplace_back(work);
  work();
  for (auto& t : ts) t.join();
}

// ------------------------------------------------------------------------------------------
// Level grids: shadow_grid.rs:13-178, bits.rs:43-132
// ------------------------------------------------------------------------------------------
static inline uint8_t level(float fraction, uint32_t bits) {
  float mx = (float)((1u << bits) - 1);
  return as_u8(std::round(clampf(fraction, 0.0f, 1.0f) * mx));
}
struct Grid {
  uint32_t w = 0, h = 0, bits = 1;
  std::vector<uint8_t> lv;
  Grid() {}
  Grid(uint32_t w_, uint32_t h_, uint32_t b) : w(w_), h(h_), bits(b <= 1 ? 1 : std::min(b, 8u)), lv((size_t)w_ * h_, 0) {}
  uint8_t get(uint32_t x, uint32_t y) const { return lv[(size_t)x + (size_t)y * w]; }
  uint8_t get_or_zero(int64_t x, int64_t y) const {
    if (x >= 0 && y >= 0 && x < (int64_t)w && y < (int64_t)h) return get((uint32_t)x, (uint32_t)y);
    return 0;
  }
  void set(uint32_t x, uint32_t y, uint8_t v) { lv[(size_t)x + (size_t)y * w] = bits == 1 ? (v != 0) : v; }
  std::vector<uint32_t> words() const {
    uint32_t per = 32 / bits;
    std::vector<uint32_t> out((lv.size() + per - 1) / per, 0);
    for (size_t i = 0; i < lv.size(); ++i) out[i / per] |= (uint32_t)lv[i] << ((i % per) * bits);
    return out;
  }
};
struct AtomicGrid {
  uint32_t w, h, bits;
  std::unique_ptr<std::atomic<uint8_t>[]> lv;
  AtomicGrid(uint32_t w_, uint32_t h_, uint32_t b) : w(w_), h(h_), bits(b <= 1 ? 1 : std::min(b, 8u)), lv(new std::atomic<uint8_t>[(size_t)w_ * h_]) {
    for (size_t i = 0; i < (size_t)w * h; ++i) lv[i].store(0, std::memory_order_relaxed);
  }
  void raise(uint32_t x, uint32_t y, uint8_t v) {
    if (bits == 1 && v == 0) return;
    std::atomic<uint8_t>& c = lv[(size_t)x + (size_t)y * w];
    uint8_t cur = c.load(std::memory_order_relaxed);
    while (cur < v && !c.compare_exchange_weak(cur, v, std::memory_order_relaxed)) {}
  }
  Grid into_grid() const {
    Grid g(w, h, bits);
    for (size_t i = 0; i < (size_t)w * h; ++i) g.lv[i] = lv[i].load(std::memory_order_relaxed);
    return g;
  }
};
struct Region {
  Rect rect;
  Grid grid;
  uint8_t get(int64_t x, int64_t y) const { return grid.get_or_zero(x - rect.min_x, y - rect.min_y); }
};
struct BitGrid {
  uint32_t w = 0, h = 0;
  std::vector<uint8_t> b;
  bool get_or_false(int64_t x, int64_t y) const {
    return x >= 0 && y >= 0 && x < (int64_t)w && y < (int64_t)h && b[(size_t)x + (size_t)y * w];
  }
};

// seams.rs:15-39
static void pad_seams(Region& region, const BitGrid& visited) {
  Rect r = region.rect;
  for (uint32_t yy = r.min_y; yy <= r.max_y; ++yy)
    for (uint32_t xx = r.min_x; xx <= r.max_x; ++xx) {
      int64_t x = xx, y = yy;
      if (visited.get_or_false(x, y)) continue;
      uint8_t lv = 0;
      const int d[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
      for (auto& dd : d) {
        int64_t nx = x + dd[0], ny = y + dd[1];
        if (visited.get_or_false(nx, ny)) lv = std::max(lv, region.get(nx, ny));
        else if (visited.get_or_false(x + 2 * dd[0], y + 2 * dd[1]))
          lv = std::max(lv, region.get(x + 2 * dd[0], y + 2 * dd[1]));
      }
      region.grid.set(xx - r.min_x, yy - r.min_y, lv);
    }
}
// seams.rs:43-95
static uint8_t dilated(const Grid& c, uint32_t x, uint32_t y) {
  uint8_t m = 0;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) m = std::max(m, c.get_or_zero((int64_t)x + dx, (int64_t)y + dy));
  return m;
}
static void dilate_row(Grid& g, const Grid& c, uint32_t y) { for (uint32_t x = 0; x < g.w; ++x) g.set(x, y, dilated(c, x, y)); }
static void dilate_col(Grid& g, const Grid& c, uint32_t x) { for (uint32_t y = 0; y < g.h; ++y) g.set(x, y, dilated(c, x, y)); }
static void dilate_shadows(Grid& g) {
  uint32_t w = g.w, h = g.h;
  Grid copy = g;
  if (h >= 2) dilate_row(g, copy, 1);
  if (h >= 4) dilate_row(g, copy, h - 2);
  if (w >= 2) dilate_col(g, copy, 1);
  if (w >= 4) dilate_col(g, copy, w - 2);
  copy = g;
  if (h >= 1) dilate_row(g, copy, 0);
  if (h >= 3) dilate_row(g, copy, h - 1);
  if (w >= 1) dilate_col(g, copy, 0);
  if (w >= 3) dilate_col(g, copy, w - 1);
}
// shadows.rs:148-160
static Grid crop_shadows(const Region& region, Rect crop) {
  Grid g(crop.width() + 4, crop.height() + 4, region.grid.bits);
  for (uint32_t y = crop.min_y; y <= crop.max_y; ++y)
    for (uint32_t x = crop.min_x; x <= crop.max_x; ++x) g.set(x - crop.min_x + 2, y - crop.min_y + 2, region.get(x, y));
  dilate_shadows(g);
  return g;
}
// shadows.rs:136-146
static bool any_light(const Region& region, Rect rect) {
  Rect r = region.rect;
  uint32_t min_x = std::max(rect.min_x, r.min_x), max_x = std::min(rect.max_x, r.max_x);
  uint32_t min_y = std::max(rect.min_y, r.min_y), max_y = std::min(rect.max_y, r.max_y);
  if (min_x > max_x || min_y > max_y) return false;
  for (uint32_t y = min_y; y <= max_y; ++y)
    for (uint32_t x = min_x; x <= max_x; ++x)
      if (region.grid.get(x - r.min_x, y - r.min_y) != 0) return true;
  return false;
}

// ------------------------------------------------------------------------------------------
// Bounce packing: bounce_pack.rs:18-86
// ------------------------------------------------------------------------------------------
static inline uint32_t bounce_bits(int32_t c) { return c == DL_BOUNCE_BITS8 ? 8 : c == DL_BOUNCE_BITS6 ? 6 : c == DL_BOUNCE_BITS5 ? 5 : 4; }
static inline float dither_amplitude(int32_t c) {
  return c == DL_BOUNCE_BITS8 ? 0.0f : c == DL_BOUNCE_BITS6 ? 0.004f : c == DL_BOUNCE_BITS5 ? 0.008f : 0.02f;
}
static inline uint32_t quantize(float value, int32_t c, Rng& rng) {
  float mx = (float)((1u << bounce_bits(c)) - 1);
  float amplitude = dither_amplitude(c);
  float color = std::isnan(value) ? 0.0f : clampf(value, 0.0f, 1.0f);
  if (amplitude > 0.0f && color > 0.001f) color = fmin_(color + rng.next_f32() * amplitude, 1.0f);
  return as_u32(std::sqrt(color) * mx);
}
static std::vector<uint32_t> pack_bounce(const std::vector<float>& values, int32_t c, Rng& rng) {
  uint32_t bits = bounce_bits(c), per = 32 / bits;
  std::vector<uint32_t> out((values.size() + per - 1) / per, 0);
  for (size_t i = 0; i < values.size(); ++i) out[i / per] |= quantize(values[i], c, rng) << ((i % per) * bits);
  return out;
}
static std::vector<uint32_t> pack_bounce_rgb(const std::vector<V3>& values, int32_t c, Rng& rng) {
  uint32_t bits = bounce_bits(c), pixel_bits = 3 * bits;
  std::vector<uint32_t> words((values.size() * pixel_bits + 31) / 32, 0);
  for (size_t i = 0; i < values.size(); ++i) {
    uint32_t pixel = 0;
    float ch[3] = {values[i].x, values[i].y, values[i].z};
    for (int k = 0; k < 3; ++k) pixel |= quantize(ch[k], c, rng) << (k * bits);
    size_t offset = i * pixel_bits;
    size_t word = offset / 32;
    uint32_t shift = (uint32_t)(offset % 32);
    words[word] |= pixel << shift;
    if (shift + pixel_bits > 32) words[word + 1] |= pixel >> (32 - shift);
  }
  return words;
}

// ------------------------------------------------------------------------------------------
// Cube conventions: shared/cube.rs:29-105; f16 packing: distance_cubes.rs:23-26 (half 2.7.1)
// ------------------------------------------------------------------------------------------
static void direction_to_face_uv(V3 d, uint32_t* face, V2* uv) {
  V3 a = vabs(d);
  float sc, tc, ma;
  if (a.x > a.y && a.x > a.z) {
    if (d.x > 0.0f) { *face = 0; sc = -d.z; tc = -d.y; } else { *face = 1; sc = d.z; tc = -d.y; }
    ma = a.x;
  } else if (a.y > a.z) {
    if (d.y > 0.0f) { *face = 2; sc = d.x; tc = d.z; } else { *face = 3; sc = d.x; tc = -d.z; }
    ma = a.y;
  } else {
    if (d.z > 0.0f) { *face = 4; sc = d.x; tc = -d.y; } else { *face = 5; sc = -d.x; tc = -d.y; }
    ma = a.z;
  }
  if (ma == 0.0f) { *uv = v2(0.5f, 0.5f); return; }
  *uv = v2((sc / ma) * 0.5f + 0.5f, (tc / ma) * 0.5f + 0.5f);
}
static V3 face_uv_to_direction(uint32_t face, V2 uv) {
  float sc = uv.x * 2.0f - 1.0f, tc = uv.y * 2.0f - 1.0f;
  V3 d;
  switch (face) {
    case 0: d = v3(1.0f, -tc, -sc); break;
    case 1: d = v3(-1.0f, -tc, sc); break;
    case 2: d = v3(sc, 1.0f, tc); break;
    case 3: d = v3(sc, -1.0f, -tc); break;
    case 4: d = v3(sc, -tc, 1.0f); break;
    default: d = v3(-sc, -tc, -1.0f); break;
  }
  return normalize(d);
}
static V3 texel_direction(uint32_t face, uint32_t x, uint32_t y, uint32_t res) {
  float size = (float)res;
  return face_uv_to_direction(face, v2(((float)x + 0.5f) / size, ((float)y + 0.5f) / size));
}
static void direction_to_texel(V3 d, uint32_t res, uint32_t* face, uint32_t* x, uint32_t* y) {
  V2 uv;
  direction_to_face_uv(d, face, &uv);
  float mx = (float)(res == 0 ? 0 : res - 1);
  *x = as_u32(clampf(std::floor(uv.x * (float)res), 0.0f, mx));
  *y = as_u32(clampf(std::floor(uv.y * (float)res), 0.0f, mx));
}
static uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t exp = (x >> 23) & 0xffu;
  uint32_t man = x & 0x7fffffu;
  if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (man ? (0x200u | (man >> 13)) : 0u));
  int32_t e = (int32_t)exp - 127 + 15;
  if (e >= 0x1f) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - e);
    uint32_t half = man >> shift;
    uint32_t rem = man & ((1u << shift) - 1);
    uint32_t mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (half & 1))) half++;
    return (uint16_t)(sign | half);
  }
  uint32_t half = ((uint32_t)e << 10) | (man >> 13);
  uint32_t rem = man & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (half & 1))) half++;
  return (uint16_t)(sign | half);
}
static inline uint32_t pack_f16_pair(float lower, float upper) {
  return ((uint32_t)f32_to_f16(upper) << 16) | (uint32_t)f32_to_f16(lower);
}

// ------------------------------------------------------------------------------------------
// Scene and per-mesh bake: baker.rs:27-203, mesh_bake.rs:73-267
// ------------------------------------------------------------------------------------------
struct Prepared {
  const dl_mesh* mesh;
  uint32_t resolution;
  std::vector<V2> uv1;
  uint32_t tri_count() const { return mesh->index_count / 3; }
  void vertices(uint32_t t, V3* out) const {
    const uint32_t* c = mesh->indices + t * 3;
    for (int k = 0; k < 3; ++k) out[k] = v3p(mesh->positions + c[k] * 3);
  }
  void normals(uint32_t t, V3* out) const {
    if (!mesh->normals || mesh->normal_count == 0) {
      V3 v[3];
      vertices(t, v);
      V3 f = normalize_or_zero(cross(v[1] - v[0], v[2] - v[0]));
      out[0] = out[1] = out[2] = f;
      return;
    }
    const uint32_t* c = mesh->indices + t * 3;
    for (int k = 0; k < 3; ++k) out[k] = v3p(mesh->normals + c[k] * 3);
  }
  void uvs(uint32_t t, V2* out) const {
    const uint32_t* c = mesh->indices + t * 3;
    for (int k = 0; k < 3; ++k) out[k] = uv1[c[k]];
  }
};
static Prepared prepare(const dl_mesh* m, const dl_bake_options& opt) {
  Prepared p;
  p.mesh = m;
  float area = -0.0f;
  for (uint32_t t = 0; t < m->index_count / 3; ++t) {
    const uint32_t* c = m->indices + t * 3;
    area = area + triangle_area(v3p(m->positions + c[0] * 3), v3p(m->positions + c[1] * 3), v3p(m->positions + c[2] * 3));
  }
  p.uv1.resize(m->vertex_count);
  for (uint32_t i = 0; i < m->vertex_count; ++i) {
    V2 uv = v2(m->uv1[i * 2], m->uv1[i * 2 + 1]);
    if (!(std::isfinite(uv.x) && std::isfinite(uv.y))) uv = v2(0, 0);
    p.uv1[i] = uv;
  }
  p.resolution = lightmap_size(area, opt.pixel_density, opt.max_lightmap_size);
  return p;
}

struct PhotonCube {
  uint32_t res;
  std::vector<float> dist;
  std::vector<int8_t> nrm;   // 3 per photon
  std::vector<uint8_t> alb;  // 3 per photon
};

struct Scene {
  const dl_bake_options* opt;
  std::vector<Light> lights;  // BVH order, then suns, then skies
  size_t local_lights;
  LightBvh light_bvh;
  Occluders occ;
  bool bounce_in_scene;
  std::vector<std::unique_ptr<PhotonCube>> cubes;  // by slot; null when none
};

struct Setup {
  std::vector<uint32_t> lights;
  V3 face;
  Rect crop;
};

// mesh_bake.rs:141-173
static Setup associate(const Scene& sc, const Prepared& pm, uint32_t t) {
  V3 v[3], n[3];
  pm.vertices(t, v);
  pm.normals(t, n);
  Setup s;
  s.face = normalize_or_zero(cross(v[1] - v[0], v[2] - v[0]));
  bool valid = s.face != v3(0, 0, 0);
  bool flat = approx_eq(n[0], n[1], 1e-5f) && approx_eq(n[1], n[2], 1e-5f);
  Aabb box = Aabb::empty().grow(v[0]).grow(v[1]).grow(v[2]);
  std::vector<uint32_t> cand;
  for (uint32_t slot = 0; slot < sc.local_lights; ++slot)
    if (sc.light_bvh.bounds[slot].intersects(box)) cand.push_back(slot);
  for (uint32_t slot = (uint32_t)sc.local_lights; slot < sc.lights.size(); ++slot) cand.push_back(slot);
  for (uint32_t slot : cand) {
    const Light& L = sc.lights[slot];
    if (valid && flat && !L.bakes_bounce && !may_face(L, v, s.face)) continue;
    if (L.distant || sphere_intersects_triangle(L.position, L.radius, v[0], v[1], v[2])) s.lights.push_back(slot);
  }
  V2 uv[3];
  pm.uvs(t, uv);
  s.crop = crop_rect(uv, pm.resolution);
  return s;
}

struct MeshTexels {
  BitGrid owned;
  std::vector<std::vector<uint8_t>> buried;
};
static void centered(const Prepared& pm, uint32_t t, V3 face, std::vector<SurfaceTexel>& out) {
  V3 v[3], n[3];
  V2 uv[3];
  pm.vertices(t, v);
  pm.normals(t, n);
  pm.uvs(t, uv);
  triangle_texels(v, n, uv, face, pm.resolution, out);
}
static void traced(const Prepared& pm, uint32_t t, V3 face, const BitGrid& owned, std::vector<SurfaceTexel>& out) {
  centered(pm, t, face, out);
  V3 v[3], n[3];
  V2 uv[3];
  pm.vertices(t, v);
  pm.normals(t, n);
  pm.uvs(t, uv);
  uint32_t size = pm.resolution;
  touched_texels(v, n, uv, face, size, [&](uint32_t x, uint32_t y) { return owned.b[(size_t)x + (size_t)y * size] != 0; }, out);
}
// mesh_texels.rs:111-135
template <class BackFace>
static bool is_buried(V3 normal, float turn, BackFace back_face) {
  V3 tangent, bitangent;
  any_orthonormal_pair(normal, &tangent, &bitangent);
  uint32_t back = 0, front = 0;
  for (uint32_t i = 0; i < 8; ++i) {
    float ring = (float)((i * 5) % 8) + 0.5f;
    float sin_theta = std::sqrt(ring / 8.0f);
    float cos_theta = std::sqrt(1.0f - sin_theta * sin_theta);
    float phi = (((float)i + turn) * TAU) / 8.0f;
    V3 direction = (tangent * cosf(phi) + bitangent * sinf(phi)) * sin_theta + normal * cos_theta;
    if (back_face(direction)) back++;
    else front++;
    if (back >= 5) return true;
    if (front > 8 - 5) return false;
  }
  return false;
}
static const uint64_t BURIAL_STREAM = 0x627572696564ull;

static MeshTexels mesh_texels(const Scene& sc, const Prepared& pm, const std::vector<Setup>& setups) {
  MeshTexels mt;
  uint32_t size = pm.resolution;
  mt.owned.w = mt.owned.h = size;
  std::vector<std::atomic<uint8_t>> owned((size_t)size * size);
  for (auto& b : owned) b.store(0, std::memory_order_relaxed);
  par_for(setups.size(), [&](size_t t) {
    if (setups[t].face == v3(0, 0, 0)) return;
    std::vector<SurfaceTexel> tx;
    centered(pm, (uint32_t)t, setups[t].face, tx);
    for (auto& s : tx) owned[(size_t)s.x + (size_t)s.y * size].store(1, std::memory_order_relaxed);
  });
  mt.owned.b.resize((size_t)size * size);
  for (size_t i = 0; i < owned.size(); ++i) mt.owned.b[i] = owned[i].load();
  mt.buried.resize(setups.size());
  uint64_t seed = sc.opt->seed, id = pm.mesh->id;
  par_for(setups.size(), [&](size_t t) {
    V3 normal = setups[t].face;
    if (normal == v3(0, 0, 0)) return;
    std::vector<SurfaceTexel> tx;
    traced(pm, (uint32_t)t, normal, mt.owned, tx);
    auto& out = mt.buried[t];
    out.resize(tx.size());
    for (size_t k = 0; k < tx.size(); ++k) {
      V3 origin = tx[k].position + normal * RAY_OFFSET;
      Rng rng = stream({seed, id, BURIAL_STREAM, (uint64_t)t, (uint64_t)tx[k].x, (uint64_t)tx[k].y});
      auto back_face = [&](V3 d) {
        float dist;
        uint32_t occ;
        if (!sc.occ.closest_hit(origin, d, INFINITY, false, &dist, &occ, nullptr)) return false;
        return dot(sc.occ.face_normal(occ), d) > 0.0f;
      };
      out[k] = is_buried(normal, rng.next_f32(), back_face) ? 1 : 0;
    }
  });
  return mt;
}
static void texels_of(const MeshTexels& mt, const Prepared& pm, uint32_t t, V3 face, std::vector<SurfaceTexel>& out) {
  std::vector<SurfaceTexel> tx;
  traced(pm, t, face, mt.owned, tx);
  const auto& b = mt.buried[t];
  for (size_t k = 0; k < tx.size(); ++k)
    if (!(k < b.size() && b[k])) out.push_back(tx[k]);
}

// area.rs:98-126
template <class Trace, class Point>
static bool adaptive_share(const V2* probes, size_t nprobes, uint32_t samples, uint32_t minimum,
                           Point spread_point, Trace trace, float* fraction) {
  bool any_lit = false, any_blocked = false;
  for (size_t k = 0; k < nprobes; ++k) {
    float w;
    bool lit;
    if (trace(probes[k], &w, &lit)) { any_lit |= lit; any_blocked |= !lit; }
  }
  float lit = 0.0f, total = 0.0f;
  for (uint32_t taken = 0; taken < samples; ++taken) {
    if (taken == minimum && any_lit != any_blocked) break;
    float w;
    bool unblocked;
    if (trace(spread_point(taken), &w, &unblocked)) {
      total += w;
      lit += unblocked ? w : 0.0f;
      any_lit |= unblocked;
      any_blocked |= !unblocked;
    }
  }
  if (any_lit != any_blocked) { *fraction = any_lit ? 1.0f : 0.0f; return true; }
  if (total > 0.0f) { *fraction = lit / total; return true; }
  return false;
}
static const V2 RECT_PROBES[5] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}, {0.5f, 0.5f}};
static const V2 TUBE_PROBES[3] = {{0, 0.5f}, {1, 0.5f}, {0.5f, 0.5f}};
static const V2 DISK_PROBES[5] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1}};
static const float END_GAP = 1.0e-3f;

// area.rs:61-90 with toward area.rs:130-146
template <class Occ>
static bool texel_visibility(const AreaShape& shape, float reach, V3 position, V3 normal, uint32_t samples,
                             uint32_t minimum, Rng& rng, Occ occluded, float* fraction, uint32_t* rays_out) {
  if (reach <= 0.0f || length_sq(position - shape.nearest_point(position)) > reach * reach) return false;
  if (shape.rectangle && dot(shape.forward, position - shape.center) <= 0.0f) return false;
  uint32_t rays = 0;
  auto trace = [&](V2 uv, float* weight, bool* lit) {
    V3 to_point = shape.point_at(uv) - position;
    float dsq = length_sq(to_point);
    if (dsq < 1.0e-12f) return false;
    float distance = std::sqrt(dsq);
    V3 direction = to_point / distance;
    float receiver = dot(normal, direction) - NDOTL_TOLERANCE;
    float emitter = shape.rectangle ? -dot(shape.forward, direction) : 1.0f;
    if (!(receiver > 0.0f && emitter > 0.0f)) return false;
    *weight = (receiver * emitter) / dsq;
    rays++;
    *lit = !occluded(direction, fmax_(distance - END_GAP, 0.0f));
    return true;
  };
  V2 offset;
  offset.x = rng.next_f32();
  offset.y = rng.next_f32();
  bool rect = shape.rectangle;
  auto sp = [&](uint32_t taken) { return spread(taken, offset, rect); };
  bool ok = adaptive_share(rect ? RECT_PROBES : TUBE_PROBES, rect ? 5 : 3, samples, minimum, sp, trace, fraction);
  *rays_out = rays;
  return ok;
}
// sun.rs:21-56
template <class Occ>
static bool sun_visibility(const Light& L, V3 normal, uint32_t samples, uint32_t minimum, Rng& rng, Occ occluded,
                           float* fraction, uint32_t* rays_out) {
  V3 toward = L.sun_toward;
  float radius = L.sun_radius;
  if (radius <= 0.0f) {
    if (dot(normal, toward) <= NDOTL_TOLERANCE) return false;
    bool lit = !occluded(toward);
    *fraction = lit ? 1.0f : 0.0f;
    *rays_out = 1;
    return true;
  }
  V3 side, up;
  any_orthonormal_pair(toward, &side, &up);
  float tn = L.sun_tan;
  uint32_t rays = 0;
  auto trace = [&](V2 disk, float* weight, bool* lit) {
    V3 direction = normalize(toward + (side * disk.x + up * disk.y) * tn);
    float receiver = dot(normal, direction) - NDOTL_TOLERANCE;
    if (receiver <= 0.0f) return false;
    rays++;
    *weight = receiver;
    *lit = !occluded(direction);
    return true;
  };
  V2 offset;
  offset.x = rng.next_f32();
  offset.y = rng.next_f32();
  auto sp = [&](uint32_t taken) { return concentric_disk(spread(taken, offset, true)); };
  bool ok = adaptive_share(DISK_PROBES, 5, samples, minimum, sp, trace, fraction);
  *rays_out = rays;
  return ok;
}
// sky.rs:20-48
static V3 cosine_direction(V3 normal, V2 square) {
  V2 disk = concentric_disk(square);
  float up = std::sqrt(fmax_(1.0f - (disk.x * disk.x + disk.y * disk.y), 0.0f));
  V3 side, across;
  any_orthonormal_pair(normal, &side, &across);
  return normalize((side * disk.x + across * disk.y) + normal * up);
}
template <class Occ>
static float sky_visibility(V3 normal, uint32_t samples, Rng& rng, Occ occluded) {
  V2 offset;
  offset.x = rng.next_f32();
  offset.y = rng.next_f32();
  uint32_t open = 0;
  for (uint32_t taken = 0; taken < samples; ++taken)
    if (!occluded(cosine_direction(normal, spread(taken, offset, true)))) open++;
  return (float)open / (float)std::max(samples, 1u);
}

// texel_shadow.rs:20-90; returns false for None
static bool texel_shadow(const Scene& sc, const Prepared& pm, const Light& L, const SurfaceTexel& tx, V3 origin,
                         uint8_t* lit, uint32_t* rays) {
  auto occluded = [&](V3 d, float dist) { return sc.occ.occluded(origin, d, dist, L.alpha_tested); };
  if (L.shape == DL_SHAPE_POINT) {
    V3 to_light = L.position - tx.position;
    float dsq = length_sq(to_light);
    if (dsq > L.radius * L.radius) return false;
    float distance = std::sqrt(dsq);
    if (distance == 0.0f) { *lit = 1; *rays = 0; return true; }
    V3 direction = to_light / distance;
    if (dot(tx.normal, direction) <= NDOTL_TOLERANCE) return false;
    *lit = occluded(direction, distance) ? 0 : 1;
    *rays = 1;
    return true;
  }
  Rng rng = stream({sc.opt->seed, pm.mesh->id, L.id, (uint64_t)tx.x, (uint64_t)tx.y});
  float fraction;
  uint32_t r = 0;
  if (L.area) {
    if (!texel_visibility(L.as, L.reach, tx.position, tx.normal, sc.opt->area_light_samples,
                          sc.opt->area_light_min_samples, rng, occluded, &fraction, &r))
      return false;
  } else if (L.sun) {
    auto towards = [&](V3 d) { return occluded(d, INFINITY); };
    if (!sun_visibility(L, tx.normal, sc.opt->area_light_samples, sc.opt->area_light_min_samples, rng, towards, &fraction, &r))
      return false;
  } else {
    auto towards = [&](V3 d) { return occluded(d, INFINITY); };
    fraction = sky_visibility(tx.normal, sc.opt->sky_samples, rng, towards);
    r = sc.opt->sky_samples;
  }
  *lit = level(fraction, L.bits);
  *rays = r;
  return true;
}

// bounce.rs:48-115 photon cube
static void build_photon_cube(const Scene& sc, const Light& L, uint32_t res, PhotonCube& cube) {
  cube.res = res;
  size_t n = (size_t)6 * res * res;
  cube.dist.assign(n, 0.0f);
  cube.nrm.assign(n * 3, 0);
  cube.alb.assign(n * 3, 0);
  par_for((size_t)6 * res, [&](size_t row) {
    uint32_t face = (uint32_t)(row / res), y = (uint32_t)(row % res);
    for (uint32_t x = 0; x < res; ++x) {
      size_t i = (size_t)face * res * res + (size_t)y * res + x;
      V3 direction = texel_direction(face, x, y, res);
      float dist;
      uint32_t occ;
      V3 w;
      if (!sc.occ.closest_hit(L.position, direction, L.radius, L.alpha_tested, &dist, &occ, &w)) {
        cube.dist[i] = L.radius;
        continue;
      }
      V3 normal = sc.occ.surface_normal(occ, w);
      V3 point = L.position + direction * dist;
      float bias = fmax_(dist * 0.001f, 0.001f);
      V3 nudged = (point - direction * bias) + normal * bias;
      bool lit_side = dot(sc.occ.face_normal(occ), direction) < 0.0f;
      cube.dist[i] = length(nudged - L.position);
      if (lit_side) {
        cube.nrm[i * 3 + 0] = as_i8(std::round(normal.x * 127.0f));
        cube.nrm[i * 3 + 1] = as_i8(std::round(normal.y * 127.0f));
        cube.nrm[i * 3 + 2] = as_i8(std::round(normal.z * 127.0f));
      }
      V3 a = sc.occ.albedo(occ);
      cube.alb[i * 3 + 0] = as_u8(std::round(clampf(a.x, 0.0f, 1.0f) * 255.0f));
      cube.alb[i * 3 + 1] = as_u8(std::round(clampf(a.y, 0.0f, 1.0f) * 255.0f));
      cube.alb[i * 3 + 2] = as_u8(std::round(clampf(a.z, 0.0f, 1.0f) * 255.0f));
    }
  });
}
static void cube_sample(const PhotonCube& c, V3 light, V3 direction, V3* pos, V3* nrm, V3* alb) {
  uint32_t face, x, y;
  direction_to_texel(direction, c.res, &face, &x, &y);
  size_t index = (size_t)face * c.res * c.res + (size_t)y * c.res + x;
  if (index >= c.dist.size()) { *pos = light; *nrm = v3(0, 0, 0); *alb = v3(0, 0, 0); return; }
  V3 n = v3((float)c.nrm[index * 3] / 127.0f, (float)c.nrm[index * 3 + 1] / 127.0f, (float)c.nrm[index * 3 + 2] / 127.0f);
  *pos = light + direction * c.dist[index];
  *nrm = normalize_or_zero(n);
  *alb = v3((float)c.alb[index * 3] / 255.0f, (float)c.alb[index * 3 + 1] / 255.0f, (float)c.alb[index * 3 + 2] / 255.0f);
}
// glam acos_approx (DirectXMath XMScalarAcos) -- parity unpinned
static float acos_approx(float v) {
  bool nonneg = v >= 0.0f;
  float x = std::fabs(v);
  float omx = 1.0f - x;
  if (omx < 0.0f) omx = 0.0f;
  float root = std::sqrt(omx);
  float r = ((((((-0.0012624911f * x + 0.0066700901f) * x - 0.0170881256f) * x + 0.0308918810f) * x -
               0.0501743046f) * x + 0.0889789874f) * x - 0.2145988016f) * x + 1.5707963050f;
  r *= root;
  return nonneg ? r : PI_F - r;
}
static inline float lerpf(float a, float b, float s) { return a + (b - a) * s; }
static inline V3 lerp3(V3 a, V3 b, float s) { return a + (b - a) * s; }
// glam Vec3::slerp -- parity unpinned (exact glam formula not available here)
static V3 slerp(V3 a, V3 b, float s) {
  float la = length(a), lb = length(b);
  float d = dot(a, b) / (la * lb);
  if (std::fabs(d) < 1.0f - 3.0e-7f) {
    float theta = acos_approx(d);
    float sin_theta = sinf(theta);
    float t1 = sinf(theta * (1.0f - s));
    float t2 = sinf(theta * s);
    float rl = lerpf(la, lb, s);
    return ((a * (rl / la)) * t1 + (b * (rl / lb)) * t2) * (1.0f / sin_theta);
  }
  if (d < 0.0f) {
    // almost opposite: rotate about any orthogonal axis by PI*s
    V3 axis;
    {
      // Vec3::any_orthogonal_vector
      V3 o = std::fabs(a.x) > std::fabs(a.y) ? v3(-a.z, 0.0f, a.x) : v3(0.0f, a.z, -a.y);
      axis = normalize(o);
    }
    float half = PI_F * s * 0.5f;
    float sn = sinf(half), cs = cosf(half);
    float q[4] = {axis.x * sn, axis.y * sn, axis.z * sn, cs};
    float rl = lerpf(la, lb, s);
    return quat_mul(q, a) * (rl / la);
  }
  return lerp3(a, b, s);
}

// BounceMap (bounce_map.rs:22-137) over a value type T: V3 for the bake, D3 for the FP64 checker.
template <class T>
struct BounceMapT {
  using S = decltype(T{}.x);
  uint32_t size;
  Rect rect;
  std::vector<T> values;
  BounceMapT(uint32_t size_, Rect crops) : size(size_) {
    uint32_t last = size - 1, m = 4;
    rect = {crops.min_x >= m ? crops.min_x - m : 0, crops.min_y >= m ? crops.min_y - m : 0,
            std::min(crops.max_x + m, last), std::min(crops.max_y + m, last)};
    values.assign((size_t)rect.width() * rect.height(), T{});
  }
  size_t index(uint32_t x, uint32_t y) const { return (size_t)(y - rect.min_y) * rect.width() + (x - rect.min_x); }
  T get(int64_t x, int64_t y) const { return rect.contains(x, y) ? values[index((uint32_t)x, (uint32_t)y)] : T{}; }
  void set(uint32_t x, uint32_t y, T v) { values[index(x, y)] = v; }
  void dilate() {
    static const int ORD[9][2] = {{0, 0}, {-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}};
    std::vector<T> out(values.size());
    size_t k = 0;
    for (uint32_t y = rect.min_y; y <= rect.max_y; ++y)
      for (uint32_t x = rect.min_x; x <= rect.max_x; ++x) {
        T r{};
        for (auto& d : ORD) {
          T v = get((int64_t)x + d[0], (int64_t)y + d[1]);
          if (max_element(v) > 0) { r = v; break; }
        }
        out[k++] = r;
      }
    values.swap(out);
  }
  void blur(const S* weights, S weight_sum) {
    int64_t last = (int64_t)size - 1;
    for (int vertical = 0; vertical < 2; ++vertical) {
      std::vector<T> out(values.size());
      size_t k = 0;
      for (uint32_t y = rect.min_y; y <= rect.max_y; ++y)
        for (uint32_t x = rect.min_x; x <= rect.max_x; ++x) {
          T sum{};
          for (int kk = -3; kk <= 3; ++kk) {
            T s = vertical ? get(x, std::min(std::max((int64_t)y + kk, (int64_t)0), last))
                           : get(std::min(std::max((int64_t)x + kk, (int64_t)0), last), y);
            sum = sum + s * weights[kk + 3];
          }
          out[k++] = sum / weight_sum;
        }
      values.swap(out);
    }
  }
  bool crop(Rect c, std::vector<T>& grid) const {
    uint32_t width = c.width() + 4;
    grid.assign((size_t)width * (c.height() + 4), T{});
    bool lit = false;
    for (uint32_t y = c.min_y; y <= c.max_y; ++y)
      for (uint32_t x = c.min_x; x <= c.max_x; ++x) {
        T v = get(x, y);
        lit |= max_element(v) > 0;
        grid[(size_t)(y - c.min_y + 2) * width + (x - c.min_x + 2)] = v;
      }
    return lit;
  }
};
using BounceMap = BounceMapT<V3>;
static void gaussian_weights(float* w, float* sum_out) {
  float sigma2 = 5.0f * 5.0f;
  float sum = 0.0f;
  for (int i = -3; i <= 3; ++i) {
    w[i + 3] = (1.0f / std::sqrt((2.0f * PI_F) * sigma2)) * expf(-(float)(i * i) / (2.0f * sigma2));
    sum += w[i + 3];
  }
  float s2 = 0.0f;
  for (int i = 0; i < 7; ++i) { w[i] = w[i] / sum; s2 += w[i]; }
  *sum_out = s2;
}

struct Gathered {
  std::vector<std::pair<std::pair<uint32_t, uint32_t>, V3>> texels;
  uint64_t rays = 0;
};
static inline V3 reflected(const Scene& sc, V3 albedo) { return sc.opt->colored_bounce ? albedo : v3(1, 1, 1); }

// bounce.rs:235-292
static Gathered gather_triangle(const Scene& sc, const std::vector<SurfaceTexel>& texels, const Setup& setup,
                                const Prepared& pm, const Light& L, const PhotonCube& cube) {
  Gathered g;
  if (setup.face == v3(0, 0, 0)) return g;
  uint32_t samples = std::max(L.src->bounce_samples, 2u);
  float step = 0.9f / (float)(samples - 1);
  V3 normal = setup.face;
  for (const auto& tx : texels) {
    V3 to_texel = tx.position - L.position;
    if (length_sq(to_texel) > L.radius * L.radius) continue;
    V3 direction = normalize_or(to_texel, v3(0, 1, 0));
    V3 origin = tx.position + normal * RAY_OFFSET;
    Rng rng = stream({sc.opt->seed, pm.mesh->id, L.id, (uint64_t)tx.x, (uint64_t)tx.y});
    V3 sum = v3(0, 0, 0);
    for (uint32_t i = 0; i < samples; ++i) {
      float sp = 0.1f + (float)i * step;
      V3 rnd = on_unit_sphere(rng);
      V3 sample = normalize_or(slerp(direction, rnd, sp), direction);
      V3 photon, pn, alb;
      cube_sample(cube, L.position, sample, &photon, &pn, &alb);
      V3 to_photon = photon - origin;
      float distance = length(to_photon);
      if (pn == v3(0, 0, 0) || distance <= 0.0f) continue;
      V3 d = to_photon / distance;
      g.rays++;
      if (!sc.occ.occluded(origin, d, distance, L.alpha_tested)) {
        float weight = fmax_(dot(normal, d), 0.0f) * fmax_(dot(pn, -d), 0.0f);
        sum = sum + reflected(sc, alb) * weight;
      }
    }
    V3 v = (sum * L.src->bounce_intensity) / (float)samples;
    g.texels.push_back({{tx.x, tx.y}, v});
  }
  return g;
}
// sun_bounce.rs:24-80
static Gathered gather_sun(const Scene& sc, const std::vector<SurfaceTexel>& texels, const Setup& setup,
                           const Prepared& pm, const Light& L) {
  Gathered g;
  if (setup.face == v3(0, 0, 0)) return g;
  uint32_t samples = std::max(L.src->bounce_samples, 2u);
  V3 normal = setup.face, toward = L.sun_toward;
  for (const auto& tx : texels) {
    V3 origin = tx.position + normal * RAY_OFFSET;
    Rng rng = stream({sc.opt->seed, pm.mesh->id, L.id, (uint64_t)tx.x, (uint64_t)tx.y});
    V2 offset;
    offset.x = rng.next_f32();
    offset.y = rng.next_f32();
    V3 sum = v3(0, 0, 0);
    for (uint32_t i = 0; i < samples; ++i) {
      V3 direction = cosine_direction(normal, spread(i, offset, true));
      g.rays++;
      float dist;
      uint32_t occ;
      V3 w;
      if (!sc.occ.closest_hit(origin, direction, INFINITY, L.alpha_tested, &dist, &occ, &w)) continue;
      V3 face = sc.occ.face_normal(occ);
      if (dot(face, direction) >= 0.0f) continue;
      float lit = dot(sc.occ.surface_normal(occ, w), toward);
      if (lit <= 0.0f || dot(face, toward) <= 0.0f) continue;
      V3 point = (origin + direction * dist) + face * RAY_OFFSET;
      g.rays++;
      if (!sc.occ.occluded(point, toward, INFINITY, L.alpha_tested)) sum = sum + reflected(sc, sc.occ.albedo(occ)) * lit;
    }
    V3 v = (sum * L.src->bounce_intensity) / (float)samples;
    g.texels.push_back({{tx.x, tx.y}, v});
  }
  return g;
}

// ------------------------------------------------------------------------------------------
// FP64 bounce: the same per-texel samples as gather_triangle / gather_sun (same random streams,
// same photon cubes, same sample counts), with every later operation in double: directions,
// slerp by its closed form, cube lookup, visibility by a double Moller-Trumbore scan of every
// occluder, weights, sums, and the window filter with double Gaussian weights.  It is the
// reference the f32 bounce's rounding error is measured against; nothing else reads it.
// ------------------------------------------------------------------------------------------
struct D3 { double x, y, z; };
static inline D3 d3(double x, double y, double z) { return {x, y, z}; }
static inline D3 d3(V3 v) { return {v.x, v.y, v.z}; }
static inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline D3 operator-(D3 a) { return {-a.x, -a.y, -a.z}; }
static inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline D3 operator/(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline double length(D3 a) { return std::sqrt(dot(a, a)); }
static inline double max_element(D3 a) { return std::fmax(a.x, std::fmax(a.y, a.z)); }
static inline D3 normalize_or(D3 a, D3 fallback) {
  double l = length(a);
  return (l > 0.0 && std::isfinite(l)) ? a / l : fallback;
}
static const double TAU_D = 6.28318530717958647692;
static const double PI_D = 3.14159265358979323846;

// on_unit_sphere (random.rs:22-28) from the same two f32 draws
static inline D3 on_unit_sphere_d(Rng& rng) {
  double z = (double)rng.next_f32() * 2.0 - 1.0;
  double angle = (double)rng.next_f32() * TAU_D;
  double r = std::sqrt(std::fmax(1.0 - z * z, 0.0));
  return d3(r * std::cos(angle), r * std::sin(angle), z);
}
// spherical interpolation of directions with lengths interpolated linearly:
// (sin((1-s)t) a/|a| + sin(s t) b/|b|) / sin t * lerp(|a|, |b|, s)
static D3 slerp_d(D3 a, D3 b, double s) {
  double la = length(a), lb = length(b);
  double c = dot(a, b) / (la * lb);
  double rl = la + (lb - la) * s;
  if (std::fabs(c) < 1.0 - 1e-12) {
    double th = std::acos(c), sn = std::sin(th);
    return (a * (std::sin((1.0 - s) * th) / la) + b * (std::sin(s * th) / lb)) * (rl / sn);
  }
  if (c < 0.0) {  // opposite directions: a half turn about an axis orthogonal to a
    D3 o = std::fabs(a.x) > std::fabs(a.y) ? d3(-a.z, 0.0, a.x) : d3(0.0, a.z, -a.y);
    D3 k = o / length(o), u = a / la;
    double ang = PI_D * s;
    D3 r = u * std::cos(ang) + cross(k, u) * std::sin(ang) + k * (dot(k, u) * (1.0 - std::cos(ang)));
    return r * rl;
  }
  return a + (b - a) * s;
}
static void direction_to_texel_d(D3 d, uint32_t res, uint32_t* face, uint32_t* x, uint32_t* y) {
  double ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z), sc, tc, ma;
  if (ax > ay && ax > az) {
    if (d.x > 0) { *face = 0; sc = -d.z; tc = -d.y; } else { *face = 1; sc = d.z; tc = -d.y; }
    ma = ax;
  } else if (ay > az) {
    if (d.y > 0) { *face = 2; sc = d.x; tc = d.z; } else { *face = 3; sc = d.x; tc = -d.z; }
    ma = ay;
  } else {
    if (d.z > 0) { *face = 4; sc = d.x; tc = -d.y; } else { *face = 5; sc = -d.x; tc = -d.y; }
    ma = az;
  }
  double u = 0.5, v = 0.5;
  if (ma != 0.0) { u = (sc / ma) * 0.5 + 0.5; v = (tc / ma) * 0.5 + 0.5; }
  double mx = (double)(res == 0 ? 0 : res - 1);
  *x = (uint32_t)std::fmin(std::fmax(std::floor(u * res), 0.0), mx);
  *y = (uint32_t)std::fmin(std::fmax(std::floor(v * res), 0.0), mx);
}
static void cube_sample_d(const PhotonCube& c, D3 light, D3 direction, D3* pos, D3* nrm, D3* alb) {
  uint32_t face, x, y;
  direction_to_texel_d(direction, c.res, &face, &x, &y);
  size_t index = (size_t)face * c.res * c.res + (size_t)y * c.res + x;
  D3 n = d3(c.nrm[index * 3] / 127.0, c.nrm[index * 3 + 1] / 127.0, c.nrm[index * 3 + 2] / 127.0);
  *pos = light + direction * (double)c.dist[index];
  *nrm = normalize_or(n, d3(0, 0, 0));
  *alb = d3(c.alb[index * 3] / 255.0, c.alb[index * 3 + 1] / 255.0, c.alb[index * 3 + 2] / 255.0);
}
// two-sided Moller-Trumbore in double (occluder.rs:225-251)
static bool intersect_d(const OccTri& o, D3 origin, D3 dir, double max, double* t_out, D3* bary) {
  D3 v0 = d3(o.v[0]), e1 = d3(o.v[1]) - v0, e2 = d3(o.v[2]) - v0;
  D3 p = cross(dir, e2);
  double det = dot(e1, p);
  if (det == 0.0 || !std::isfinite(det)) return false;
  double inv = 1.0 / det;
  D3 s = origin - v0;
  double u = dot(s, p) * inv;
  if (!(u >= 0.0 && u <= 1.0)) return false;
  D3 q = cross(s, e1);
  double v = dot(dir, q) * inv;
  if (v < 0.0 || u + v > 1.0) return false;
  double t = dot(e2, q) * inv;
  if (!(t >= 0.0 && t <= max)) return false;
  *t_out = t;
  if (bary) *bary = d3((1.0 - u) - v, u, v);
  return true;
}
static bool alpha_blocks_d(const Occluders& oc, const OccTri& o, D3 w, bool alpha) {
  return !alpha || oc.blocks(o, v3((float)w.x, (float)w.y, (float)w.z));
}
static bool occluded_d(const Occluders& oc, D3 origin, D3 dir, double max, bool alpha) {
  for (const OccTri& o : oc.tris) {
    double t;
    D3 w;
    if (intersect_d(o, origin, dir, max, &t, &w) && alpha_blocks_d(oc, o, w, alpha)) return true;
  }
  return false;
}
static bool closest_hit_d(const Occluders& oc, D3 origin, D3 dir, bool alpha, double* dist, uint32_t* occ, D3* bary) {
  bool hit = false;
  for (uint32_t i = 0; i < oc.tris.size(); ++i) {
    double t;
    D3 w;
    if (intersect_d(oc.tris[i], origin, dir, INFINITY, &t, &w) && alpha_blocks_d(oc, oc.tris[i], w, alpha) && (!hit || t < *dist)) {
      hit = true;
      *dist = t;
      *occ = i;
      *bary = w;
    }
  }
  return hit;
}
static D3 face_normal_d(const Occluders& oc, uint32_t i) {
  const OccTri& o = oc.tris[i];
  return normalize_or(cross(d3(o.v[1]) - d3(o.v[0]), d3(o.v[2]) - d3(o.v[0])), d3(0, 0, 0));
}
static D3 surface_normal_d(const Occluders& oc, uint32_t i, D3 w) {
  const OccTri& o = oc.tris[i];
  D3 face = face_normal_d(oc, i);
  const dl_mesh* m = oc.meshes[o.mesh];
  if (!m->normals || m->normal_count == 0) return face;
  const uint32_t* c = m->indices + o.triangle * 3;
  D3 n = d3(v3p(m->normals + c[0] * 3)) * w.x + d3(v3p(m->normals + c[1] * 3)) * w.y + d3(v3p(m->normals + c[2] * 3)) * w.z;
  return normalize_or(n, face);
}
static D3 cosine_direction_d(D3 normal, V2 square) {
  double px = square.x * 2.0 - 1.0, py = square.y * 2.0 - 1.0, dx = 0.0, dy = 0.0;
  if (px != 0.0 || py != 0.0) {
    double radius, angle;
    if (std::fabs(px) > std::fabs(py)) { radius = px; angle = (PI_D / 4) * (py / px); }
    else { radius = py; angle = PI_D / 2 - (PI_D / 4) * (px / py); }
    dx = std::cos(angle) * radius;
    dy = std::sin(angle) * radius;
  }
  double up = std::sqrt(std::fmax(1.0 - (dx * dx + dy * dy), 0.0));
  double sign = std::signbit(normal.z) ? -1.0 : 1.0;
  double a = -1.0 / (sign + normal.z), b = normal.x * normal.y * a;
  D3 side = d3(1.0 + sign * normal.x * normal.x * a, sign * b, -sign * normal.x);
  D3 across = d3(b, sign + normal.y * normal.y * a, -normal.y);
  D3 d = side * dx + across * dy + normal * up;
  return d / length(d);
}
struct GatheredD {
  std::vector<std::pair<std::pair<uint32_t, uint32_t>, D3>> texels;
};
static inline D3 reflected_d(const Scene& sc, D3 albedo) { return sc.opt->colored_bounce ? albedo : d3(1, 1, 1); }
static GatheredD gather_triangle_d(const Scene& sc, const std::vector<SurfaceTexel>& texels, const Setup& setup,
                                   const Prepared& pm, const Light& L, const PhotonCube& cube) {
  GatheredD g;
  if (setup.face == v3(0, 0, 0)) return g;
  uint32_t samples = std::max(L.src->bounce_samples, 2u);
  double step = 0.9 / (double)(samples - 1);
  D3 normal = d3(setup.face), light = d3(L.position);
  for (const auto& tx : texels) {
    D3 pos = d3(tx.position), to_texel = pos - light;
    if (dot(to_texel, to_texel) > (double)L.radius * (double)L.radius) continue;
    D3 direction = normalize_or(to_texel, d3(0, 1, 0));
    D3 origin = pos + normal * (double)RAY_OFFSET;
    Rng rng = stream({sc.opt->seed, pm.mesh->id, L.id, (uint64_t)tx.x, (uint64_t)tx.y});
    D3 sum = d3(0, 0, 0);
    for (uint32_t i = 0; i < samples; ++i) {
      double sp = 0.1 + (double)i * step;
      D3 rnd = on_unit_sphere_d(rng);
      D3 sample = normalize_or(slerp_d(direction, rnd, sp), direction);
      D3 photon, pn, alb;
      cube_sample_d(cube, light, sample, &photon, &pn, &alb);
      D3 to_photon = photon - origin;
      double distance = length(to_photon);
      if ((pn.x == 0 && pn.y == 0 && pn.z == 0) || distance <= 0.0) continue;
      D3 d = to_photon / distance;
      if (!occluded_d(sc.occ, origin, d, distance, L.alpha_tested))
        sum = sum + reflected_d(sc, alb) * (std::fmax(dot(normal, d), 0.0) * std::fmax(dot(pn, -d), 0.0));
    }
    g.texels.push_back({{tx.x, tx.y}, sum * (double)L.src->bounce_intensity / (double)samples});
  }
  return g;
}
static GatheredD gather_sun_d(const Scene& sc, const std::vector<SurfaceTexel>& texels, const Setup& setup,
                              const Prepared& pm, const Light& L) {
  GatheredD g;
  if (setup.face == v3(0, 0, 0)) return g;
  uint32_t samples = std::max(L.src->bounce_samples, 2u);
  D3 normal = d3(setup.face), toward = d3(L.sun_toward);
  for (const auto& tx : texels) {
    D3 origin = d3(tx.position) + normal * (double)RAY_OFFSET;
    Rng rng = stream({sc.opt->seed, pm.mesh->id, L.id, (uint64_t)tx.x, (uint64_t)tx.y});
    V2 offset;
    offset.x = rng.next_f32();
    offset.y = rng.next_f32();
    D3 sum = d3(0, 0, 0);
    for (uint32_t i = 0; i < samples; ++i) {
      D3 direction = cosine_direction_d(normal, spread(i, offset, true));
      double dist;
      uint32_t occ;
      D3 w;
      if (!closest_hit_d(sc.occ, origin, direction, L.alpha_tested, &dist, &occ, &w)) continue;
      D3 face = face_normal_d(sc.occ, occ);
      if (dot(face, direction) >= 0.0) continue;
      double lit = dot(surface_normal_d(sc.occ, occ, w), toward);
      if (lit <= 0.0 || dot(face, toward) <= 0.0) continue;
      D3 point = origin + direction * dist + face * (double)RAY_OFFSET;
      if (!occluded_d(sc.occ, point, toward, INFINITY, L.alpha_tested)) sum = sum + reflected_d(sc, d3(sc.occ.albedo(occ))) * lit;
    }
    g.texels.push_back({{tx.x, tx.y}, sum * (double)L.src->bounce_intensity / (double)samples});
  }
  return g;
}
static void gaussian_weights_d(double* w, double* sum_out) {
  double sigma2 = 25.0, sum = 0.0;
  for (int i = -3; i <= 3; ++i) {
    w[i + 3] = (1.0 / std::sqrt(2.0 * PI_D * sigma2)) * std::exp(-(double)(i * i) / (2.0 * sigma2));
    sum += w[i + 3];
  }
  double s2 = 0.0;
  for (int i = 0; i < 7; ++i) { w[i] /= sum; s2 += w[i]; }
  *sum_out = s2;
}

// Bounce capture (test hook): the crops the bounce packing reads, per (mesh, triangle, light).
// Mode 1 keeps the f32 crops of the bake, mode 2 the FP64 crops of the same entries.
struct CapturedCrop {
  uint64_t mesh_id, light_id;
  uint32_t triangle, width, height;
  std::vector<double> rgb;
};
static uint32_t g_capture_mode = 0;
static std::mutex g_capture_mu;
static std::vector<CapturedCrop> g_captured;
static std::vector<std::pair<uint64_t, PhotonCube>> g_cubes;  // the last bake's photon cubes, by light id
// Texel window (test hook): in mesh `g_window_mesh` only triangles whose crop meets the rectangle
// [x0, x1] x [y0, y1] are baked; the others keep neither texels nor lights.
static bool g_window_on = false;
static uint64_t g_window_mesh = 0;
static Rect g_window{0, 0, 0, 0};
static void capture_crop(uint64_t mesh_id, uint64_t light_id, uint32_t triangle, Rect crop, const double* rgb, size_t n) {
  CapturedCrop c{mesh_id, light_id, triangle, crop.width() + 4, crop.height() + 4, std::vector<double>(rgb, rgb + n)};
  std::lock_guard<std::mutex> lock(g_capture_mu);
  g_captured.push_back(std::move(c));
}

struct Entry {
  uint32_t light;
  bool has_bounce = false;
  std::vector<V3> bounce;
};
struct MeshOut {
  const dl_mesh* mesh;
  uint32_t resolution;
  std::vector<uint32_t> words;
  uint64_t texels = 0, rays = 0, bounce_rays = 0, lights_removed = 0, samples = 0;
};

// triangle_words.rs:47-89
struct TriLight {
  uint32_t light;
  bool has_shadow;
  std::vector<uint32_t> shadow;
  bool has_bounce;
  std::vector<uint32_t> bounce;
};
struct TriRecord {
  Rect crop;
  std::vector<TriLight> lights;
};
static std::vector<uint32_t> serialize_triangles(const std::vector<TriRecord>& tris, bool bounce_layout) {
  size_t stride = bounce_layout ? 3 : 2;
  std::vector<uint32_t> words;
  for (const auto& t : tris) {
    words.push_back(0);
    words.push_back(t.crop.min_x - 2u);
    words.push_back(t.crop.min_y - 2u);
    words.push_back(t.crop.width() + 4);
  }
  std::vector<size_t> light_data;
  for (size_t i = 0; i < tris.size(); ++i) {
    uint32_t offset = (uint32_t)words.size();
    words[i * 4] = offset;
    light_data.push_back(offset);
    words.push_back((uint32_t)tris[i].lights.size());
    for (const auto& l : tris[i].lights) {
      words.push_back(l.light);
      words.push_back(0);
      if (bounce_layout) words.push_back(0);
    }
  }
  for (size_t i = 0; i < tris.size(); ++i)
    for (size_t k = 0; k < tris[i].lights.size(); ++k) {
      size_t entry = light_data[i] + 1 + k * stride;
      const auto& l = tris[i].lights[k];
      if (l.has_shadow) {
        words[entry + 1] = (uint32_t)words.size();
        words.insert(words.end(), l.shadow.begin(), l.shadow.end());
      }
      if (bounce_layout && l.has_bounce) {
        words[entry + 2] = (uint32_t)words.size();
        words.insert(words.end(), l.bounce.begin(), l.bounce.end());
      }
    }
  return words;
}

static MeshOut bake_mesh(const Scene& sc, const Prepared& pm) {
  MeshOut out;
  out.mesh = pm.mesh;
  out.resolution = pm.resolution;
  uint32_t nt = pm.tri_count();
  uint32_t size = pm.resolution;
  size_t nl = sc.lights.size();
  std::vector<Setup> setups(nt);
  par_for(nt, [&](size_t t) { setups[t] = associate(sc, pm, (uint32_t)t); });
  if (g_window_on && pm.mesh->id == g_window_mesh)
    for (auto& st : setups)
      if (st.crop.max_x < g_window.min_x || st.crop.min_x > g_window.max_x || st.crop.max_y < g_window.min_y ||
          st.crop.min_y > g_window.max_y) {
        st.face = v3(0, 0, 0);
        st.lights.clear();
      }
  // light_regions mesh_bake.rs:175-192
  std::vector<int> have(nl, 0);
  std::vector<Rect> rects(nl);
  for (auto& s : setups)
    for (uint32_t l : s.lights) {
      if (!have[l]) { have[l] = 1; rects[l] = s.crop; }
      else rects[l] = rects[l].unite(s.crop);
    }
  std::vector<uint32_t> slot(nl, 0xffffffffu);
  std::vector<std::pair<uint32_t, Rect>> list;
  for (size_t l = 0; l < nl; ++l)
    if (have[l]) { slot[l] = (uint32_t)list.size(); list.push_back({(uint32_t)l, rects[l]}); }

  MeshTexels mt = mesh_texels(sc, pm, setups);

  // trace mesh_bake.rs:203-267
  std::vector<std::unique_ptr<AtomicGrid>> grids;
  for (auto& pr : list) grids.emplace_back(new AtomicGrid(pr.second.width(), pr.second.height(), sc.lights[pr.first].bits));
  std::vector<std::atomic<uint8_t>> visited((size_t)size * size);
  for (auto& v : visited) v.store(0, std::memory_order_relaxed);
  std::vector<uint64_t> tri_texels(nt, 0), tri_rays(nt, 0), tri_samples(nt, 0);
  par_for(nt, [&](size_t t) {
    const Setup& s = setups[t];
    if (s.face == v3(0, 0, 0)) return;
    std::vector<SurfaceTexel> tx;
    texels_of(mt, pm, (uint32_t)t, s.face, tx);
    V3 offset = s.face * RAY_OFFSET;
    uint64_t rays = 0, samples = 0;
    for (uint32_t l : s.lights) {
      const Light& L = sc.lights[l];
      uint32_t index = slot[l];
      Rect rect = list[index].second;
      AtomicGrid& grid = *grids[index];
      for (const auto& texel : tx) {
        V3 origin = texel.position + offset;
        uint8_t lit;
        uint32_t r;
        if (!texel_shadow(sc, pm, L, texel, origin, &lit, &r)) continue;
        rays += r;
        samples++;
        visited[(size_t)texel.x + (size_t)texel.y * size].store(1, std::memory_order_relaxed);
        grid.raise(texel.x - rect.min_x, texel.y - rect.min_y, lit);
      }
    }
    tri_texels[t] = tx.size();
    tri_rays[t] = rays;
    tri_samples[t] = samples;
  });
  for (uint32_t t = 0; t < nt; ++t) { out.texels += tri_texels[t]; out.rays += tri_rays[t]; out.samples += tri_samples[t]; }
  BitGrid vis;
  vis.w = vis.h = size;
  vis.b.resize((size_t)size * size);
  for (size_t i = 0; i < vis.b.size(); ++i) vis.b[i] = visited[i].load();
  std::vector<Region> bits;
  for (size_t i = 0; i < list.size(); ++i) bits.push_back({list[i].second, grids[i]->into_grid()});
  grids.clear();

  // bounce bounce.rs:155-217
  std::vector<std::vector<Entry>> entries(nt);
  for (uint32_t t = 0; t < nt; ++t)
    for (uint32_t l : setups[t].lights) entries[t].push_back({l, false, {}});
  if (sc.bounce_in_scene) {
    float gw[7], gsum;
    gaussian_weights(gw, &gsum);
    for (size_t l = 0; l < nl; ++l) {
      const Light& L = sc.lights[l];
      bool is_cube = sc.cubes[l] != nullptr;
      bool is_sun = L.bakes_bounce && L.sun;
      if (!is_cube && !is_sun) continue;
      std::vector<std::pair<uint32_t, uint32_t>> receivers;
      for (uint32_t t = 0; t < nt; ++t)
        for (uint32_t k = 0; k < setups[t].lights.size(); ++k)
          if (setups[t].lights[k] == l) { receivers.push_back({t, k}); break; }
      if (receivers.empty()) continue;
      Rect crops = setups[receivers[0].first].crop;
      for (auto& r : receivers) crops = crops.unite(setups[r.first].crop);
      std::vector<Gathered> gathered(receivers.size());
      par_for(receivers.size(), [&](size_t i) {
        uint32_t t = receivers[i].first;
        std::vector<SurfaceTexel> tx;
        texels_of(mt, pm, t, setups[t].face, tx);
        gathered[i] = is_cube ? gather_triangle(sc, tx, setups[t], pm, L, *sc.cubes[l]) : gather_sun(sc, tx, setups[t], pm, L);
      });
      BounceMap map(size, crops);
      for (auto& g : gathered) {
        out.bounce_rays += g.rays;
        for (auto& e : g.texels) map.set(e.first.first, e.first.second, e.second);
      }
      map.dilate();
      map.blur(gw, gsum);
      for (auto& r : receivers) {
        Entry& e = entries[r.first][r.second];
        e.has_bounce = map.crop(setups[r.first].crop, e.bounce);
        if (!e.has_bounce) e.bounce.clear();
      }
      if (g_capture_mode == 1) {
        for (auto& r : receivers) {
          const Entry& e = entries[r.first][r.second];
          if (!e.has_bounce) continue;
          std::vector<double> rgb(e.bounce.size() * 3);
          for (size_t k = 0; k < e.bounce.size(); ++k) {
            rgb[3 * k] = e.bounce[k].x;
            rgb[3 * k + 1] = e.bounce[k].y;
            rgb[3 * k + 2] = e.bounce[k].z;
          }
          capture_crop(pm.mesh->id, L.id, r.first, setups[r.first].crop, rgb.data(), rgb.size());
        }
      } else if (g_capture_mode == 2) {
        std::vector<GatheredD> gd(receivers.size());
        par_for(receivers.size(), [&](size_t i) {
          uint32_t t = receivers[i].first;
          std::vector<SurfaceTexel> tx;
          texels_of(mt, pm, t, setups[t].face, tx);
          gd[i] = is_cube ? gather_triangle_d(sc, tx, setups[t], pm, L, *sc.cubes[l]) : gather_sun_d(sc, tx, setups[t], pm, L);
        });
        BounceMapT<D3> md(size, crops);
        for (auto& g : gd)
          for (auto& e : g.texels) md.set(e.first.first, e.first.second, e.second);
        md.dilate();
        double gwd[7], gsumd;
        gaussian_weights_d(gwd, &gsumd);
        md.blur(gwd, gsumd);
        for (auto& r : receivers) {
          if (!entries[r.first][r.second].has_bounce) continue;
          std::vector<D3> grid;
          md.crop(setups[r.first].crop, grid);
          capture_crop(pm.mesh->id, L.id, r.first, setups[r.first].crop, &grid[0].x, grid.size() * 3);
        }
      }
    }
  }

  // optimize shadows.rs:54-74
  std::vector<uint64_t> removed(nt, 0);
  par_for(nt, [&](size_t t) {
    V2 uv[3];
    pm.uvs((uint32_t)t, uv);
    Rect rect = optimize_rect(uv, size);
    auto& es = entries[t];
    size_t before = es.size();
    std::vector<Entry> kept;
    for (auto& e : es)
      if (e.has_bounce || any_light(bits[slot[e.light]], rect)) kept.push_back(std::move(e));
    es.swap(kept);
    removed[t] = before - es.size();
  });
  for (auto r : removed) out.lights_removed += r;
  std::vector<int> used(nl, 0);
  for (auto& es : entries)
    for (auto& e : es) used[e.light] = 1;
  par_for(list.size(), [&](size_t i) {
    if (used[list[i].first]) pad_seams(bits[i], vis);
  });
  // into_records shadows.rs:87-133
  std::vector<TriRecord> records(nt);
  par_for(nt, [&](size_t t) {
    records[t].crop = setups[t].crop;
    for (auto& e : entries[t]) {
      const Region& region = bits[slot[e.light]];
      const Light& L = sc.lights[e.light];
      TriLight tl;
      tl.light = e.light;
      tl.has_shadow = true;
      tl.shadow = crop_shadows(region, setups[t].crop).words();
      tl.has_bounce = e.has_bounce;
      if (e.has_bounce) {
        int32_t comp = L.src->bounce_compression >= 0 ? L.src->bounce_compression : sc.opt->bounce_compression;
        Rng rng = stream({sc.opt->seed, pm.mesh->id, L.id, (uint64_t)t});
        if (sc.opt->colored_bounce) tl.bounce = pack_bounce_rgb(e.bounce, comp, rng);
        else {
          std::vector<float> vals(e.bounce.size());
          for (size_t k = 0; k < vals.size(); ++k) vals[k] = e.bounce[k].x;
          tl.bounce = pack_bounce(vals, comp, rng);
        }
      }
      records[t].lights.push_back(std::move(tl));
    }
  });
  out.words = serialize_triangles(records, sc.bounce_in_scene);
  return out;
}

// distance_cubes.rs:29-72
static std::vector<uint32_t> bake_distance_cubes(const Scene& sc) {
  const uint32_t res = 64;
  size_t nl = sc.local_lights;
  std::vector<uint32_t> words(nl * 6 * res * res / 2);
  par_for(nl * 6, [&](size_t job) {
    const Light& L = sc.lights[job / 6];
    uint32_t face = (uint32_t)(job % 6);
    std::vector<float> d(res * res);
    for (uint32_t texel = 0; texel < res * res; ++texel) {
      V3 dir = texel_direction(face, texel % res, texel / res, res);
      float dist;
      uint32_t occ;
      d[texel] = sc.occ.closest_hit(L.position, dir, L.radius, L.alpha_tested, &dist, &occ, nullptr) ? dist : L.radius;
    }
    for (uint32_t k = 0; k < res * res / 2; ++k) words[job * (res * res / 2) + k] = pack_f16_pair(d[2 * k], d[2 * k + 1]);
  });
  return words;
}

// ---- sun maps: sun_map.rs:25-150 ---------------------------------------------------------------
struct SunMapOut {
  float rows[12];
  uint32_t resolution;
  std::vector<uint32_t> texels;
};
// pack_sun_depth (sun_map.rs:77-80): two unorm16 depths, `lower` in bits 0-15.
static uint32_t pack_sun_depth(float lower, float upper) {
  auto unorm = [](float d) { return as_u32(std::round(clampf(d, 0.0f, 1.0f) * 65535.0f)); };
  return unorm(upper) << 16 | unorm(lower);
}
static SunMapOut bake_sun_map(const Scene& sc, const Light& L, uint32_t resolution) {
  SunMapOut out{};
  // SunFrame::new (sun_map.rs:100-109)
  const V3 toward = L.sun_toward;
  V3 side, up;
  any_orthonormal_pair(toward, &side, &up);
  const V3 along = -toward;
  // SunFrame::bounds (sun_map.rs:119-142) over every occluder corner
  V3 mn = v3(INFINITY, INFINITY, INFINITY), mx = v3(-INFINITY, -INFINITY, -INFINITY);
  for (const OccTri& t : sc.occ.tris)
    for (int k = 0; k < 3; ++k) {
      V3 l = v3(dot(t.v[k], side), dot(t.v[k], up), dot(t.v[k], along));
      mn = vmin(mn, l);
      mx = vmax(mx, l);
    }
  bool finite = std::isfinite(mn.x) && std::isfinite(mn.y) && std::isfinite(mn.z) && std::isfinite(mx.x) && std::isfinite(mx.y) &&
                std::isfinite(mx.z);
  if (!finite) {
    out.resolution = 1;
    out.texels.push_back(pack_sun_depth(1.0f, 1.0f));
    return out;
  }
  V3 size = mx - mn;
  float margin = size.z * 0.01f + 0.05f;
  float extent = fmax_(size.x, size.y) + 2.0f * margin;
  V3 center = (mn + mx) * 0.5f;
  V3 bmin = v3(center.x - extent * 0.5f, center.y - extent * 0.5f, mn.z - margin);
  float depth = size.z + 2.0f * margin;
  // SunFrame::rows (sun_map.rs:144-151)
  auto row = [&](float* r, V3 axis, float lo, float length) {
    V3 a = axis / length;
    r[0] = a.x;
    r[1] = a.y;
    r[2] = a.z;
    r[3] = -lo / length;
  };
  row(out.rows, side, bmin.x, extent);
  row(out.rows + 4, up, bmin.y, extent);
  row(out.rows + 8, along, bmin.z, depth);
  out.resolution = resolution;
  const float texel = extent / (float)resolution;
  std::vector<float> depths((size_t)resolution * resolution);
  par_for(resolution, [&](size_t y) {
    for (uint32_t x = 0; x < resolution; ++x) {
      float across = bmin.x + ((float)x + 0.5f) * texel;
      float down = bmin.y + ((float)y + 0.5f) * texel;
      V3 origin = (side * across + up * down) + along * bmin.z;
      float dist;
      uint32_t occ;
      depths[y * resolution + x] = sc.occ.closest_hit(origin, along, depth, L.alpha_tested, &dist, &occ, nullptr) ? dist / depth : 1.0f;
    }
  });
  for (size_t i = 0; i < depths.size(); i += 2)
    out.texels.push_back(pack_sun_depth(depths[i], i + 1 < depths.size() ? depths[i + 1] : 1.0f));
  return out;
}

// ---- sky probes: sky_probes.rs:46-191 ----------------------------------------------------------
struct SkyProbesOut {
  bool present = false;
  V3 origin;
  float spacing;
  uint32_t counts[3];
  std::vector<uint32_t> words;
};
static const uint64_t PROBE_STREAM = 0x736b7970726f6265ull;
static const uint32_t MAX_SKY_PROBES = 65536;
static const V3 PROBE_AXES[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// ProbeGrid::over (sky_probes.rs:46-62)
static void probe_grid(V3 mn, V3 mx, float spacing, V3* origin, float* spacing_out, uint32_t counts[3]) {
  V3 extent = vmax(mx - mn, v3(0, 0, 0)) + v3(spacing, spacing, spacing);
  auto counts_at = [&](float s, uint32_t* c) {
    c[0] = as_u32(std::ceil(extent.x / s)) + 1;
    c[1] = as_u32(std::ceil(extent.y / s)) + 1;
    c[2] = as_u32(std::ceil(extent.z / s)) + 1;
  };
  auto count = [](const uint32_t* c) { return (uint64_t)c[0] * c[1] * c[2]; };
  counts_at(spacing, counts);
  while (count(counts) > MAX_SKY_PROBES) {
    float ratio = (float)count(counts) / (float)MAX_SKY_PROBES;
    spacing *= fmax_(std::cbrt(ratio), 1.01f);
    counts_at(spacing, counts);
  }
  V3 span = v3((float)(counts[0] - 1), (float)(counts[1] - 1), (float)(counts[2] - 1)) * spacing;
  *origin = (mn + mx) * 0.5f - span * 0.5f;
  *spacing_out = spacing;
}

// SkyProbes::pack (format/types.rs:105-112)
static void pack_probe(const float* open, uint32_t* words) {
  words[0] = words[1] = 0;
  for (int a = 0; a < 6; ++a) {
    uint32_t byte = as_u32(std::round(clampf(open[a], 0.0f, 1.0f) * 255.0f));
    words[a / 4] |= byte << (8 * (a % 4));
  }
}

// fill_inside (sky_probes.rs:110-144): layer by layer from the previous pass.
static void fill_inside(const uint32_t counts[3], std::vector<float>& open, std::vector<uint8_t>& valid) {
  const int64_t cx = counts[0], cy = counts[1], cz = counts[2];
  const size_t n = valid.size();
  for (;;) {
    std::vector<float> prev_open = open;
    std::vector<uint8_t> prev_valid = valid;
    bool changed = false;
    for (size_t index = 0; index < n; ++index) {
      if (valid[index]) continue;
      int64_t at[3] = {(int64_t)(index % cx), (int64_t)(index % (cx * cy) / cx), (int64_t)(index / (cx * cy))};
      float sum[6] = {0, 0, 0, 0, 0, 0};
      float found = 0.0f;
      for (int a = 0; a < 6; ++a) {
        int64_t nx = at[0] + (int64_t)PROBE_AXES[a].x, ny = at[1] + (int64_t)PROBE_AXES[a].y, nz = at[2] + (int64_t)PROBE_AXES[a].z;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= cx || ny >= cy || nz >= cz) continue;
        size_t ni = (size_t)((nz * cy + ny) * cx + nx);
        if (!prev_valid[ni]) continue;
        for (int k = 0; k < 6; ++k) sum[k] += prev_open[ni * 6 + k];
        found += 1.0f;
      }
      if (found > 0.0f) {
        for (int k = 0; k < 6; ++k) open[index * 6 + k] = sum[k] / found;
        valid[index] = 1;
        changed = true;
      }
    }
    if (!changed) break;
  }
}

// bake_sky_probes (sky_probes.rs:146-191) with trace_probe (sky_probes.rs:80-106)
static SkyProbesOut bake_sky_probes(const Scene& sc, float spacing, uint32_t samples, bool alpha, uint64_t seed) {
  SkyProbesOut out;
  if (sc.occ.tris.empty()) return out;
  V3 mn = sc.occ.tris[0].v[0], mx = mn;
  for (const OccTri& t : sc.occ.tris)
    for (int k = 0; k < 3; ++k) {
      mn = vmin(mn, t.v[k]);
      mx = vmax(mx, t.v[k]);
    }
  probe_grid(mn, mx, spacing, &out.origin, &out.spacing, out.counts);
  const size_t n = (size_t)out.counts[0] * out.counts[1] * out.counts[2];
  std::vector<float> open(n * 6);
  std::vector<uint8_t> valid(n);
  par_for(n, [&](size_t index) {
    uint32_t px = (uint32_t)(index % out.counts[0]);
    uint32_t py = (uint32_t)(index % ((size_t)out.counts[0] * out.counts[1]) / out.counts[0]);
    uint32_t pz = (uint32_t)(index / ((size_t)out.counts[0] * out.counts[1]));
    V3 position = out.origin + v3((float)px, (float)py, (float)pz) * out.spacing;
    uint32_t back = 0;
    for (int a = 0; a < 6; ++a) {
      Rng rng = stream({seed, PROBE_STREAM, px, py, pz, (uint64_t)a});
      auto occluded = [&](V3 direction) {
        float dist;
        uint32_t occ;
        if (!sc.occ.closest_hit(position, direction, INFINITY, alpha, &dist, &occ, nullptr)) return false;
        if (!(dist > 0.0f)) return false;
        if (dot(sc.occ.face_normal(occ), direction) > 0.0f) back++;
        return true;
      };
      open[index * 6 + a] = sky_visibility(PROBE_AXES[a], samples, rng, occluded);
    }
    float rays = (float)(std::max(samples, 1u) * 6u);
    valid[index] = (float)back <= 0.25f * rays ? 1 : 0;
  });
  fill_inside(out.counts, open, valid);
  out.words.resize(n * 2);
  const float ones[6] = {1, 1, 1, 1, 1, 1};
  for (size_t i = 0; i < n; ++i) pack_probe(valid[i] ? &open[i * 6] : ones, &out.words[i * 2]);
  out.present = true;
  return out;
}

}  // namespace o

// ------------------------------------------------------------------------------------------
// C API
// ------------------------------------------------------------------------------------------
struct dlo_scene {
  dl_baked_settings settings;
  dl_bake_stats stats;
  std::vector<dl_baked_light> lights;
  std::vector<dl_light_bvh_node> bvh;
  std::vector<o::MeshOut> meshes;
  std::vector<uint32_t> distance_cubes;
  std::vector<o::SunMapOut> sun_maps;
  o::SkyProbesOut sky_probes;
};

static thread_local std::string g_err;
static thread_local double g_timing[3];  // photon cubes, meshes, distance cubes (ms) of the last bake
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
static std::string hexid(uint64_t id) {
  char b[32];
  snprintf(b, sizeof b, "0x%016llx", (unsigned long long)id);
  return b;
}

// input.rs:336-373
static int validate_options(const dl_bake_options& o) {
  if (!(std::isfinite(o.pixel_density) && o.pixel_density > 0.0f)) return fail(DL_E_INVALID_OPTIONS, "invalid bake options: pixel density must be positive");
  if (o.max_lightmap_size < 2) return fail(DL_E_INVALID_OPTIONS, "invalid bake options: the maximum shadow map size must be at least 2");
  if (o.photon_cube_resolution == 0) return fail(DL_E_INVALID_OPTIONS, "invalid bake options: the photon cube resolution must be positive");
  if (o.area_light_samples == 0) return fail(DL_E_INVALID_OPTIONS, "invalid bake options: area lights need at least one sample");
  if (!(o.sun_map_resolution >= 1 && o.sun_map_resolution <= 8192)) return fail(DL_E_INVALID_OPTIONS, "invalid bake options: the sun map resolution must be 1 to 8192");
  if (o.sky_samples == 0) return fail(DL_E_INVALID_OPTIONS, "invalid bake options: skies need at least one sample");
  if (!(std::isfinite(o.sky_probe_spacing) && o.sky_probe_spacing > 0.0f)) return fail(DL_E_INVALID_OPTIONS, "invalid bake options: the sky probe spacing must be positive");
  return DL_OK;
}
// input.rs:233-294
static int validate_mesh(const dl_mesh& m) {
  std::string id = hexid(m.id);
  uint32_t vertices = m.vertex_count;
  auto bad = [&](const std::string& r) { return fail(DL_E_INVALID_MESH, "mesh " + id + ": " + r); };
  if (m.index_count % 3 != 0) return bad(std::to_string(m.index_count) + " indices is not a multiple of 3");
  for (uint32_t i = 0; i < m.index_count; ++i)
    if (m.indices[i] >= vertices) return bad("index " + std::to_string(m.indices[i]) + " is out of range for " + std::to_string(vertices) + " vertices");
  if (m.uv1_count != vertices) return bad(std::to_string(m.uv1_count) + " UV1 coordinates for " + std::to_string(vertices) + " vertices");
  uint32_t nn = m.normals ? m.normal_count : 0;
  if (nn != 0 && nn != vertices) return bad(std::to_string(nn) + " normals for " + std::to_string(vertices) + " vertices");
  for (uint32_t i = 0; i < vertices * 3; ++i)
    if (!std::isfinite(m.positions[i])) return bad("non-finite vertex position");
  if (m.uv0 && m.uv0_count != vertices) return bad(std::to_string(m.uv0_count) + " UV0 coordinates for " + std::to_string(vertices) + " vertices");
  if (m.alpha) {
    if (!m.uv0) return bad("an alpha texture needs UV0");
    if (m.alpha->rgba8_len != (uint64_t)m.alpha->width * m.alpha->height * 4) return bad("alpha texture size doesn't match its data");
  }
  return DL_OK;
}
static bool quat_ok(const float* q) {
  for (int i = 0; i < 4; ++i)
    if (!std::isfinite(q[i])) return false;
  float len2 = (q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]);
  return std::fabs(len2 - 1.0f) <= 2e-4f;
}
// input.rs:296-321
static int validate_light(const dl_light& l) {
  std::string id = hexid(l.id);
  auto bad = [&](const char* r) { return fail(DL_E_INVALID_LIGHT, "light " + id + ": " + r); };
  if (!(std::isfinite(l.position[0]) && std::isfinite(l.position[1]) && std::isfinite(l.position[2]))) return bad("non-finite position");
  bool distant = l.shape == DL_SHAPE_SUN || l.shape == DL_SHAPE_SKY;
  if (!distant && !(std::isfinite(l.radius) && l.radius > 0.0f)) return bad("radius must be positive and finite");
  if (l.shape == DL_SHAPE_RECTANGLE || l.shape == DL_SHAPE_TUBE) {
    if (!(std::isfinite(l.size[0]) && std::isfinite(l.size[1])) || l.size[0] < 0.0f || l.size[1] < 0.0f) return bad("area light size must be finite and not negative");
    if (!quat_ok(l.rotation)) return bad("area light rotation must be a unit quaternion");
  } else if (l.shape == DL_SHAPE_SUN) {
    if (!quat_ok(l.rotation)) return bad("sun rotation must be a unit quaternion");
    if (!(l.angular_diameter >= 0.0f && l.angular_diameter <= 10.0f)) return bad("sun angular diameter must be 0 to 10 degrees");
  }
  return DL_OK;
}

extern "C" {

const char* dlo_last_error(void) { return g_err.c_str(); }
void dlo_last_timings(double* out3) {
  for (int i = 0; i < 3; ++i) out3[i] = g_timing[i];
}

static int bake_impl(const dl_bake_input* input, const dl_bake_options* options, uint32_t threads, uint32_t brute_force,
                     const uint8_t* receivers, dlo_scene** out);

int dlo_bake(const dl_bake_input* input, const dl_bake_options* options, uint32_t threads, uint32_t brute_force,
             dlo_scene** out) {
  return bake_impl(input, options, threads, brute_force, nullptr, out);
}

int dlo_bake_receivers(const dl_bake_input* input, const dl_bake_options* options, uint32_t threads,
                       const uint8_t* receivers, dlo_scene** out) {
  return bake_impl(input, options, threads, 0, receivers, out);
}

static int bake_impl(const dl_bake_input* input, const dl_bake_options* options, uint32_t threads, uint32_t brute_force,
                     const uint8_t* receivers, dlo_scene** out) {
  using namespace o;
  if (!input || !options || !out) return fail(DL_E_ARGUMENT, "null argument");
  *out = nullptr;
  g_threads = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
  const dl_bake_options& opt = *options;
  int st = validate_options(opt);
  if (st) return st;
  for (uint32_t i = 0; i < input->light_count; ++i)
    if (input->lights[i].shape < 0 || input->lights[i].shape > 4) return fail(DL_E_ARGUMENT, "unknown light shape");
  // sorted_meshes / sorted_lights (baker.rs:205-233)
  std::vector<const dl_mesh*> meshes;
  for (uint32_t i = 0; i < input->mesh_count; ++i) meshes.push_back(&input->meshes[i]);
  std::stable_sort(meshes.begin(), meshes.end(), [](const dl_mesh* a, const dl_mesh* b) { return a->id < b->id; });
  for (size_t i = 1; i < meshes.size(); ++i)
    if (meshes[i - 1]->id == meshes[i]->id) return fail(DL_E_DUPLICATE_ID, "two meshes share BakeId " + hexid(meshes[i]->id));
  for (auto* m : meshes)
    if ((st = validate_mesh(*m))) return st;
  std::vector<const dl_light*> lights;
  for (uint32_t i = 0; i < input->light_count; ++i) lights.push_back(&input->lights[i]);
  std::stable_sort(lights.begin(), lights.end(), [](const dl_light* a, const dl_light* b) { return a->id < b->id; });
  for (size_t i = 1; i < lights.size(); ++i)
    if (lights[i - 1]->id == lights[i]->id) return fail(DL_E_DUPLICATE_ID, "two lights share BakeId " + hexid(lights[i]->id));
  for (auto* l : lights)
    if ((st = validate_light(*l))) return st;

  std::unique_ptr<Scene> sc(new Scene());
  sc->opt = &opt;
  std::vector<const dl_light*> local, suns, skies;
  for (auto* l : lights) {
    if (l->shape == DL_SHAPE_SUN) suns.push_back(l);
    else if (l->shape == DL_SHAPE_SKY) skies.push_back(l);
    else local.push_back(l);
  }
  std::vector<V3> positions;
  std::vector<float> radii;
  for (auto* l : local) { positions.push_back(v3p(l->position)); radii.push_back(l->radius); }
  sc->light_bvh = build_light_bvh(positions, radii);
  for (uint32_t i : sc->light_bvh.order) sc->lights.push_back(make_light(local[i], opt));
  for (auto* l : suns) sc->lights.push_back(make_light(l, opt));
  for (auto* l : skies) sc->lights.push_back(make_light(l, opt));
  sc->local_lights = local.size();
  sc->bounce_in_scene = false;
  for (auto& L : sc->lights) sc->bounce_in_scene |= L.bakes_bounce;
  // Occluders::build (occluder.rs:80-115)
  sc->occ.meshes = meshes;
  sc->occ.brute = brute_force != 0;
  for (uint32_t mi = 0; mi < meshes.size(); ++mi) {
    const dl_mesh* m = meshes[mi];
    if (!m->occludes) continue;
    for (uint32_t t = 0; t < m->index_count / 3; ++t) {
      OccTri ot;
      for (int k = 0; k < 3; ++k) ot.v[k] = v3p(m->positions + m->indices[t * 3 + k] * 3);
      V3 n = cross(ot.v[1] - ot.v[0], ot.v[2] - ot.v[0]);
      if (length_sq(n) > 0.0f && finite3(n)) { ot.mesh = mi; ot.triangle = t; sc->occ.tris.push_back(ot); }
    }
  }
  if (!sc->occ.brute) sc->occ.build_bvh();
  auto t_cubes = std::chrono::steady_clock::now();
  // photon cubes (bounce.rs:126-151)
  sc->cubes.resize(sc->lights.size());
  // With a receiver sample only the lights whose sphere meets a sampled mesh's bounds can be
  // gathered from; the other cubes would never be read.
  std::vector<Aabb> sampled;
  if (receivers)
    for (uint32_t i = 0; i < input->mesh_count; ++i) {
      if (!receivers[i]) continue;
      const dl_mesh& m = input->meshes[i];
      Aabb b = Aabb::empty();
      for (uint32_t v = 0; v < m.vertex_count; ++v) b = b.grow(v3p(m.positions + 3 * (size_t)v));
      sampled.push_back(b);
    }
  for (size_t l = 0; l < sc->lights.size(); ++l) {
    const Light& L = sc->lights[l];
    if (!L.bakes_bounce || L.sun) continue;
    if (receivers) {
      const V3 r = v3(L.radius, L.radius, L.radius);
      const Aabb sphere = Aabb::empty().grow(L.position - r).grow(L.position + r);
      bool near = false;
      for (const Aabb& b : sampled) near |= b.intersects(sphere);
      if (!near) continue;
    }
    sc->cubes[l].reset(new PhotonCube());
    build_photon_cube(*sc, L, opt.photon_cube_resolution, *sc->cubes[l]);
  }
  if (g_capture_mode) {
    std::lock_guard<std::mutex> lock(g_capture_mu);
    g_cubes.clear();
    for (size_t l = 0; l < sc->lights.size(); ++l)
      if (sc->cubes[l]) g_cubes.push_back({sc->lights[l].id, *sc->cubes[l]});
  }
  auto t_meshes = std::chrono::steady_clock::now();
  std::vector<Prepared> prepared;
  for (auto* m : meshes)
    if (!receivers || receivers[m - input->meshes]) prepared.push_back(prepare(m, opt));
  std::unique_ptr<dlo_scene> res(new dlo_scene());
  res->meshes.resize(prepared.size());
  if (prepared.size() >= g_threads) {
    par_for(prepared.size(), [&](size_t i) { res->meshes[i] = bake_mesh(*sc, prepared[i]); });
  } else {
    for (size_t i = 0; i < prepared.size(); ++i) res->meshes[i] = bake_mesh(*sc, prepared[i]);
  }
  auto t_maps = std::chrono::steady_clock::now();
  if (opt.dynamic_geometry != DL_DYNAMIC_LIGHTING_ONLY) {
    // baker.rs:105-152: distance cubes, a sun map per sun (slot order), sky probes if a sky bakes
    res->distance_cubes = bake_distance_cubes(*sc);
    bool sky = false, sky_alpha = false;
    for (size_t l = sc->local_lights; l < sc->lights.size(); ++l) {
      const o::Light& L = sc->lights[l];
      if (L.sun) res->sun_maps.push_back(o::bake_sun_map(*sc, L, opt.sun_map_resolution));
      if (L.sky) {
        sky = true;
        sky_alpha |= L.alpha_tested;
      }
    }
    if (sky) res->sky_probes = o::bake_sky_probes(*sc, opt.sky_probe_spacing, opt.sky_samples, sky_alpha, opt.seed);
  }
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  g_timing[0] = ms(t_cubes, t_meshes);
  g_timing[1] = ms(t_meshes, t_maps);
  g_timing[2] = ms(t_maps, std::chrono::steady_clock::now());
  // settings, lights, stats (baker.rs:154-200)
  res->settings.bounce = sc->bounce_in_scene;
  res->settings.colored_bounce = sc->bounce_in_scene && opt.colored_bounce;
  res->settings.dynamic_geometry = opt.dynamic_geometry;
  res->settings.pixel_density = opt.pixel_density;
  res->settings.max_lightmap_size = opt.max_lightmap_size;
  res->settings.distance_cube_resolution = 64;
  for (auto& L : sc->lights) {
    dl_baked_light b{};
    b.id = L.id;
    memcpy(b.origin, L.src->position, 12);
    b.radius = L.radius;
    b.shape = L.shape;
    // BakedLight keeps only what its LightShape holds (format/types.rs, light/shape.rs): an area
    // light's rotation and size, a sun's rotation and angular diameter, nothing for the rest
    const float identity[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    memcpy(b.rotation, identity, 16);
    if (L.shape == DL_SHAPE_RECTANGLE || L.shape == DL_SHAPE_TUBE) {
      memcpy(b.rotation, L.src->rotation, 16);
      memcpy(b.size, L.src->size, 8);
    } else if (L.shape == DL_SHAPE_SUN) {
      memcpy(b.rotation, L.src->rotation, 16);
      b.angular_diameter = L.src->angular_diameter;
    }
    b.shadow_bits_per_texel = L.bits;
    b.bounce_compression = L.bakes_bounce ? (L.src->bounce_compression >= 0 ? L.src->bounce_compression : opt.bounce_compression) : -1;
    res->lights.push_back(b);
  }
  res->bvh = sc->light_bvh.nodes;
  memset(&res->stats, 0, sizeof res->stats);
  res->stats.meshes = (uint32_t)res->meshes.size();
  for (auto& p : prepared) res->stats.triangles += p.tri_count();
  for (auto& m : res->meshes) {
    res->stats.texels += m.texels;
    res->stats.rays += m.rays;
    res->stats.bounce_rays += m.bounce_rays;
    res->stats.lights_removed += m.lights_removed;
    res->stats.triangle_words += m.words.size();
    res->stats.samples += m.samples;
  }
  *out = res.release();
  return DL_OK;
}

void dlo_scene_free(dlo_scene* s) { delete s; }
void dlo_scene_settings(const dlo_scene* s, dl_baked_settings* out) { *out = s->settings; }
void dlo_scene_stats(const dlo_scene* s, dl_bake_stats* out) { *out = s->stats; }
uint32_t dlo_scene_light_count(const dlo_scene* s) { return (uint32_t)s->lights.size(); }
int dlo_scene_light(const dlo_scene* s, uint32_t i, dl_baked_light* out) {
  if (i >= s->lights.size()) return DL_E_ARGUMENT;
  *out = s->lights[i];
  return DL_OK;
}
uint32_t dlo_scene_light_bvh(const dlo_scene* s, const dl_light_bvh_node** nodes) {
  *nodes = s->bvh.data();
  return (uint32_t)s->bvh.size();
}
uint32_t dlo_scene_mesh_count(const dlo_scene* s) { return (uint32_t)s->meshes.size(); }
int dlo_scene_mesh(const dlo_scene* s, uint32_t i, dl_baked_mesh_view* out) {
  if (i >= s->meshes.size()) return DL_E_ARGUMENT;
  const auto& m = s->meshes[i];
  out->id = m.mesh->id;
  out->content_hash = m.mesh->content_hash;
  out->resolution = m.resolution;
  out->occludes = m.mesh->occludes;
  out->triangles = m.words.data();
  out->triangle_word_count = m.words.size();
  return DL_OK;
}
uint64_t dlo_scene_distance_cubes(const dlo_scene* s, const uint32_t** words) {
  *words = s->distance_cubes.data();
  return s->distance_cubes.size();
}
uint32_t dlo_scene_sun_map_count(const dlo_scene* s) { return (uint32_t)s->sun_maps.size(); }
int dlo_scene_sun_map(const dlo_scene* s, uint32_t i, dl_sun_map_view* out) {
  if (i >= s->sun_maps.size()) return DL_E_ARGUMENT;
  const auto& m = s->sun_maps[i];
  memcpy(out->rows, m.rows, sizeof m.rows);
  out->resolution = m.resolution;
  out->texels = m.texels.data();
  out->texel_word_count = m.texels.size();
  return DL_OK;
}
void dlo_scene_sky_probes(const dlo_scene* s, dl_sky_probes_view* out) {
  memset(out, 0, sizeof *out);
  const auto& p = s->sky_probes;
  if (!p.present) return;
  out->present = 1;
  out->origin[0] = p.origin.x;
  out->origin[1] = p.origin.y;
  out->origin[2] = p.origin.z;
  out->spacing = p.spacing;
  memcpy(out->counts, p.counts, sizeof p.counts);
  out->words = p.words.data();
  out->word_count = p.words.size();
}

// ---- unit-level entry points -----------------------------------------------------------------
uint64_t dlo_splitmix64(uint64_t v) { return o::splitmix64(v); }
uint64_t dlo_seed(const uint64_t* parts, uint32_t n) { return o::seed_of(parts, n); }
void dlo_pcg32(uint64_t seed, uint64_t stream, uint32_t n, uint32_t* out) {
  o::Rng r = o::rng_new(seed, stream);
  for (uint32_t i = 0; i < n; ++i) out[i] = r.next_u32();
}
void dlo_stream_f32(const uint64_t* parts, uint32_t nparts, uint32_t n, float* out) {
  o::Rng r = o::stream_n(parts, nparts);
  for (uint32_t i = 0; i < n; ++i) out[i] = r.next_f32();
}
uint32_t dlo_lightmap_size(float area, float density, uint32_t max_size) { return o::lightmap_size(area, density, max_size); }
void dlo_rect(uint32_t kind, const float* uvs6, uint32_t size, uint32_t* out4) {
  o::V2 uv[3] = {{uvs6[0], uvs6[1]}, {uvs6[2], uvs6[3]}, {uvs6[4], uvs6[5]}};
  o::Rect r = kind == 0 ? o::crop_rect(uv, size) : kind == 1 ? o::trace_rect(uv, size) : o::optimize_rect(uv, size);
  out4[0] = r.min_x; out4[1] = r.min_y; out4[2] = r.max_x; out4[3] = r.max_y;
}
void dlo_closest_point(const float* p, const float* a, const float* b, const float* c, float* out) {
  o::V3 r = o::closest_point_on_triangle(o::v3p(p), o::v3p(a), o::v3p(b), o::v3p(c));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int dlo_uv_barycentrics(const float* t6, const float* p2, float* out3) {
  o::V2 t[3] = {{t6[0], t6[1]}, {t6[2], t6[3]}, {t6[4], t6[5]}};
  o::V3 w;
  if (!o::uv_barycentrics(t, o::v2(p2[0], p2[1]), &w)) return 0;
  out3[0] = w.x; out3[1] = w.y; out3[2] = w.z;
  return 1;
}
static void tri_args(const float* v9, const float* n9, const float* uv6, o::V3* v, o::V3* n, o::V2* uv) {
  for (int k = 0; k < 3; ++k) { v[k] = o::v3p(v9 + 3 * k); n[k] = o::v3p(n9 + 3 * k); uv[k] = o::v2(uv6[2 * k], uv6[2 * k + 1]); }
}
static uint32_t dump_texels(const std::vector<o::SurfaceTexel>& tx, float* out, uint32_t cap) {
  for (size_t i = 0; i < tx.size() && i < cap; ++i) {
    float* r = out + i * 8;
    r[0] = (float)tx[i].x; r[1] = (float)tx[i].y;
    r[2] = tx[i].position.x; r[3] = tx[i].position.y; r[4] = tx[i].position.z;
    r[5] = tx[i].normal.x; r[6] = tx[i].normal.y; r[7] = tx[i].normal.z;
  }
  return (uint32_t)tx.size();
}
uint32_t dlo_triangle_texels(const float* v9, const float* n9, const float* uv6, const float* face3, uint32_t size, float* out, uint32_t cap) {
  o::V3 v[3], n[3];
  o::V2 uv[3];
  tri_args(v9, n9, uv6, v, n, uv);
  std::vector<o::SurfaceTexel> tx;
  o::triangle_texels(v, n, uv, o::v3p(face3), size, tx);
  return dump_texels(tx, out, cap);
}
uint32_t dlo_touched_texels(const float* v9, const float* n9, const float* uv6, const float* face3, uint32_t size, const uint32_t* owned, float* out, uint32_t cap) {
  o::V3 v[3], n[3];
  o::V2 uv[3];
  tri_args(v9, n9, uv6, v, n, uv);
  std::vector<o::SurfaceTexel> tx;
  auto own = [&](uint32_t x, uint32_t y) { size_t i = (size_t)x + (size_t)y * size; return owned && ((owned[i / 32] >> (i % 32)) & 1); };
  o::touched_texels(v, n, uv, o::v3p(face3), size, own, tx);
  return dump_texels(tx, out, cap);
}
uint8_t dlo_level(float fraction, uint32_t bits) { return o::level(fraction, bits); }
void dlo_spread(uint32_t index, const float* off, uint32_t rect, float* out2) {
  o::V2 r = o::spread(index, o::v2(off[0], off[1]), rect != 0);
  out2[0] = r.x; out2[1] = r.y;
}
void dlo_concentric_disk(const float* sq, float* out2) {
  o::V2 r = o::concentric_disk(o::v2(sq[0], sq[1]));
  out2[0] = r.x; out2[1] = r.y;
}
int dlo_is_buried(const float* normal3, float turn, const uint8_t* answers, uint32_t* rays, float* dirs24) {
  uint32_t k = 0;
  bool b = o::is_buried(o::v3p(normal3), turn, [&](o::V3 d) {
    if (k < 8) { dirs24[k * 3] = d.x; dirs24[k * 3 + 1] = d.y; dirs24[k * 3 + 2] = d.z; }
    return answers[k++] != 0;
  });
  *rays = k;
  return b ? 1 : 0;
}
// Blockers for the sampler KATs: kind 0 none, 1 all, 2 x-interval of the reached point
// [blocker[0], blocker[1]) (area) or direction.x < blocker[0] (sun/sky), 3 point.x < 0.
static bool blocked_point(int32_t kind, const float* b, o::V3 p) {
  if (kind == 0) return false;
  if (kind == 1) return true;
  if (kind == 2) return p.x >= b[0] && p.x < b[1];
  return p.x < 0.0f;
}
int dlo_area_visibility(const dl_light* light, float reach_override, const float* pos3, const float* normal3, uint32_t samples,
                        uint32_t minimum, const uint64_t* key, uint32_t nkey, int32_t kind, const float* blocker, float* fraction, uint32_t* rays) {
  dl_bake_options opt;
  memset(&opt, 0, sizeof opt);
  o::Light L = o::make_light(light, opt);
  o::Rng rng = o::stream_n(key, nkey);
  o::V3 pos = o::v3p(pos3);
  auto occ = [&](o::V3 d, float dist) { return blocked_point(kind, blocker, pos + d * dist); };
  float reach = std::isnan(reach_override) ? L.reach : reach_override;
  return o::texel_visibility(L.as, reach, pos, o::v3p(normal3), samples, minimum, rng, occ, fraction, rays) ? 1 : 0;
}
int dlo_sun_visibility(const dl_light* light, const float* normal3, uint32_t samples, uint32_t minimum, const uint64_t* key, uint32_t nkey,
                       int32_t kind, const float* blocker, float* fraction, uint32_t* rays) {
  dl_bake_options opt;
  memset(&opt, 0, sizeof opt);
  o::Light L = o::make_light(light, opt);
  o::Rng rng = o::stream_n(key, nkey);
  auto occ = [&](o::V3 d) {
    if (kind == 2) return d.x < blocker[0];
    if (kind == 4) return d.x / d.y < blocker[0];
    return kind == 1;
  };
  return o::sun_visibility(L, o::v3p(normal3), samples, minimum, rng, occ, fraction, rays) ? 1 : 0;
}
float dlo_sky_visibility(const float* normal3, uint32_t samples, const uint64_t* key, uint32_t nkey, int32_t kind, const float* blocker) {
  o::Rng rng = o::stream_n(key, nkey);
  auto occ = [&](o::V3 d) {
    if (kind == 2) return d.x < blocker[0];
    if (kind == 5) return d.y < blocker[0];
    return kind == 1;
  };
  return o::sky_visibility(o::v3p(normal3), samples, rng, occ);
}
uint32_t dlo_light_bvh(const float* positions, const float* radii, uint32_t n, dl_light_bvh_node* nodes, uint32_t cap, uint32_t* order) {
  std::vector<o::V3> p;
  std::vector<float> r;
  for (uint32_t i = 0; i < n; ++i) { p.push_back(o::v3p(positions + 3 * i)); r.push_back(radii[i]); }
  o::LightBvh b = o::build_light_bvh(p, r);
  for (size_t i = 0; i < b.nodes.size() && i < cap; ++i) nodes[i] = b.nodes[i];
  for (size_t i = 0; i < b.order.size(); ++i) order[i] = b.order[i];
  return (uint32_t)b.nodes.size();
}
void dlo_pad_seams(uint32_t w, uint32_t h, uint8_t* levels, const uint8_t* visited) {
  o::Region r;
  r.rect = {0, 0, w - 1, h - 1};
  r.grid = o::Grid(w, h, 8);
  memcpy(r.grid.lv.data(), levels, (size_t)w * h);
  o::BitGrid v;
  v.w = w; v.h = h;
  v.b.assign(visited, visited + (size_t)w * h);
  o::pad_seams(r, v);
  memcpy(levels, r.grid.lv.data(), (size_t)w * h);
}
void dlo_dilate(uint32_t w, uint32_t h, uint8_t* levels) {
  o::Grid g(w, h, 8);
  memcpy(g.lv.data(), levels, (size_t)w * h);
  o::dilate_shadows(g);
  memcpy(levels, g.lv.data(), (size_t)w * h);
}
uint32_t dlo_pack_levels(const uint8_t* levels, uint32_t count, uint32_t bits, uint32_t* out) {
  o::Grid g(count, 1, bits);
  for (uint32_t i = 0; i < count; ++i) g.lv[i] = levels[i];
  auto w = g.words();
  memcpy(out, w.data(), w.size() * 4);
  return (uint32_t)w.size();
}
uint32_t dlo_serialize(uint32_t triangles, const uint32_t* crops, const uint32_t* light_counts, const uint32_t* lights,
                       const uint32_t* shadow_lens, const uint32_t* shadow_words, const uint32_t* bounce_lens,
                       const uint32_t* bounce_words, uint32_t bounce_layout, uint32_t* out, uint32_t cap) {
  std::vector<o::TriRecord> recs(triangles);
  size_t li = 0, sw = 0, bw = 0;
  for (uint32_t t = 0; t < triangles; ++t) {
    recs[t].crop = {crops[t * 4], crops[t * 4 + 1], crops[t * 4 + 2], crops[t * 4 + 3]};
    for (uint32_t k = 0; k < light_counts[t]; ++k, ++li) {
      o::TriLight tl;
      tl.light = lights[li];
      uint32_t sl = shadow_lens[li], bl = bounce_lens[li];
      tl.has_shadow = sl != 0xffffffffu;
      if (tl.has_shadow) { tl.shadow.assign(shadow_words + sw, shadow_words + sw + sl); sw += sl; }
      tl.has_bounce = bl != 0xffffffffu;
      if (tl.has_bounce) { tl.bounce.assign(bounce_words + bw, bounce_words + bw + bl); bw += bl; }
      recs[t].lights.push_back(tl);
    }
  }
  auto w = o::serialize_triangles(recs, bounce_layout != 0);
  for (size_t i = 0; i < w.size() && i < cap; ++i) out[i] = w[i];
  return (uint32_t)w.size();
}
uint32_t dlo_pack_bounce(const float* values, uint32_t count, uint32_t rgb, int32_t compression, const uint64_t* key, uint32_t nkey,
                         uint32_t* out, uint32_t cap) {
  // nkey == 0: EffectRng::default() = for_light(0, 0) (effects/rng.rs:39-41, 74-78)
  o::Rng rng = nkey ? o::stream_n(key, nkey) : o::rng_new(o::splitmix64(0 ^ o::splitmix64(0)), 0);
  std::vector<uint32_t> w;
  if (rgb) {
    std::vector<o::V3> v(count);
    for (uint32_t i = 0; i < count; ++i) v[i] = o::v3p(values + 3 * i);
    w = o::pack_bounce_rgb(v, compression, rng);
  } else {
    std::vector<float> v(values, values + count);
    w = o::pack_bounce(v, compression, rng);
  }
  for (size_t i = 0; i < w.size() && i < cap; ++i) out[i] = w[i];
  return (uint32_t)w.size();
}
uint32_t dlo_f16_pair(float lower, float upper) { return o::pack_f16_pair(lower, upper); }
uint32_t dlo_sun_depth_pair(float lower, float upper) { return o::pack_sun_depth(lower, upper); }
void dlo_probe_grid(const float* mn, const float* mx, float spacing, float* origin, float* spacing_out, uint32_t* counts) {
  o::V3 org;
  o::probe_grid(o::v3p(mn), o::v3p(mx), spacing, &org, spacing_out, counts);
  origin[0] = org.x;
  origin[1] = org.y;
  origin[2] = org.z;
}
void dlo_fill_inside(const uint32_t* counts, float* open, uint8_t* valid) {
  size_t n = (size_t)counts[0] * counts[1] * counts[2];
  std::vector<float> o6(open, open + n * 6);
  std::vector<uint8_t> v(valid, valid + n);
  o::fill_inside(counts, o6, v);
  memcpy(open, o6.data(), n * 6 * sizeof(float));
  memcpy(valid, v.data(), n);
}
void dlo_pack_probe(const float* open6, uint32_t* words2) { o::pack_probe(open6, words2); }
static o::Occluders soup(const float* tris, uint32_t n, uint32_t brute) {
  o::Occluders occ;
  for (uint32_t i = 0; i < n; ++i) {
    o::OccTri t;
    for (int k = 0; k < 3; ++k) t.v[k] = o::v3p(tris + i * 9 + k * 3);
    t.mesh = 0;
    t.triangle = i;
    occ.tris.push_back(t);
  }
  occ.brute = brute != 0;
  if (!occ.brute) occ.build_bvh();
  return occ;
}
int dlo_ray_any(const float* tris, uint32_t n, const float* origin, const float* dir, float max, uint32_t brute) {
  o::Occluders occ = soup(tris, n, brute);
  return occ.occluded(o::v3p(origin), o::v3p(dir), max, false) ? 1 : 0;
}
int dlo_ray_closest(const float* tris, uint32_t n, const float* origin, const float* dir, float max, uint32_t brute, float* t,
                    uint32_t* index) {
  o::Occluders occ = soup(tris, n, brute);
  return occ.closest_hit(o::v3p(origin), o::v3p(dir), max, false, t, index, nullptr) ? 1 : 0;
}
void dlo_rays_batch(const float* tris, uint32_t n, const float* rays, uint32_t count, uint32_t closest, uint32_t brute, float* t_out,
                    uint32_t* idx_out) {
  o::Occluders occ = soup(tris, n, brute);
  for (uint32_t i = 0; i < count; ++i) {
    const float* r = rays + i * 7;
    if (closest) {
      float t;
      uint32_t idx;
      if (occ.closest_hit(o::v3p(r), o::v3p(r + 3), r[6], false, &t, &idx, nullptr)) { t_out[i] = t; idx_out[i] = idx; }
      else { t_out[i] = -1.0f; idx_out[i] = 0xffffffffu; }
    } else {
      idx_out[i] = occ.occluded(o::v3p(r), o::v3p(r + 3), r[6], false) ? 1 : 0;
    }
  }
}
float dlo_sinf(float x) { return sinf(x); }
float dlo_cosf(float x) { return cosf(x); }

// ---- bounce-side hooks: cube conventions, bounce map, slerp, photons, captured crops ----------
void dlo_cube_face_uv(const float* d3, uint32_t* face, float* uv2) {
  o::V2 uv;
  o::direction_to_face_uv(o::v3p(d3), face, &uv);
  uv2[0] = uv.x;
  uv2[1] = uv.y;
}
void dlo_cube_direction(uint32_t face, const float* uv2, float* out3) {
  o::V3 d = o::face_uv_to_direction(face, o::v2(uv2[0], uv2[1]));
  out3[0] = d.x; out3[1] = d.y; out3[2] = d.z;
}
void dlo_cube_texel_direction(uint32_t face, uint32_t x, uint32_t y, uint32_t res, float* out3) {
  o::V3 d = o::texel_direction(face, x, y, res);
  out3[0] = d.x; out3[1] = d.y; out3[2] = d.z;
}
void dlo_cube_texel(const float* d3, uint32_t res, uint32_t* out3) { o::direction_to_texel(o::v3p(d3), res, &out3[0], &out3[1], &out3[2]); }
uint32_t dlo_cube_texel_index(uint32_t face, uint32_t x, uint32_t y, uint32_t res) { return face * res * res + y * res + x; }
void dlo_gaussian_weights(float* w7, float* sum) { o::gaussian_weights(w7, sum); }
static o::BounceMap bounce_map_of(uint32_t size, const uint32_t* crops4, const float* texels5, uint32_t n, uint32_t ops) {
  o::BounceMap map(size, o::Rect{crops4[0], crops4[1], crops4[2], crops4[3]});
  for (uint32_t i = 0; i < n; ++i) {
    const float* t = texels5 + 5 * (size_t)i;
    map.set((uint32_t)t[0], (uint32_t)t[1], o::v3(t[2], t[3], t[4]));
  }
  if (ops & 1) map.dilate();
  if (ops & 2) {
    float w[7], sum;
    o::gaussian_weights(w, &sum);
    map.blur(w, sum);
  }
  return map;
}
void dlo_bounce_map(uint32_t size, const uint32_t* crops4, const float* texels5, uint32_t n, uint32_t ops, float* out) {
  o::BounceMap map = bounce_map_of(size, crops4, texels5, n, ops);
  for (uint32_t y = 0; y < size; ++y)
    for (uint32_t x = 0; x < size; ++x) {
      o::V3 v = map.get(x, y);
      float* q = out + 3 * ((size_t)y * size + x);
      q[0] = v.x; q[1] = v.y; q[2] = v.z;
    }
}
int dlo_bounce_crop(uint32_t size, const uint32_t* crops4, const float* texels5, uint32_t n, const uint32_t* crop4, float* out) {
  o::BounceMap map = bounce_map_of(size, crops4, texels5, n, 0);
  std::vector<o::V3> grid;
  bool lit = map.crop(o::Rect{crop4[0], crop4[1], crop4[2], crop4[3]}, grid);
  for (size_t k = 0; k < grid.size(); ++k) {
    out[3 * k] = grid[k].x; out[3 * k + 1] = grid[k].y; out[3 * k + 2] = grid[k].z;
  }
  return lit ? 1 : 0;
}
void dlo_slerp(const float* a3, const float* b3, float s, float* out3) {
  o::V3 r = o::slerp(o::v3p(a3), o::v3p(b3), s);
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
float dlo_acos_approx(float v) { return o::acos_approx(v); }
void dlo_on_unit_sphere(const uint64_t* key, uint32_t nkey, uint32_t n, float* out3n) {
  o::Rng rng = o::stream_n(key, nkey);
  for (uint32_t i = 0; i < n; ++i) {
    o::V3 v = o::on_unit_sphere(rng);
    out3n[3 * i] = v.x; out3n[3 * i + 1] = v.y; out3n[3 * i + 2] = v.z;
  }
}
void dlo_set_window(uint64_t mesh_id, const uint32_t* rect4) {
  o::g_window_on = rect4 != nullptr;
  o::g_window_mesh = mesh_id;
  if (rect4) o::g_window = o::Rect{rect4[0], rect4[1], rect4[2], rect4[3]};
}
void dlo_set_bounce_capture(uint32_t mode) {
  std::lock_guard<std::mutex> lock(o::g_capture_mu);
  o::g_capture_mode = mode;
  o::g_captured.clear();
  o::g_cubes.clear();
}
uint64_t dlo_bounce_capture_count(void) {
  std::lock_guard<std::mutex> lock(o::g_capture_mu);
  std::sort(o::g_captured.begin(), o::g_captured.end(), [](const o::CapturedCrop& a, const o::CapturedCrop& b) {
    if (a.mesh_id != b.mesh_id) return a.mesh_id < b.mesh_id;
    if (a.triangle != b.triangle) return a.triangle < b.triangle;
    return a.light_id < b.light_id;
  });
  return o::g_captured.size();
}
int dlo_bounce_capture_get(uint64_t i, uint64_t* meta5, double* rgb, uint64_t cap) {
  std::lock_guard<std::mutex> lock(o::g_capture_mu);
  if (i >= o::g_captured.size()) return DL_E_ARGUMENT;
  const o::CapturedCrop& c = o::g_captured[i];
  meta5[0] = c.mesh_id;
  meta5[1] = c.triangle;
  meta5[2] = c.light_id;
  meta5[3] = c.width;
  meta5[4] = c.height;
  if (rgb) std::memcpy(rgb, c.rgb.data(), std::min<uint64_t>(cap, c.rgb.size()) * sizeof(double));
  return DL_OK;
}
int dlo_photon_cube(uint64_t light_id, uint32_t* res, const float** dist, const int8_t** nrm, const uint8_t** alb) {
  std::lock_guard<std::mutex> lock(o::g_capture_mu);
  for (auto& c : o::g_cubes)
    if (c.first == light_id) {
      *res = c.second.res;
      *dist = c.second.dist.data();
      *nrm = c.second.nrm.data();
      *alb = c.second.alb.data();
      return DL_OK;
    }
  return DL_E_ARGUMENT;
}

}  // extern "C"

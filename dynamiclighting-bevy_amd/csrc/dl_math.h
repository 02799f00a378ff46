// f32 arithmetic of the bake, shared by host code and gfx950 kernels.
//
// Every function keeps the reference's operation order (glam 0.33 scalar formulas, Rust
// conversion semantics).  The library is compiled with -ffp-contract=off on both sides, f32
// division and sqrt are correctly rounded on gfx950 (hipcc default), and denormals are kept, so
// host and device produce the same bits.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DLI __host__ __device__ inline
#else
#include <cmath>
#define DLI inline
#endif

namespace dl {

struct f2 { float x, y; };
struct f3 { float x, y, z; };

DLI f2 mk2(float x, float y) { f2 r; r.x = x; r.y = y; return r; }
DLI f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
DLI f3 add(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
DLI f3 sub(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
DLI f3 neg(f3 a) { return mk3(-a.x, -a.y, -a.z); }
DLI f3 mul(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
DLI f3 divs(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
DLI bool eq(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
DLI bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
// glam dot: (x*x' + y*y') + z*z'
DLI float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glam cross: (y z' - y' z, z x' - z' x, x y' - x' y)
DLI f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
DLI float len_sq(f3 a) { return dot(a, a); }
DLI float len(f3 a) { return sqrtf(dot(a, a)); }
// Rust f32::min / f32::max: a NaN operand yields the other one.
DLI float minf_(float a, float b) { return fminf(a, b); }
DLI float maxf_(float a, float b) { return fmaxf(a, b); }
DLI f3 vmin(f3 a, f3 b) { return mk3(minf_(a.x, b.x), minf_(a.y, b.y), minf_(a.z, b.z)); }
DLI f3 vmax(f3 a, f3 b) { return mk3(maxf_(a.x, b.x), maxf_(a.y, b.y), maxf_(a.z, b.z)); }
DLI f2 vmin2(f2 a, f2 b) { return mk2(minf_(a.x, b.x), minf_(a.y, b.y)); }
DLI f2 vmax2(f2 a, f2 b) { return mk2(maxf_(a.x, b.x), maxf_(a.y, b.y)); }
DLI float max_elem(f3 a) { return maxf_(a.x, maxf_(a.y, a.z)); }
DLI float min_elem(f3 a) { return minf_(a.x, minf_(a.y, a.z)); }
DLI bool finite_(float v) { return v - v == 0.0f; }
// glam normalize family: v * (1 / length)
DLI f3 normalize(f3 a) { return mul(a, 1.0f / len(a)); }
DLI bool try_normalize(f3 a, f3* out) {
  float rcp = 1.0f / len(a);
  if (finite_(rcp) && rcp > 0.0f) { *out = mul(a, rcp); return true; }
  return false;
}
DLI f3 normalize_or(f3 a, f3 fallback) {
  f3 r;
  return try_normalize(a, &r) ? r : fallback;
}
DLI f3 normalize_or_zero(f3 a) { return normalize_or(a, mk3(0.0f, 0.0f, 0.0f)); }
DLI float signum_(float v) {
  if (v != v) return v;
  uint32_t b;
  __builtin_memcpy(&b, &v, 4);
  return (b >> 31) ? -1.0f : 1.0f;
}
// Vec3::any_orthonormal_pair (Duff et al. 2017)
DLI void orthonormal_pair(f3 n, f3* t, f3* b) {
  float sign = signum_(n.z);
  float a = -1.0f / (sign + n.z);
  float bb = (n.x * n.y) * a;
  *t = mk3(1.0f + ((sign * n.x) * n.x) * a, sign * bb, (-sign) * n.x);
  *b = mk3(bb, sign + (n.y * n.y) * a, -n.y);
}
// Quat * Vec3 = v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w)
DLI f3 quat_rotate(const float* q, f3 v) {
  f3 b = mk3(q[0], q[1], q[2]);
  float w = q[3];
  float b2 = dot(b, b);
  return add(add(mul(v, w * w - b2), mul(b, dot(v, b) * 2.0f)), mul(cross(b, v), w * 2.0f));
}

// Rust saturating casts
DLI uint32_t sat_u32(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)v;
}
DLI uint32_t sat_u8(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 255.0f) return 255u;
  return (uint32_t)v;
}
DLI int32_t sat_i8(float v) {
  if (v != v) return 0;
  if (v <= -128.0f) return -128;
  if (v >= 127.0f) return 127;
  return (int32_t)v;
}
// Rust f32::clamp keeps NaN
DLI float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

static constexpr float kTau = 6.28318530717958647692f;
static constexpr float kPi = 3.14159265358979323846f;
static constexpr float kFracPi4 = 0.785398163397448309616f;
static constexpr float kNdotlTolerance = -0.1f;  // mesh_bake.rs:25
static constexpr float kRayOffset = 0.001f;       // mesh_bake.rs:27
static constexpr float kEndGap = 1.0e-3f;         // area.rs:34

// ---- keyed random streams: effects/rng.rs:10-71, bake/random.rs:9-20 ----------------------
DLI uint64_t splitmix64(uint64_t v) {
  v += 0x9e3779b97f4a7c15ull;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
  return v ^ (v >> 31);
}
struct Pcg {
  uint64_t state, inc;
};
DLI uint32_t pcg_next(Pcg& r) {
  uint64_t old = r.state;
  r.state = old * 6364136223846793005ull + r.inc;
  uint32_t x = (uint32_t)(((old >> 18) ^ old) >> 27);
  uint32_t rot = (uint32_t)(old >> 59);
  return (x >> rot) | (x << ((32u - rot) & 31u));
}
DLI float pcg_f32(Pcg& r) { return (float)(pcg_next(r) >> 8) * (1.0f / 16777216.0f); }
DLI Pcg pcg_new(uint64_t seed, uint64_t stream) {
  Pcg r;
  r.state = 0;
  r.inc = (stream << 1) | 1ull;
  pcg_next(r);
  r.state += seed;
  pcg_next(r);
  return r;
}
DLI uint64_t key_fold(uint64_t acc, uint64_t part) { return splitmix64(acc ^ splitmix64(part)); }
static constexpr uint64_t kKeyInit = 0x0b4ce5eed0000000ull;
DLI Pcg pcg_from_key(uint64_t s) { return pcg_new(s, (s << 17) | (s >> 47)); }
DLI Pcg stream5(uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e) {
  uint64_t s = key_fold(key_fold(key_fold(key_fold(key_fold(kKeyInit, a), b), c), d), e);
  return pcg_from_key(s);
}
DLI Pcg stream6(uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e, uint64_t f) {
  uint64_t s = key_fold(key_fold(key_fold(key_fold(key_fold(key_fold(kKeyInit, a), b), c), d), e), f);
  return pcg_from_key(s);
}
DLI Pcg stream4(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  uint64_t s = key_fold(key_fold(key_fold(key_fold(kKeyInit, a), b), c), d);
  return pcg_from_key(s);
}
static constexpr uint64_t kBurialStream = 0x627572696564ull;  // mesh_texels.rs:28
static constexpr uint64_t kProbeStream = 0x736b7970726f6265ull;  // sky_probes.rs:21

// ---- texel rectangles: mesh_builder.rs:83-137 ------------------------------------------------
struct Rect { uint32_t x0, y0, x1, y1; };
DLI void uv_minmax(const f2* uv, f2* mn, f2* mx) {
  *mn = vmin2(vmin2(uv[0], uv[1]), uv[2]);
  *mx = vmax2(vmax2(uv[0], uv[1]), uv[2]);
}
DLI uint32_t texel_clamp(float v, uint32_t size) { return sat_u32(clampf_(v, 0.0f, (float)(size - 1))); }
DLI Rect crop_rect(const f2* uv, uint32_t size) {
  f2 mn, mx;
  uv_minmax(uv, &mn, &mx);
  float s = (float)size;
  Rect r;
  r.x0 = texel_clamp(floorf(mn.x * s) - 2.0f, size);
  r.y0 = texel_clamp(floorf(mn.y * s) - 2.0f, size);
  r.x1 = texel_clamp(ceilf(mx.x * s) + 2.0f, size);
  r.y1 = texel_clamp(ceilf(mx.y * s) + 2.0f, size);
  return r;
}
DLI Rect center_rect(const f2* uv, uint32_t size, float grow) {
  f2 mn, mx;
  uv_minmax(uv, &mn, &mx);
  float s = (float)size;
  Rect r;
  r.x0 = texel_clamp(floorf(mn.x * s - 0.5f) - grow, size);
  r.y0 = texel_clamp(floorf(mn.y * s - 0.5f) - grow, size);
  r.x1 = texel_clamp(ceilf(mx.x * s - 0.5f) + grow, size);
  r.y1 = texel_clamp(ceilf(mx.y * s - 0.5f) + grow, size);
  return r;
}
// touched_texels scan range: texels.rs:86-93 (returns false when empty)
DLI bool touched_range(const f2* uv, uint32_t size, Rect* r) {
  float s = (float)size;
  f2 mn, mx;
  uv_minmax(uv, &mn, &mx);
  f2 first = vmax2(mk2(floorf(mn.x * s), floorf(mn.y * s)), mk2(0.0f, 0.0f));
  f2 last = vmin2(mk2(floorf(mx.x * s), floorf(mx.y * s)), mk2(s - 1.0f, s - 1.0f));
  r->x0 = sat_u32(first.x);
  r->y0 = sat_u32(first.y);
  r->x1 = sat_u32(last.x);
  r->y1 = sat_u32(last.y);
  return r->x0 <= r->x1 && r->y0 <= r->y1;
}

// ---- geometry: geometry.rs:59-140 ------------------------------------------------------------
DLI float tri_area(f3 a, f3 b, f3 c) { return len(cross(sub(b, a), sub(c, a))) * 0.5f; }
DLI f3 closest_point(f3 p, f3 a, f3 b, f3 c) {
  f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  f3 bp = sub(p, b);
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return add(a, mul(ab, d1 / (d1 - d3)));
  f3 cp = sub(p, c);
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return add(a, mul(ac, d2 / (d2 - d6)));
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
    return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
  float denom = 1.0f / ((va + vb) + vc);
  return add(add(a, mul(ab, vb * denom)), mul(ac, vc * denom));
}
DLI float signed_area2(f2 a, f2 b, f2 c) {
  float ax = a.x - c.x, ay = a.y - c.y, bx = b.x - c.x, by = b.y - c.y;
  return (ax * by - ay * bx) * 0.5f;
}
DLI bool uv_bary(const f2* t, f2 p, f3* w) {
  float area = signed_area2(t[0], t[1], t[2]);
  if (area == 0.0f) return false;
  float a1 = signed_area2(t[1], t[2], p) / area;
  float a2 = signed_area2(t[2], t[0], p) / area;
  float a3 = signed_area2(t[0], t[1], p) / area;
  if (a1 >= 0.0f && a2 >= 0.0f && a3 >= 0.0f) { *w = mk3(a1, a2, a3); return true; }
  return false;
}
DLI f2 texel_center(uint32_t x, uint32_t y, uint32_t size) {
  return mk2(((float)x + 0.5f) / (float)size, ((float)y + 0.5f) / (float)size);
}
DLI f3 lerp3w(const f3* v, f3 w) { return add(add(mul(v[0], w.x), mul(v[1], w.y)), mul(v[2], w.z)); }
// texels.rs:125-141
DLI bool nearest_weights(const f2* uv, f2 p, f3* out) {
  f3 a = mk3(uv[0].x, uv[0].y, 0.0f), b = mk3(uv[1].x, uv[1].y, 0.0f), c = mk3(uv[2].x, uv[2].y, 0.0f);
  f3 n3 = closest_point(mk3(p.x, p.y, 0.0f), a, b, c);
  f3 nl = mk3(n3.x, n3.y, 0.0f);
  float area = cross(sub(b, a), sub(c, a)).z;
  if (area == 0.0f || !finite_(area)) return false;
  f3 w = mk3(cross(sub(b, nl), sub(c, nl)).z, cross(sub(c, nl), sub(a, nl)).z, cross(sub(a, nl), sub(b, nl)).z);
  w = divs(w, area);
  w = vmax(w, mk3(0.0f, 0.0f, 0.0f));
  float sum = (w.x + w.y) + w.z;
  *out = divs(w, maxf_(sum, 1.17549435e-38f));
  return true;
}
DLI bool approx_eq3(f3 a, f3 b, float eps) {
  f3 d = sub(a, b);
  return max_elem(mk3(fabsf(d.x), fabsf(d.y), fabsf(d.z))) < eps;
}
DLI bool sphere_touches(f3 center, float radius, f3 a, f3 b, f3 c) {
  f3 cp = closest_point(center, a, b, c);
  return len_sq(sub(cp, center)) <= radius * radius;
}

// ---- two-sided Moller-Trumbore: occluder.rs:225-251 -------------------------------------------
// e1 = v1 - v0 and e2 = v2 - v0 are stored precomputed (same bits).  Returns t, or a negative
// value for a miss; `u`, `v` give the barycentrics (1 - u - v, u, v).
DLI bool mt_hit(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float max, float* t_out, float* u_out, float* v_out) {
  f3 p = cross(d, e2);
  float det = dot(e1, p);
  if (det == 0.0f || !finite_(det)) return false;
  float inv = 1.0f / det;
  f3 s = sub(o, v0);
  float u = dot(s, p) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  f3 q = cross(s, e1);
  float v = dot(d, q) * inv;
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = dot(e2, q) * inv;
  if (!(t >= 0.0f && t <= max)) return false;
  *t_out = t;
  *u_out = u;
  *v_out = v;
  return true;
}

// ---- low-discrepancy points: spread.rs:16-58 -------------------------------------------------
DLI uint32_t bitrev32(uint32_t v) {
  v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
  v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
  v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
  v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
  return (v >> 16) | (v << 16);
}
DLI uint32_t sobol2(uint32_t index) {
  uint32_t value = 0, dir = 1u << 31;
  while (index != 0) {
    if (index & 1u) value ^= dir;
    index >>= 1;
    dir ^= dir >> 1;
  }
  return value;
}
DLI float unit24(uint32_t bits) { return (float)(bits >> 8) / 16777216.0f; }
DLI f2 spread_pt(uint32_t index, f2 offset, bool rectangle) {
  float a = unit24(bitrev32(index)) + offset.x;
  float b = (rectangle ? unit24(sobol2(index)) : 0.0f) + offset.y;
  return mk2(a - truncf(a), b - truncf(b));
}

// ---- shadow levels: shadow_grid.rs:13-16 -----------------------------------------------------
DLI uint32_t level_of(float fraction, uint32_t bits) {
  float mx = (float)((1u << bits) - 1u);
  return sat_u8(roundf(clampf_(fraction, 0.0f, 1.0f) * mx));
}

// ---- IEEE binary16, round to nearest even (half 2.7.1 from_f32) ------------------------------
DLI uint32_t f32_to_f16_bits(float f) {
  uint32_t x;
  __builtin_memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u, e8 = (x >> 23) & 0xffu, man = x & 0x7fffffu;
  if (e8 == 0xffu) return sign | 0x7c00u | (man ? (0x200u | (man >> 13)) : 0u);
  int32_t e = (int32_t)e8 - 112;
  if (e >= 31) return sign | 0x7c00u;
  if (e <= 0) {
    if (e < -10) return sign;
    man |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - e);
    uint32_t h = man >> shift, rem = man & ((1u << shift) - 1u), mid = 1u << (shift - 1u);
    if (rem > mid || (rem == mid && (h & 1u))) h++;
    return sign | h;
  }
  uint32_t h = ((uint32_t)e << 10) | (man >> 13), rem = man & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return sign | h;
}

// ---- cube conventions: shared/cube.rs:29-105 -------------------------------------------------
DLI f3 cube_texel_dir(uint32_t face, uint32_t x, uint32_t y, uint32_t res) {
  float size = (float)res;
  float u = ((float)x + 0.5f) / size, v = ((float)y + 0.5f) / size;
  float sc = u * 2.0f - 1.0f, tc = v * 2.0f - 1.0f;
  f3 d;
  switch (face) {
    case 0: d = mk3(1.0f, -tc, -sc); break;
    case 1: d = mk3(-1.0f, -tc, sc); break;
    case 2: d = mk3(sc, 1.0f, tc); break;
    case 3: d = mk3(sc, -1.0f, -tc); break;
    case 4: d = mk3(sc, -tc, 1.0f); break;
    default: d = mk3(-sc, -tc, -1.0f); break;
  }
  return normalize(d);
}
DLI void cube_texel_parts(f3 d, uint32_t res, uint32_t* face_out, uint32_t* x_out, uint32_t* y_out) {
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  uint32_t face;
  float sc, tc, ma;
  if (ax > ay && ax > az) {
    if (d.x > 0.0f) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
    ma = ax;
  } else if (ay > az) {
    if (d.y > 0.0f) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
    ma = ay;
  } else {
    if (d.z > 0.0f) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
    ma = az;
  }
  float u = 0.5f, v = 0.5f;
  if (ma != 0.0f) { u = (sc / ma) * 0.5f + 0.5f; v = (tc / ma) * 0.5f + 0.5f; }
  float mx = (float)(res == 0 ? 0u : res - 1u);
  *face_out = face;
  *x_out = sat_u32(clampf_(floorf(u * (float)res), 0.0f, mx));
  *y_out = sat_u32(clampf_(floorf(v * (float)res), 0.0f, mx));
}
DLI uint32_t cube_texel_of(f3 d, uint32_t res) {
  uint32_t face, x, y;
  cube_texel_parts(d, res, &face, &x, &y);
  return face * res * res + y * res + x;
}

}  // namespace dl

// gfx950 kernels of the bake.  Each kernel cites the reference function it computes; the f32
// arithmetic it shares with the host lives in dl_math.h.
//
// Ray queries traverse the 4-wide occluder BVH one ray per lane, with the traversal stack in
// LDS (stack[i * blockDim + lane], conflict-free); the nearest hit child is taken next without
// a push.  Hot loops are wave-uniform in their (triangle, light) work unit so neighbouring lanes
// trace neighbouring texels toward the same light.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>

#include "dl_kernels.h"

namespace dl {

static constexpr float kBoxEps = 1.0f / 65536.0f;
#ifndef DL_GATHER_GROUP
#define DL_GATHER_GROUP 2
#endif
static constexpr int kGatherGroup = DL_GATHER_GROUP;  // samples generated together per job
static constexpr uint32_t kThreads = 256;

__device__ __forceinline__ uint64_t owner_in(const uint64_t* __restrict__ start, uint64_t lo, uint64_t hi, uint64_t idx) {
  while (hi - lo > 1) {
    uint64_t mid = (lo + hi) >> 1;
    if (start[mid] <= idx) lo = mid;
    else hi = mid;
  }
  return lo;
}
// The segment s < n with start[s] <= idx < start[s + 1] (the last such s when segments are empty).
__device__ __forceinline__ uint64_t owner_of(const uint64_t* __restrict__ start, uint64_t n, uint64_t idx) {
  return owner_in(start, 0, n, idx);
}
// owner_of for grid-stride loops whose lanes hold increasing indices: one search per wave, for its
// first active lane, and a search of its own only for a lane past that segment.
__device__ __forceinline__ uint64_t owner_of_wave(const uint64_t* __restrict__ start, uint64_t n, uint64_t idx) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)idx), hi = __builtin_amdgcn_readfirstlane((uint32_t)(idx >> 32));
  const uint64_t g = owner_in(start, 0, n, ((uint64_t)hi << 32) | lo);
  if (g + 1 >= n || idx < start[g + 1]) return g;
  return owner_in(start, g + 1, n, idx);
}

// Reciprocal of a direction component for the slab tests (hardware reciprocal, about 1 ulp: the
// box tests only cull, and their padding covers it).
__device__ __forceinline__ float inv_dir(float d) {
  float c = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
  return __builtin_amdgcn_rcpf(c);
}

struct Hit {
  float t, u, v;
  uint32_t occ;  // occluder index (lowest wins ties)
  uint32_t pos;  // leaf position of the triangle
};

// One ray against the occluder BVH.  ANY: true when some occluder has 0 <= t <= tmax (t finite).
// !ANY: the lowest (t, occluder index) hit with t <= tmax in `hit`.
// Node fetches and triangle tests of one lane, summed in registers and flushed per wave when a
// kernel runs instrumented (DevScene::counters).
struct Tally {
  uint32_t nodes = 0, tris = 0;
};

// Texel index of a repeating texture coordinate (input.rs:90-97: rem_euclid, floor, clamp).
__device__ __forceinline__ uint32_t wrap_texel(float c, uint32_t size) {
  float r = fmodf(c, 1.0f);
  if (r < 0.0f) r = r + 1.0f;
  const float texel = floorf(r * (float)size);
  if (!finite_(texel)) return 0u;
  const uint32_t v = sat_u32(texel);
  return v < size - 1 ? v : size - 1;
}
// Whether a hit with barycentrics ((1-u)-v, u, v) on occluder `occ` lands on an opaque texel of
// its alpha texture (occluder.rs:147-160, AlphaTexture::is_opaque input.rs:79-88).
__device__ __forceinline__ bool alpha_blocks(const DevScene& sc, uint32_t occ, float u, float v) {
  if (!sc.occ_tex) return true;
  const uint32_t t = sc.occ_tex[occ];
  if (t == 0xffffffffu) return true;
  const float* uv = sc.occ_uv0 + 6 * (size_t)occ;
  const float wx = (1.0f - u) - v;
  const float ux = (uv[0] * wx + uv[2] * u) + uv[4] * v;
  const float uy = (uv[1] * wx + uv[3] * u) + uv[5] * v;
  const uint32_t* ti = sc.tex + 4 * (size_t)t;
  const uint64_t bit = (uint64_t)wrap_texel(uy, ti[1]) * ti[0] + wrap_texel(ux, ti[0]);
  return (sc.tex_bits[ti[2] + bit / 32] >> (bit % 32)) & 1u;
}


// A ray prepared for slab tests: origin, reciprocal direction and origin * reciprocal.
struct SlabRay {
  float ix, iy, iz, ox, oy, oz;
};
__device__ __forceinline__ SlabRay slab_ray(f3 o, f3 d) {
  SlabRay r;
  r.ix = inv_dir(d.x);
  r.iy = inv_dir(d.y);
  r.iz = inv_dir(d.z);
  r.ox = o.x * r.ix;
  r.oy = o.y * r.iy;
  r.oz = o.z * r.iz;
  return r;
}

// The child slots of a node (> 0 inner node, < 0 leaf, 0 unused), held as int4 vectors: a select
// over the elements of an array would be turned back into an indexed load from scratch memory.
struct Kids {
  int4 lo;
#if DL_BVH_WIDTH == 8
  int4 hi;
#endif
};
__device__ __forceinline__ int32_t sel4(const int4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
// Child slot k of a node (k known at compile time or not).
__device__ __forceinline__ int32_t kid(const Kids& ch, int k) {
#if DL_BVH_WIDTH == 8
  return k < 4 ? sel4(ch.lo, k) : sel4(ch.hi, k - 4);
#else
  return sel4(ch.lo, k);
#endif
}

// Box tests of the (up to kBvhWidth) children of `node`: bit k of the result is set when child
// k's decoded slab interval meets [0, limit], with tn[k] its entry distance.  A child bound is
// origin + q * step (8-bit q, power-of-two step); its plane distance is evaluated as
// q * (step / d) + (origin / d - o / d) with fused multiply-adds.  step / d is exact, and the
// roundings of the other terms are relative to the node's coordinates, the ray origin and the
// distance itself: the extra quantization step (at least the float spacing of the node's
// coordinates), the boxes' padding (2^-15 of the coordinate) and the interval's widening
// (2^-16 of the distance) cover them, so the test only culls boxes the ray misses.  The node
// record is dead once this returns, which keeps it out of the registers of the leaf tests.
__device__ __forceinline__ uint32_t node_hits(const DevScene& sc, int32_t node, const SlabRay& r, float limit,
                                              float tn[kBvhWidth], Kids* ch) {
  constexpr int kQv = (kNodeQWords + kNodeQPad) / 4;  // float4 rows of quantized bounds
  constexpr int kw = kBvhWidth / 4;                   // words per bound row
  const uint4* nb = reinterpret_cast<const uint4*>(sc.nodes + node);
  const uint4 q0 = nb[0];
  uint32_t qw[4 * kQv];
#pragma unroll
  for (int v = 0; v < kQv; ++v) {
    const uint4 t = nb[1 + v];
    qw[4 * v] = t.x;
    qw[4 * v + 1] = t.y;
    qw[4 * v + 2] = t.z;
    qw[4 * v + 3] = t.w;
  }
  // 0 marks an unused child slot (the root is no child)
  ch->lo = reinterpret_cast<const int4*>(nb + 1 + kQv)[0];
#if DL_BVH_WIDTH == 8
  ch->hi = reinterpret_cast<const int4*>(nb + 1 + kQv)[1];
#endif
  const float ax = __uint_as_float((q0.w & 0xffu) << 23) * r.ix;
  const float ay = __uint_as_float(((q0.w >> 8) & 0xffu) << 23) * r.iy;
  const float az = __uint_as_float(((q0.w >> 16) & 0xffu) << 23) * r.iz;
  const float bx = __builtin_fmaf(__uint_as_float(q0.x), r.ix, -r.ox);
  const float by = __builtin_fmaf(__uint_as_float(q0.y), r.iy, -r.oy);
  const float bz = __builtin_fmaf(__uint_as_float(q0.z), r.iz, -r.oz);
  uint32_t mask = 0;
#pragma unroll
  for (int k = 0; k < kBvhWidth; ++k) {
    const int sh = 8 * (k % 4), w = k / 4;
    const float t0x = __builtin_fmaf((float)((qw[w] >> sh) & 0xffu), ax, bx);
    const float t1x = __builtin_fmaf((float)((qw[kw + w] >> sh) & 0xffu), ax, bx);
    const float t0y = __builtin_fmaf((float)((qw[2 * kw + w] >> sh) & 0xffu), ay, by);
    const float t1y = __builtin_fmaf((float)((qw[3 * kw + w] >> sh) & 0xffu), ay, by);
    const float t0z = __builtin_fmaf((float)((qw[4 * kw + w] >> sh) & 0xffu), az, bz);
    const float t1z = __builtin_fmaf((float)((qw[5 * kw + w] >> sh) & 0xffu), az, bz);
    float n = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    float f = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    n = __builtin_fmaf(fabsf(n), -kBoxEps, n);
    f = __builtin_fmaf(fabsf(f), kBoxEps, f);
    tn[k] = n;
    if (kid(*ch, k) != 0 && n <= f && f >= 0.0f && n <= limit) mask |= 1u << k;
  }
  return mask;
}

// Sum of v over the wave, added to *dst by its first lane (all 64 lanes must call it).
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63u) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// base[key] += v over the active lanes of a wave: one atomic when they share a key, one per lane
// otherwise (all 64 lanes must call it).
__device__ __forceinline__ void wave_add_at(unsigned long long* base, uint32_t key, uint32_t v, bool active) {
  const uint64_t act = __ballot(active);
  if (!act) return;
  const uint32_t k0 = __shfl(key, (int)__builtin_ctzll(act), 64);
  if (__ballot(active && key != k0) == 0ull) wave_add(&base[k0], active ? v : 0u);
  else if (active && v) atomicAdd(&base[key], (unsigned long long)v);
}

// Bit k set when child k is a leaf (a negative child reference).
__device__ __forceinline__ uint32_t leaf_bits(const Kids& ch) {
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < kBvhWidth; ++k) m |= kid(ch, k) < 0 ? 1u << k : 0u;
  return m;
}

// Pushes the hit inner children of a node except the nearest, which it returns (-1 if none).
__device__ __forceinline__ int32_t push_children(const DevScene& sc, uint32_t mask, const Kids& ch, const float tn[kBvhWidth],
                                                 uint32_t* __restrict__ stack, uint32_t stride, uint32_t& sp) {
  uint16_t* const stack16 = reinterpret_cast<uint16_t*>(stack - threadIdx.x) + threadIdx.x;
  const bool s16 = sc.stack16 != 0;
  int32_t next = -1;
  float next_t = INFINITY;
#pragma unroll
  for (int k = 0; k < kBvhWidth; ++k) {
    const int32_t c = kid(ch, k);
    if (!((mask >> k) & 1u) || c < 0) continue;
    if (next < 0) {
      next = c;
      next_t = tn[k];
      continue;
    }
    uint32_t push = (uint32_t)c;
    if (tn[k] < next_t) {
      push = (uint32_t)next;
      next = c;
      next_t = tn[k];
    }
    if (s16) stack16[sp * stride] = (uint16_t)push;
    else stack[sp * stride] = push;
    sp++;
  }
  if (sp > sc.stack) atomicOr(sc.overflow, 1u);
  return next;
}

// Closest-hit order: pushes the hit inner children of a node except the nearest, farthest first
// (so that pops come nearest first), and returns the nearest (-1 if none).
__device__ __forceinline__ int32_t push_children_sorted(const DevScene& sc, uint32_t mask, const Kids& ch,
                                                        const float tn[kBvhWidth], uint32_t* __restrict__ stack,
                                                        uint32_t stride, uint32_t& sp) {
  uint16_t* const stack16 = reinterpret_cast<uint16_t*>(stack - threadIdx.x) + threadIdx.x;
  const bool s16 = sc.stack16 != 0;
  float key[kBvhWidth];
  int32_t val[kBvhWidth];
#pragma unroll
  for (int k = 0; k < kBvhWidth; ++k) {
    const int32_t c = kid(ch, k);
    const bool push = ((mask >> k) & 1u) && c > 0;
    val[k] = push ? c : 0;  // 0: nothing to push (the order is a heuristic; no child is dropped)
    key[k] = push ? tn[k] : INFINITY;
  }
  // sorting network on (key, val), ascending
#define DL_CSWAP(a, b)                                   \
  if (key[b] < key[a]) {                                 \
    const float tk = key[a];                             \
    key[a] = key[b];                                     \
    key[b] = tk;                                         \
    const int32_t tv = val[a];                           \
    val[a] = val[b];                                     \
    val[b] = tv;                                         \
  }
  if constexpr (kBvhWidth == 4) {
    DL_CSWAP(0, 1) DL_CSWAP(2, 3) DL_CSWAP(0, 2) DL_CSWAP(1, 3) DL_CSWAP(1, 2)
  } else {  // 19 compare-exchanges (Batcher's network for eight)
    DL_CSWAP(0, 2) DL_CSWAP(1, 3) DL_CSWAP(4, 6) DL_CSWAP(5, 7) DL_CSWAP(0, 4) DL_CSWAP(1, 5) DL_CSWAP(2, 6)
    DL_CSWAP(3, 7) DL_CSWAP(0, 1) DL_CSWAP(2, 3) DL_CSWAP(4, 5) DL_CSWAP(6, 7) DL_CSWAP(2, 4) DL_CSWAP(3, 5)
    DL_CSWAP(1, 4) DL_CSWAP(3, 6) DL_CSWAP(1, 2) DL_CSWAP(3, 4) DL_CSWAP(5, 6)
  }
#undef DL_CSWAP
#pragma unroll
  for (int k = kBvhWidth - 1; k >= 1; --k) {
    if (val[k] == 0) continue;
    if (s16) stack16[sp * stride] = (uint16_t)val[k];
    else stack[sp * stride] = (uint32_t)val[k];
    sp++;
  }
  if (sp > sc.stack) atomicOr(sc.overflow, 1u);
  return val[0] == 0 ? -1 : val[0];
}

// The triangles of the hit leaf children of a closest-hit node visit, nearest box first; a leaf
// whose box starts beyond the best hit so far (the bound of node_hits with the hit's distance)
// holds no hit at or below it and is skipped.  Keeps the lowest (t, occluder index) hit.
__device__ __forceinline__ void closest_leaves(const DevScene& sc, uint32_t leaves, const Kids& ch, const float tn[kBvhWidth], f3 o, f3 d,
                                               float tmax, uint32_t alpha, bool& have, Hit& hit, Tally& tl) {
  while (leaves) {
    int kb = 0;
    float tb = INFINITY;
    bool any = false;
#pragma unroll
    for (int k = 0; k < kBvhWidth; ++k)
      if (((leaves >> k) & 1u) && (!any || tn[k] < tb)) {
        kb = k;
        tb = tn[k];
        any = true;
      }
    leaves &= ~(1u << kb);
    if (have && tb > hit.t * (1.0f + kBoxEps)) continue;
    const uint32_t v = (uint32_t)(-1 - kid(ch, kb));
    for (uint32_t i = v >> 2, end = (v >> 2) + (v & 3u) + 1u; i < end; ++i) {
      const float4 a = sc.tris[3 * i], b = sc.tris[3 * i + 1], cc = sc.tris[3 * i + 2];
      float t, u, vv;
      tl.tris++;
      if (!mt_hit(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), o, d, tmax, &t, &u, &vv)) continue;
      if (!finite_(t)) continue;
      if (alpha && !alpha_blocks(sc, __float_as_uint(a.w), u, vv)) continue;
      const uint32_t occ = __float_as_uint(a.w);
      if (!have || t < hit.t || (t == hit.t && occ < hit.occ)) {
        have = true;
        hit.t = t;
        hit.u = u;
        hit.v = vv;
        hit.occ = occ;
        hit.pos = i;
      }
    }
  }
}

__device__ __forceinline__ int32_t pop_node(const DevScene& sc, const uint32_t* __restrict__ stack, uint32_t stride, uint32_t sp) {
  const uint16_t* const stack16 = reinterpret_cast<const uint16_t*>(stack - threadIdx.x) + threadIdx.x;
  return sc.stack16 ? (int32_t)stack16[sp * stride] : (int32_t)stack[sp * stride];
}

// `alpha`: the light's rays pass transparent texels of alpha-tested occluders.  Children are
// visited nearest first; leaves are tested before the inner children are pushed.  Neither order
// changes a result: any-hit reports whether some hit exists and closest-hit keeps the lowest
// (t, occluder index).
template <bool ANY>
__device__ __forceinline__ bool traverse(const DevScene& sc, uint32_t* __restrict__ stack, uint32_t stride, f3 o, f3 d,
                                         float tmax, Hit* hit, Tally& tl, uint32_t alpha = 0) {
  if (sc.ntris == 0) return false;
  const SlabRay sr = slab_ray(o, d);
  bool have = false;
  float bt = tmax;
  uint32_t sp = 0;
  int32_t node = 0;
  for (;;) {
    tl.nodes++;
    float tn[kBvhWidth];
    Kids ch;
    const float limit = (ANY || !have ? tmax : bt) * (1.0f + kBoxEps);
    const uint32_t mask = node_hits(sc, node, sr, limit, tn, &ch);
    if constexpr (!ANY) {
      closest_leaves(sc, mask & leaf_bits(ch), ch, tn, o, d, tmax, alpha, have, *hit, tl);
      if (have) bt = hit->t;
      const int32_t next = push_children_sorted(sc, mask, ch, tn, stack, stride, sp);
      if (next >= 0) {
        node = next;
        continue;
      }
      if (sp == 0) break;
      sp--;
      node = pop_node(sc, stack, stride, sp);
      continue;
    }
    // the triangles of every hit leaf child, one test per loop trip
    uint32_t leaves = mask & leaf_bits(ch), i = 0, end = 0;
    for (;;) {
      if (i == end) {
        if (!leaves) break;
        const uint32_t v = (uint32_t)(-1 - kid(ch, __builtin_ctz(leaves)));
        leaves &= leaves - 1;
        i = v >> 2;
        end = i + (v & 3u) + 1u;
      }
      const float4 a = sc.tris[3 * i], b = sc.tris[3 * i + 1], cc = sc.tris[3 * i + 2];
      const uint32_t pos = i++;
      float t, u, vv;
      tl.tris++;
      if (!mt_hit(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), o, d, tmax, &t, &u, &vv)) continue;
      if (!finite_(t)) continue;
      if (alpha && !alpha_blocks(sc, __float_as_uint(a.w), u, vv)) continue;
      if (ANY) return true;
      const uint32_t occ = __float_as_uint(a.w);
      if (!have || t < bt || (t == bt && occ < hit->occ)) {
        have = true;
        bt = t;
        hit->t = t;
        hit->u = u;
        hit->v = vv;
        hit->occ = occ;
        hit->pos = pos;
      }
    }
    const int32_t next = push_children(sc, mask, ch, tn, stack, stride, sp);
    if (next >= 0) {
      node = next;
      continue;
    }
    if (sp == 0) break;
    sp--;
    node = pop_node(sc, stack, stride, sp);
  }
  return have;
}

__device__ __forceinline__ f3 tri_face_normal(const DevScene& sc, uint32_t pos) {
  const float4 b = sc.tris[3 * pos + 1], c = sc.tris[3 * pos + 2];
  return normalize_or_zero(cross(mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z)));
}

__device__ __forceinline__ f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// ---- light association and per-triangle rectangles: mesh_bake.rs:141-173 ----------------------
// light_shape.rs:47-65
__device__ bool may_face(const DevLight& L, const f3* v, f3 normal) {
  f3 center = divs(add(add(v[0], v[1]), v[2]), 3.0f);
  if (L.shape == 4) return true;  // sky
  if (L.shape == 3) return dot(normal, ld3(L.toward)) + L.ssin > kNdotlTolerance;
  if (L.shape == 0) return dot(normal, normalize_or_zero(sub(ld3(L.pos), center))) > kNdotlTolerance;
  f3 c = ld3(L.pos), x = ld3(L.ax), y = ld3(L.ay), fw = ld3(L.af);
  if (L.rect) {
    bool lit_side = false;
    for (int k = 0; k < 3; ++k) lit_side |= dot(fw, sub(v[k], c)) > 0.0f;
    if (!lit_side) return false;
  }
  f3 hx = mul(x, L.hx), hy = mul(y, L.hy);
  f3 corners[4] = {sub(sub(c, hx), hy), sub(add(c, hx), hy), add(add(c, hx), hy), add(sub(c, hx), hy)};
  for (int k = 0; k < 4; ++k)
    if (dot(normal, normalize_or_zero(sub(corners[k], center))) > kNdotlTolerance) return true;
  return false;
}

__device__ __forceinline__ bool light_keeps(const DevLight& L, uint32_t slot, uint32_t local_lights, const f3* v,
                                            f3 face, bool valid, bool flat, f3 bmin, f3 bmax) {
  if (slot < local_lights) {
    // light BVH query: sphere bounds overlap the triangle bounds (light_bvh.rs:196-220)
    if (!(L.bmin[0] <= bmax.x && L.bmin[1] <= bmax.y && L.bmin[2] <= bmax.z && bmin.x <= L.bmax[0] &&
          bmin.y <= L.bmax[1] && bmin.z <= L.bmax[2]))
      return false;
  }
  if (valid && flat && !L.bakes_bounce && !may_face(L, v, face)) return false;
  return L.distant || sphere_touches(ld3(L.pos), L.radius, v[0], v[1], v[2]);
}

__device__ __forceinline__ void load_tri(const DevTris& tr, uint32_t t, f3* v, f3* n, f2* uv) {
  for (int k = 0; k < 3; ++k) {
    v[k] = ld3(tr.v + (size_t)t * 9 + 3 * k);
    n[k] = ld3(tr.nrm + (size_t)t * 9 + 3 * k);
    uv[k] = mk2(tr.uv[(size_t)t * 6 + 2 * k], tr.uv[(size_t)t * 6 + 2 * k + 1]);
  }
}

__global__ void k_setup(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, uint32_t nlights,
                        uint32_t local_lights, uint64_t* light_count, uint64_t* trace_count, uint64_t* touched_count) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tr.n; t += gridDim.x * blockDim.x) {
    f3 v[3], n[3];
    f2 uv[3];
    load_tri(tr, t, v, n, uv);
    f3 face = normalize_or_zero(cross(sub(v[1], v[0]), sub(v[2], v[0])));
    bool valid = !is_zero(face);
    bool flat = approx_eq3(n[0], n[1], 1e-5f) && approx_eq3(n[1], n[2], 1e-5f);
    f3 bmin = vmin(vmin(v[0], v[1]), v[2]), bmax = vmax(vmax(v[0], v[1]), v[2]);
    uint32_t cnt = 0;
    for (uint32_t s = 0; s < nlights; ++s) cnt += light_keeps(lights[s], s, local_lights, v, face, valid, flat, bmin, bmax);
    light_count[t] = cnt;
    uint32_t res = ms.res[tr.mesh[t]];
    Rect crop = crop_rect(uv, res);
    tr.crop[4 * t] = crop.x0;
    tr.crop[4 * t + 1] = crop.y0;
    tr.crop[4 * t + 2] = crop.x1;
    tr.crop[4 * t + 3] = crop.y1;
    tr.face[3 * t] = face.x;
    tr.face[3 * t + 1] = face.y;
    tr.face[3 * t + 2] = face.z;
    uint64_t a = 0, b = 0;
    if (valid) {
      Rect r = center_rect(uv, res, 0.0f);
      a = (uint64_t)(r.x1 - r.x0 + 1) * (r.y1 - r.y0 + 1);
      Rect q;
      if (touched_range(uv, res, &q)) b = (uint64_t)(q.x1 - q.x0 + 1) * (q.y1 - q.y0 + 1);
    }
    trace_count[t] = a;
    touched_count[t] = a + b;
  }
}

__global__ void k_assoc_fill(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, uint32_t nlights,
                             uint32_t local_lights, const uint64_t* __restrict__ ent_start, uint32_t* ent_light,
                             uint32_t* ent_tri) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tr.n; t += gridDim.x * blockDim.x) {
    f3 v[3], n[3];
    f2 uv[3];
    load_tri(tr, t, v, n, uv);
    f3 face = ld3(tr.face + 3 * (size_t)t);
    bool valid = !is_zero(face);
    bool flat = approx_eq3(n[0], n[1], 1e-5f) && approx_eq3(n[1], n[2], 1e-5f);
    f3 bmin = vmin(vmin(v[0], v[1]), v[2]), bmax = vmax(vmax(v[0], v[1]), v[2]);
    uint64_t w = ent_start[t];
    for (uint32_t s = 0; s < nlights; ++s)
      if (light_keeps(lights[s], s, local_lights, v, face, valid, flat, bmin, bmax)) {
        ent_light[w] = s;
        ent_tri[w] = t;
        w++;
      }
  }
}

// ---- owned texels: mesh_texels.rs:47-55 over triangle_texels (texels.rs:48-72) ------------------
__global__ void k_owned(DevTris tr, DevMeshes ms, const uint64_t* __restrict__ start, uint64_t total) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t t = (uint32_t)owner_of_wave(start, tr.n, c);
    uint64_t local = c - start[t];
    uint32_t m = tr.mesh[t], res = ms.res[m];
    f2 uv[3];
    for (int k = 0; k < 3; ++k) uv[k] = mk2(tr.uv[(size_t)t * 6 + 2 * k], tr.uv[(size_t)t * 6 + 2 * k + 1]);
    Rect r = center_rect(uv, res, 0.0f);
    uint32_t w = r.x1 - r.x0 + 1;
    uint32_t x = r.x0 + (uint32_t)(local % w), y = r.y0 + (uint32_t)(local / w);
    f3 bw;
    if (uv_bary(uv, texel_center(x, y, res), &bw)) {
      uint64_t i = (uint64_t)x + (uint64_t)y * res;
      atomicOr(&ms.owned[ms.owned_off[m] + i / 32], 1u << (i % 32));
    }
  }
}

// ---- traced texels: centered then touched (texels.rs:48-121) -----------------------------------
__global__ void k_texels(DevTris tr, DevMeshes ms, const uint64_t* __restrict__ trace_count,
                         const uint64_t* __restrict__ start, uint64_t total, int fill, unsigned long long* counts,
                         const uint64_t* __restrict__ tex_start, DevTexels out) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t t = (uint32_t)owner_of_wave(start, tr.n, c);
    uint64_t local = c - start[t];
    uint32_t m = tr.mesh[t], res = ms.res[m];
    f3 v[3], n[3];
    f2 uv[3];
    load_tri(tr, t, v, n, uv);
    f3 face = ld3(tr.face + 3 * (size_t)t);
    uint32_t x, y;
    f3 w;
    uint64_t a = trace_count[t];
    if (local < a) {
      Rect r = center_rect(uv, res, 0.0f);
      uint32_t rw = r.x1 - r.x0 + 1;
      x = r.x0 + (uint32_t)(local % rw);
      y = r.y0 + (uint32_t)(local / rw);
      if (!uv_bary(uv, texel_center(x, y, res), &w)) continue;
    } else {
      local -= a;
      Rect q;
      touched_range(uv, res, &q);
      uint32_t qw = q.x1 - q.x0 + 1;
      x = q.x0 + (uint32_t)(local % qw);
      y = q.y0 + (uint32_t)(local / qw);
      uint64_t i = (uint64_t)x + (uint64_t)y * res;
      if ((ms.owned[ms.owned_off[m] + i / 32] >> (i % 32)) & 1u) continue;
      f2 center = texel_center(x, y, res);
      if (!nearest_weights(uv, center, &w)) continue;
      float nx = (uv[0].x * w.x + uv[1].x * w.y) + uv[2].x * w.z;
      float ny = (uv[0].y * w.x + uv[1].y * w.y) + uv[2].y * w.z;
      float dx = nx - center.x, dy = ny - center.y;
      if (maxf_(fabsf(dx), fabsf(dy)) > 0.5f / (float)res) continue;
    }
    if (!fill) {
      atomicAdd(&counts[t], 1ull);
      continue;
    }
    uint64_t slot = tex_start[t] + atomicAdd(&counts[t], 1ull);
    f3 p = lerp3w(v, w);
    f3 nn;
    if (!try_normalize(lerp3w(n, w), &nn)) nn = face;
    out.xy[slot] = x | (y << 16);
    out.pos[slot] = make_float4(p.x, p.y, p.z, 0.0f);
    out.nrm[slot] = make_float4(nn.x, nn.y, nn.z, 0.0f);
    out.tri[slot] = t;
  }
}

// ---- burial: mesh_texels.rs:57-87, is_buried mesh_texels.rs:111-135 ----------------------------
// is_buried casts up to eight closest-hit rays from a texel (offset along the face normal) over
// the hemisphere of the face: ray k at sin^2(theta) = ((5k mod 8) + 0.5) / 8 and azimuth
// (k + turn) * tau / 8, turn drawn once from the texel's burial stream.  A hit on a face that
// faces along the ray is a back face; five back faces bury the texel, four front faces (or
// misses) keep it, whichever comes first.
// Here the rays are streamed: k_burial_rays writes rays k0 .. k0+nk-1 as closest-hit rays, the
// closest-hit ray stream traces them, and k_burial_decide tallies them in ray order.  Ray q of
// texel j of a pass of n texels sits at slot q * n + j, so that the stream's waves take the same
// ray of neighbouring texels together (near-parallel rays from nearby origins).  The first
// pass casts rays 0-3 of every texel, which decide every texel whose four rays all reach front
// faces (the common case); the second casts rays 4-7 of the rest.  The verdict is the one the
// early exits give: rays after the deciding one are cast but not counted.
// Direction of burial ray k of a texel.
__device__ __forceinline__ f3 burial_dir(const TrigTables& trig, f3 normal, f3 tangent, f3 bitangent, float turn, uint32_t k) {
  const float ring = (float)((k * 5u) % 8u) + 0.5f;
  const float sin_theta = sqrtf(ring / 8.0f);
  const float cos_theta = sqrtf(1.0f - sin_theta * sin_theta);
  const float phi = (((float)k + turn) * kTau) / 8.0f;
  float sp, cp;
  lsincosf(trig, phi, &sp, &cp);
  return add(mul(add(mul(tangent, cp), mul(bitangent, sp)), sin_theta), mul(normal, cos_theta));
}

// The burial test with one lane per texel casting its rays in order (the early exits cut the
// rays cast; lanes of a wave end at different rays).
__global__ void k_burial_lanes(DevScene sc, DevTris tr, DevMeshes ms, DevTexels tx, uint64_t ntex, uint64_t seed,
                               TrigTables trig, uint8_t* buried, unsigned long long* kept) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntex; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t = tx.tri[i];
    const f3 normal = ld3(tr.face + 3 * (size_t)t);
    const float4 p4 = tx.pos[i];
    const uint32_t xy = tx.xy[i];
    const f3 origin = add(mk3(p4.x, p4.y, p4.z), mul(normal, kRayOffset));
    Pcg rng = stream6(seed, ms.id[tr.mesh[t]], kBurialStream, tr.local[t], xy & 0xffffu, xy >> 16);
    const float turn = pcg_f32(rng);
    f3 tangent, bitangent;
    orthonormal_pair(normal, &tangent, &bitangent);
    uint32_t back = 0, front = 0;
    bool is_buried = false;
    for (uint32_t k = 0; k < 8; ++k) {
      const f3 dir = burial_dir(trig, normal, tangent, bitangent, turn, k);
      Hit h;
      const bool bf = traverse<false>(sc, stack, blockDim.x, origin, dir, INFINITY, &h, tl) && dot(tri_face_normal(sc, h.pos), dir) > 0.0f;
      if (bf) back++;
      else front++;
      if (back >= 5) {
        is_buried = true;
        break;
      }
      if (front > 3) break;
    }
    buried[i] = is_buried ? 1 : 0;
    if (!is_buried) atomicAdd(&kept[t], 1ull);
  }
  if (sc.counters) {
    wave_add(&sc.counters[0], tl.nodes);
    wave_add(&sc.counters[1], tl.tris);
  }
}

// Texel j of a pass is tx index first + j, or list[j] with a list (whose length is *count).
__global__ void k_burial_rays(DevTris tr, DevMeshes ms, DevTexels tx, uint64_t first, const uint32_t* __restrict__ list,
                              const unsigned long long* __restrict__ count, uint64_t n, uint32_t k0, uint32_t nk,
                              uint64_t seed, TrigTables trig, float4* rays, float4* hits, uint64_t* total) {
  if (count) n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = n * nk;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = list ? (uint64_t)list[j] : first + j;
    const uint32_t t = tx.tri[i];
    const f3 normal = ld3(tr.face + 3 * (size_t)t);
    const float4 p4 = tx.pos[i];
    const uint32_t xy = tx.xy[i];
    const f3 origin = add(mk3(p4.x, p4.y, p4.z), mul(normal, kRayOffset));
    Pcg rng = stream6(seed, ms.id[tr.mesh[t]], kBurialStream, tr.local[t], xy & 0xffffu, xy >> 16);
    const float turn = pcg_f32(rng);
    f3 tangent, bitangent;
    orthonormal_pair(normal, &tangent, &bitangent);
    for (uint32_t q = 0; q < nk; ++q) {
      const f3 dir = burial_dir(trig, normal, tangent, bitangent, turn, k0 + q);
      const uint64_t r = q * n + j;
      rays[2 * r] = make_float4(origin.x, origin.y, origin.z, INFINITY);
      rays[2 * r + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(2u));
      hits[r] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0xffffffffu));
    }
  }
}

// The rays of a burial pass that hit nothing (slots below *total), appended to `out` as 8 floats
// (origin, direction, +inf, 0) while fewer than `cap` are there; *count counts them all.
__global__ void k_burial_misses(const float4* __restrict__ rays, const float4* __restrict__ hits, const uint64_t* __restrict__ total,
                                float* out, uint64_t cap, unsigned long long* count) {
  const uint64_t n = *total;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    if (__float_as_uint(hits[r].w) != 0xffffffffu) continue;
    const unsigned long long k = atomicAdd(count, 1ull);
    if (k >= cap) continue;
    const float4 a = rays[2 * r], b = rays[2 * r + 1];
    float* o = out + 8 * k;
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
    o[3] = b.x;
    o[4] = b.y;
    o[5] = b.z;
    o[6] = INFINITY;
    o[7] = 0.0f;
  }
}

// *count += the nonzero flags among n.
__global__ void k_count_nonzero(const uint32_t* __restrict__ flags, uint64_t n, unsigned long long* count) {
  uint32_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) c += flags[i] != 0u;
  wave_add(count, c);
}

// Tallies rays k0 .. k0+nk-1 of each texel of a pass on the streamed hits: a hit whose triangle faces along the ray is a back face.  back[i] carries the back-face
// count of the earlier rays (their front faces are the rest).  Decided texels get buried[i] and,
// when kept, a count in kept[triangle]; undecided ones are appended to `next` (*next_n).
__global__ void k_burial_decide(DevScene sc, DevTexels tx, uint64_t first, const uint32_t* __restrict__ list,
                                const unsigned long long* __restrict__ count, uint64_t n, uint32_t k0, uint32_t nk,
                                const float4* __restrict__ rays, const float4* __restrict__ hits, uint8_t* back_count,
                                uint8_t* buried, unsigned long long* kept, uint32_t* next, unsigned long long* next_n) {
  if (count) n = *count;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = list ? (uint64_t)list[j] : first + j;
    uint32_t back = k0 ? back_count[i] : 0u, front = k0 - back;
    int verdict = -1;  // 1 buried, 0 kept
    for (uint32_t q = 0; q < nk && verdict < 0; ++q) {
      const uint64_t r = q * n + j;
      const uint32_t pos = __float_as_uint(hits[r].w);
      bool bf = false;
      if (pos != 0xffffffffu) {
        const float4 d4 = rays[2 * r + 1];
        bf = dot(tri_face_normal(sc, pos), mk3(d4.x, d4.y, d4.z)) > 0.0f;
      }
      if (bf) back++;
      else front++;
      if (back >= 5) verdict = 1;
      else if (front > 3) verdict = 0;
    }
    if (verdict < 0) {
      back_count[i] = (uint8_t)back;
      next[atomicAdd(next_n, 1ull)] = (uint32_t)i;
      continue;
    }
    buried[i] = (uint8_t)verdict;
    if (!verdict) atomicAdd(&kept[tx.tri[i]], 1ull);
  }
}

__global__ void k_compact(DevTexels in, uint64_t ntex, const uint8_t* __restrict__ buried,
                          const uint64_t* __restrict__ start, unsigned long long* cursor, DevTexels out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntex; i += (uint64_t)gridDim.x * blockDim.x) {
    if (buried[i]) continue;
    uint32_t t = in.tri[i];
    uint64_t slot = start[t] + atomicAdd(&cursor[t], 1ull);
    out.xy[slot] = in.xy[i];
    out.pos[slot] = in.pos[i];
    out.nrm[slot] = in.nrm[i];
    out.tri[slot] = t;
  }
}

// ---- per-texel shadows: texel_shadow.rs:20-90, area.rs:61-146, sun.rs:21-56, sky.rs:20-48 ------
__device__ __forceinline__ f2 concentric_disk(const TrigTables& trig, f2 sq) {
  float px = sq.x * 2.0f - 1.0f, py = sq.y * 2.0f - 1.0f;
  if (px == 0.0f && py == 0.0f) return mk2(0.0f, 0.0f);
  float radius, angle;
  if (fabsf(px) > fabsf(py)) {
    radius = px;
    angle = kFracPi4 * (py / px);
  } else {
    radius = py;
    angle = 2.0f * kFracPi4 - kFracPi4 * (px / py);
  }
  float sa, ca;
  lsincosf(trig, angle, &sa, &ca);
  return mk2(ca * radius, sa * radius);
}

__device__ __forceinline__ f3 cosine_dir(const TrigTables& trig, f3 normal, f2 sq) {
  f2 disk = concentric_disk(trig, sq);
  float up = sqrtf(maxf_(1.0f - (disk.x * disk.x + disk.y * disk.y), 0.0f));
  f3 side, across;
  orthonormal_pair(normal, &side, &across);
  return normalize(add(add(mul(side, disk.x), mul(across, disk.y)), mul(normal, up)));
}

struct Share {
  bool any_lit = false, any_blocked = false;
  float lit = 0.0f, total = 0.0f;
};
__device__ __forceinline__ bool share_result(const Share& s, float* fraction) {
  if (s.any_lit != s.any_blocked) {
    *fraction = s.any_lit ? 1.0f : 0.0f;
    return true;
  }
  if (s.total > 0.0f) {
    *fraction = s.lit / s.total;
    return true;
  }
  return false;
}

// area.rs:61-90 with adaptive_share area.rs:98-126 and toward area.rs:130-146
__device__ bool area_visibility(const DevScene& sc, uint32_t* stack, uint32_t stride, const DevLight& L, f3 origin, f3 pos,
                                f3 normal, Pcg& rng, uint32_t samples, uint32_t minimum, float* fraction, uint32_t* rays, Tally& tl) {
  f3 c = ld3(L.pos), ax = ld3(L.ax), ay = ld3(L.ay), fw = ld3(L.af);
  f3 local = sub(pos, c);
  float npx = clampf_(dot(local, ax), -L.hx, L.hx), npy = clampf_(dot(local, ay), -L.hy, L.hy);
  f3 nearest = add(add(c, mul(ax, npx)), mul(ay, npy));
  if (L.reach <= 0.0f || len_sq(sub(pos, nearest)) > L.reach * L.reach) return false;
  if (L.rect && dot(fw, sub(pos, c)) <= 0.0f) return false;
  uint32_t nr = 0;
  Share s;
  auto trace = [&](f2 uv, bool probe) {
    float ox = (uv.x * 2.0f - 1.0f) * L.hx, oy = (uv.y * 2.0f - 1.0f) * L.hy;
    f3 point = add(add(c, mul(ax, ox)), mul(ay, oy));
    f3 to_point = sub(point, pos);
    float dsq = len_sq(to_point);
    if (dsq < 1.0e-12f) return;
    float distance = sqrtf(dsq);
    f3 dir = divs(to_point, distance);
    float receiver = dot(normal, dir) - kNdotlTolerance;
    float emitter = L.rect ? -dot(fw, dir) : 1.0f;
    if (!(receiver > 0.0f && emitter > 0.0f)) return;
    float weight = (receiver * emitter) / dsq;
    nr++;
    Hit h;
    bool lit = !traverse<true>(sc, stack, stride, origin, dir, maxf_(distance - kEndGap, 0.0f), &h, tl, L.alpha);
    if (!probe) {
      s.total += weight;
      s.lit += lit ? weight : 0.0f;
    }
    s.any_lit |= lit;
    s.any_blocked |= !lit;
  };
  if (L.rect) {
    trace(mk2(0.0f, 0.0f), true);
    trace(mk2(1.0f, 0.0f), true);
    trace(mk2(1.0f, 1.0f), true);
    trace(mk2(0.0f, 1.0f), true);
    trace(mk2(0.5f, 0.5f), true);
  } else {
    trace(mk2(0.0f, 0.5f), true);
    trace(mk2(1.0f, 0.5f), true);
    trace(mk2(0.5f, 0.5f), true);
  }
  f2 offset;
  offset.x = pcg_f32(rng);
  offset.y = pcg_f32(rng);
  for (uint32_t taken = 0; taken < samples; ++taken) {
    if (taken == minimum && s.any_lit != s.any_blocked) break;
    trace(spread_pt(taken, offset, L.rect != 0), false);
  }
  *rays = nr;
  return share_result(s, fraction);
}

// sun.rs:21-56
__device__ bool sun_visibility(const DevScene& sc, uint32_t* stack, uint32_t stride, const TrigTables& trig,
                               const DevLight& L, f3 origin, f3 normal, Pcg& rng, uint32_t samples, uint32_t minimum,
                               float* fraction, uint32_t* rays, Tally& tl) {
  f3 toward = ld3(L.toward);
  Hit h;
  if (L.srad <= 0.0f) {
    if (dot(normal, toward) <= kNdotlTolerance) return false;
    *fraction = traverse<true>(sc, stack, stride, origin, toward, INFINITY, &h, tl, L.alpha) ? 0.0f : 1.0f;
    *rays = 1;
    return true;
  }
  f3 side, up;
  orthonormal_pair(toward, &side, &up);
  uint32_t nr = 0;
  Share s;
  auto trace = [&](f2 disk, bool probe) {
    f3 dir = normalize(add(toward, mul(add(mul(side, disk.x), mul(up, disk.y)), L.stan)));
    float receiver = dot(normal, dir) - kNdotlTolerance;
    if (receiver <= 0.0f) return;
    nr++;
    bool lit = !traverse<true>(sc, stack, stride, origin, dir, INFINITY, &h, tl, L.alpha);
    if (!probe) {
      s.total += receiver;
      s.lit += lit ? receiver : 0.0f;
    }
    s.any_lit |= lit;
    s.any_blocked |= !lit;
  };
  trace(mk2(0.0f, 0.0f), true);
  trace(mk2(1.0f, 0.0f), true);
  trace(mk2(0.0f, 1.0f), true);
  trace(mk2(-1.0f, 0.0f), true);
  trace(mk2(0.0f, -1.0f), true);
  f2 offset;
  offset.x = pcg_f32(rng);
  offset.y = pcg_f32(rng);
  for (uint32_t taken = 0; taken < samples; ++taken) {
    if (taken == minimum && s.any_lit != s.any_blocked) break;
    trace(concentric_disk(trig, spread_pt(taken, offset, true)), false);
  }
  *rays = nr;
  return share_result(s, fraction);
}

__device__ __forceinline__ void raise_level(uint32_t* pool, uint64_t off, uint32_t bits, uint64_t i, uint32_t level) {
  if (level == 0) return;
  if (bits == 1) {
    atomicOr(&pool[off + i / 32], 1u << (i % 32));
    return;
  }
  uint32_t* w = &pool[off + i / 4];
  uint32_t shift = 8u * (uint32_t)(i % 4);
  uint32_t cur = *w;
  for (;;) {
    uint32_t have = (cur >> shift) & 0xffu;
    if (have >= level) break;
    uint32_t want = (cur & ~(0xffu << shift)) | (level << shift);
    uint32_t prev = atomicCAS(w, cur, want);
    if (prev == cur) break;
    cur = prev;
  }
}

__device__ __forceinline__ uint32_t region_level(const DevRegions& rg, uint32_t r, int64_t x, int64_t y) {
  const uint32_t* rc = rg.rect + 4 * (size_t)r;
  if (x < (int64_t)rc[0] || y < (int64_t)rc[1] || x > (int64_t)rc[2] || y > (int64_t)rc[3]) return 0;
  uint64_t w = rc[2] - rc[0] + 1;
  uint64_t i = (uint64_t)(x - rc[0]) + (uint64_t)(y - rc[1]) * w;
  uint64_t off = rg.off[r];
  if (rg.bits[r] == 1) return (rg.pool[off + i / 32] >> (i % 32)) & 1u;
  return (rg.pool[off + i / 4] >> (8u * (uint32_t)(i % 4))) & 0xffu;
}

// The (triangle x light x texel) loop of mesh_bake.rs:203-267; one wave per (triangle, light,
// 64-texel chunk).
// Lights that cast several rays per texel: area lights, soft suns, skies.
__device__ __forceinline__ bool vis_multi(const DevLight& L) {
  return L.shape == 1 || L.shape == 2 || L.shape == 4 || (L.shape == 3 && L.srad > 0.0f);
}

// MULTI: the per-lane area, soft-sun and sky branches are compiled in; without them (when those
// lights are streamed) the kernel holds far fewer registers.
template <bool MULTI>
__global__ void k_trace(DevScene sc, DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx,
                        const uint64_t* __restrict__ tex_start, DevEntries ent, DevRegions rg,
                        const uint64_t* __restrict__ unit_start, uint64_t units, TraceParams p, TrigTables trig) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t waves = blockDim.x >> 6;
  for (uint64_t unit = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); unit < units; unit += (uint64_t)gridDim.x * waves) {
    uint32_t t = (uint32_t)owner_of(unit_start, tr.n, unit);
    uint64_t local = unit - unit_start[t];
    uint64_t t0 = tex_start[t], nt = tex_start[t + 1] - t0;
    uint64_t chunks = (nt + 63) / 64;
    uint64_t k = local / chunks, chunk = local % chunks;
    uint64_t e = ent.start[t] + k;
    uint32_t slot = ent.light[e], r = ent.region[e];
    const DevLight& L = lights[slot];
    uint64_t i = chunk * 64 + lane;
    uint32_t rays = 0, some_n = 0;
    if (i < nt) {
      uint64_t j = t0 + i;
      uint32_t xy = tx.xy[j];
      uint32_t x = xy & 0xffffu, y = xy >> 16;
      float4 p4 = tx.pos[j], n4 = tx.nrm[j];
      f3 pos = mk3(p4.x, p4.y, p4.z), nrm = mk3(n4.x, n4.y, n4.z);
      f3 face = ld3(tr.face + 3 * (size_t)t);
      f3 origin = add(pos, mul(face, kRayOffset));
      uint32_t m = tr.mesh[t];
      bool some = false;
      uint32_t level = 0;
      Hit h;
      if (L.shape == 0) {
        f3 to_light = sub(ld3(L.pos), pos);
        float dsq = len_sq(to_light);
        if (dsq <= L.radius * L.radius) {
          float distance = sqrtf(dsq);
          if (distance == 0.0f) {
            some = true;
            level = 1;
          } else {
            f3 dir = divs(to_light, distance);
            if (dot(nrm, dir) > kNdotlTolerance) {
              some = true;
              rays = 1;
              level = traverse<true>(sc, stack, blockDim.x, origin, dir, distance, &h, tl, L.alpha) ? 0u : 1u;
            }
          }
        }
      } else if (!MULTI) {
        if (L.shape == 3 && !(L.srad > 0.0f)) {  // a sun of zero angular size (sun.rs:21-30)
          const f3 toward = ld3(L.toward);
          if (!(dot(nrm, toward) <= kNdotlTolerance)) {
            const float fraction = traverse<true>(sc, stack, blockDim.x, origin, toward, INFINITY, &h, tl, L.alpha) ? 0.0f : 1.0f;
            some = true;
            rays = 1;
            level = level_of(fraction, L.bits);
          }
        }  // the other lights are traced by the k_vis_* kernels
      } else if (!(p.stream_multi && vis_multi(L))) {  // streamed lights are traced by the k_vis_* kernels
        Pcg rng = stream5(p.seed, ms.id[m], L.id, x, y);
        float fraction = 0.0f;
        if (L.shape == 1 || L.shape == 2) {
          some = area_visibility(sc, stack, blockDim.x, L, origin, pos, nrm, rng, p.area_samples, p.area_min, &fraction, &rays, tl);
        } else if (L.shape == 3) {
          some = sun_visibility(sc, stack, blockDim.x, trig, L, origin, nrm, rng, p.area_samples, p.area_min, &fraction, &rays, tl);
        } else {
          f2 offset;
          offset.x = pcg_f32(rng);
          offset.y = pcg_f32(rng);
          uint32_t open = 0;
          for (uint32_t taken = 0; taken < p.sky_samples; ++taken) {
            f3 dir = cosine_dir(trig, nrm, spread_pt(taken, offset, true));
            open += traverse<true>(sc, stack, blockDim.x, origin, dir, INFINITY, &h, tl, L.alpha) ? 0u : 1u;
          }
          fraction = (float)open / (float)(p.sky_samples > 1 ? p.sky_samples : 1u);
          rays = p.sky_samples;
          some = true;
        }
        if (some) level = level_of(fraction, L.bits);
        else rays = 0;  // rays of a texel the light does not reach are not counted (mesh_bake.rs:241-245)
      }
      if (some) {
        some_n = 1;
        uint32_t res = ms.res[m];
        uint64_t vi = (uint64_t)x + (uint64_t)y * res;
        atomicOr(&ms.visited[ms.visited_off[m] + vi / 32], 1u << (vi % 32));
        const uint32_t* rc = rg.rect + 4 * (size_t)r;
        uint64_t ri = (uint64_t)(x - rc[0]) + (uint64_t)(y - rc[1]) * (rc[2] - rc[0] + 1);
        raise_level(rg.pool, rg.off[r], rg.bits[r], ri, level);
        if (rg.vis) atomicOr(&rg.vis[rg.vis_off[r] + ri / 32], 1u << (ri % 32));
      }
    }
    uint32_t m = tr.mesh[t];
    wave_add(&ms.rays[m], rays);
    wave_add(&ms.samples[m], some_n);
    if (rg.rays) {  // one entry per wave: its region
      wave_add(&rg.rays[r], rays);
      wave_add(&rg.samples[r], some_n);
    }
    if (sc.counters) {
      wave_add(&sc.counters[0], tl.nodes);
      wave_add(&sc.counters[1], tl.tris);
      tl = Tally{};
    }
  }
}

// ---- streamed multi-ray visibility: area lights, soft suns, skies -----------------------------
// The same results as k_trace's area_visibility, sun_visibility and sky branches, reorganised so
// the traversal runs in the persistent any-hit kernel.  A job is one (texel, light entry) pair of
// a light that casts several rays per texel.  Phase A writes every job's probe rays and its first
// min(minimum, samples) sample rays (sample-major slots; a slot whose ray fails the receiver or
// emitter test holds no ray), k_any_stream clears the "lit" word of every occluded ray, and
// k_vis_mid folds them in order into the job's running share.  Jobs still undecided after
// `minimum` samples (area.rs:82-88, sun.rs:47-53) get phase B: the remaining samples, traced the
// same way; k_vis_finish folds those in and writes the level, the visited bit and the stats.
// Sums run over the samples in their order, so every f32 result equals the per-lane loop's.

struct VisJob {
  uint32_t e, t;
  uint64_t j;
};
// Job q: entry me[k] for the k with js[k] <= q < js[k + 1], texel q - js[k] of its triangle.
__device__ __forceinline__ VisJob vis_job(const uint64_t* __restrict__ js, const uint32_t* __restrict__ me, uint32_t nme,
                                          const DevEntries& ent, const uint64_t* __restrict__ tex_start, uint64_t q) {
  const uint32_t k = (uint32_t)owner_of_wave(js, nme, q);
  VisJob v;
  v.e = me[k];
  v.t = ent.tri[v.e];
  v.j = tex_start[v.t] + (q - js[k]);
  return v;
}

// Probe rays before the samples: 5 for rectangles and soft suns, 3 for tubes, none for skies.
__device__ __forceinline__ uint32_t vis_probes(const DevLight& L) {
  return L.shape == 4 ? 0u : (L.shape == 2 ? 3u : 5u);
}
// Phase A slots of a job: its probes and min(minimum, samples) samples; a sky's samples.
__device__ __forceinline__ uint32_t vis_slots_a(const DevLight& L, const TraceParams& p) {
  if (L.shape == 4) return p.sky_samples;
  return vis_probes(L) + (p.area_min < p.area_samples ? p.area_min : p.area_samples);
}

struct VisRay {
  bool ray;
  f3 dir;
  float tmax, weight;
};
// Ray number k of a job (probes first, then samples k - probes), as area_visibility, sun_visibility
// and the sky loop build it.  `offset` is the job's spread offset.
__device__ __forceinline__ VisRay vis_ray(const DevLight& L, const TrigTables& trig, f3 pos, f3 nrm, f2 offset, uint32_t k) {
  VisRay r{false, mk3(0.0f, 0.0f, 0.0f), 0.0f, 0.0f};
  if (L.shape == 4) {
    r.dir = cosine_dir(trig, nrm, spread_pt(k, offset, true));
    r.ray = true;
    r.tmax = INFINITY;
    r.weight = 1.0f;
    return r;
  }
  const uint32_t probes = vis_probes(L);
  if (L.shape == 3) {
    f2 disk;
    if (k < probes) {
      const float px[5] = {0.0f, 1.0f, 0.0f, -1.0f, 0.0f}, py[5] = {0.0f, 0.0f, 1.0f, 0.0f, -1.0f};
      disk = mk2(px[k], py[k]);
    } else {
      disk = concentric_disk(trig, spread_pt(k - probes, offset, true));
    }
    const f3 toward = ld3(L.toward);
    f3 side, up;
    orthonormal_pair(toward, &side, &up);
    const f3 dir = normalize(add(toward, mul(add(mul(side, disk.x), mul(up, disk.y)), L.stan)));
    const float receiver = dot(nrm, dir) - kNdotlTolerance;
    if (receiver <= 0.0f) return r;
    r.ray = true;
    r.dir = dir;
    r.tmax = INFINITY;
    r.weight = receiver;
    return r;
  }
  f2 uv;
  if (k < probes) {
    const float rx[5] = {0.0f, 1.0f, 1.0f, 0.0f, 0.5f}, ry[5] = {0.0f, 0.0f, 1.0f, 1.0f, 0.5f};
    const float tx_[3] = {0.0f, 1.0f, 0.5f};
    uv = L.rect ? mk2(rx[k], ry[k]) : mk2(tx_[k], 0.5f);
  } else {
    uv = spread_pt(k - probes, offset, L.rect != 0);
  }
  const f3 c = ld3(L.pos), ax = ld3(L.ax), ay = ld3(L.ay), fw = ld3(L.af);
  const float ox = (uv.x * 2.0f - 1.0f) * L.hx, oy = (uv.y * 2.0f - 1.0f) * L.hy;
  const f3 point = add(add(c, mul(ax, ox)), mul(ay, oy));
  const f3 to_point = sub(point, pos);
  const float dsq = len_sq(to_point);
  if (dsq < 1.0e-12f) return r;
  const float distance = sqrtf(dsq);
  const f3 dir = divs(to_point, distance);
  const float receiver = dot(nrm, dir) - kNdotlTolerance;
  const float emitter = L.rect ? -dot(fw, dir) : 1.0f;
  if (!(receiver > 0.0f && emitter > 0.0f)) return r;
  r.ray = true;
  r.dir = dir;
  r.tmax = maxf_(distance - kEndGap, 0.0f);
  r.weight = (receiver * emitter) / dsq;
  return r;
}

// area_visibility's reach test (area.rs:61-70): false when the texel is out of the light's reach.
__device__ __forceinline__ bool vis_reach(const DevLight& L, f3 pos) {
  if (L.shape != 1 && L.shape != 2) return true;
  const f3 c = ld3(L.pos), ax = ld3(L.ax), ay = ld3(L.ay), fw = ld3(L.af);
  const f3 local = sub(pos, c);
  const float npx = clampf_(dot(local, ax), -L.hx, L.hx), npy = clampf_(dot(local, ay), -L.hy, L.hy);
  const f3 nearest = add(add(c, mul(ax, npx)), mul(ay, npy));
  if (L.reach <= 0.0f || len_sq(sub(pos, nearest)) > L.reach * L.reach) return false;
  if (L.rect && dot(fw, sub(pos, c)) <= 0.0f) return false;
  return true;
}

struct VisTexel {
  f3 pos, nrm, origin;
  uint32_t x, y, m;
  f2 offset;
};
__device__ __forceinline__ VisTexel vis_texel(const DevTris& tr, const DevMeshes& ms, const DevTexels& tx, const DevLight& L,
                                              uint64_t seed, uint32_t t, uint64_t j) {
  VisTexel v;
  const float4 p4 = tx.pos[j], n4 = tx.nrm[j];
  v.pos = mk3(p4.x, p4.y, p4.z);
  v.nrm = mk3(n4.x, n4.y, n4.z);
  v.origin = add(v.pos, mul(ld3(tr.face + 3 * (size_t)t), kRayOffset));
  const uint32_t xy = tx.xy[j];
  v.x = xy & 0xffffu;
  v.y = xy >> 16;
  v.m = tr.mesh[t];
  Pcg rng = stream5(seed, ms.id[v.m], L.id, v.x, v.y);
  v.offset.x = pcg_f32(rng);  // drawn after the probes, which take no random numbers
  v.offset.y = pcg_f32(rng);
  return v;
}


__global__ void k_vis_count(DevEntries ent, const DevLight* __restrict__ lights, const uint64_t* __restrict__ js,
                            const uint32_t* __restrict__ me, uint32_t nme, const uint64_t* __restrict__ tex_start,
                            uint64_t q0, uint64_t n, TraceParams p, uint64_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i == n) {
      out[i] = 0;
      continue;
    }
    const VisJob v = vis_job(js, me, nme, ent, tex_start, q0 + i);
    out[i] = vis_slots_a(lights[ent.light[v.e]], p);
  }
}

// Phase A rays.  Slot of ray k of job i: k * n + i.  contrib[3 * slot] is 1 for a traced ray (the
// any-hit kernel clears it when the ray is blocked); wts[slot] is its sample weight.
__global__ void k_vis_rays_a(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                             const uint64_t* __restrict__ js, const uint32_t* __restrict__ me, uint32_t nme,
                             const uint64_t* __restrict__ tex_start, uint64_t q0, uint64_t n, TraceParams p, TrigTables trig,
                             float4* rays, float* contrib, float* wts, VisState* state) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const VisJob v = vis_job(js, me, nme, ent, tex_start, q0 + i);
    const DevLight& L = lights[ent.light[v.e]];
    const VisTexel vt = vis_texel(tr, ms, tx, L, p.seed, v.t, v.j);
    const bool reach = vis_reach(L, vt.pos);
    const uint32_t slots = vis_slots_a(L, p);
    const uint32_t alpha = L.alpha ? 1u : 0u;
    for (uint32_t k = 0; k < slots; ++k) {
      const uint64_t slot = (uint64_t)k * n + i;
      VisRay r{false, mk3(0.0f, 0.0f, 0.0f), 0.0f, 0.0f};
      if (reach) r = vis_ray(L, trig, vt.pos, vt.nrm, vt.offset, k);
      rays[2 * slot] = make_float4(vt.origin.x, vt.origin.y, vt.origin.z, r.tmax);
      rays[2 * slot + 1] = make_float4(r.dir.x, r.dir.y, r.dir.z, __uint_as_float(alpha | (r.ray ? 2u : 0u)));
      contrib[3 * slot] = r.ray ? 1.0f : 0.0f;
      contrib[3 * slot + 1] = 0.0f;
      contrib[3 * slot + 2] = 0.0f;
      wts[slot] = r.weight;
    }
    state[i] = VisState{0.0f, 0.0f, 0u, reach ? 4u : 0u};
  }
}

// Folds ray k of a job into its share (the `trace` closure of area_visibility / sun_visibility).
__device__ __forceinline__ void vis_fold(VisState& s, bool ray, bool lit, bool probe, float w) {
  if (!ray) return;
  s.nr++;
  if (!probe) {
    s.total += w;
    s.lit += lit ? w : 0.0f;
  }
  s.flags |= lit ? 1u : 2u;
}

__global__ void k_vis_mid(DevEntries ent, const DevLight* __restrict__ lights, const uint64_t* __restrict__ js,
                          const uint32_t* __restrict__ me, uint32_t nme, const uint64_t* __restrict__ tex_start, uint64_t q0,
                          uint64_t n, TraceParams p, const float4* __restrict__ rays, const float* __restrict__ contrib,
                          const float* __restrict__ wts, VisState* state, uint64_t* need) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i == n) {
      need[i] = 0;
      continue;
    }
    const VisJob v = vis_job(js, me, nme, ent, tex_start, q0 + i);
    const DevLight& L = lights[ent.light[v.e]];
    VisState s = state[i];
    const uint32_t slots = vis_slots_a(L, p), probes = vis_probes(L);
    for (uint32_t k = 0; k < slots; ++k) {
      const uint64_t slot = (uint64_t)k * n + i;
      const bool ray = (__float_as_uint(rays[2 * slot + 1].w) & 2u) != 0u;
      vis_fold(s, ray, contrib[3 * slot] > 0.0f, k < probes, wts[slot]);
    }
    // area.rs:82-88: after `minimum` samples, stop once the rays agree
    bool more = false;
    if (L.shape != 4 && (s.flags & 4u) && p.area_min < p.area_samples) {
      const bool decided = ((s.flags & 1u) != 0u) != ((s.flags & 2u) != 0u);
      more = !decided;
    }
    if (more) s.flags |= 8u;
    state[i] = s;
    need[i] = more ? 1u : 0u;
  }
}

// Phase B rays of the jobs listed in `blist` (nb of them): samples minimum .. samples - 1, sample
// k in slot (k - minimum) * nb + b.
__global__ void k_vis_rays_b(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                             const uint64_t* __restrict__ js, const uint32_t* __restrict__ me, uint32_t nme,
                             const uint64_t* __restrict__ tex_start, uint64_t q0, const uint32_t* __restrict__ blist, uint64_t nb,
                             TraceParams p, TrigTables trig, float4* rays, float* contrib, float* wts) {
  for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x) {
    const VisJob v = vis_job(js, me, nme, ent, tex_start, q0 + blist[b]);
    const DevLight& L = lights[ent.light[v.e]];
    const VisTexel vt = vis_texel(tr, ms, tx, L, p.seed, v.t, v.j);
    const uint32_t probes = vis_probes(L), alpha = L.alpha ? 1u : 0u;
    for (uint32_t k = p.area_min; k < p.area_samples; ++k) {
      const uint64_t slot = (uint64_t)(k - p.area_min) * nb + b;
      const VisRay r = vis_ray(L, trig, vt.pos, vt.nrm, vt.offset, probes + k);
      rays[2 * slot] = make_float4(vt.origin.x, vt.origin.y, vt.origin.z, r.tmax);
      rays[2 * slot + 1] = make_float4(r.dir.x, r.dir.y, r.dir.z, __uint_as_float(alpha | (r.ray ? 2u : 0u)));
      contrib[3 * slot] = r.ray ? 1.0f : 0.0f;
      contrib[3 * slot + 1] = 0.0f;
      contrib[3 * slot + 2] = 0.0f;
      wts[slot] = r.weight;
    }
  }
}

__global__ void k_vis_list(const uint64_t* __restrict__ need, const uint64_t* __restrict__ scan, uint64_t n, uint32_t* blist) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (need[i]) blist[scan[i]] = (uint32_t)i;
}

// Phase B folded in, then the texel's level, visited bit and stats, as k_trace writes them.
__global__ void k_vis_finish(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                             DevRegions rg, const uint64_t* __restrict__ js, const uint32_t* __restrict__ me, uint32_t nme,
                             const uint64_t* __restrict__ tex_start, uint64_t q0, uint64_t n, TraceParams p,
                             const VisState* __restrict__ state, const uint64_t* __restrict__ bscan, uint64_t nb,
                             const float4* __restrict__ rays, const float* __restrict__ contrib, const float* __restrict__ wts) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
    const uint64_t i = i0 + lane;
    uint32_t m = 0xffffffffu, rays_n = 0, some_n = 0, rr = 0;
    if (i < n) {
      const VisJob v = vis_job(js, me, nme, ent, tex_start, q0 + i);
      const DevLight& L = lights[ent.light[v.e]];
      VisState s = state[i];
      if (s.flags & 8u) {
        const uint64_t b = bscan[i];
        for (uint32_t k = p.area_min; k < p.area_samples; ++k) {
          const uint64_t slot = (uint64_t)(k - p.area_min) * nb + b;
          const bool ray = (__float_as_uint(rays[2 * slot + 1].w) & 2u) != 0u;
          vis_fold(s, ray, contrib[3 * slot] > 0.0f, false, wts[slot]);
        }
      }
      bool some = false;
      float fraction = 0.0f;
      if (L.shape == 4) {
        // sky.rs:30-48: every sample traced; the fraction is the open share
        some = true;
        fraction = s.lit / (float)(p.sky_samples > 1 ? p.sky_samples : 1u);
        rays_n = p.sky_samples;
      } else if (s.flags & 4u) {
        Share sh;
        sh.any_lit = (s.flags & 1u) != 0u;
        sh.any_blocked = (s.flags & 2u) != 0u;
        sh.lit = s.lit;
        sh.total = s.total;
        some = share_result(sh, &fraction);
        rays_n = some ? s.nr : 0u;
      }
      const uint32_t t = v.t;
      m = tr.mesh[t];
      if (some) {
        some_n = 1;
        const uint32_t xy = tx.xy[v.j];
        const uint32_t x = xy & 0xffffu, y = xy >> 16;
        const uint32_t res = ms.res[m];
        const uint64_t vi = (uint64_t)x + (uint64_t)y * res;
        atomicOr(&ms.visited[ms.visited_off[m] + vi / 32], 1u << (vi % 32));
        const uint32_t r = ent.region[v.e];
        const uint32_t* rc = rg.rect + 4 * (size_t)r;
        const uint64_t ri = (uint64_t)(x - rc[0]) + (uint64_t)(y - rc[1]) * (rc[2] - rc[0] + 1);
        raise_level(rg.pool, rg.off[r], rg.bits[r], ri, level_of(fraction, L.bits));
        if (rg.vis) atomicOr(&rg.vis[rg.vis_off[r] + ri / 32], 1u << (ri % 32));
      }
      if (rg.rays) rr = ent.region[v.e];
    }
    if (rg.rays) {
      wave_add_at(rg.rays, rr, rays_n, i < n);
      wave_add_at(rg.samples, rr, some_n, i < n);
    }
    const uint32_t m0 = __shfl(m, 0, 64);
    if (__ballot(i < n && m != m0) == 0ull) {
      if (m0 != 0xffffffffu) {
        wave_add(&ms.rays[m0], rays_n);
        wave_add(&ms.samples[m0], some_n);
      }
    } else if (i < n) {
      if (rays_n) atomicAdd(&ms.rays[m], (unsigned long long)rays_n);
      if (some_n) atomicAdd(&ms.samples[m], 1ull);
    }
  }
}

// ---- optimize: shadows.rs:54-74 with any_light shadows.rs:136-146 -------------------------------
// One wave per entry: whether any texel of the triangle's centred rect (within its region) holds a
// non-zero level, 64 texels per step, stopping at the first step that finds one.
__global__ void k_optimize(DevTris tr, DevMeshes ms, DevEntries ent, DevRegions rg, const uint8_t* __restrict__ has_bounce,
                           uint8_t* keep) {
  const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
  for (uint64_t e = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); e < ent.n; e += (uint64_t)gridDim.x * waves) {
    if (has_bounce && has_bounce[e]) {
      if (lane == 0) keep[e] = 1;
      continue;
    }
    uint32_t t = ent.tri[e], r = ent.region[e];
    f2 uv[3];
    for (int k = 0; k < 3; ++k) uv[k] = mk2(tr.uv[(size_t)t * 6 + 2 * k], tr.uv[(size_t)t * 6 + 2 * k + 1]);
    Rect o = center_rect(uv, ms.res[tr.mesh[t]], 1.0f);
    const uint32_t* rc = rg.rect + 4 * (size_t)r;
    uint32_t x0 = max(o.x0, rc[0]), x1 = min(o.x1, rc[2]), y0 = max(o.y0, rc[1]), y1 = min(o.y1, rc[3]);
    bool any = false;
    if (x0 <= x1 && y0 <= y1) {
      const uint32_t w = x1 - x0 + 1;
      const uint64_t cells = (uint64_t)w * (y1 - y0 + 1);
      for (uint64_t c0 = 0; c0 < cells && !any; c0 += 64) {
        const uint64_t c = c0 + lane;
        bool lit = false;
        if (c < cells) lit = region_level(rg, r, x0 + (uint32_t)(c % w), y0 + (uint32_t)(c / w)) != 0;
        any = __ballot(lit) != 0;
      }
    }
    if (lane == 0) keep[e] = any ? 1 : 0;
  }
}

// ---- pad_seams: seams.rs:15-39 over every used region ---------------------------------------
__device__ __forceinline__ bool visited_at(const DevMeshes& ms, uint32_t m, uint32_t res, int64_t x, int64_t y) {
  if (x < 0 || y < 0 || x >= (int64_t)res || y >= (int64_t)res) return false;
  uint64_t i = (uint64_t)x + (uint64_t)y * res;
  return (ms.visited[ms.visited_off[m] + i / 32] >> (i % 32)) & 1u;
}
__global__ void k_pad_seams(DevMeshes ms, DevRegions rg, const uint32_t* __restrict__ used, const uint64_t* __restrict__ start,
                            uint32_t nused, uint64_t total) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = (uint32_t)owner_of_wave(start, nused, c);
    uint32_t r = used[u];
    uint64_t local = c - start[u];
    const uint32_t* rc = rg.rect + 4 * (size_t)r;
    uint32_t w = rc[2] - rc[0] + 1;
    int64_t x = rc[0] + (int64_t)(local % w), y = rc[1] + (int64_t)(local / w);
    uint32_t m = rg.mesh[r], res = ms.res[m];
    if (visited_at(ms, m, res, x, y)) continue;
    uint32_t level = 0;
    const int dx[4] = {-1, 1, 0, 0}, dy[4] = {0, 0, -1, 1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int64_t nx = x + dx[k], ny = y + dy[k];
      if (visited_at(ms, m, res, nx, ny)) level = max(level, region_level(rg, r, nx, ny));
      else if (visited_at(ms, m, res, x + 2 * dx[k], y + 2 * dy[k]))
        level = max(level, region_level(rg, r, x + 2 * dx[k], y + 2 * dy[k]));
    }
    raise_level(rg.pool, rg.off[r], rg.bits[r], (uint64_t)(x - rc[0]) + (uint64_t)(y - rc[1]) * w, level);
  }
}

// ---- crop_shadows + dilate_shadows + into_words: shadows.rs:148-160, seams.rs:43-95 ------------
struct CropView {
  const DevRegions* rg;
  uint32_t r, cx0, cy0, W, H;
  __device__ uint32_t orig(int32_t px, int32_t py) const {
    if (px < 2 || py < 2 || px >= (int32_t)W - 2 || py >= (int32_t)H - 2) return 0;
    return region_level(*rg, r, (int64_t)cx0 + px - 2, (int64_t)cy0 + py - 2);
  }
  __device__ uint32_t m1(int32_t px, int32_t py) const {
    uint32_t v = 0;
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) v = max(v, orig(px + dx, py + dy));
    return v;
  }
  __device__ bool ring0(int32_t px, int32_t py) const { return px == 0 || py == 0 || px == (int32_t)W - 1 || py == (int32_t)H - 1; }
  __device__ bool ring1(int32_t px, int32_t py) const { return px == 1 || py == 1 || px == (int32_t)W - 2 || py == (int32_t)H - 2; }
  // the grid after the first dilation pass
  __device__ uint32_t p1(int32_t px, int32_t py) const {
    if (px < 0 || py < 0 || px >= (int32_t)W || py >= (int32_t)H) return 0;
    return ring1(px, py) ? m1(px, py) : orig(px, py);
  }
  __device__ uint32_t final_(int32_t px, int32_t py) const {
    if (ring0(px, py)) {
      uint32_t v = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) v = max(v, p1(px + dx, py + dy));
      return v;
    }
    return p1(px, py);
  }
};

__global__ void k_crop(DevTris tr, DevEntries ent, DevRegions rg, const uint64_t* __restrict__ kept_entry,
                       const uint64_t* __restrict__ word_start, uint64_t nkept, uint64_t total,
                       const uint64_t* __restrict__ kept_out, uint32_t* out) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t j = owner_of_wave(word_start, nkept, c);
    uint64_t wi = c - word_start[j];
    uint64_t e = kept_entry[j];
    uint32_t t = ent.tri[e], r = ent.region[e];
    CropView cv;
    cv.rg = &rg;
    cv.r = r;
    cv.cx0 = tr.crop[4 * (size_t)t];
    cv.cy0 = tr.crop[4 * (size_t)t + 1];
    cv.W = tr.crop[4 * (size_t)t + 2] - cv.cx0 + 1 + 4;
    cv.H = tr.crop[4 * (size_t)t + 3] - cv.cy0 + 1 + 4;
    uint32_t bits = rg.bits[r];
    uint32_t per = 32 / bits;
    uint64_t n = (uint64_t)cv.W * cv.H;
    uint32_t word = 0;
    for (uint32_t k = 0; k < per; ++k) {
      uint64_t i = wi * per + k;
      if (i >= n) break;
      int32_t px = (int32_t)(i % cv.W), py = (int32_t)(i / cv.W);
      uint32_t v = cv.final_(px, py);
      word |= v << (k * bits);
    }
    out[kept_out[j] + wi] = word;
  }
}

// ---- serialize_triangles: triangle_words.rs:47-89 (headers and light data) ----------------------
__global__ void k_serialize(DevTris tr, const uint64_t* __restrict__ mesh_base, const uint32_t* __restrict__ ld_off,
                            const uint64_t* __restrict__ tri_kept_start, const uint64_t* __restrict__ kept_entry,
                            const uint32_t* __restrict__ shadow_off, const uint32_t* __restrict__ bounce_off,
                            const uint32_t* __restrict__ ent_light, uint32_t stride, const uint32_t* __restrict__ first_in_mesh,
                            uint32_t* out) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tr.n; t += gridDim.x * blockDim.x) {
    uint32_t m = tr.mesh[t];
    uint32_t lt = t - first_in_mesh[m];
    uint32_t* w = out + mesh_base[m];
    uint32_t off = ld_off[t];
    uint32_t cx0 = tr.crop[4 * (size_t)t], cy0 = tr.crop[4 * (size_t)t + 1], cx1 = tr.crop[4 * (size_t)t + 2];
    w[4 * lt] = off;
    w[4 * lt + 1] = cx0 - 2u;
    w[4 * lt + 2] = cy0 - 2u;
    w[4 * lt + 3] = cx1 - cx0 + 1 + 4;
    uint64_t k0 = tri_kept_start[t], k1 = tri_kept_start[t + 1];
    w[off] = (uint32_t)(k1 - k0);
    for (uint64_t j = k0; j < k1; ++j) {
      uint32_t* le = w + off + 1 + (j - k0) * stride;
      le[0] = ent_light[kept_entry[j]];
      le[1] = shadow_off[j];
      if (stride == 3) le[2] = bounce_off ? bounce_off[j] : 0u;
    }
  }
}

// ---- distance cubes: distance_cubes.rs:29-72 ---------------------------------------------------
__global__ void k_distance_cubes(DevScene sc, const DevLight* __restrict__ lights, uint32_t nlocal, uint32_t* out) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint32_t res = 64, per_face = res * res / 2;
  uint64_t total = (uint64_t)nlocal * 6 * per_face;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t slot = (uint32_t)(c / (6 * per_face));
    uint32_t rem = (uint32_t)(c % (6 * per_face));
    uint32_t face = rem / per_face, pair = rem % per_face;
    const DevLight& L = lights[slot];
    f3 o = ld3(L.pos);
    float d[2];
    for (int k = 0; k < 2; ++k) {
      uint32_t texel = pair * 2 + k;
      f3 dir = cube_texel_dir(face, texel % res, texel / res, res);
      Hit h;
      d[k] = traverse<false>(sc, stack, blockDim.x, o, dir, L.radius, &h, tl, L.alpha) ? h.t : L.radius;
    }
    out[c] = (f32_to_f16_bits(d[1]) << 16) | f32_to_f16_bits(d[0]);
  }
}

// ---- sun maps: sun_map.rs:25-75; one thread per word (two texels, row-major) ------------------
__device__ __forceinline__ uint32_t unorm16(float d) {
  float v = roundf(clampf_(d, 0.0f, 1.0f) * 65535.0f);
  return v > 0.0f ? (uint32_t)v : 0u;  // NaN -> 0
}
__global__ void k_sun_map(DevScene sc, SunMapParams p, uint32_t* out, uint64_t words) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const f3 side = mk3(p.side[0], p.side[1], p.side[2]), up = mk3(p.up[0], p.up[1], p.up[2]);
  const f3 along = mk3(p.along[0], p.along[1], p.along[2]);
  const uint64_t texels = (uint64_t)p.res * p.res;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x) {
    float d[2];
    for (int k = 0; k < 2; ++k) {
      const uint64_t idx = 2 * w + k;
      d[k] = 1.0f;
      if (idx >= texels) continue;
      const uint32_t x = (uint32_t)(idx % p.res), y = (uint32_t)(idx / p.res);
      float across = p.min[0] + ((float)x + 0.5f) * p.texel;
      float down = p.min[1] + ((float)y + 0.5f) * p.texel;
      f3 origin = add(add(mul(side, across), mul(up, down)), mul(along, p.min[2]));
      Hit h;
      if (traverse<false>(sc, stack, blockDim.x, origin, along, p.depth, &h, tl, p.alpha)) d[k] = h.t / p.depth;
    }
    out[w] = unorm16(d[1]) << 16 | unorm16(d[0]);
  }
}

// ---- sky probes: trace_probe (sky_probes.rs:80-106); one thread per (probe, axis) -------------
__global__ void k_sky_probes(DevScene sc, TrigTables trig, SkyProbeParams p, float* open, uint32_t* back) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint64_t n = (uint64_t)p.counts[0] * p.counts[1] * p.counts[2] * 6;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t index = i / 6;
    const uint32_t axis = (uint32_t)(i % 6);
    const uint32_t px = (uint32_t)(index % p.counts[0]);
    const uint32_t py = (uint32_t)(index % ((uint64_t)p.counts[0] * p.counts[1]) / p.counts[0]);
    const uint32_t pz = (uint32_t)(index / ((uint64_t)p.counts[0] * p.counts[1]));
    f3 position = add(mk3(p.origin[0], p.origin[1], p.origin[2]), mul(mk3((float)px, (float)py, (float)pz), p.spacing));
    f3 normal = mk3(0.0f, 0.0f, 0.0f);
    if (axis == 0) normal.x = 1.0f;
    else if (axis == 1) normal.x = -1.0f;
    else if (axis == 2) normal.y = 1.0f;
    else if (axis == 3) normal.y = -1.0f;
    else if (axis == 4) normal.z = 1.0f;
    else normal.z = -1.0f;
    Pcg rng = stream6(p.seed, kProbeStream, px, py, pz, axis);
    f2 offset;
    offset.x = pcg_f32(rng);
    offset.y = pcg_f32(rng);
    uint32_t nopen = 0, nback = 0;
    for (uint32_t taken = 0; taken < p.samples; ++taken) {
      f3 dir = cosine_dir(trig, normal, spread_pt(taken, offset, true));
      Hit h;
      bool occluded = traverse<false>(sc, stack, blockDim.x, position, dir, INFINITY, &h, tl, p.alpha) && h.t > 0.0f;
      if (occluded && dot(tri_face_normal(sc, h.pos), dir) > 0.0f) nback++;
      if (!occluded) nopen++;
    }
    open[i] = (float)nopen / (float)(p.samples > 1 ? p.samples : 1u);
    back[i] = nback;
  }
}

__global__ void k_ray_queries(DevScene sc, const float* __restrict__ rays, uint64_t n, int closest, float* t_out, uint32_t* idx_out) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const float* r = rays + 8 * i;
    Hit h;
    if (closest) {
      bool hit = traverse<false>(sc, stack, blockDim.x, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], &h, tl);
      t_out[i] = hit ? h.t : -1.0f;
      idx_out[i] = hit ? h.occ : 0xffffffffu;
    } else {
      idx_out[i] = traverse<true>(sc, stack, blockDim.x, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], &h, tl) ? 1u : 0u;
    }
  }
}

// ---- ray queries without device allocation ------------------------------------------------
// One fixed scratch per GPU for the process lifetime: the persistent query kernel's queue head
// (u64, words 0-1), its hit count (u64, words 2-3) and the traversal's stack-overflow flag (word
// 4).  Every call on a GPU shares it, so calls that overlap on one GPU race.
constexpr uint32_t kNoAllocScratchWords = 64;  // 256 bytes
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

__global__ void k_scratch_clear(uint32_t* s, uint32_t words) {
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) s[i] = 0u;
}

// Persistent any-hit / closest-hit queries: each wave takes the next 64 rays from the queue head
// until the head passes n (every wave reaches that exit, the head only grows).
__global__ void k_ray_queries_queue(DevScene sc, const float* __restrict__ rays, uint64_t n, int closest, float* t_out,
                                    uint32_t* idx_out, unsigned long long* head, unsigned long long* hits) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(head, 64ull);
    base = __shfl(base, 0);
    if (base >= n) break;
    const uint64_t i = base + lane;
    bool hit = false;
    if (i < n) {
      const float* r = rays + 8 * i;
      Hit h;
      if (closest) {
        hit = traverse<false>(sc, stack, blockDim.x, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], &h, tl);
        t_out[i] = hit ? h.t : -1.0f;
        idx_out[i] = hit ? h.occ : 0xffffffffu;
      } else {
        hit = traverse<true>(sc, stack, blockDim.x, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), r[6], &h, tl);
        idx_out[i] = hit ? 1u : 0u;
      }
    }
    const unsigned long long mask = __ballot(hit);
    if (lane == 0 && mask) atomicAdd(hits, (unsigned long long)__popcll(mask));
  }
}

__global__ void k_trig_eval(TrigTables trig, const float* __restrict__ x, uint64_t n, float* s, float* c) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    s[i] = lsinf(trig, x[i]);
    c[i] = lcosf(trig, x[i]);
  }
}

// Region-local visited bits into the mesh visited grids: one thread per visited word.
__global__ void k_region_visited_or(DevMeshes ms, DevRegions rg, uint32_t nreg, const uint64_t* __restrict__ start, uint64_t total) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = (uint32_t)owner_of(start, nreg, w);
    const uint32_t bits = rg.vis[rg.vis_off[r] + (w - start[r])];
    if (!bits) continue;
    const uint32_t* rc = rg.rect + 4 * (size_t)r;
    const uint64_t rw = (uint64_t)(rc[2] - rc[0] + 1), texels = rw * (rc[3] - rc[1] + 1);
    const uint32_t m = rg.mesh[r];
    const uint64_t res = ms.res[m];
    for (uint32_t b = 0; b < 32; ++b) {
      const uint64_t ri = (w - start[r]) * 32 + b;
      if (ri >= texels || !((bits >> b) & 1u)) continue;
      const uint64_t vi = (rc[0] + ri % rw) + (rc[1] + ri / rw) * res;
      atomicOr(&ms.visited[ms.visited_off[m] + vi / 32], 1u << (vi % 32));
    }
  }
}

// Per-texel visit counts of a rebake service's meshes (16 bits per texel, two per word): for
// each visited bit of n pairs' region-local bits (pair p's words at vis[p], region rect[4p..],
// mesh resolution res[p], counts from word count_off[p]), adds (sign > 0) or subtracts one.
// One thread per visited word; start holds the pairs' word offsets (n + 1).
__global__ void k_vis_count(const uint64_t* __restrict__ vis, const uint32_t* __restrict__ rect, const uint32_t* __restrict__ res,
                            const uint64_t* __restrict__ count_off, uint32_t n, const uint64_t* __restrict__ start, uint64_t total,
                            uint32_t* counts, int sign) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t p = (uint32_t)owner_of(start, n, w);
    const uint32_t bits = reinterpret_cast<const uint32_t*>(vis[p])[w - start[p]];
    if (!bits) continue;
    const uint32_t* rc = rect + 4 * (size_t)p;
    const uint64_t rw = (uint64_t)(rc[2] - rc[0] + 1), texels = rw * (rc[3] - rc[1] + 1);
    for (uint32_t b = 0; b < 32; ++b) {
      const uint64_t ri = (w - start[p]) * 32 + b;
      if (ri >= texels || !((bits >> b) & 1u)) continue;
      const uint64_t vi = (rc[0] + ri % rw) + (rc[1] + ri / rw) * (uint64_t)res[p];
      const uint32_t one = 1u << (16u * (uint32_t)(vi & 1u));
      if (sign > 0) atomicAdd(&counts[count_off[p] + vi / 2], one);
      else atomicSub(&counts[count_off[p] + vi / 2], one);
    }
  }
}

// Visited grids of n meshes from their visit counts: bit i of mesh m's grid (from word
// visited_off[m]) is set when texel i's count (from word count_off[m]) is nonzero.  One thread
// per grid word; start holds the meshes' grid word offsets (n + 1).
__global__ void k_visited_from_counts(const uint32_t* __restrict__ res, const uint64_t* __restrict__ count_off, uint32_t n,
                                      const uint64_t* __restrict__ start, uint64_t total, const uint32_t* __restrict__ counts,
                                      uint32_t* visited) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t m = (uint32_t)owner_of(start, n, w);
    const uint64_t texels = (uint64_t)res[m] * res[m], i0 = (w - start[m]) * 32;
    uint32_t out = 0;
    for (uint32_t b = 0; b < 32 && i0 + b < texels; b += 2) {
      const uint32_t c = counts[count_off[m] + (i0 + b) / 2];
      if (c & 0xffffu) out |= 1u << b;
      if ((c >> 16) && i0 + b + 1 < texels) out |= 2u << b;
    }
    visited[w] = out;
  }
}

__global__ void k_copy_ranges(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, const uint64_t* __restrict__ ranges,
                              const uint64_t* __restrict__ start, uint64_t n, uint64_t total) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = owner_of(start, n, w);
    const uint64_t off = w - start[k];
    dst[ranges[3 * k + 1] + off] = src[ranges[3 * k] + off];
  }
}

// Renumbers the light slots of meshes' words in place: mesh j's words start at base[j] with nt[j]
// triangle headers (4 words: the offset of its light data, then the crop), light data = entry
// count, then per entry `stride` words whose first is the slot (triangle_words.rs:47-89).  One wave
// per mesh, lanes over its triangles.
__global__ void k_remap_slots(uint32_t* words, const uint64_t* __restrict__ base, const uint32_t* __restrict__ nt, uint32_t n,
                              uint32_t stride, const uint32_t* __restrict__ table) {
  const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
  for (uint64_t j = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); j < n; j += (uint64_t)gridDim.x * waves) {
    uint32_t* w = words + base[j];
    for (uint32_t t = lane; t < nt[j]; t += 64) {
      uint32_t* ld = w + w[4 * (size_t)t];
      const uint32_t count = ld[0];
      for (uint32_t e = 0; e < count; ++e) ld[1 + (size_t)stride * e] = table[ld[1 + (size_t)stride * e]];
    }
  }
}

__global__ void k_add_u32(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n, uint32_t delta) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[i] + delta;
}

// ---- launchers ------------------------------------------------------------------------------
static inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  return (uint32_t)(g < cap ? g : cap);
}
static inline size_t stack_bytes(const DevScene& sc, uint32_t block) {
  return (size_t)(sc.stack + 4) * block * (sc.stack16 ? 2 : 4);
}

hipError_t launch_setup(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, uint32_t nlights,
                        uint32_t local_lights, uint64_t* light_count, uint64_t* trace_count, uint64_t* touched_count) {
  if (tris.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup, dim3(grid_for(tris.n, kThreads)), dim3(kThreads), 0, s, tris, meshes, lights, nlights,
                     local_lights, light_count, trace_count, touched_count);
  return hipGetLastError();
}
hipError_t launch_assoc_fill(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, uint32_t nlights,
                             uint32_t local_lights, const uint64_t* ent_start, uint32_t* ent_light, uint32_t* ent_tri) {
  if (tris.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_assoc_fill, dim3(grid_for(tris.n, kThreads)), dim3(kThreads), 0, s, tris, meshes, lights, nlights,
                     local_lights, ent_start, ent_light, ent_tri);
  return hipGetLastError();
}
hipError_t launch_owned(hipStream_t s, DevTris tris, DevMeshes meshes, const uint64_t* trace_start, uint64_t total) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_owned, dim3(grid_for(total, kThreads)), dim3(kThreads), 0, s, tris, meshes, trace_start, total);
  return hipGetLastError();
}
hipError_t launch_texels(hipStream_t s, DevTris tris, DevMeshes meshes, const uint64_t* trace_count,
                         const uint64_t* cand_start, uint64_t total, int fill, unsigned long long* counts,
                         const uint64_t* tex_start, DevTexels out) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_texels, dim3(grid_for(total, kThreads)), dim3(kThreads), 0, s, tris, meshes, trace_count, cand_start,
                     total, fill, counts, tex_start, out);
  return hipGetLastError();
}
hipError_t launch_burial_lanes(hipStream_t s, DevScene scene, DevTris tris, DevMeshes meshes, DevTexels tx, uint64_t ntex,
                               uint64_t seed, TrigTables trig, uint8_t* buried, unsigned long long* kept, uint32_t block) {
  if (ntex == 0) return hipSuccess;
  hipLaunchKernelGGL(k_burial_lanes, dim3(grid_for(ntex, block)), dim3(block), stack_bytes(scene, block), s, scene, tris, meshes,
                     tx, ntex, seed, trig, buried, kept);
  return hipGetLastError();
}
hipError_t launch_burial_rays(hipStream_t s, DevTris tris, DevMeshes meshes, DevTexels tx, uint64_t first, const uint32_t* list,
                              const unsigned long long* count, uint64_t max_n, uint32_t k0, uint32_t nk, uint64_t seed,
                              TrigTables trig, float4* rays, float4* hits, uint64_t* total) {
  if (max_n == 0 || nk == 0 || k0 + nk > 8) return max_n == 0 ? hipSuccess : hipErrorInvalidValue;
  hipLaunchKernelGGL(k_burial_rays, dim3(grid_for(max_n, kThreads)), dim3(kThreads), 0, s, tris, meshes, tx, first, list, count,
                     max_n, k0, nk, seed, trig, rays, hits, total);
  return hipGetLastError();
}
hipError_t launch_burial_misses(hipStream_t s, const float4* rays, const float4* hits, const uint64_t* total, uint64_t max_rays,
                                float* out, uint64_t cap, unsigned long long* count) {
  if (max_rays == 0) return hipSuccess;
  hipLaunchKernelGGL(k_burial_misses, dim3(grid_for(max_rays, kThreads)), dim3(kThreads), 0, s, rays, hits, total, out, cap, count);
  return hipGetLastError();
}
hipError_t launch_count_nonzero(hipStream_t s, const uint32_t* flags, uint64_t n, unsigned long long* count) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_nonzero, dim3((uint32_t)std::min<uint64_t>(grid_for(n, kThreads), 4096)), dim3(kThreads), 0, s, flags, n,
                     count);
  return hipGetLastError();
}
hipError_t launch_burial_decide(hipStream_t s, DevScene scene, DevTexels tx, uint64_t first, const uint32_t* list,
                                const unsigned long long* count, uint64_t max_n, uint32_t k0, uint32_t nk, const float4* rays,
                                const float4* hits, uint8_t* back_count, uint8_t* buried, unsigned long long* kept, uint32_t* next,
                                unsigned long long* next_n) {
  if (max_n == 0 || nk == 0 || k0 + nk > 8) return max_n == 0 ? hipSuccess : hipErrorInvalidValue;
  hipLaunchKernelGGL(k_burial_decide, dim3(grid_for(max_n, kThreads)), dim3(kThreads), 0, s, scene, tx, first, list, count, max_n,
                     k0, nk, rays, hits, back_count, buried, kept, next, next_n);
  return hipGetLastError();
}
// A fingerprint of n words: the sum (mod 2^64) of splitmix64(index << 32 | word) over the words,
// added to *out (an order-free sum, so the result does not depend on the launch).
__global__ void k_words_hash(const uint32_t* __restrict__ w, uint64_t n, unsigned long long* out) {
  unsigned long long acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = (i << 32) ^ (uint64_t)w[i];
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    acc += z ^ (z >> 31);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63u) == 0) atomicAdd(out, acc);
}
hipError_t launch_words_hash(hipStream_t s, const uint32_t* words, uint64_t n, unsigned long long* out) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_words_hash, dim3((uint32_t)std::min<uint64_t>(grid_for(n, kThreads), 4096)), dim3(kThreads), 0, s, words, n,
                     out);
  return hipGetLastError();
}
__global__ void k_iota_u64(uint64_t* out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) out[i] = i;
}
hipError_t launch_iota_u64(hipStream_t s, uint64_t* out, uint64_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota_u64, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, out, n);
  return hipGetLastError();
}
hipError_t launch_remap_slots(hipStream_t s, uint32_t* words, const uint64_t* base, const uint32_t* nt, uint32_t n, uint32_t stride,
                              const uint32_t* table) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_remap_slots, dim3(grid_for(n, kThreads / 64)), dim3(kThreads), 0, s, words, base, nt, n, stride, table);
  return hipGetLastError();
}
hipError_t launch_add_u32(hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t delta) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_add_u32, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, in, out, n, delta);
  return hipGetLastError();
}
hipError_t launch_compact(hipStream_t s, DevTexels in, uint64_t ntex, const uint8_t* buried, const uint64_t* start,
                          unsigned long long* cursor, DevTexels out) {
  if (ntex == 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact, dim3(grid_for(ntex, kThreads)), dim3(kThreads), 0, s, in, ntex, buried, start, cursor, out);
  return hipGetLastError();
}
hipError_t launch_trace(hipStream_t s, DevScene scene, DevTris tris, DevMeshes meshes, const DevLight* lights,
                        DevTexels tx, const uint64_t* tex_start, DevEntries ent, DevRegions reg,
                        const uint64_t* unit_start, uint64_t units, TraceParams p, TrigTables trig, uint32_t block, bool multi) {
  if (units == 0) return hipSuccess;
  uint32_t waves = block / 64;
  hipLaunchKernelGGL(multi ? k_trace<true> : k_trace<false>, dim3(grid_for(units, waves)), dim3(block), stack_bytes(scene, block), s,
                     scene, tris, meshes, lights, tx, tex_start, ent, reg, unit_start, units, p, trig);
  return hipGetLastError();
}
hipError_t launch_vis_count(hipStream_t s, DevEntries ent, const DevLight* lights, const uint64_t* js, const uint32_t* me,
                            uint32_t nme, const uint64_t* tex_start, uint64_t q0, uint64_t n, TraceParams p, uint64_t* out) {
  hipLaunchKernelGGL(k_vis_count, dim3(grid_for(n + 1, kThreads)), dim3(kThreads), 0, s, ent, lights, js, me, nme, tex_start, q0,
                     n, p, out);
  return hipGetLastError();
}
hipError_t launch_vis_rays_a(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                             const uint64_t* js, const uint32_t* me, uint32_t nme, const uint64_t* tex_start, uint64_t q0,
                             uint64_t n, TraceParams p, TrigTables trig, float4* rays, float* contrib, float* wts,
                             VisState* state) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vis_rays_a, dim3(grid_for(n, 128)), dim3(128), 0, s, tris, meshes, lights, tx, ent, js, me, nme, tex_start,
                     q0, n, p, trig, rays, contrib, wts, state);
  return hipGetLastError();
}
hipError_t launch_vis_mid(hipStream_t s, DevEntries ent, const DevLight* lights, const uint64_t* js, const uint32_t* me,
                          uint32_t nme, const uint64_t* tex_start, uint64_t q0, uint64_t n, TraceParams p, const float4* rays,
                          const float* contrib, const float* wts, VisState* state, uint64_t* need) {
  hipLaunchKernelGGL(k_vis_mid, dim3(grid_for(n + 1, kThreads)), dim3(kThreads), 0, s, ent, lights, js, me, nme, tex_start, q0,
                     n, p, rays, contrib, wts, state, need);
  return hipGetLastError();
}
hipError_t launch_vis_list(hipStream_t s, const uint64_t* need, const uint64_t* scan, uint64_t n, uint32_t* blist) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vis_list, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, need, scan, n, blist);
  return hipGetLastError();
}
hipError_t launch_vis_rays_b(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                             const uint64_t* js, const uint32_t* me, uint32_t nme, const uint64_t* tex_start, uint64_t q0,
                             const uint32_t* blist, uint64_t nb, TraceParams p, TrigTables trig, float4* rays, float* contrib,
                             float* wts) {
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vis_rays_b, dim3(grid_for(nb, 128)), dim3(128), 0, s, tris, meshes, lights, tx, ent, js, me, nme,
                     tex_start, q0, blist, nb, p, trig, rays, contrib, wts);
  return hipGetLastError();
}
hipError_t launch_vis_finish(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                             DevRegions rg, const uint64_t* js, const uint32_t* me, uint32_t nme, const uint64_t* tex_start,
                             uint64_t q0, uint64_t n, TraceParams p, const VisState* state, const uint64_t* bscan, uint64_t nb,
                             const float4* rays, const float* contrib, const float* wts) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vis_finish, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, tris, meshes, lights, tx, ent, rg, js, me,
                     nme, tex_start, q0, n, p, state, bscan, nb, rays, contrib, wts);
  return hipGetLastError();
}
hipError_t launch_optimize(hipStream_t s, DevTris tris, DevMeshes meshes, DevEntries ent, DevRegions reg,
                           const uint8_t* has_bounce, uint8_t* keep) {
  if (ent.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_optimize, dim3(grid_for(ent.n, kThreads / 64)), dim3(kThreads), 0, s, tris, meshes, ent, reg, has_bounce, keep);
  return hipGetLastError();
}
hipError_t launch_pad_seams(hipStream_t s, DevMeshes meshes, DevRegions reg, const uint32_t* used_regions,
                            const uint64_t* area_start, uint32_t nused, uint64_t total) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pad_seams, dim3(grid_for(total, kThreads)), dim3(kThreads), 0, s, meshes, reg, used_regions,
                     area_start, nused, total);
  return hipGetLastError();
}
hipError_t launch_crop(hipStream_t s, DevTris tris, DevEntries ent, DevRegions reg, const uint64_t* kept_entry,
                       const uint64_t* kept_word_start, uint64_t nkept, uint64_t total_words, const uint64_t* kept_out,
                       uint32_t* out) {
  if (total_words == 0) return hipSuccess;
  hipLaunchKernelGGL(k_crop, dim3(grid_for(total_words, kThreads)), dim3(kThreads), 0, s, tris, ent, reg, kept_entry,
                     kept_word_start, nkept, total_words, kept_out, out);
  return hipGetLastError();
}
hipError_t launch_serialize(hipStream_t s, DevTris tris, const uint64_t* mesh_base, const uint32_t* ld_off,
                            const uint64_t* tri_kept_start, const uint64_t* kept_entry, const uint32_t* kept_shadow_off,
                            const uint32_t* kept_bounce_off, const uint32_t* ent_light, uint32_t stride,
                            const uint32_t* tri_first_in_mesh, uint32_t* out) {
  if (tris.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_serialize, dim3(grid_for(tris.n, kThreads)), dim3(kThreads), 0, s, tris, mesh_base, ld_off,
                     tri_kept_start, kept_entry, kept_shadow_off, kept_bounce_off, ent_light, stride, tri_first_in_mesh, out);
  return hipGetLastError();
}
hipError_t launch_distance_cubes(hipStream_t s, DevScene scene, const DevLight* lights, uint32_t nlocal, uint32_t* out) {
  if (nlocal == 0) return hipSuccess;
  uint64_t total = (uint64_t)nlocal * 6 * 2048;
  hipLaunchKernelGGL(k_distance_cubes, dim3(grid_for(total, 128)), dim3(128), stack_bytes(scene, 128), s, scene, lights,
                     nlocal, out);
  return hipGetLastError();
}
hipError_t launch_ray_queries(hipStream_t s, DevScene scene, const float* rays, uint64_t n, int closest, float* t_out,
                              uint32_t* idx_out) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ray_queries, dim3(grid_for(n, 128)), dim3(128), stack_bytes(scene, 128), s, scene, rays, n, closest,
                     t_out, idx_out);
  return hipGetLastError();
}
namespace {
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[64] = {};  // the scratch's device address, per device
int g_scratch_cus[64] = {};
}  // namespace

hipError_t launch_ray_queries_noalloc(hipStream_t s, int device, DevScene scene, const float* rays, uint64_t n, int closest,
                                      float* t_out, uint32_t* idx_out, uint64_t* hits, uint32_t* overflow) {
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  uint32_t* ptr = nullptr;
  int cus = 0;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {  // once per device, with the device current
      void* p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return e;
      e = hipDeviceGetAttribute(&g_scratch_cus[device], hipDeviceAttributeMultiprocessorCount, device);
      if (e != hipSuccess) return e;
      g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    ptr = g_scratch_ptr[device];
    cus = g_scratch_cus[device];
  }
  scene.overflow = ptr + 4;
  scene.counters = nullptr;
  hipLaunchKernelGGL(k_scratch_clear, dim3(1), dim3(64), 0, s, ptr, kNoAllocScratchWords);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n) {
    const uint32_t block = 128;
    const uint32_t grid = grid_for(n, block, (uint32_t)std::max(1, cus) * 8u);
    hipLaunchKernelGGL(k_ray_queries_queue, dim3(grid), dim3(block), stack_bytes(scene, block), s, scene, rays, n, closest, t_out,
                       idx_out, reinterpret_cast<unsigned long long*>(ptr), reinterpret_cast<unsigned long long*>(ptr + 2));
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  uint32_t back[5] = {0, 0, 0, 0, 0};
  if ((e = hipMemcpyAsync(back, ptr, sizeof back, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (hits) *hits = (uint64_t)back[2] | ((uint64_t)back[3] << 32);
  if (overflow) *overflow = back[4];
  return hipSuccess;
}
hipError_t launch_region_visited_or(hipStream_t s, DevMeshes meshes, DevRegions rg, uint32_t nregions, const uint64_t* vis_words_start,
                                    uint64_t total_words) {
  if (total_words == 0) return hipSuccess;
  hipLaunchKernelGGL(k_region_visited_or, dim3(grid_for(total_words, kThreads)), dim3(kThreads), 0, s, meshes, rg, nregions,
                     vis_words_start, total_words);
  return hipGetLastError();
}
hipError_t launch_vis_count(hipStream_t s, const uint64_t* vis, const uint32_t* rect, const uint32_t* res, const uint64_t* count_off,
                            uint32_t n, const uint64_t* start, uint64_t total, uint32_t* counts, int sign) {
  if (total == 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vis_count, dim3(grid_for(total, kThreads)), dim3(kThreads), 0, s, vis, rect, res, count_off, n, start, total,
                     counts, sign);
  return hipGetLastError();
}
hipError_t launch_visited_from_counts(hipStream_t s, const uint32_t* res, const uint64_t* count_off, uint32_t n, const uint64_t* start,
                                      uint64_t total, const uint32_t* counts, uint32_t* visited) {
  if (total == 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_visited_from_counts, dim3(grid_for(total, kThreads)), dim3(kThreads), 0, s, res, count_off, n, start, total,
                     counts, visited);
  return hipGetLastError();
}
hipError_t launch_copy_ranges(hipStream_t s, const uint32_t* src, uint32_t* dst, const uint64_t* ranges, const uint64_t* range_start,
                              uint64_t n, uint64_t total_words) {
  if (total_words == 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_ranges, dim3(grid_for(total_words, kThreads)), dim3(kThreads), 0, s, src, dst, ranges, range_start, n,
                     total_words);
  return hipGetLastError();
}
hipError_t launch_trig_eval(hipStream_t s, TrigTables trig, const float* x, uint64_t n, float* sin_out, float* cos_out) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_trig_eval, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, trig, x, n, sin_out, cos_out);
  return hipGetLastError();
}

hipError_t scan_u64(hipStream_t s, const uint64_t* in, uint64_t* out, uint64_t n, void** tmp, size_t* tmp_bytes) {
  size_t need = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)(n + 1), s);
  if (e != hipSuccess) return e;
  if (need > *tmp_bytes) {
    if (*tmp) (void)hipFree(*tmp);
    *tmp = nullptr;
    e = hipMalloc(tmp, need);
    if (e != hipSuccess) { *tmp_bytes = 0; return e; }
    *tmp_bytes = need;
  }
  return hipcub::DeviceScan::ExclusiveSum(*tmp, need, in, out, (int)(n + 1), s);
}

// =============================================================================================
// Bounce
// =============================================================================================

// Occluders::surface_normal (occluder.rs:131-145)
__device__ __forceinline__ f3 surface_normal(const DevScene& sc, const DevOcc& occ, const Hit& h) {
  f3 face = tri_face_normal(sc, h.pos);
  if (!(occ.flags[h.occ] & 1u)) return face;
  const float* n = occ.nrm + 9 * (size_t)h.occ;
  float wx = (1.0f - h.u) - h.v;
  f3 s = add(add(mul(ld3(n), wx), mul(ld3(n + 3), h.u)), mul(ld3(n + 6), h.v));
  f3 r;
  return try_normalize(s, &r) ? r : face;
}

// The photon of texel (face, x, y) of a light's cube (PhotonCube::photon, bounce.rs:48-115): the
// 16-byte record (distance, packed lit-side normal, packed albedo, 0).
__device__ __forceinline__ uint4 photon_record(const DevScene& sc, const DevOcc& occ, const DevLight& L, uint32_t face, uint32_t x,
                                               uint32_t y, uint32_t res, uint32_t* stack, Tally& tl) {
  f3 lp = ld3(L.pos);
  f3 dir = cube_texel_dir(face, x, y, res);
  Hit h;
  uint32_t w0, w1 = 0, w2 = 0;
  if (!traverse<false>(sc, stack, blockDim.x, lp, dir, L.radius, &h, tl, L.alpha)) {
    w0 = __float_as_uint(L.radius);
  } else {
    f3 normal = surface_normal(sc, occ, h);
    f3 point = add(lp, mul(dir, h.t));
    float bias = maxf_(h.t * 0.001f, 0.001f);
    f3 nudged = add(sub(point, mul(dir, bias)), mul(normal, bias));
    bool lit_side = dot(tri_face_normal(sc, h.pos), dir) < 0.0f;
    w0 = __float_as_uint(len(sub(nudged, lp)));
    if (lit_side)
      w1 = ((uint32_t)sat_i8(roundf(normal.x * 127.0f)) & 0xffu) | (((uint32_t)sat_i8(roundf(normal.y * 127.0f)) & 0xffu) << 8) |
           (((uint32_t)sat_i8(roundf(normal.z * 127.0f)) & 0xffu) << 16);
    const float* a = occ.albedo + 3 * (size_t)h.occ;
    w2 = sat_u8(roundf(clampf_(a[0], 0.0f, 1.0f) * 255.0f)) | (sat_u8(roundf(clampf_(a[1], 0.0f, 1.0f) * 255.0f)) << 8) |
         (sat_u8(roundf(clampf_(a[2], 0.0f, 1.0f) * 255.0f)) << 16);
  }
  return make_uint4(w0, w1, w2, 0u);
}

// PhotonCube::build (bounce.rs:126-151): every texel of every cube.
__global__ void k_photons(DevScene sc, DevOcc occ, const DevLight* __restrict__ lights, const uint32_t* __restrict__ cube_slot,
                          uint32_t ncubes, uint32_t res, uint32_t* photons) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint64_t per = 6ull * res * res;
  const uint64_t total = per * ncubes;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = (uint32_t)(i / per);
    uint64_t rem = i % per;
    uint32_t face = (uint32_t)(rem / ((uint64_t)res * res));
    uint32_t yx = (uint32_t)(rem % ((uint64_t)res * res));
    uint32_t y = yx / res, x = yx % res;
    reinterpret_cast<uint4*>(photons)[i] = photon_record(sc, occ, lights[cube_slot[c]], face, x, y, res, stack, tl);
  }
}

// Lazy photon cubes: only the texels a gather reads.  need holds one byte per texel in tile order
// (byte 64 * tile + bit, tile and bit as in the tile masks of k_photon_tiles); the listed form
// traces texels list[0 .. total) of that numbering (64 * tile + bit, below 2^32).
__global__ void k_photons_listed(DevScene sc, DevOcc occ, const DevLight* __restrict__ lights,
                                 const uint32_t* __restrict__ cube_slot, uint32_t res, uint32_t tpf,
                                 const uint32_t* __restrict__ list, uint64_t total, uint32_t* photons) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint32_t per_face = tpf * tpf, per_cube = 6 * per_face;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = list[i], tile = v >> 6, bit = v & 63u;
    const uint32_t c = tile / per_cube, rem = tile % per_cube, face = rem / per_face, t = rem % per_face;
    const uint32_t x = (t % tpf) * kPhotonTile + (bit & 7u), y = (t / tpf) * kPhotonTile + (bit >> 3);
    reinterpret_cast<uint4*>(photons)[(uint64_t)c * 6 * res * res + (uint64_t)face * res * res + (uint64_t)y * res + x] =
        photon_record(sc, occ, lights[cube_slot[c]], face, x, y, res, stack, tl);
  }
}

// Per tile, the mask of its needed texels (need_mask) and their count (cnt; entry `tiles` is 0,
// for the scan).  One thread per tile, reading its 64 need bytes.
__global__ void k_need_count(const uint4* __restrict__ need, uint64_t tiles, unsigned long long* need_mask, uint64_t* cnt) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t <= tiles; t += (uint64_t)gridDim.x * blockDim.x) {
    if (t == tiles) {
      cnt[t] = 0;
      continue;
    }
    unsigned long long m = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint4 v = need[4 * t + q];
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if ((w[k >> 2] >> (8 * (k & 3))) & 0xffu) m |= 1ull << (16 * q + k);
    }
    need_mask[t] = m;
    cnt[t] = (uint64_t)__popcll(m);
  }
}

// The needed texels of each tile, in bit order, at its offset.
__global__ void k_need_list(const unsigned long long* __restrict__ need_mask, const uint64_t* __restrict__ off, uint64_t tiles,
                            uint32_t* list) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tiles; t += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long m = need_mask[t];
    uint64_t o = off[t];
    while (m) {
      const uint32_t bit = (uint32_t)__ffsll(m) - 1u;
      list[o++] = (uint32_t)(t << 6) | bit;
      m &= m - 1ull;
    }
  }
}

// Photon tiles: what a gather sample can learn about its photon without reading it.  A sample
// whose direction s falls in cube texel i reconstructs the photon as lp + s * distance_i
// (bounce.rs:262-286), so every such point lies in the set {lp + w_i * f / |f|} over the texel's
// square f of the face plane (major axis 1), widened by 1e-6 in face coordinates for the f32
// roundings that pick the texel.  That set lies in the box of the eight points
// lp + (w_i / |f|max) c and lp + (w_i / |f|min) c over the square's corners c (|f| over the
// square lies in [|f|min, |f|max]).  Per tile (a step of 8 texels) the summary covers the 9x9
// texels from its corner, so that the at most 2x2 candidate texels of a fast-pass sample always
// lie in one summary: the box spans those of its texels whose photon has a lit surface (non-zero
// normal), padded by 1e-5 of its coordinates and of |lp| + distance for the f32 roundings of
// lp + s * distance.  One wave per tile, two texels per lane (index lane and 64 + lane, bit
// dy * 9 + dx of the tile's two live words).  With need_mask (lazy cubes) only the needed texels
// count: the others hold no photon and are left out of the mask and the box, which then describe
// the needed texels only (the gather asks only about those).
__device__ __forceinline__ void face_plane(uint32_t face, double sc, double tc, double* f) {
  switch (face) {
    case 0: f[0] = 1.0; f[1] = -tc; f[2] = -sc; break;
    case 1: f[0] = -1.0; f[1] = -tc; f[2] = sc; break;
    case 2: f[0] = sc; f[1] = 1.0; f[2] = tc; break;
    case 3: f[0] = sc; f[1] = -1.0; f[2] = -tc; break;
    case 4: f[0] = sc; f[1] = -tc; f[2] = 1.0; break;
    default: f[0] = -sc; f[1] = -tc; f[2] = -1.0; break;
  }
}
__device__ __forceinline__ float f32_down(double v) {
  float f = (float)v;
  return (double)f > v ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float f32_up(double v) {
  float f = (float)v;
  return (double)f < v ? nextafterf(f, INFINITY) : f;
}
__global__ void k_photon_tiles(const DevLight* __restrict__ lights, const uint32_t* __restrict__ cube_slot, uint32_t ncubes,
                               uint32_t res, uint32_t tpf, const uint32_t* __restrict__ photons,
                               const unsigned long long* __restrict__ need_mask, unsigned long long* live, float4* box) {
  const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
  const uint64_t per = 6ull * tpf * tpf, total = per * ncubes;
  for (uint64_t q = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); q < total; q += (uint64_t)gridDim.x * waves) {
    const uint32_t c = (uint32_t)(q / per), rem = (uint32_t)(q % per);
    const uint32_t face = rem / (tpf * tpf), t = rem % (tpf * tpf);
    const DevLight& L = lights[cube_slot[c]];
    const double lp[3] = {(double)L.pos[0], (double)L.pos[1], (double)L.pos[2]};
    // the tile's kTileSpan^2 texels, two per lane: index lane and 64 + lane (dy * kTileSpan + dx)
    bool lit2[2] = {false, false};
    uint32_t code2[2] = {0u, 0u};
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, wmax = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t idx = (uint32_t)h * 64u + lane;
      if (idx >= kTileSpan * kTileSpan) continue;
      const uint32_t x = (t % tpf) * kPhotonTile + idx % kTileSpan, y = (t / tpf) * kPhotonTile + idx / kTileSpan;
      if (x >= res || y >= res) continue;
      if (need_mask) {  // lazy cubes: the texel's need bit, in its own tile's 8x8 mask
        const uint64_t own = (uint64_t)c * per + (uint64_t)face * tpf * tpf + (y / kPhotonTile) * tpf + x / kPhotonTile;
        if (!((need_mask[own] >> ((y % kPhotonTile) * kPhotonTile + x % kPhotonTile)) & 1ull)) continue;
      }
      const uint4 ph = reinterpret_cast<const uint4*>(photons)[(uint64_t)c * 6 * res * res + (uint64_t)face * res * res +
                                                               (uint64_t)y * res + x];
      const uint32_t code = ph.y & 0xffffffu;
      code2[h] = code;
      lit2[h] = code != 0u;
      if (lit2[h]) {
        const double w = (double)__uint_as_float(ph.x), e = 1.0e-6;
        wmax = fmax(wmax, w);
        const double s0 = ((double)x / res - e) * 2.0 - 1.0, s1 = ((double)(x + 1) / res + e) * 2.0 - 1.0;
        const double t0 = ((double)y / res - e) * 2.0 - 1.0, t1 = ((double)(y + 1) / res + e) * 2.0 - 1.0;
        const double sn = fmin(fmax(0.0, s0), s1), tn = fmin(fmax(0.0, t0), t1);
        const double fmin_ = sqrt(1.0 + sn * sn + tn * tn);
        double fmax2 = 0.0;
        for (int k = 0; k < 4; ++k) {
          const double sc = (k & 1) ? s1 : s0, tc = (k & 2) ? t1 : t0;
          fmax2 = fmax(fmax2, 1.0 + sc * sc + tc * tc);
        }
        const double near_ = w / sqrt(fmax2), far_ = w / fmin_;
        for (int k = 0; k < 4; ++k) {
          double f[3];
          face_plane(face, (k & 1) ? s1 : s0, (k & 2) ? t1 : t0, f);
          for (int a = 0; a < 3; ++a) {
            const double p0 = lp[a] + f[a] * near_, p1 = lp[a] + f[a] * far_;
            lo[a] = fmin(lo[a], fmin(p0, p1));
            hi[a] = fmax(hi[a], fmax(p0, p1));
          }
        }
      }
    }
    const uint64_t mask0 = __ballot(lit2[0]), mask1 = __ballot(lit2[1]);
    // the first lit texel's normal code; uniform when every lit texel has it
    const uint32_t first = mask0 ? (uint32_t)__ffsll((unsigned long long)mask0) - 1u
                                 : mask1 ? (uint32_t)__ffsll((unsigned long long)mask1) - 1u : 0u;
    const uint32_t code0 = mask0 ? __shfl(code2[0], (int)first) : __shfl(code2[1], (int)first);
    const bool uniform =
        (mask0 | mask1) != 0 && __ballot((lit2[0] && code2[0] != code0) || (lit2[1] && code2[1] != code0)) == 0;
    const bool any = (mask0 | mask1) != 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      for (int a = 0; a < 3; ++a) {
        lo[a] = fmin(lo[a], __shfl_xor(lo[a], off));
        hi[a] = fmax(hi[a], __shfl_xor(hi[a], off));
      }
      wmax = fmax(wmax, __shfl_xor(wmax, off));
    }
    if (lane == 0) {
      live[2 * q] = (unsigned long long)mask0;
      live[2 * q + 1] = (unsigned long long)mask1;  // texels 64 .. kTileSpan^2 - 1
      float4 b0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b1 = b0, b2 = b0;
      if (any) {
        const f3 n = normalize_or_zero(mk3((float)(int8_t)(code0 & 0xffu) / 127.0f, (float)(int8_t)((code0 >> 8) & 0xffu) / 127.0f,
                                           (float)(int8_t)((code0 >> 16) & 0xffu) / 127.0f));
        b0 = make_float4(n.x, n.y, n.z, uniform ? 1.0f : 0.0f);
        const double lpm = fmax(fabs(lp[0]), fmax(fabs(lp[1]), fabs(lp[2])));
        double scale = 0.0;
        float l[3], h[3];
        for (int a = 0; a < 3; ++a) {
          const double pad = 1.0e-5 * (fmax(fabs(lo[a]), fabs(hi[a])) + lpm + wmax) + 1.0e-6;
          l[a] = f32_down(lo[a] - pad);
          h[a] = f32_up(hi[a] + pad);
          scale += fmax(fabs((double)l[a]), fabs((double)h[a]));
        }
        b1 = make_float4(l[0], l[1], l[2], f32_up(scale));
        b2 = make_float4(h[0], h[1], h[2], 0.0f);
      }
      box[3 * q] = b0;
      box[3 * q + 1] = b1;
      box[3 * q + 2] = b2;
    }
  }
}

// Whether a sample in a lit texel of tile `tb` (three float4 of tile_box) certainly brings no
// light: with o the ray origin, n the receiver's face normal and m a margin well above the f32
// roundings of the gather's own arithmetic (1e-5 of the coordinates involved), either the tile's
// photons share the receiver's normal (the weight max(n.d, 0) max(pn.-d, 0) is then zero for
// every d) and o is outside the box (so the distance is positive), or the whole box lies behind
// the receiver's plane (n.d < 0), or the tile's photons share a normal whose plane has o behind it
// (pn.-d < 0).  In all three cases the ray counts as cast and its contribution is zero.
__device__ __forceinline__ bool tile_dark(const float4& b0, const float4& b1, const float4& b2, f3 o, f3 n, float o_scale) {
  const float m = 1.0e-5f * (o_scale + b1.w) + 1.0e-6f;
  const bool uniform = b0.w != 0.0f;
  const bool inside = o.x >= b1.x && o.x <= b2.x && o.y >= b1.y && o.y <= b2.y && o.z >= b1.z && o.z <= b2.z;
  if (uniform && b0.x == n.x && b0.y == n.y && b0.z == n.z && !inside) return true;
  const float front = n.x * ((n.x >= 0.0f ? b2.x : b1.x) - o.x) + n.y * ((n.y >= 0.0f ? b2.y : b1.y) - o.y) +
                      n.z * ((n.z >= 0.0f ? b2.z : b1.z) - o.z);
  if (front < -m) return true;
  if (!uniform) return false;
  const float back = b0.x * ((b0.x >= 0.0f ? b1.x : b2.x) - o.x) + b0.y * ((b0.y >= 0.0f ? b1.y : b2.y) - o.y) +
                     b0.z * ((b0.z >= 0.0f ? b1.z : b2.z) - o.z);
  return back > m;
}
__device__ __forceinline__ bool tile_outside(const float4& b1, const float4& b2, f3 o) {
  return !(o.x >= b1.x && o.x <= b2.x && o.y >= b1.y && o.y <= b2.y && o.z >= b1.z && o.z <= b2.z);
}
// The tile holding cube texel (face, x, y) and the texel's bit in its summary (dy * kTileSpan + dx).
__device__ __forceinline__ uint32_t photon_tile(uint32_t face, uint32_t x, uint32_t y, uint32_t tpf, uint32_t* bit) {
  *bit = (y % kPhotonTile) * kTileSpan + (x % kPhotonTile);
  return face * tpf * tpf + (y / kPhotonTile) * tpf + x / kPhotonTile;
}
// The same tile and the texel's bit in the tile's own 8x8 texels (need marks, lazy cubes).
__device__ __forceinline__ uint32_t photon_tile8(uint32_t face, uint32_t x, uint32_t y, uint32_t tpf, uint32_t* bit) {
  *bit = (y % kPhotonTile) * kPhotonTile + (x % kPhotonTile);
  return face * tpf * tpf + (y / kPhotonTile) * tpf + x / kPhotonTile;
}
// Bit b of a tile's summary (two words).
__device__ __forceinline__ bool tile_bit(const unsigned long long* live, uint64_t tile, uint32_t b) {
  return (live[2 * tile + (b >> 6)] >> (b & 63u)) & 1ull;
}

// glam acos_approx (DirectXMath XMScalarAcos polynomial)
__device__ __forceinline__ float acos_approx(float v) {
  bool nonneg = v >= 0.0f;
  float x = fabsf(v);
  float omx = 1.0f - x;
  if (omx < 0.0f) omx = 0.0f;
  float root = sqrtf(omx);
  float r = ((((((-0.0012624911f * x + 0.0066700901f) * x - 0.0170881256f) * x + 0.0308918810f) * x - 0.0501743046f) * x +
              0.0889789874f) * x - 0.2145988016f) * x + 1.5707963050f;
  r *= root;
  return nonneg ? r : kPi - r;
}
// glam Vec3::slerp
__device__ f3 slerp3(const TrigTables& trig, f3 a, f3 b, float s) {
  float la = len(a), lb = len(b);
  float d = dot(a, b) / (la * lb);
  if (fabsf(d) < 1.0f - 3.0e-7f) {
    float theta = acos_approx(d);
    float sin_theta = lsinf(trig, theta);
    float t1 = lsinf(trig, theta * (1.0f - s));
    float t2 = lsinf(trig, theta * s);
    float rl = la + (lb - la) * s;
    return mul(add(mul(mul(a, rl / la), t1), mul(mul(b, rl / lb), t2)), 1.0f / sin_theta);
  }
  if (d < 0.0f) {
    f3 o = fabsf(a.x) > fabsf(a.y) ? mk3(-a.z, 0.0f, a.x) : mk3(0.0f, a.z, -a.y);
    f3 axis = normalize(o);
    float half = kPi * s * 0.5f;
    float sn, cs;
    lsincosf(trig, half, &sn, &cs);
    float q[4] = {axis.x * sn, axis.y * sn, axis.z * sn, cs};
    float rl = la + (lb - la) * s;
    return mul(quat_rotate(q, a), rl / la);
  }
  return add(a, mul(sub(b, a), s));
}
// random::on_unit_sphere (random.rs:22-28)
__device__ __forceinline__ f3 on_unit_sphere(const TrigTables& trig, Pcg& rng) {
  float z = pcg_f32(rng) * 2.0f - 1.0f;
  float angle = pcg_f32(rng) * kTau;
  float r = sqrtf(maxf_(1.0f - z * z, 0.0f));
  float sa, ca;
  lsincosf(trig, angle, &sa, &ca);
  return mk3(r * ca, r * sa, z);
}

struct BounceUnit {
  uint32_t t;      // batch-local triangle
  uint64_t e;      // entry
  uint64_t i;      // texel index within the triangle
  uint64_t nt;     // texels of the triangle
  uint64_t t0;     // first texel
};
__device__ __forceinline__ BounceUnit bounce_unit(const DevBounce& b, const DevEntries& ent, const uint64_t* tex_start, uint64_t unit,
                                                  uint32_t lane) {
  BounceUnit u;
  uint64_t j = owner_of(b.unit_start, b.n, unit);
  u.e = b.entry[j];
  u.t = ent.tri[u.e];
  u.t0 = tex_start[u.t];
  u.nt = tex_start[u.t + 1] - u.t0;
  u.i = (unit - b.unit_start[j]) * 64 + lane;
  return u;
}

// Marks the filter tile of window w that holds window texel (lx, ly) when the value v written there
// has a positive component: only such values survive the dilation, so the blur of a tile whose
// 3x3 tile neighbourhood holds none is zero everywhere.
__device__ __forceinline__ void mark_tile(const DevBounce& b, uint32_t w, uint64_t lx, uint64_t ly, f3 v) {
  if (!(max_elem(v) > 0.0f)) return;
  const uint32_t* wr = b.win_rect + 4 * (size_t)w;
  const uint64_t tiles_x = ((uint64_t)(wr[2] - wr[0]) + kFilterTile) / kFilterTile;
  b.tile_lit[b.tile_start[w] + (ly / kFilterTile) * tiles_x + lx / kFilterTile] = 1;
}

// "Later triangles overwrite shared texels" (bounce.rs:202-209): the highest triangle that
// gathers a texel owns it.
__global__ void k_bounce_owner(DevTris tr, const DevLight* __restrict__ lights, DevTexels tx, const uint64_t* __restrict__ tex_start,
                               DevEntries ent, DevBounce b, uint64_t units) {
  const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
  for (uint64_t unit = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); unit < units; unit += (uint64_t)gridDim.x * waves) {
    BounceUnit u = bounce_unit(b, ent, tex_start, unit, lane);
    if (u.i >= u.nt) continue;
    const DevLight& L = lights[ent.light[u.e]];
    uint64_t j = u.t0 + u.i;
    if (L.shape != 3) {
      float4 p4 = tx.pos[j];
      f3 to_texel = sub(mk3(p4.x, p4.y, p4.z), ld3(L.pos));
      if (len_sq(to_texel) > L.radius * L.radius) continue;
    }
    uint32_t w = b.ent_window[u.e];
    const uint32_t* wr = b.win_rect + 4 * (size_t)w;
    uint32_t xy = tx.xy[j];
    uint64_t k = (uint64_t)((xy & 0xffffu) - wr[0]) + (uint64_t)((xy >> 16) - wr[1]) * (wr[2] - wr[0] + 1);
    unsigned int* owner = reinterpret_cast<unsigned int*>(b.map_b) + b.win_off[w] * 3;
    atomicMax(&owner[k], u.t + 1);
  }
}

// gather_triangle (bounce.rs:235-292) and gather_sun (sun_bounce.rs:24-80)
__global__ void k_bounce_gather(DevScene sc, DevOcc occ, DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx,
                                const uint64_t* __restrict__ tex_start, DevEntries ent, DevBounce b, uint64_t units,
                                BounceParams p, TrigTables trig) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
  for (uint64_t unit = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); unit < units; unit += (uint64_t)gridDim.x * waves) {
    BounceUnit u = bounce_unit(b, ent, tex_start, unit, lane);
    uint32_t m = tr.mesh[u.t];
    uint32_t rays = 0;
    if (u.i < u.nt) {
      const DevLight& L = lights[ent.light[u.e]];
      uint64_t j = u.t0 + u.i;
      uint32_t xy = tx.xy[j];
      uint32_t x = xy & 0xffffu, y = xy >> 16;
      float4 p4 = tx.pos[j];
      f3 pos = mk3(p4.x, p4.y, p4.z);
      f3 normal = ld3(tr.face + 3 * (size_t)u.t);
      uint32_t samples = L.bounce_samples > 2 ? L.bounce_samples : 2u;
      bool write = true;
      f3 sum = mk3(0.0f, 0.0f, 0.0f);
      Hit h;
      if (L.shape != 3) {
        f3 lp = ld3(L.pos);
        f3 to_texel = sub(pos, lp);
        if (p.sun_only || len_sq(to_texel) > L.radius * L.radius) {
          write = false;
        } else {
          const uint32_t* cube = p.photons + (size_t)b.cube_of_slot[ent.light[u.e]] * 6 * p.cube_res * p.cube_res * 4;
          float step = 0.9f / (float)(samples - 1);
          f3 direction = normalize_or(to_texel, mk3(0.0f, 1.0f, 0.0f));
          f3 origin = add(pos, mul(normal, kRayOffset));
          Pcg rng = stream5(p.seed, ms.id[m], L.id, x, y);
          for (uint32_t k = 0; k < samples; ++k) {
            float spread = 0.1f + (float)k * step;
            f3 rnd = on_unit_sphere(trig, rng);
            f3 sample = normalize_or(slerp3(trig, direction, rnd, spread), direction);
            uint32_t ci = cube_texel_of(sample, p.cube_res);
            const uint4 ph = reinterpret_cast<const uint4*>(cube)[ci];
            const uint32_t w0 = ph.x, w1 = ph.y, w2 = ph.z;
            f3 pn = normalize_or_zero(mk3((float)(int8_t)(w1 & 0xffu) / 127.0f, (float)(int8_t)((w1 >> 8) & 0xffu) / 127.0f,
                                          (float)(int8_t)((w1 >> 16) & 0xffu) / 127.0f));
            f3 photon = add(lp, mul(sample, __uint_as_float(w0)));
            f3 to_photon = sub(photon, origin);
            float distance = len(to_photon);
            if (is_zero(pn) || distance <= 0.0f) continue;
            f3 d = divs(to_photon, distance);
            rays++;
            if (!traverse<true>(sc, stack, blockDim.x, origin, d, distance, &h, tl, L.alpha)) {
              float weight = maxf_(dot(normal, d), 0.0f) * maxf_(dot(pn, neg(d)), 0.0f);
              f3 alb = p.colored ? mk3((float)(w2 & 0xffu) / 255.0f, (float)((w2 >> 8) & 0xffu) / 255.0f,
                                       (float)((w2 >> 16) & 0xffu) / 255.0f)
                                 : mk3(1.0f, 1.0f, 1.0f);
              sum = add(sum, mul(alb, weight));
            }
          }
        }
      } else {
        f3 toward = ld3(L.toward);
        f3 origin = add(pos, mul(normal, kRayOffset));
        Pcg rng = stream5(p.seed, ms.id[m], L.id, x, y);
        f2 offset;
        offset.x = pcg_f32(rng);
        offset.y = pcg_f32(rng);
        for (uint32_t k = 0; k < samples; ++k) {
          f3 direction = cosine_dir(trig, normal, spread_pt(k, offset, true));
          rays++;
          if (!traverse<false>(sc, stack, blockDim.x, origin, direction, INFINITY, &h, tl, L.alpha)) continue;
          f3 face = tri_face_normal(sc, h.pos);
          if (dot(face, direction) >= 0.0f) continue;
          float lit = dot(surface_normal(sc, occ, h), toward);
          if (lit <= 0.0f || dot(face, toward) <= 0.0f) continue;
          f3 point = add(add(origin, mul(direction, h.t)), mul(face, kRayOffset));
          rays++;
          Hit h2;
          if (!traverse<true>(sc, stack, blockDim.x, point, toward, INFINITY, &h2, tl, L.alpha)) {
            f3 alb = mk3(1.0f, 1.0f, 1.0f);
            if (p.colored) {
              const float* a = occ.albedo + 3 * (size_t)h.occ;
              alb = mk3(a[0], a[1], a[2]);
            }
            sum = add(sum, mul(alb, lit));
          }
        }
      }
      if (write) {
        uint32_t w = b.ent_window[u.e];
        const uint32_t* wr = b.win_rect + 4 * (size_t)w;
        uint64_t k = (uint64_t)(x - wr[0]) + (uint64_t)(y - wr[1]) * (wr[2] - wr[0] + 1);
        const unsigned int* owner = reinterpret_cast<const unsigned int*>(b.map_b) + b.win_off[w] * 3;
        if (owner[k] == u.t + 1) {
          f3 v = divs(mul(sum, L.bounce_intensity), (float)samples);
          float* dst = b.map_a + (b.win_off[w] + k) * 3;
          dst[0] = v.x;
          dst[1] = v.y;
          dst[2] = v.z;
          mark_tile(b, w, (uint64_t)(x - wr[0]), (uint64_t)(y - wr[1]), v);
        }
      }
    }
    wave_add(&b.rays[m], rays);
    if (b.wrays) wave_add(&b.wrays[b.ent_window[u.e]], rays);  // one entry per wave
    if (sc.counters) {
      wave_add(&sc.counters[0], tl.nodes);
      wave_add(&sc.counters[1], tl.tris);
      tl = Tally{};
    }
  }
}

// ---- streamed point-light gather ----------------------------------------------------------------
// The same gather as k_bounce_gather's point-light branch (bounce.rs:235-292), reorganised so a
// wave's lanes do not wait for one another: every lane owns one (texel, light) job and walks its
// samples in order, one BVH node per loop iteration; a lane whose ray ends sets up its next sample
// (or takes a new job from a global queue) while the other lanes keep traversing.  Each lane
// still sums its own samples in sample order from its own stream, so results are bit-identical.
//
// Jobs are (texel, entry) pairs of point bounce lights within the light's radius; those whose
// texel the triangle owns in the window map are traced, the rest (their result would be
// overwritten by a later triangle, bounce.rs:202-209) only count their rays for BakeStats.

// One BVH node of an any-hit traversal (the loop body of traverse<true>).  Returns 0 while the
// traversal continues (node/sp updated), 1 on a hit, 2 when the ray is unoccluded.
__device__ __forceinline__ uint32_t visit_any(const DevScene& sc, uint32_t* __restrict__ stack, uint32_t stride, f3 o, f3 d,
                                              const SlabRay& sr, float tmax, int32_t& node, uint32_t& sp, Tally& tl,
                                              uint32_t alpha) {
  tl.nodes++;
  float tn[kBvhWidth];
  Kids ch;
  const uint32_t mask = node_hits(sc, node, sr, tmax * (1.0f + kBoxEps), tn, &ch);
  uint32_t leaves = mask & leaf_bits(ch), i = 0, end = 0;
  for (;;) {
    if (i == end) {
      if (!leaves) break;
      const uint32_t v = (uint32_t)(-1 - kid(ch, __builtin_ctz(leaves)));
      leaves &= leaves - 1;
      i = v >> 2;
      end = i + (v & 3u) + 1u;
    }
    const float4 a = sc.tris[3 * i], b = sc.tris[3 * i + 1], cc = sc.tris[3 * i + 2];
    i++;
    float t, u, vv;
    tl.tris++;
    if (!mt_hit(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), o, d, tmax, &t, &u, &vv)) continue;
    if (!finite_(t)) continue;
    if (alpha && !alpha_blocks(sc, __float_as_uint(a.w), u, vv)) continue;
    return 1u;
  }
  const int32_t next = push_children(sc, mask, ch, tn, stack, stride, sp);
  if (next >= 0) {
    node = next;
    return 0u;
  }
  if (sp == 0) return 2u;
  sp--;
  node = pop_node(sc, stack, stride, sp);
  return 0u;
}

// The hit leaf children of the last node a closest-hit stream step visited, still to be tested:
// `mask` over the node's children (refs `ch`, box entries `tn`), and triangles [i, end) of the
// leaf under test.
struct PendingLeaves {
  Kids ch;
  float tn[kBvhWidth];
  uint32_t mask, i, end;
};

// One step of a closest-hit traversal in a ray stream: one triangle of the pending leaves (the
// nearest leaf first; a leaf whose box starts beyond the best hit is dropped), or, with none
// left, one BVH node, whose hit leaf children become pending and whose inner children are pushed
// (the nearest is visited next; node -1: pop).  The node and leaf order is that of traverse<false>
// (a node's leaves before the next node), one triangle per step, so that the lanes of a wave do
// similar work per step.  Keeps the lowest (t, occluder index) hit.  Returns 0 while the
// traversal continues, 2 once it has ended.
__device__ __forceinline__ uint32_t step_closest(const DevScene& sc, uint32_t* __restrict__ stack, uint32_t stride, f3 o, f3 d,
                                                 const SlabRay& sr, float tmax, int32_t& node, uint32_t& sp, Tally& tl,
                                                 uint32_t alpha, bool& have, Hit& hit, PendingLeaves& q) {
  while (q.i == q.end && q.mask) {
    int kb = 0;
    float tb = INFINITY;
    bool any = false;
#pragma unroll
    for (int k = 0; k < kBvhWidth; ++k)
      if (((q.mask >> k) & 1u) && (!any || q.tn[k] < tb)) {
        kb = k;
        tb = q.tn[k];
        any = true;
      }
    q.mask &= ~(1u << kb);
    if (have && tb > hit.t * (1.0f + kBoxEps)) continue;
    const uint32_t v = (uint32_t)(-1 - kid(q.ch, kb));
    q.i = v >> 2;
    q.end = q.i + (v & 3u) + 1u;
  }
  if (q.i < q.end) {
    const uint32_t i = q.i++;
    const float4 a = sc.tris[3 * i], b = sc.tris[3 * i + 1], cc = sc.tris[3 * i + 2];
    float t, u, vv;
    tl.tris++;
    if (mt_hit(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(cc.x, cc.y, cc.z), o, d, tmax, &t, &u, &vv) && finite_(t) &&
        (!alpha || alpha_blocks(sc, __float_as_uint(a.w), u, vv))) {
      const uint32_t occ = __float_as_uint(a.w);
      if (!have || t < hit.t || (t == hit.t && occ < hit.occ)) {
        have = true;
        hit.t = t;
        hit.u = u;
        hit.v = vv;
        hit.occ = occ;
        hit.pos = i;
      }
    }
    return 0u;
  }
  if (node < 0) {
    if (sp == 0) return 2u;
    sp--;
    node = pop_node(sc, stack, stride, sp);
  }
  tl.nodes++;
  const uint32_t mask = node_hits(sc, node, sr, (have ? hit.t : tmax) * (1.0f + kBoxEps), q.tn, &q.ch);
  q.mask = mask & leaf_bits(q.ch);
  node = push_children_sorted(sc, mask, q.ch, q.tn, stack, stride, sp);
  return 0u;
}

// Per-job constants of a point-light gather (bounce.rs:247-262).
struct GatherJob {
  f3 origin, normal, dir0, lp;
  float step;
  uint32_t samples;
  const uint32_t* cube;
  uint64_t tile0;  // the cube's first photon tile
  float o_scale;   // |origin.x| + |origin.y| + |origin.z|
};

__device__ __forceinline__ GatherJob gather_job(const DevTris& tr, const DevMeshes& ms, const DevLight* __restrict__ lights,
                                                const DevTexels& tx, const DevEntries& ent, const DevBounce& b,
                                                const BounceParams& p, uint32_t j, uint32_t e, Pcg* rng, uint32_t* mesh) {
  // rng: the job's sample stream (null when the caller brings the draws)
  const uint32_t t = ent.tri[e], slot = ent.light[e];
  const DevLight& L = lights[slot];
  GatherJob g;
  float4 p4 = tx.pos[j];
  f3 pos = mk3(p4.x, p4.y, p4.z);
  g.normal = ld3(tr.face + 3 * (size_t)t);
  g.lp = ld3(L.pos);
  g.samples = L.bounce_samples > 2 ? L.bounce_samples : 2u;
  g.step = 0.9f / (float)(g.samples - 1);
  g.dir0 = normalize_or(sub(pos, g.lp), mk3(0.0f, 1.0f, 0.0f));
  g.origin = add(pos, mul(g.normal, kRayOffset));
  g.cube = p.photons + (size_t)b.cube_of_slot[slot] * 6 * p.cube_res * p.cube_res * 4;
  g.tile0 = (uint64_t)b.cube_of_slot[slot] * 6 * p.tiles_per_face * p.tiles_per_face;
  g.o_scale = fabsf(g.origin.x) + fabsf(g.origin.y) + fabsf(g.origin.z);
  const uint32_t m = tr.mesh[t];
  if (rng) {
    const uint32_t xy = tx.xy[j];
    *rng = stream5(p.seed, ms.id[m], L.id, xy & 0xffffu, xy >> 16);
  }
  *mesh = m;
  return g;
}

// ---- batched sample generation ------------------------------------------------------------
// The exception lookups of N values of one function (see trig_finish), with the N lookups'
// loads issued together: the run words first, then the results of the actual exceptions.
template <int N>
__device__ __forceinline__ void trig_finish_n(const TrigTables& t, const TrigTable& tab, const float* x, const double* v, float* out) {
  bool need[N];
  uint2 run[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    out[i] = (float)v[i];
    need[i] = trig_in_domain(f2u(x[i])) && trig_near_midpoint(v[i], t.tol);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {  // runs without an exception are cleared by the cached bitmap
    if (need[i]) {
      const uint32_t di = trig_domain_index(f2u(x[i]));
      need[i] = ((tab.any[di >> 10] >> ((di >> 5) & 31u)) & 1u) != 0u;
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) run[i] = need[i] ? tab.runs[trig_domain_index(f2u(x[i])) >> 5] : make_uint2(0u, 0u);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t bit = trig_domain_index(f2u(x[i])) & 31u;
    if (need[i] && ((run[i].x >> bit) & 1u))
      out[i] = u2f(tab.vals[run[i].y + (uint32_t)__builtin_popcount(run[i].x & ((1u << bit) - 1u))]);
  }
}

// The random draws of on_unit_sphere (random.rs:22-28) for samples k0 .. k0 + G - 1 (those below
// g.samples; the others get 0): z in [-1, 1) and the azimuth, in sample order.
template <int G>
__device__ __forceinline__ void gather_draws(const GatherJob& g, Pcg& rng, uint32_t k0, float* z, float* ang) {
#pragma unroll
  for (int j = 0; j < G; ++j) {
    z[j] = 0.0f;
    ang[j] = 0.0f;
    if (k0 + j < g.samples) {
      z[j] = pcg_f32(rng) * 2.0f - 1.0f;
      ang[j] = pcg_f32(rng) * kTau;
    }
  }
}

// The exact sample directions of G samples from their draws (random.rs:22-28, bounce.rs:262-270:
// on_unit_sphere, glam's slerp from the direction to the texel, normalize): sample j is sample
// k[j] of job *gj[j] (la[j] = |gj[j]->dir0|).
template <int G>
__device__ __forceinline__ void exact_directions_of(const GatherJob* const* gj, const TrigTables& trig, const float* z,
                                                    const float* ang, const uint32_t* k, const float* la, f3* sample) {
  // on_unit_sphere: sin/cos of the azimuths
  double vs[G], vc[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    int q;
    const double r = trig_reduce(ang[j], &q);
    vs[j] = trig_select(r, q, 0);
    vc[j] = trig_select(r, q, 1);
  }
  float sa[G], ca[G];
  trig_finish_n<G>(trig, trig.sin, ang, vs, sa);
  trig_finish_n<G>(trig, trig.cos, ang, vc, ca);
  // slerp (glam Vec3::slerp) up to its sines
  f3 rnd[G];
  float lb[G], sp[G], arg[3 * G];
  bool common[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const f3 a = gj[j]->dir0;
    const float r = sqrtf(maxf_(1.0f - z[j] * z[j], 0.0f));
    rnd[j] = mk3(r * ca[j], r * sa[j], z[j]);
    sp[j] = 0.1f + (float)k[j] * gj[j]->step;
    lb[j] = len(rnd[j]);
    const float dd = dot(a, rnd[j]) / (la[j] * lb[j]);
    common[j] = fabsf(dd) < 1.0f - 3.0e-7f;
    const float theta = common[j] ? acos_approx(dd) : 0.0f;
    arg[3 * j] = theta;
    arg[3 * j + 1] = theta * (1.0f - sp[j]);
    arg[3 * j + 2] = theta * sp[j];
  }
  double va[3 * G];
#pragma unroll
  for (int i = 0; i < 3 * G; ++i) va[i] = trig_value(arg[i], 0);
  float sv[3 * G];
  trig_finish_n<3 * G>(trig, trig.sin, arg, va, sv);
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const f3 a = gj[j]->dir0;
    f3 sl;
    if (common[j]) {
      const float rl = la[j] + (lb[j] - la[j]) * sp[j];
      sl = mul(add(mul(mul(a, rl / la[j]), sv[3 * j + 1]), mul(mul(rnd[j], rl / lb[j]), sv[3 * j + 2])), 1.0f / sv[3 * j]);
    } else {
      sl = slerp3(trig, a, rnd[j], sp[j]);
    }
    sample[j] = normalize_or(sl, a);
  }
}
// Samples k0 .. k0 + G - 1 of one job (la = |g.dir0|).
template <int G>
__device__ __forceinline__ void exact_directions(const GatherJob& g, const TrigTables& trig, const float* z, const float* ang,
                                                 uint32_t k0, float la, f3* sample) {
  const GatherJob* gj[G];
  uint32_t k[G];
  float las[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    gj[j] = &g;
    k[j] = k0 + (uint32_t)j;
    las[j] = la;
  }
  exact_directions_of<G>(gj, trig, z, ang, k, las, sample);
}

// Whether sample k of a job casts a ray (bounce.rs:262-286: its photon has a lit surface and a
// positive distance), reading the photon only when its tile leaves it open: an unlit texel casts
// none, a lit one whose tile box excludes the origin casts one.
__device__ __forceinline__ bool gather_sample_casts(const GatherJob& g, const TrigTables& trig, Pcg& rng, uint32_t k,
                                                    const BounceParams& p) {
  float z, ang;
  gather_draws<1>(g, rng, k, &z, &ang);
  f3 sample;
  exact_directions<1>(g, trig, &z, &ang, k, len(g.dir0), &sample);
  uint32_t face, x, y;
  cube_texel_parts(sample, p.cube_res, &face, &x, &y);
  const uint32_t ci = face * p.cube_res * p.cube_res + y * p.cube_res + x;
  if (p.tile_live) {
    uint32_t bit;
    const uint64_t tile = g.tile0 + photon_tile(face, x, y, p.tiles_per_face, &bit);
    if (!tile_bit(p.tile_live, tile, bit)) return false;
    if (tile_outside(p.tile_box[3 * tile + 1], p.tile_box[3 * tile + 2], g.origin)) return true;
  }
  const uint4 ph = reinterpret_cast<const uint4*>(g.cube)[ci];
  f3 pn = normalize_or_zero(mk3((float)(int8_t)(ph.y & 0xffu) / 127.0f, (float)(int8_t)((ph.y >> 8) & 0xffu) / 127.0f,
                                (float)(int8_t)((ph.y >> 16) & 0xffu) / 127.0f));
  f3 photon = add(g.lp, mul(sample, __uint_as_float(ph.x)));
  return !is_zero(pn) && len(sub(photon, g.origin)) > 0.0f;
}

// G consecutive samples k0 .. k0 + G - 1 of a job (those below g.samples) from their draws, each
// as in bounce.rs:262-286 (spread, random direction, slerp, photon, distance, weight): the sin/cos
// searches, the photon tile reads and the photon fetches of the G samples issued together.  A
// sample whose tile shows that it casts no ray, or a ray that brings no light, skips its photon
// fetch; its outputs are those the photon would have given (ray false, or ray true with a zero
// contribution, which is never traced).  la = |g.dir0|.
template <int G>
__device__ __forceinline__ void gather_exact_of(const GatherJob* const* gj, const TrigTables& trig, const float* z,
                                                const float* ang, const uint32_t* k, const float* la, const BounceParams& p,
                                                f3* dout, float* dist, f3* cout, bool* ray, uint32_t* reads) {
  const uint32_t res = p.cube_res, colored = p.colored;
  // sample direction, cube texel, photon
  uint32_t ci[G], tile[G], bit[G];
  f3 sample[G];
  exact_directions_of<G>(gj, trig, z, ang, k, la, sample);
#pragma unroll
  for (int j = 0; j < G; ++j) {
    uint32_t face, x, y;
    cube_texel_parts(sample[j], res, &face, &x, &y);
    ci[j] = face * res * res + y * res + x;
    tile[j] = photon_tile(face, x, y, p.tiles_per_face, &bit[j]);
  }
  // what the photon tiles decide without the photon: 0 read the photon, 1 unlit texel (no ray),
  // 2 a ray that brings no light
  uint32_t known[G];
#pragma unroll
  for (int j = 0; j < G; ++j) known[j] = 0u;
  if (p.tile_live) {
    bool lw[G];
    float4 b0[G], b1[G], b2[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const uint64_t t = gj[j]->tile0 + tile[j];
      lw[j] = tile_bit(p.tile_live, t, bit[j]);
      b0[j] = p.tile_box[3 * t];
      b1[j] = p.tile_box[3 * t + 1];
      b2[j] = p.tile_box[3 * t + 2];
    }
#pragma unroll
    for (int j = 0; j < G; ++j)
      known[j] = !lw[j] ? 1u : tile_dark(b0[j], b1[j], b2[j], gj[j]->origin, gj[j]->normal, gj[j]->o_scale) ? 2u : 0u;
  }
  uint4 ph[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const bool read = k[j] < gj[j]->samples && known[j] == 0u;
    ph[j] = read ? reinterpret_cast<const uint4*>(gj[j]->cube)[ci[j]] : make_uint4(0u, 0u, 0u, 0u);
    *reads += read ? 1u : 0u;
  }
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const GatherJob& g = *gj[j];
    if (known[j] != 0u) {
      ray[j] = k[j] < g.samples && known[j] == 2u;
      dout[j] = mk3(0.0f, 0.0f, 0.0f);
      cout[j] = mk3(0.0f, 0.0f, 0.0f);
      dist[j] = 0.0f;
      continue;
    }
    const uint32_t w1 = ph[j].y, w2 = ph[j].z;
    const f3 pn = normalize_or_zero(mk3((float)(int8_t)(w1 & 0xffu) / 127.0f, (float)(int8_t)((w1 >> 8) & 0xffu) / 127.0f,
                                        (float)(int8_t)((w1 >> 16) & 0xffu) / 127.0f));
    const f3 photon = add(g.lp, mul(sample[j], __uint_as_float(ph[j].x)));
    const f3 to_photon = sub(photon, g.origin);
    const float dl = len(to_photon);
    ray[j] = false;
    dout[j] = mk3(0.0f, 0.0f, 0.0f);
    cout[j] = mk3(0.0f, 0.0f, 0.0f);
    dist[j] = 0.0f;
    if (k[j] >= g.samples || is_zero(pn) || dl <= 0.0f) continue;
    const f3 dir = divs(to_photon, dl);
    const float weight = maxf_(dot(g.normal, dir), 0.0f) * maxf_(dot(pn, neg(dir)), 0.0f);
    const f3 alb = colored ? mk3((float)(w2 & 0xffu) / 255.0f, (float)((w2 >> 8) & 0xffu) / 255.0f,
                                 (float)((w2 >> 16) & 0xffu) / 255.0f)
                           : mk3(1.0f, 1.0f, 1.0f);
    ray[j] = true;
    dout[j] = dir;
    dist[j] = dl;
    cout[j] = mul(alb, weight);
  }
}
// Samples k0 .. k0 + G - 1 of one job (la = |g.dir0|).
template <int G>
__device__ __forceinline__ void gather_exact(const GatherJob& g, const TrigTables& trig, const float* z, const float* ang,
                                             uint32_t k0, float la, const BounceParams& p, f3* dout, float* dist, f3* cout,
                                             bool* ray, uint32_t* reads) {
  const GatherJob* gj[G];
  uint32_t k[G];
  float las[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    gj[j] = &g;
    k[j] = k0 + (uint32_t)j;
    las[j] = la;
  }
  gather_exact_of<G>(gj, trig, z, ang, k, las, p, dout, dist, cout, ray, reads);
}
// Draws and evaluates samples k0 .. k0 + G - 1 in order.
template <int G>
__device__ __forceinline__ void gather_group(const GatherJob& g, const TrigTables& trig, Pcg& rng, uint32_t k0, float la,
                                             const BounceParams& p, f3* dout, float* dist, f3* cout, bool* ray,
                                             uint32_t* reads) {
  float z[G], ang[G];
  gather_draws<G>(g, rng, k0, z, ang);
  gather_exact<G>(g, trig, z, ang, k0, la, p, dout, dist, cout, ray, reads);
}

// Job lists: pass 0 counts, pass 1 writes.  counts[0..1]: traced / count-only jobs (pass 0),
// cursors[0..1] (pass 1).
// Pass 0 writes each bounce entry's traced and count-only job counts (cnt[q], cnt[n + 1 + q]);
// pass 1 writes the jobs at the scanned offsets (off[q], off[n + 1 + q]), in entry order and
// texel order within an entry.
__global__ void k_gather_jobs(DevTris tr, const DevLight* __restrict__ lights, DevTexels tx, const uint64_t* __restrict__ tex_start,
                              DevEntries ent, DevBounce b, int pass, uint64_t* cnt, const uint64_t* __restrict__ off,
                              uint2* trace_jobs, uint2* count_jobs, uint2* sun_jobs) {
  // one wave per bounce entry, 64 of its texels per trip
  const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
  const uint64_t n = b.n;
  for (uint64_t q = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); q < n; q += (uint64_t)gridDim.x * waves) {
    const uint64_t e = b.entry[q];
    const uint32_t t = ent.tri[e];
    const DevLight& L = lights[ent.light[e]];
    uint64_t nt_traced = 0, nt_counted = 0, nt_sun = 0;
    uint64_t base_t = pass == 1 ? off[q] : 0, base_c = pass == 1 ? off[n + 1 + q] : 0;
    const uint64_t t0 = tex_start[t], nt = tex_start[t + 1] - t0;
    if (L.shape == 3) {  // a sun gathers at every texel of the triangle
      nt_sun = nt;
      if (pass == 1)
        for (uint64_t c = lane; c < nt; c += 64) sun_jobs[off[2 * (n + 1) + q] + c] = make_uint2((uint32_t)(t0 + c), (uint32_t)e);
    } else {
      const f3 lp = ld3(L.pos);
      const float r2 = L.radius * L.radius;
      const uint32_t w = b.ent_window[e];
      const uint32_t* wr = b.win_rect + 4 * (size_t)w;
      const uint32_t wx0 = wr[0], wy0 = wr[1], ww = wr[2] - wr[0] + 1;
      const unsigned int* owner = reinterpret_cast<const unsigned int*>(b.map_b) + b.win_off[w] * 3;
      for (uint64_t c = 0; c < nt; c += 64) {
        const uint64_t j = t0 + c + lane;
        bool traced = false, counted = false;
        if (c + lane < nt) {
          const float4 p4 = tx.pos[j];
          if (len_sq(sub(mk3(p4.x, p4.y, p4.z), lp)) <= r2) {
            const uint32_t xy = tx.xy[j];
            const uint64_t k = (uint64_t)((xy & 0xffffu) - wx0) + (uint64_t)((xy >> 16) - wy0) * ww;
            traced = owner[k] == t + 1;
            counted = !traced;
          }
        }
        const uint64_t mt = __ballot(traced), mc = __ballot(counted);
        if (pass == 1) {
          const uint32_t rt = __builtin_amdgcn_mbcnt_hi((uint32_t)(mt >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mt, 0u));
          const uint32_t rc = __builtin_amdgcn_mbcnt_hi((uint32_t)(mc >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mc, 0u));
          if (traced) trace_jobs[base_t + rt] = make_uint2((uint32_t)j, (uint32_t)e);
          if (counted) count_jobs[base_c + rc] = make_uint2((uint32_t)j, (uint32_t)e);
          base_t += (uint64_t)__popcll(mt);
          base_c += (uint64_t)__popcll(mc);
        }
        nt_traced += (uint64_t)__popcll(mt);
        nt_counted += (uint64_t)__popcll(mc);
      }
    }
    if (pass == 0 && lane == 0) {
      cnt[q] = nt_traced;
      cnt[n + 1 + q] = nt_counted;
      cnt[2 * (n + 1) + q] = nt_sun;
    }
  }
  if (pass == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    cnt[n] = 0;
    cnt[2 * n + 1] = 0;
    cnt[3 * n + 2] = 0;
  }
}

// Rays of the jobs whose result a later triangle overwrites: the samples are set up (they decide
// which samples cast a ray) but nothing is traced.
__global__ void k_gather_count(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                               DevBounce b, const uint2* __restrict__ jobs, uint64_t njobs, BounceParams p, TrigTables trig) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < njobs; i0 += stride) {
    const uint64_t i = i0 + lane;
    uint32_t m = 0, rays = 0, win = 0;
    if (i < njobs) {
      const uint2 jb = jobs[i];
      Pcg rng;
      GatherJob g = gather_job(tr, ms, lights, tx, ent, b, p, jb.x, jb.y, &rng, &m);
      for (uint32_t k = 0; k < g.samples; ++k) rays += gather_sample_casts(g, trig, rng, k, p) ? 1u : 0u;
      win = b.ent_window[jb.y];
    }
    wave_add_at(b.rays, m, rays, i < njobs);
    if (b.wrays) wave_add_at(b.wrays, win, rays, i < njobs);
  }
}

// Where sample k of job i of a chunk of n jobs keeps its ray and contribution: job-major (order 0:
// the job's samples together, from rs[i]) or sample-major (order 1: sample k of every job
// together, so a wave of jobs writes and reads each sample's slots coalesced).  The any-hit kernel
// hands out a job's samples together in either order: they share an origin, so a wave's lanes
// walk the same part of the BVH.
__device__ __forceinline__ uint64_t ray_slot(uint32_t order, uint64_t job_base, uint64_t i, uint64_t n, uint32_t k) {
  return order ? (uint64_t)k * n + i : job_base + k;
}

// Ray offsets of a chunk of jobs: samples per job (n + 1 entries, the last one 0, for a scan).
__global__ void k_gather_nrays(DevEntries ent, const DevLight* __restrict__ lights, const uint2* __restrict__ jobs, uint64_t n,
                               uint64_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (i == n) {
      out[i] = 0;
      continue;
    }
    const uint32_t bs = lights[ent.light[jobs[i].y]].bounce_samples;
    out[i] = bs > 2 ? bs : 2u;
  }
}

// Phase 1 of the point-light gather: every sample of every job in order, written as a ray
// (origin | length, direction | flags: bit 0 alpha test, bit 1 casts a ray) and the light it
// brings if unoccluded (zero when it casts none).  Also counts each mesh's bounce rays.
__global__ void k_gather_rays(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* __restrict__ jobs, uint64_t n, const uint64_t* __restrict__ rs,
                              BounceParams p, TrigTables trig, float4* rays, float* contrib, uint32_t order, uint32_t* tmask,
                              uint32_t words, uint64_t* tcount) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) tcount[n] = 0;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
    const uint64_t i = i0 + lane;
    uint32_t m = 0xffffffffu, nr = 0, win = 0, reads = 0;
    if (i < n) {
      const uint2 jb = jobs[i];
      win = b.ent_window[jb.y];
      Pcg rng;
      GatherJob g = gather_job(tr, ms, lights, tx, ent, b, p, jb.x, jb.y, &rng, &m);
      const uint32_t alpha = lights[ent.light[jb.y]].alpha ? 1u : 0u;
      const uint64_t base = rs[i];
      const float la = len(g.dir0);
      constexpr int G = kGatherGroup;
      uint32_t word = 0, traced = 0;  // bits of the traced samples, 32 per word of tmask
      for (uint32_t k0 = 0; k0 < g.samples; k0 += G) {
        f3 d[G], c[G];
        float dist[G];
        bool ray[G];
        gather_group<G>(g, trig, rng, k0, la, p, d, dist, c, ray, &reads);
#pragma unroll
        for (int j = 0; j < G; ++j) {
          if (k0 + j >= g.samples) continue;
          nr += ray[j] ? 1u : 0u;
          // a ray whose contribution is exactly zero adds +0 to the sum whether or not it is
          // occluded: it counts as cast but is not traced (NaN contributions are traced)
          // Only traced samples keep a ray and a contribution: the others add +0 to the sum,
          // which k_gather_sum skips by the same bits.
          const bool trace = ray[j] && !(c[j].x == 0.0f && c[j].y == 0.0f && c[j].z == 0.0f);
          const uint32_t k = k0 + j;
          if (trace) {
            const uint64_t r = ray_slot(order, base, i, n, k);
            rays[2 * r] = make_float4(g.origin.x, g.origin.y, g.origin.z, dist[j]);
            rays[2 * r + 1] = make_float4(d[j].x, d[j].y, d[j].z, __uint_as_float(alpha | 2u));
            contrib[3 * r] = c[j].x;
            contrib[3 * r + 1] = c[j].y;
            contrib[3 * r + 2] = c[j].z;
            word |= 1u << (k & 31u);
            traced++;
          }
          if ((k & 31u) == 31u) {
            tmask[i * words + (k >> 5)] = word;
            word = 0;
          }
        }
      }
      for (uint32_t w = g.samples >> 5; w < words; ++w) tmask[i * words + w] = w == (g.samples >> 5) ? word : 0u;
      tcount[i] = traced;
    }
    wave_add_at(b.rays, m, nr, i < n);
    if (b.wrays) wave_add_at(b.wrays, win, nr, i < n);
    if (p.photon_reads) wave_add(p.photon_reads, reads);
  }
}

// ---- split sample setup: a fast pass over every sample, the exact chain for the few it leaves --
// Phase A (k_gather_split) follows each sample with fast f32 arithmetic (hardware sin, cos,
// reciprocal and square roots): where that direction's cube texel is certain (away from the face
// and tile edges by more than its error can move it) and the photon tiles decide the sample (no
// lit photon in any texel it can fall in, or a ray that brings no light), the sample needs nothing
// more.  Every other sample is open: phase A keeps its random draws in its ray slot, and phase B
// (k_gather_open) evaluates it with the exact chain (gather_exact), so every result is that of
// gather_group.
//
// kSplitDirErr bounds the difference of each component of the fast unit direction from the
// exact one wherever the cosine of the slerp angle lies in [-kSplitCos, kSplitCosNear] (samples
// beyond are open: near -1 the slerp's 1 / sin(angle) amplifies the fast sines' errors; near +1
// the slerp tends to the lerp and stays well conditioned, up to a margin from glam's lerp branch
// at 1 - 3e-7).  The GPU test of dl_test_gather_directions measures that difference on 16M
// samples drawn as the bake draws them and requires it below a tenth of the bound (about 6e-6 is
// measured, 5e-7 for cosines above 0.99).
static constexpr float kSplitCos = 0.99f;
static constexpr float kSplitCosNear = 0.9999f;
static constexpr float kSplitDirErr = 1.0e-4f;

// glam's slerp and normalize of gather_exact in fast arithmetic; false outside
// [-kSplitCos, kSplitCosNear].
__device__ __forceinline__ bool fast_direction(f3 a, float la, float inv_la, float z, float ang, float sp, f3* out) {
  float sa, ca;
  __sincosf(ang, &sa, &ca);
  const float r = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-z, z, 1.0f), 0.0f));
  const f3 rnd = mk3(r * ca, r * sa, z);
  const float lb2 = fmaf(rnd.x, rnd.x, fmaf(rnd.y, rnd.y, rnd.z * rnd.z));
  const float ilb = __builtin_amdgcn_rsqf(lb2);
  const float lb = lb2 * ilb;
  const float dd = fmaf(a.x, rnd.x, fmaf(a.y, rnd.y, a.z * rnd.z)) * (inv_la * ilb);
  if (!(dd <= kSplitCosNear && dd >= -kSplitCos)) return false;
  const float x = fabsf(dd);
  float t = fmaf(fmaf(fmaf(fmaf(fmaf(fmaf(fmaf(-0.0012624911f, x, 0.0066700901f), x, -0.0170881256f), x, 0.0308918810f), x,
                                -0.0501743046f), x, 0.0889789874f), x, -0.2145988016f), x, 1.5707963050f);
  t *= __builtin_amdgcn_sqrtf(1.0f - x);
  const float theta = dd >= 0.0f ? t : kPi - t;
  const float inv = __builtin_amdgcn_rcpf(__sinf(theta));
  const float rl = fmaf(lb - la, sp, la);
  const float wa = rl * inv_la * __sinf(theta * (1.0f - sp)) * inv, wb = rl * ilb * __sinf(theta * sp) * inv;
  const f3 sl = mk3(fmaf(a.x, wa, rnd.x * wb), fmaf(a.y, wa, rnd.y * wb), fmaf(a.z, wa, rnd.z * wb));
  const float il = __builtin_amdgcn_rsqf(fmaf(sl.x, sl.x, fmaf(sl.y, sl.y, sl.z * sl.z)));
  *out = mk3(sl.x * il, sl.y * il, sl.z * il);
  return true;
}

// The face of unit direction d (cube_texel_parts' choice) and its continuous texel coordinates
// (u * res, v * res); *gap: how far d is from choosing another face (the largest absolute
// component minus the next one).
__device__ __forceinline__ void fast_face(f3 d, float res, uint32_t* face, float* tu, float* tv, float* gap) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  float sc, tc, ma;
  if (ax > ay && ax > az) {
    if (d.x > 0.0f) { *face = 0; sc = -d.z; tc = -d.y; } else { *face = 1; sc = d.z; tc = -d.y; }
    ma = ax;
    *gap = ax - fmaxf(ay, az);
  } else if (ay > az) {
    if (d.y > 0.0f) { *face = 2; sc = d.x; tc = d.z; } else { *face = 3; sc = d.x; tc = -d.z; }
    ma = ay;
    *gap = ay - fmaxf(ax, az);
  } else {
    if (d.z > 0.0f) { *face = 4; sc = d.x; tc = -d.y; } else { *face = 5; sc = -d.x; tc = -d.y; }
    ma = az;
    *gap = az - fmaxf(ax, ay);
  }
  const float im = __builtin_amdgcn_rcpf(ma);
  *tu = fmaf(sc * im, 0.5f, 0.5f) * res;
  *tv = fmaf(tc * im, 0.5f, 0.5f) * res;
}

// The texels a sample's exact direction can fall in, from its fast direction: {x0, x1} x {y0, y1}
// of one face (their coordinates widened by the margin, clamped as cube_texel_parts clamps), all
// in the summary of the tile holding (x0, y0).
struct FastCand {
  bool sure;      // the face is certain and the candidates span at most 2x2 texels (else the sample is open)
  uint64_t tile;  // the tile whose summary holds every candidate
  uint32_t bit[4];
  uint32_t face, x0, x1, y0, y1;
};
__device__ __forceinline__ FastCand fast_candidates(const GatherJob& g, const BounceParams& p, f3 d, float fres, float margin) {
  FastCand c;
  uint32_t face;
  float tu, tv, gap;
  fast_face(d, fres, &face, &tu, &tv, &gap);
  const int32_t hi = (int32_t)p.cube_res - 1;
  const uint32_t x0 = (uint32_t)min(max((int32_t)floorf(tu - margin), 0), hi), x1 = (uint32_t)min(max((int32_t)floorf(tu + margin), 0), hi);
  const uint32_t y0 = (uint32_t)min(max((int32_t)floorf(tv - margin), 0), hi), y1 = (uint32_t)min(max((int32_t)floorf(tv + margin), 0), hi);
  // (at cube resolutions where the margin reaches half a texel the span can exceed 2x2: open)
  c.sure = gap > 2.5f * kSplitDirErr && x1 <= x0 + 1u && y1 <= y0 + 1u;
  const uint32_t tpf = p.tiles_per_face, tx = x0 / kPhotonTile, ty = y0 / kPhotonTile;
  c.tile = g.tile0 + (uint64_t)(face * tpf * tpf + ty * tpf + tx);
  const uint32_t dx0 = x0 - tx * kPhotonTile, dx1 = x1 - tx * kPhotonTile;
  const uint32_t dy0 = y0 - ty * kPhotonTile, dy1 = y1 - ty * kPhotonTile;
  c.bit[0] = dy0 * kTileSpan + dx0;
  c.bit[1] = dy0 * kTileSpan + dx1;
  c.bit[2] = dy1 * kTileSpan + dx0;
  c.bit[3] = dy1 * kTileSpan + dx1;
  c.face = face;
  c.x0 = x0;
  c.x1 = x1;
  c.y0 = y0;
  c.y1 = y1;
  return c;
}
// Whether the tile decides a sample whose exact texel is one of its candidates: 1 no ray (none of
// them holds a lit photon), 2 a ray that brings no light (all of them hold one and the tile is
// dark for this receiver), 0 open.
__device__ __forceinline__ uint32_t fast_decide(const GatherJob& g, const BounceParams& p, const FastCand& c) {
  if (!c.sure) return 0u;
  const uint4 w = reinterpret_cast<const uint4*>(p.tile_live)[c.tile];
  const unsigned long long w0 = (unsigned long long)w.x | ((unsigned long long)w.y << 32);
  const unsigned long long w1 = (unsigned long long)w.z | ((unsigned long long)w.w << 32);
  uint32_t lit = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t b = c.bit[k];
    lit += (uint32_t)(((b < 64u ? w0 : w1) >> (b & 63u)) & 1ull);
  }
  if (lit == 0u) return 1u;
  if (lit != 4u) return 0u;
  return tile_dark(p.tile_box[3 * c.tile], p.tile_box[3 * c.tile + 1], p.tile_box[3 * c.tile + 2], g.origin, g.normal, g.o_scale)
             ? 2u
             : 0u;
}

// Phase A: per job its samples in order; open samples keep (z, azimuth) in their ray slot and set
// their bit of omask (ocount: open samples per job); tmask and tcount start at zero for phase B.
// Counts the rays of the samples it decides.  Needs the photon tiles (p.tile_live).
__global__ void k_gather_split(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                               DevBounce b, const uint2* __restrict__ jobs, uint64_t n, const uint64_t* __restrict__ rs,
                               BounceParams p, float4* rays, uint32_t order, uint32_t* omask, uint32_t* tmask, uint32_t words,
                               uint64_t* ocount, uint64_t* tcount) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const float fres = (float)p.cube_res;
  const float margin = 3.0f * kSplitDirErr * fres + 1.0e-3f;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ocount[n] = 0;
    tcount[n] = 0;
  }
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
    const uint64_t i = i0 + lane;
    uint32_t m = 0xffffffffu, nr = 0, win = 0;
    if (i < n) {
      const uint2 jb = jobs[i];
      win = b.ent_window[jb.y];
      Pcg rng;
      const GatherJob g = gather_job(tr, ms, lights, tx, ent, b, p, jb.x, jb.y, &rng, &m);
      const uint64_t base = rs[i];
      const float la = len(g.dir0), inv_la = 1.0f / la;
      uint32_t word = 0, open = 0;
      // two samples per trip: their directions, candidate texels and tile reads side by side
      for (uint32_t k0 = 0; k0 < g.samples; k0 += 2) {
        float z[2], ang[2];
        gather_draws<2>(g, rng, k0, z, ang);
        FastCand c[2];
        bool fast[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          f3 d = mk3(0.0f, 0.0f, 1.0f);
          fast[j] = fast_direction(g.dir0, la, inv_la, z[j], ang[j], 0.1f + (float)(k0 + j) * g.step, &d);
          c[j] = fast_candidates(g, p, d, fres, margin);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const uint32_t k = k0 + j;
          if (k >= g.samples) break;
          const uint32_t v = fast[j] ? fast_decide(g, p, c[j]) : 0u;
          nr += v == 2u ? 1u : 0u;
          if (v == 0u) {
            rays[2 * ray_slot(order, base, i, n, k)] = make_float4(z[j], ang[j], 0.0f, 0.0f);
            word |= 1u << (k & 31u);
            open++;
          }
          if ((k & 31u) == 31u) {
            omask[i * words + (k >> 5)] = word;
            tmask[i * words + (k >> 5)] = 0u;
            word = 0;
          }
        }
      }
      for (uint32_t w = g.samples >> 5; w < words; ++w) {
        omask[i * words + w] = w == (g.samples >> 5) ? word : 0u;
        tmask[i * words + w] = 0u;
      }
      ocount[i] = open;
      tcount[i] = 0;
    }
    wave_add_at(b.rays, m, nr, i < n);
    if (b.wrays) wave_add_at(b.wrays, win, nr, i < n);
  }
}

// Lazy photon cubes: the texels the gather of a list of n jobs can read (traced jobs: k_gather_split
// and k_gather_open; count-only jobs: k_gather_count).  Per sample, when its fast direction's
// candidate texels are sure, all of them (its exact texel is one of them, and phase A asks the
// tiles about no other); else its exact texel (the only one phase B and the count read).  Sets
// need[64 * tile + bit] = 1 for each, tile and bit as in the photon tiles; p.tiles_per_face must be
// set, the tiles themselves need not exist yet.
__global__ void k_photon_mark(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* __restrict__ jobs, uint64_t n, BounceParams p, TrigTables trig,
                              uint8_t* need) {
  const float fres = (float)p.cube_res;
  const float margin = 3.0f * kSplitDirErr * fres + 1.0e-3f;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 jb = jobs[i];
    Pcg rng;
    uint32_t m;
    const GatherJob g = gather_job(tr, ms, lights, tx, ent, b, p, jb.x, jb.y, &rng, &m);
    const float la = len(g.dir0), inv_la = 1.0f / la;
    for (uint32_t k = 0; k < g.samples; ++k) {
      float z, ang;
      gather_draws<1>(g, rng, k, &z, &ang);
      f3 d = mk3(0.0f, 0.0f, 1.0f);
      const bool fast = fast_direction(g.dir0, la, inv_la, z, ang, 0.1f + (float)k * g.step, &d);
      const FastCand c = fast_candidates(g, p, d, fres, margin);
      if (fast && c.sure) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          uint32_t bit;
          const uint32_t t = photon_tile8(c.face, (j & 1) ? c.x1 : c.x0, (j & 2) ? c.y1 : c.y0, p.tiles_per_face, &bit);
          need[64 * (g.tile0 + t) + bit] = 1;
        }
      } else {
        f3 sample;
        exact_directions<1>(g, trig, &z, &ang, k, la, &sample);
        uint32_t face, x, y, bit;
        cube_texel_parts(sample, p.cube_res, &face, &x, &y);
        need[64 * (g.tile0 + photon_tile8(face, x, y, p.tiles_per_face, &bit)) + bit] = 1;
      }
    }
  }
}

// Phase B: the open samples (slots list[0 .. *count)), each with the exact chain from its draws;
// traced ones write their ray and contribution, set their bit of tmask and add to tcount.
__global__ void k_gather_open(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* __restrict__ jobs, uint64_t n, const uint64_t* __restrict__ rs,
                              BounceParams p, TrigTables trig, float4* rays, float* contrib, uint32_t order,
                              const uint32_t* __restrict__ list, const uint64_t* __restrict__ count, uint32_t* tmask,
                              uint32_t words, unsigned long long* tcount) {
  // two list entries per lane (2q and 2q + 1, often samples of one job): their exact chains run
  // side by side, so that each one's dependent loads overlap the other's
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x, total = *count, pairs = (total + 1) / 2;
  for (uint64_t q0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); q0 < pairs; q0 += stride) {
    const uint64_t q = q0 + lane;
    uint32_t m[2] = {0xffffffffu, 0xffffffffu}, nr[2] = {0u, 0u}, win[2] = {0u, 0u}, reads = 0;
    bool valid[2] = {false, false};
    if (q < pairs) {
      uint32_t r[2], k[2];
      uint64_t i[2];
      uint2 jb[2];
      GatherJob g[2];
      float z[2], ang[2], la[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        valid[h] = 2 * q + h < total;
        r[h] = list[valid[h] ? 2 * q + h : 2 * q];
        if (order) {  // slot k * n + i (slots and n are below 2^32)
          k[h] = r[h] / (uint32_t)n;
          i[h] = r[h] - k[h] * (uint32_t)n;
        } else {
          i[h] = owner_of(rs, n, r[h]);
          k[h] = (uint32_t)(r[h] - rs[i[h]]);
        }
        jb[h] = jobs[i[h]];
        win[h] = b.ent_window[jb[h].y];
        g[h] = gather_job(tr, ms, lights, tx, ent, b, p, jb[h].x, jb[h].y, nullptr, &m[h]);
        if (!valid[h]) g[h].samples = 0u;  // a copy of entry 2q past the end: reads and casts nothing
        const float4 dr = rays[2 * r[h]];
        z[h] = dr.x;
        ang[h] = dr.y;
        la[h] = len(g[h].dir0);
      }
      const GatherJob* gj[2] = {&g[0], &g[1]};
      f3 d[2], c[2];
      float dist[2];
      bool ray[2];
      gather_exact_of<2>(gj, trig, z, ang, k, la, p, d, dist, c, ray, &reads);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (!valid[h]) continue;
        nr[h] = ray[h] ? 1u : 0u;
        if (ray[h] && !(c[h].x == 0.0f && c[h].y == 0.0f && c[h].z == 0.0f)) {
          const uint32_t alpha = lights[ent.light[jb[h].y]].alpha ? 1u : 0u;
          rays[2 * r[h]] = make_float4(g[h].origin.x, g[h].origin.y, g[h].origin.z, dist[h]);
          rays[2 * r[h] + 1] = make_float4(d[h].x, d[h].y, d[h].z, __uint_as_float(alpha | 2u));
          contrib[3 * r[h]] = c[h].x;
          contrib[3 * r[h] + 1] = c[h].y;
          contrib[3 * r[h] + 2] = c[h].z;
          atomicOr(&tmask[i[h] * words + (k[h] >> 5)], 1u << (k[h] & 31u));
          atomicAdd(&tcount[i[h]], 1ull);
        }
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      wave_add_at(b.rays, m[h], nr[h], valid[h]);
      if (b.wrays) wave_add_at(b.wrays, win[h], nr[h], valid[h]);
    }
    if (p.photon_reads) wave_add(p.photon_reads, reads);
  }
}

// The exact and the fast direction of gather samples (a3: the job's direction to the texel, 3
// floats; z, ang: the draws; sp: the spread), for the test of kSplitDirErr.
__global__ void k_test_directions(TrigTables trig, const float* __restrict__ a3, const float* __restrict__ z,
                                  const float* __restrict__ ang, const float* __restrict__ sp, uint64_t n, float* exact3,
                                  float* fast3, uint32_t* fast_ok) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const f3 a = mk3(a3[3 * i], a3[3 * i + 1], a3[3 * i + 2]);
    const float r = sqrtf(maxf_(1.0f - z[i] * z[i], 0.0f));
    float sa, ca;
    lsincosf(trig, ang[i], &sa, &ca);
    const f3 rnd = mk3(r * ca, r * sa, z[i]);
    const f3 e = normalize_or(slerp3(trig, a, rnd, sp[i]), a);
    const float la = len(a);
    f3 f = mk3(0.0f, 0.0f, 0.0f);
    fast_ok[i] = fast_direction(a, la, 1.0f / la, z[i], ang[i], sp[i], &f) ? 1u : 0u;
    exact3[3 * i] = e.x;
    exact3[3 * i + 1] = e.y;
    exact3[3 * i + 2] = e.z;
    fast3[3 * i] = f.x;
    fast3[3 * i + 1] = f.y;
    fast3[3 * i + 2] = f.z;
  }
}

// The ray slots of the traced samples of a chunk's jobs, job by job in sample order, at the
// scanned offsets of the jobs' traced counts: the any-hit stream then walks traced rays only.
__global__ void k_gather_list(const uint32_t* __restrict__ tmask, uint32_t words, const uint64_t* __restrict__ toff,
                              const uint64_t* __restrict__ rs, uint64_t n, uint32_t order, uint32_t* list) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t at = toff[i];
    const uint64_t base = rs[i];
    for (uint32_t w = 0; w < words; ++w)
      for (uint32_t m = tmask[i * words + w]; m; m &= m - 1u)
        list[at++] = (uint32_t)ray_slot(order, base, i, n, w * 32u + (uint32_t)__builtin_ctz(m));
  }
}

// The traced ray slots (flag 2) of n jobs, job by job in sample order.  Job i owns its first
// rs[i + 1] - rs[i] samples (all `slots` without rs), stored at ray_slot(order, base, i, n, k)
// with base rs[i] (i * slots without rs).  Pass 0 counts each job's traced rays into count
// (count[n] = 0, for the scan); pass 1 writes them at the scanned offsets.
__global__ void k_ray_list(const float4* __restrict__ rays, uint64_t n, uint32_t slots, const uint64_t* __restrict__ rs,
                           uint32_t order, int pass, uint64_t* count, const uint64_t* __restrict__ off, uint32_t* list) {
  if (pass == 0 && blockIdx.x == 0 && threadIdx.x == 0) count[n] = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t mine = rs ? (uint32_t)(rs[i + 1] - rs[i]) : slots;
    const uint64_t base = rs ? rs[i] : i * slots;
    uint64_t at = pass == 1 ? off[i] : 0;
    for (uint32_t k = 0; k < mine; ++k) {
      const uint64_t slot = ray_slot(order, base, i, n, k);
      if (!(__float_as_uint(rays[2 * slot + 1].w) & 2u)) continue;
      if (pass == 1) list[at] = (uint32_t)slot;
      at++;
    }
    if (pass == 0) count[i] = at;
  }
}

// Phase 2: persistent any-hit traversal of the chunk's rays.  Each lane keeps one ray in flight
// and the next one prefetched in registers; a lane whose ray ends swaps in the prefetched ray at
// once and requests another, so lanes never wait for the wave's longest ray.  `queue` (zeroed)
// hands out runs of `chunk` rays per wave.  An occluded ray's contribution is set to zero.
// CLOSEST: closest-hit traversal instead; `contrib` is then a float4 per ray slot that receives
// (t, u, v, leaf position) of the hit of every traced ray that hits (slots of rays that miss, or
// hold no ray, keep what they held).
template <int MIN_WAVES, bool CLOSEST>
__global__ __attribute__((amdgpu_waves_per_eu(MIN_WAVES, 8))) void k_any_stream(DevScene sc, const float4* __restrict__ rays, const uint64_t* __restrict__ total, float* contrib,
                             unsigned int* queue, uint32_t chunk, uint32_t refill, uint32_t order,
                             const uint64_t* __restrict__ rs, uint32_t njobs, uint32_t positions,
                             const uint32_t* __restrict__ list) {
  extern __shared__ uint32_t lds[];
  uint32_t* stack = lds + threadIdx.x;
  Tally tl;
  const uint32_t lane = threadIdx.x & 63u;
  // slots to hand out: the rays of the chunk (job-major), or every (sample, job) slot (sample-major,
  // where slots past a job's sample count hold no ray)
  // with a list: the *total listed ray slots, in list order
  const uint32_t nrays = order && !list ? positions : (uint32_t)*total;  // a chunk holds fewer than 2^32 slots
  const uint32_t per_job = order ? positions / njobs : 0;       // slots per job (the most samples)
  uint32_t cbase = 0, cend = 0;
  bool drained = false;
  bool has_next = false, live = false;
  uint32_t nidx = 0, cidx = 0;
  float4 n0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), n1 = n0;
  f3 o = mk3(0.0f, 0.0f, 0.0f), d = o;
  SlabRay sr{};
  float tmax = 0.0f;
  uint32_t alpha = 0, sp = 0;
  int32_t node = 0;
  bool have = false;
  Hit hit{};
  PendingLeaves pend{};
  for (;;) {
    // Finished lanes take their prefetched ray, and lanes without one request another, only
    // once at least `refill` lanes are idle: these steps then run with many lanes at a time.
    const uint32_t idle = (uint32_t)__popcll(__ballot(!live));
    if (idle >= refill || idle == 64u) {
    if (!live && has_next) {
      has_next = false;
      cidx = nidx;
      const uint32_t flags = __float_as_uint(n1.w);
      if ((flags & 2u) && sc.ntris != 0) {
        o = mk3(n0.x, n0.y, n0.z);
        tmax = n0.w;
        d = mk3(n1.x, n1.y, n1.z);
        alpha = flags & 1u;
        sr = slab_ray(o, d);
        node = 0;
        sp = 0;
        have = false;
        pend.mask = 0;
        pend.i = pend.end = 0;
        live = true;
      }
    }
    const uint64_t need = __ballot(!has_next);
    if (need && !drained) {
      if (cbase >= cend) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(queue, chunk);
        base = __builtin_amdgcn_readfirstlane(base);
        cbase = base;
        if (base >= nrays) drained = true;
        else cend = nrays - base > chunk ? base + chunk : nrays;
      }
      if (!drained) {
        const uint32_t avail = cend - cbase;
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        if (!has_next && rank < avail) {
          nidx = cbase + rank;
          bool slot = true;
          if (list) {
            nidx = list[nidx];
          } else if (order) {  // hand out a job's samples together, from their sample-major slots
            const uint32_t job = nidx / per_job, k = nidx - job * per_job;
            slot = rs == nullptr || (uint64_t)k < rs[job + 1] - rs[job];
            nidx = k * njobs + job;
          }
          if (slot) {
            n0 = rays[2 * (size_t)nidx];
            n1 = rays[2 * (size_t)nidx + 1];
            has_next = true;
          }
        }
        const uint32_t want = (uint32_t)__popcll(need);
        cbase += avail < want ? avail : want;
      }
    }
    if (!__ballot(live || has_next)) {
      if (drained) break;
      continue;
    }
    }
    if (live) {
      if constexpr (CLOSEST) {
        if (step_closest(sc, stack, blockDim.x, o, d, sr, tmax, node, sp, tl, alpha, have, hit, pend) != 0u) {
          live = false;
          if (have) reinterpret_cast<float4*>(contrib)[cidx] = make_float4(hit.t, hit.u, hit.v, __uint_as_float(hit.pos));
        }
      } else {
        const uint32_t r = visit_any(sc, stack, blockDim.x, o, d, sr, tmax, node, sp, tl, alpha);
        if (r != 0u) {
          live = false;
          if (r == 1u) {
            contrib[3 * (size_t)cidx] = 0.0f;
            contrib[3 * (size_t)cidx + 1] = 0.0f;
            contrib[3 * (size_t)cidx + 2] = 0.0f;
          }
        }
      }
    }
  }
  if (sc.counters) {
    wave_add(&sc.counters[0], tl.nodes);
    wave_add(&sc.counters[1], tl.tris);
  }
}

// Phase 3: each job's samples summed in sample order (bounce.rs:287-290) into its window texel.
// Only the traced samples (the bits of tmask) are read: the others contribute +0, occluded ones
// hold +0, and adding +0 leaves the sum (never -0) unchanged.
__global__ void k_gather_sum(DevEntries ent, const DevLight* __restrict__ lights, DevTexels tx, DevBounce b,
                             const uint2* __restrict__ jobs, uint64_t n, const uint64_t* __restrict__ rs,
                             const float* __restrict__ contrib, uint32_t order, const uint32_t* __restrict__ tmask,
                             uint32_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 jb = jobs[i];
    const DevLight& L = lights[ent.light[jb.y]];
    const uint32_t samples = L.bounce_samples > 2 ? L.bounce_samples : 2u;
    const uint64_t base = rs[i];
    f3 sum = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t w = 0; w < words; ++w)
      for (uint32_t m = tmask[i * words + w]; m; m &= m - 1u) {
        const float* c = contrib + 3 * ray_slot(order, base, i, n, w * 32u + (uint32_t)__builtin_ctz(m));
        sum = add(sum, mk3(c[0], c[1], c[2]));
      }
    const f3 v = divs(mul(sum, L.bounce_intensity), (float)samples);
    const uint32_t w = b.ent_window[jb.y];
    const uint32_t* wr = b.win_rect + 4 * (size_t)w;
    const uint32_t xy = tx.xy[jb.x];
    const uint64_t dst = b.win_off[w] + (uint64_t)((xy & 0xffffu) - wr[0]) + (uint64_t)((xy >> 16) - wr[1]) * (wr[2] - wr[0] + 1);
    float* out = b.map_a + dst * 3;
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
    mark_tile(b, w, (uint64_t)((xy & 0xffffu) - wr[0]), (uint64_t)((xy >> 16) - wr[1]), v);
  }
}

// ---- streamed sun bounce -----------------------------------------------------------------------
// gather_sun (sun_bounce.rs:24-80) split like the point-light gather: a job is a (texel, sun
// bounce entry) pair; its `samples` cosine-weighted rays from the texel look for the closest hit
// (k_sun_rays, then the closest-hit ray stream), a hit on a surface facing the ray and lit by the
// sun casts one ray from the hit point toward the sun (k_sun_mid, then the any-hit stream), and
// the job's texel receives the mean of albedo x cosine over the samples whose sun ray is
// unoccluded, summed in sample order (k_sun_sum).  A chunk holds n jobs of S ray slots each, at
// ray_slot(order, i * S, i, n, k).  Ray flags: 1 alpha-tested occluders let light through
// transparent texels, 2 the ray is traced, 4 the sun ray counts as cast (texels a later triangle
// overwrites, bounce.rs:202-209, count their sun rays without tracing them).

// Whether the job's triangle owns its texel in the window map.
__device__ __forceinline__ bool owns_texel(const DevBounce& b, const DevEntries& ent, const DevTexels& tx, uint2 jb) {
  const uint32_t w = b.ent_window[jb.y];
  const uint32_t* wr = b.win_rect + 4 * (size_t)w;
  const uint32_t xy = tx.xy[jb.x];
  const uint64_t k = (uint64_t)((xy & 0xffffu) - wr[0]) + (uint64_t)((xy >> 16) - wr[1]) * (wr[2] - wr[0] + 1);
  const unsigned int* owner = reinterpret_cast<const unsigned int*>(b.map_b) + b.win_off[w] * 3;
  return owner[k] == ent.tri[jb.y] + 1;
}

__global__ void k_sun_rays(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevTexels tx, DevEntries ent,
                           const uint2* __restrict__ jobs, uint64_t n, uint32_t S, uint32_t order, BounceParams p, TrigTables trig,
                           float4* rays, float4* hits, uint64_t* total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = n * S;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 jb = jobs[i];
    const uint32_t t = ent.tri[jb.y];
    const DevLight& L = lights[ent.light[jb.y]];
    const uint32_t samples = L.bounce_samples > 2 ? L.bounce_samples : 2u;
    const float4 p4 = tx.pos[jb.x];
    const f3 normal = ld3(tr.face + 3 * (size_t)t);
    const f3 origin = add(mk3(p4.x, p4.y, p4.z), mul(normal, kRayOffset));
    const uint32_t xy = tx.xy[jb.x];
    Pcg rng = stream5(p.seed, ms.id[tr.mesh[t]], L.id, xy & 0xffffu, xy >> 16);
    f2 offset;
    offset.x = pcg_f32(rng);
    offset.y = pcg_f32(rng);
    const float flags = __uint_as_float(2u | (L.alpha ? 1u : 0u));
    for (uint32_t k = 0; k < S; ++k) {
      const uint64_t r = ray_slot(order, i * S, i, n, k);
      hits[r] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(~0u));
      if (k < samples) {
        const f3 d = cosine_dir(trig, normal, spread_pt(k, offset, true));
        rays[2 * r] = make_float4(origin.x, origin.y, origin.z, INFINITY);
        rays[2 * r + 1] = make_float4(d.x, d.y, d.z, flags);
      } else {
        rays[2 * r + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
  }
}

// One thread per ray slot, in storage order: the closest hit of the slot's ray (leaf position
// ~0u: none) becomes the sun ray of the slot and its contribution if unoccluded.
__global__ void k_sun_mid(DevScene sc, DevOcc occ, const DevLight* __restrict__ lights, DevEntries ent, DevTexels tx, DevBounce b,
                          const uint2* __restrict__ jobs, uint64_t n, uint32_t S, uint32_t order, uint32_t colored, float4* rays,
                          const float4* __restrict__ hits, float* contrib) {
  const uint64_t slots = n * S;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < slots; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = order ? r % n : r / S;
    const float4 n1 = rays[2 * r + 1];
    const float4 h4 = hits[r];
    const uint32_t pos = __float_as_uint(h4.w);
    f3 c = mk3(0.0f, 0.0f, 0.0f);
    float4 out1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if ((__float_as_uint(n1.w) & 2u) && pos != ~0u) {
      const f3 dir = mk3(n1.x, n1.y, n1.z);
      const f3 face = tri_face_normal(sc, pos);
      if (!(dot(face, dir) >= 0.0f)) {
        const uint2 jb = jobs[i];
        const DevLight& L = lights[ent.light[jb.y]];
        const f3 toward = ld3(L.toward);
        Hit h;
        h.t = h4.x;
        h.u = h4.y;
        h.v = h4.z;
        h.pos = pos;
        h.occ = __float_as_uint(sc.tris[3 * (size_t)pos].w);
        const float lit = dot(surface_normal(sc, occ, h), toward);
        if (!(lit <= 0.0f || dot(face, toward) <= 0.0f)) {
          const float4 n0 = rays[2 * r];
          const f3 point = add(add(mk3(n0.x, n0.y, n0.z), mul(dir, h.t)), mul(face, kRayOffset));
          const uint32_t traced = owns_texel(b, ent, tx, jb) ? 2u | (__float_as_uint(n1.w) & 1u) : 0u;
          rays[2 * r] = make_float4(point.x, point.y, point.z, INFINITY);
          out1 = make_float4(toward.x, toward.y, toward.z, __uint_as_float(4u | traced));
          f3 alb = mk3(1.0f, 1.0f, 1.0f);
          if (colored) alb = ld3(occ.albedo + 3 * (size_t)h.occ);
          c = mul(alb, lit);
        }
      }
    }
    rays[2 * r + 1] = out1;
    contrib[3 * r] = c.x;
    contrib[3 * r + 1] = c.y;
    contrib[3 * r + 2] = c.z;
  }
}

// One thread per job: contributions summed in sample order (an occluded or absent sun ray adds
// +0, which leaves the sum unchanged), the mean written to the owned window texel, and the job's
// rays (every sample's, plus each sun ray cast) added to its mesh.
__global__ void k_sun_sum(DevTris tr, DevEntries ent, const DevLight* __restrict__ lights, DevTexels tx, DevBounce b,
                          const uint2* __restrict__ jobs, uint64_t n, uint32_t S, uint32_t order, const float4* __restrict__ rays,
                          const float* __restrict__ contrib) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
    const uint64_t i = i0 + lane;
    uint32_t m = 0xffffffffu, nr = 0, win = 0;
    if (i < n) {
      const uint2 jb = jobs[i];
      win = b.ent_window[jb.y];
      const DevLight& L = lights[ent.light[jb.y]];
      const uint32_t samples = L.bounce_samples > 2 ? L.bounce_samples : 2u;
      f3 sum = mk3(0.0f, 0.0f, 0.0f);
      nr = samples;
      for (uint32_t k = 0; k < samples; ++k) {
        const uint64_t r = ray_slot(order, i * S, i, n, k);
        sum = add(sum, mk3(contrib[3 * r], contrib[3 * r + 1], contrib[3 * r + 2]));
        nr += (__float_as_uint(rays[2 * r + 1].w) >> 2) & 1u;
      }
      m = tr.mesh[ent.tri[jb.y]];
      if (owns_texel(b, ent, tx, jb)) {
        const f3 v = divs(mul(sum, L.bounce_intensity), (float)samples);
        const uint32_t w = b.ent_window[jb.y];
        const uint32_t* wr = b.win_rect + 4 * (size_t)w;
        const uint32_t xy = tx.xy[jb.x];
        float* out = b.map_a + 3 * (b.win_off[w] + (uint64_t)((xy & 0xffffu) - wr[0]) + (uint64_t)((xy >> 16) - wr[1]) * (wr[2] - wr[0] + 1));
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
        mark_tile(b, w, (uint64_t)((xy & 0xffffu) - wr[0]), (uint64_t)((xy >> 16) - wr[1]), v);
      }
    }
    wave_add_at(b.rays, m, nr, i < n);
    if (b.wrays) wave_add_at(b.wrays, win, nr, i < n);
  }
}

// A window of the filtered map: texels outside it, and texels of filter tiles the filter left
// unwritten (all zero), read as zero.
struct WinView {
  const float* map;
  uint32_t x0, y0, x1, y1;
  const uint8_t* nb;  // the window's tile_nb flags
  __device__ f3 get(int64_t x, int64_t y) const {
    if (x < (int64_t)x0 || y < (int64_t)y0 || x > (int64_t)x1 || y > (int64_t)y1) return mk3(0.0f, 0.0f, 0.0f);
    const uint64_t tiles_x = (x1 - x0 + kFilterTile) / kFilterTile;
    if (!nb[(uint64_t)(y - y0) / kFilterTile * tiles_x + (uint64_t)(x - x0) / kFilterTile]) return mk3(0.0f, 0.0f, 0.0f);
    const float* p = map + 3 * ((uint64_t)(x - x0) + (uint64_t)(y - y0) * (x1 - x0 + 1));
    return mk3(p[0], p[1], p[2]);
  }
};
__device__ __forceinline__ WinView win_view(const DevBounce& b, uint32_t w) {
  const uint32_t* wr = b.win_rect + 4 * (size_t)w;
  return WinView{b.map_b + 3 * b.win_off[w], wr[0], wr[1], wr[2], wr[3], b.tile_nb + b.tile_start[w]};
}

// BounceMap::dilate, then blur along x, then along y (bounce_map.rs:60-125), fused: a block takes
// a kFilterTile-square tile of one window, loads the gather's values (map_a) with a halo of 4 into LDS,
// dilates the tile and a halo of 3, blurs the tile's columns and 3 rows either side along x, then
// along y into map_b.  Texels outside the window read as zero at every stage, as the three
// separate passes read them; the sums run in the same order, so every value is the same.
__global__ __launch_bounds__(256) void k_bounce_filter(DevBounce b, const uint32_t* __restrict__ win_res,
                                                       const uint64_t* __restrict__ tile_start, uint32_t nwin, uint64_t ntiles,
                                                       BlurWeights bw) {
  constexpr int T = kFilterTile, TI = T + 8, TD = T + 6;
  __shared__ float s_in[TI * TI * 3];
  __shared__ float s_d[TD * TD * 3];
  __shared__ float s_h[T * TD * 3];
  const int ox[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1}, oy[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t w = (uint32_t)owner_of(tile_start, nwin, tile);
    const uint32_t* wr = b.win_rect + 4 * (size_t)w;
    const int64_t x0 = wr[0], y0 = wr[1], x1 = wr[2], y1 = wr[3];
    const uint64_t ww = (uint64_t)(x1 - x0 + 1), tiles_x = (ww + T - 1) / T;
    const uint64_t local = tile - tile_start[w];
    const int64_t X0 = x0 + (int64_t)(local % tiles_x) * T, Y0 = y0 + (int64_t)(local / tiles_x) * T;
    const int64_t last = (int64_t)win_res[w] - 1;
    const float* in = b.map_a + 3 * b.win_off[w];
    auto inside = [&](int64_t x, int64_t y) { return x >= x0 && y >= y0 && x <= x1 && y <= y1; };
    {
      // no positive value within the tile's 3x3 tile neighbourhood (its halo of 4 lies inside it):
      // every output is zero (the same for every thread of the block)
      const int64_t tx = (int64_t)(local % tiles_x), ty = (int64_t)(local / tiles_x);
      const int64_t tiles_y = (y1 - y0 + T) / T;
      bool lit = false;
      for (int64_t v = ty - 1; v <= ty + 1; ++v)
        for (int64_t u = tx - 1; u <= tx + 1; ++u)
          if (u >= 0 && v >= 0 && u < (int64_t)tiles_x && v < tiles_y) lit |= b.tile_lit[tile_start[w] + (uint64_t)(v * (int64_t)tiles_x + u)] != 0;
      if (threadIdx.x == 0) b.tile_nb[tile] = lit ? 1 : 0;
      if (!lit) continue;  // left unwritten: readers see the flag and take zeros
    }
    for (int q = threadIdx.x; q < TI * TI; q += blockDim.x) {
      const int64_t x = X0 - 4 + q % TI, y = Y0 - 4 + q / TI;
      f3 v = mk3(0.0f, 0.0f, 0.0f);
      if (inside(x, y)) {
        const float* p = in + 3 * ((uint64_t)(x - x0) + (uint64_t)(y - y0) * ww);
        v = mk3(p[0], p[1], p[2]);
      }
      s_in[3 * q] = v.x;
      s_in[3 * q + 1] = v.y;
      s_in[3 * q + 2] = v.z;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < TD * TD; q += blockDim.x) {  // dilate: the first lit of 9 neighbours
      const int qx = q % TD, qy = q / TD;
      f3 r = mk3(0.0f, 0.0f, 0.0f);
      if (inside(X0 - 3 + qx, Y0 - 3 + qy)) {
        for (int k = 0; k < 9; ++k) {
          const int c = 3 * ((qx + 1 + ox[k]) + (qy + 1 + oy[k]) * TI);
          const f3 sv = mk3(s_in[c], s_in[c + 1], s_in[c + 2]);
          if (max_elem(sv) > 0.0f) {
            r = sv;
            break;
          }
        }
      }
      s_d[3 * q] = r.x;
      s_d[3 * q + 1] = r.y;
      s_d[3 * q + 2] = r.z;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < T * TD; q += blockDim.x) {  // blur along x, clamped to the map
      const int qx = q % T, qy = q / T;
      const int64_t x = X0 + qx, y = Y0 - 3 + qy;
      f3 r = mk3(0.0f, 0.0f, 0.0f);
      if (inside(x, y)) {
        f3 sum = mk3(0.0f, 0.0f, 0.0f);
        for (int k = -3; k <= 3; ++k) {
          int64_t xx = x + k;
          xx = xx < 0 ? 0 : (xx > last ? last : xx);
          const int c = 3 * ((int)(xx - (X0 - 3)) + qy * TD);
          sum = add(sum, mul(mk3(s_d[c], s_d[c + 1], s_d[c + 2]), bw.w[k + 3]));
        }
        r = divs(sum, bw.sum);
      }
      s_h[3 * q] = r.x;
      s_h[3 * q + 1] = r.y;
      s_h[3 * q + 2] = r.z;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < T * T; q += blockDim.x) {  // blur along y into map_b
      const int qx = q % T, qy = q / T;
      const int64_t x = X0 + qx, y = Y0 + qy;
      if (!inside(x, y)) continue;
      f3 sum = mk3(0.0f, 0.0f, 0.0f);
      for (int k = -3; k <= 3; ++k) {
        int64_t yy = y + k;
        yy = yy < 0 ? 0 : (yy > last ? last : yy);
        const int c = 3 * (qx + (int)(yy - (Y0 - 3)) * T);
        sum = add(sum, mul(mk3(s_h[c], s_h[c + 1], s_h[c + 2]), bw.w[k + 3]));
      }
      const f3 r = divs(sum, bw.sum);
      float* d = b.map_b + 3 * (b.win_off[w] + (uint64_t)(x - x0) + (uint64_t)(y - y0) * ww);
      d[0] = r.x;
      d[1] = r.y;
      d[2] = r.z;
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}

// BounceMap::crop's lit test (bounce_map.rs:128-137): any positive texel inside the crop.
__global__ void k_bounce_has(DevTris tr, DevEntries ent, DevBounce b, uint8_t* has) {
  // one wave per entry: 64 consecutive texels of the crop (row-major) per step, until one is lit
  const uint32_t lane = threadIdx.x & 63u, waves = blockDim.x >> 6;
  for (uint64_t e = (uint64_t)blockIdx.x * waves + (threadIdx.x >> 6); e < ent.n; e += (uint64_t)gridDim.x * waves) {
    const uint32_t w = b.ent_window[e];
    bool any = false;
    if (w != 0xffffffffu) {
      const uint32_t* c = tr.crop + 4 * (size_t)ent.tri[e];
      const uint32_t* wr = b.win_rect + 4 * (size_t)w;
      const WinView v = win_view(b, w);
      const uint32_t cw = c[2] >= c[0] ? c[2] - c[0] + 1u : 0u, ch = c[3] >= c[1] ? c[3] - c[1] + 1u : 0u;
      uint32_t count = cw * ch;  // crop sides are at most 65535
      // A filtered texel is positive only in a tile with a marked tile among its 3x3 neighbours:
      // without a mark on the window tiles the crop overlaps, grown by one, nothing is lit.
      const int64_t ix0 = c[0] > wr[0] ? c[0] : wr[0], ix1 = c[2] < wr[2] ? c[2] : wr[2];
      const int64_t iy0 = c[1] > wr[1] ? c[1] : wr[1], iy1 = c[3] < wr[3] ? c[3] : wr[3];
      bool marked = false;
      if (count && ix0 <= ix1 && iy0 <= iy1) {
        const int64_t T = kFilterTile, tiles_x = ((int64_t)wr[2] - wr[0] + T) / T, tiles_y = ((int64_t)wr[3] - wr[1] + T) / T;
        const int64_t tx0 = (ix0 - wr[0]) / T > 0 ? (ix0 - wr[0]) / T - 1 : 0;
        const int64_t ty0 = (iy0 - wr[1]) / T > 0 ? (iy0 - wr[1]) / T - 1 : 0;
        const int64_t tx1 = (ix1 - wr[0]) / T + 1 < tiles_x ? (ix1 - wr[0]) / T + 1 : tiles_x - 1;
        const int64_t ty1 = (iy1 - wr[1]) / T + 1 < tiles_y ? (iy1 - wr[1]) / T + 1 : tiles_y - 1;
        const uint64_t nx = (uint64_t)(tx1 - tx0 + 1), nt = nx * (uint64_t)(ty1 - ty0 + 1);
        for (uint64_t base = 0; base < nt && !marked; base += 64u) {
          const uint64_t q = base + lane;
          const bool f = q < nt && b.tile_lit[b.tile_start[w] + (uint64_t)(ty0 + (int64_t)(q / nx)) * (uint64_t)tiles_x +
                                              (uint64_t)(tx0 + (int64_t)(q % nx))] != 0;
          marked = __ballot(f) != 0ull;
        }
      }
      if (!marked) count = 0;
      for (uint32_t base = 0; base < count && !any; base += 64u) {
        const uint32_t q = base + lane;
        bool lit = false;
        if (q < count) lit = max_elem(v.get((int64_t)c[0] + q % cw, (int64_t)c[1] + q / cw)) > 0.0f;
        any = __ballot(lit) != 0ull;
      }
    }
    if (lane == 0) has[e] = any ? 1 : 0;
  }
}

__device__ __forceinline__ uint32_t comp_bits(int32_t c) { return c == 0 ? 8u : c == 1 ? 6u : c == 2 ? 5u : 4u; }
__device__ __forceinline__ float comp_amp(int32_t c) { return c == 0 ? 0.0f : c == 1 ? 0.004f : c == 2 ? 0.008f : 0.02f; }
// quantize (bounce_pack.rs:75-86) without dither
__device__ __forceinline__ uint32_t quantize0(float value, uint32_t bits) {
  float mx = (float)((1u << bits) - 1u);
  float color = value != value ? 0.0f : clampf_(value, 0.0f, 1.0f);
  return sat_u32(sqrtf(color) * mx);
}
__device__ __forceinline__ uint32_t quantize_d(float value, uint32_t bits, float amp, Pcg& rng) {
  float mx = (float)((1u << bits) - 1u);
  float color = value != value ? 0.0f : clampf_(value, 0.0f, 1.0f);
  if (amp > 0.0f && color > 0.001f) color = minf_(color + pcg_f32(rng) * amp, 1.0f);
  return sat_u32(sqrtf(color) * mx);
}

struct PackView {
  WinView v;
  uint32_t cx0, cy0, W, H;
  __device__ f3 pixel(uint64_t i) const {
    int32_t px = (int32_t)(i % W), py = (int32_t)(i / W);
    if (px < 2 || py < 2 || px >= (int32_t)W - 2 || py >= (int32_t)H - 2) return mk3(0.0f, 0.0f, 0.0f);
    return v.get((int64_t)cx0 + px - 2, (int64_t)cy0 + py - 2);
  }
};

__device__ __forceinline__ PackView pack_view(const DevTris& tr, const DevEntries& ent, const DevBounce& b, uint64_t e) {
  uint32_t t = ent.tri[e], w = b.ent_window[e];
  PackView pv;
  pv.v = win_view(b, w);
  pv.cx0 = tr.crop[4 * (size_t)t];
  pv.cy0 = tr.crop[4 * (size_t)t + 1];
  pv.W = tr.crop[4 * (size_t)t + 2] - pv.cx0 + 5;
  pv.H = tr.crop[4 * (size_t)t + 3] - pv.cy0 + 5;
  return pv;
}

__global__ void k_bounce_pack_words(DevTris tr, DevEntries ent, DevBounce b, const uint64_t* __restrict__ kept_entry,
                                    const uint64_t* __restrict__ word_start, uint64_t nkept, uint64_t total,
                                    const uint64_t* __restrict__ kept_out, const int32_t* __restrict__ comp_of_slot,
                                    uint32_t colored, uint32_t* out) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t j = owner_of_wave(word_start, nkept, c);
    uint64_t wi = c - word_start[j];
    uint64_t e = kept_entry[j];
    PackView pv = pack_view(tr, ent, b, e);
    uint32_t bits = comp_bits(comp_of_slot[ent.light[e]]);
    uint64_t n = (uint64_t)pv.W * pv.H;
    uint32_t word = 0;
    if (colored) {
      uint64_t pb = 3 * bits;
      uint64_t lo = wi * 32, hi = lo + 32;
      for (uint64_t i = lo / pb; i < n && i * pb < hi; ++i) {
        f3 v = pv.pixel(i);
        uint64_t pixel = quantize0(v.x, bits) | (quantize0(v.y, bits) << bits) | (quantize0(v.z, bits) << (2 * bits));
        int64_t shift = (int64_t)(i * pb) - (int64_t)lo;
        word |= shift >= 0 ? (uint32_t)(pixel << shift) : (uint32_t)(pixel >> (-shift));
      }
    } else {
      uint32_t per = 32 / bits;
      for (uint32_t k = 0; k < per; ++k) {
        uint64_t i = wi * per + k;
        if (i >= n) break;
        word |= quantize0(pv.pixel(i).x, bits) << (k * bits);
      }
    }
    out[kept_out[j] + wi] = word;
  }
}

// Capture (test hook): the f32 RGB values every kept bounce entry's packing reads, its crop
// with 2 texels of zero padding, at the entry's offset (3 floats per texel) of `out`.
__global__ void k_bounce_dump(DevTris tr, DevEntries ent, DevBounce b, const uint64_t* __restrict__ kept_entry, uint64_t nkept,
                              const uint64_t* __restrict__ value_start, float* out) {
  for (uint64_t j = blockIdx.x; j < nkept; j += gridDim.x) {
    const PackView pv = pack_view(tr, ent, b, kept_entry[j]);
    const uint64_t n = (uint64_t)pv.W * pv.H;
    float* o = out + value_start[j];
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const f3 v = pv.pixel(i);
      o[3 * i] = v.x;
      o[3 * i + 1] = v.y;
      o[3 * i + 2] = v.z;
    }
  }
}

// Dithered precisions draw from one stream per (mesh, light, triangle) in pixel order.
__global__ void k_bounce_pack_seq(DevTris tr, DevMeshes ms, const DevLight* __restrict__ lights, DevEntries ent, DevBounce b,
                                  const uint64_t* __restrict__ kept_entry, uint64_t nkept, const uint64_t* __restrict__ kept_out,
                                  const uint64_t* __restrict__ word_start, const int32_t* __restrict__ comp_of_slot,
                                  uint32_t colored, uint64_t seed, uint32_t* out) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nkept; j += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t e = kept_entry[j];
    PackView pv = pack_view(tr, ent, b, e);
    int32_t comp = comp_of_slot[ent.light[e]];
    uint32_t bits = comp_bits(comp);
    float amp = comp_amp(comp);
    uint32_t t = ent.tri[e];
    Pcg rng = stream4(seed, ms.id[tr.mesh[t]], lights[ent.light[e]].id, tr.local[t]);
    uint64_t n = (uint64_t)pv.W * pv.H;
    uint64_t words = word_start[j + 1] - word_start[j];
    uint32_t* o = out + kept_out[j];
    for (uint64_t k = 0; k < words; ++k) o[k] = 0;
    if (colored) {
      uint64_t pb = 3 * bits;
      for (uint64_t i = 0; i < n; ++i) {
        f3 v = pv.pixel(i);
        uint32_t px = quantize_d(v.x, bits, amp, rng);
        px |= quantize_d(v.y, bits, amp, rng) << bits;
        px |= quantize_d(v.z, bits, amp, rng) << (2 * bits);
        uint64_t off = i * pb, w = off / 32;
        uint32_t shift = (uint32_t)(off % 32);
        o[w] |= px << shift;
        if (shift + pb > 32) o[w + 1] |= px >> (32 - shift);
      }
    } else {
      uint32_t per = 32 / bits;
      for (uint64_t i = 0; i < n; ++i) o[i / per] |= quantize_d(pv.pixel(i).x, bits, amp, rng) << ((i % per) * bits);
    }
  }
}

hipError_t launch_photons(hipStream_t s, DevScene scene, DevOcc occ, const DevLight* lights, const uint32_t* cube_slot,
                          uint32_t ncubes, uint32_t res, uint32_t* photons, uint32_t block) {
  uint64_t total = 6ull * res * res * ncubes;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_photons, dim3(grid_for(total, block)), dim3(block), stack_bytes(scene, block), s, scene, occ, lights,
                     cube_slot, ncubes, res, photons);
  return hipGetLastError();
}
hipError_t launch_photon_tiles(hipStream_t s, const DevLight* lights, const uint32_t* cube_slot, uint32_t ncubes, uint32_t res,
                               const uint32_t* photons, unsigned long long* live, float4* box,
                               const unsigned long long* need_mask) {
  const uint32_t tpf = photon_tiles_per_face(res);
  const uint64_t tiles = 6ull * tpf * tpf * ncubes;
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_photon_tiles, dim3(grid_for(tiles, kThreads / 64)), dim3(kThreads), 0, s, lights, cube_slot, ncubes, res, tpf,
                     photons, need_mask, live, box);
  return hipGetLastError();
}
hipError_t launch_photon_mark(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* jobs, uint64_t n, BounceParams p, TrigTables trig, uint8_t* need) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_photon_mark, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, tris, meshes, lights, tx, ent, b, jobs, n, p,
                     trig, need);
  return hipGetLastError();
}
hipError_t launch_need_count(hipStream_t s, const uint8_t* need, uint64_t tiles, unsigned long long* need_mask, uint64_t* cnt) {
  hipLaunchKernelGGL(k_need_count, dim3(grid_for(tiles + 1, kThreads)), dim3(kThreads), 0, s,
                     reinterpret_cast<const uint4*>(need), tiles, need_mask, cnt);
  return hipGetLastError();
}
hipError_t launch_need_list(hipStream_t s, const unsigned long long* need_mask, const uint64_t* off, uint64_t tiles, uint32_t* list) {
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_need_list, dim3(grid_for(tiles, kThreads)), dim3(kThreads), 0, s, need_mask, off, tiles, list);
  return hipGetLastError();
}
hipError_t launch_photons_listed(hipStream_t s, DevScene scene, DevOcc occ, const DevLight* lights, const uint32_t* cube_slot,
                                 uint32_t res, const uint32_t* list, uint64_t total, uint32_t* photons, uint32_t block) {
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_photons_listed, dim3(grid_for(total, block)), dim3(block), stack_bytes(scene, block), s, scene, occ, lights,
                     cube_slot, res, photon_tiles_per_face(res), list, total, photons);
  return hipGetLastError();
}
hipError_t launch_bounce_owner(hipStream_t s, DevTris tris, const DevLight* lights, DevTexels tx, const uint64_t* tex_start,
                               DevEntries ent, DevBounce b, uint64_t units) {
  if (units == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bounce_owner, dim3(grid_for(units, kThreads / 64)), dim3(kThreads), 0, s, tris, lights, tx, tex_start,
                     ent, b, units);
  return hipGetLastError();
}
hipError_t launch_bounce_gather(hipStream_t s, DevScene scene, DevOcc occ, DevTris tris, DevMeshes meshes,
                                const DevLight* lights, DevTexels tx, const uint64_t* tex_start, DevEntries ent, DevBounce b,
                                uint64_t units, BounceParams p, TrigTables trig, uint32_t block) {
  if (units == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bounce_gather, dim3(grid_for(units, block / 64)), dim3(block), stack_bytes(scene, block), s, scene, occ,
                     tris, meshes, lights, tx, tex_start, ent, b, units, p, trig);
  return hipGetLastError();
}
hipError_t launch_sun_map(hipStream_t s, DevScene scene, const SunMapParams& p, uint32_t* out, uint64_t words) {
  if (words == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sun_map, dim3(grid_for(words, kThreads)), dim3(kThreads), stack_bytes(scene, kThreads), s, scene, p, out,
                     words);
  return hipGetLastError();
}
hipError_t launch_sky_probes(hipStream_t s, DevScene scene, TrigTables trig, const SkyProbeParams& p, float* open, uint32_t* back) {
  const uint64_t n = (uint64_t)p.counts[0] * p.counts[1] * p.counts[2] * 6;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sky_probes, dim3(grid_for(n, 64)), dim3(64), stack_bytes(scene, 64), s, scene, trig, p, open, back);
  return hipGetLastError();
}
hipError_t launch_gather_jobs(hipStream_t s, DevTris tris, const DevLight* lights, DevTexels tx, const uint64_t* tex_start,
                              DevEntries ent, DevBounce b, int pass, uint64_t* cnt, const uint64_t* off, uint2* trace_jobs,
                              uint2* count_jobs, uint2* sun_jobs) {
  hipLaunchKernelGGL(k_gather_jobs, dim3(grid_for(std::max<uint64_t>(b.n, 1), kThreads / 64)), dim3(kThreads), 0, s, tris, lights,
                     tx, tex_start, ent, b, pass, cnt, off, trace_jobs, count_jobs, sun_jobs);
  return hipGetLastError();
}
hipError_t launch_sun_rays(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                           const uint2* jobs, uint64_t n, uint32_t slots, uint32_t order, BounceParams p, TrigTables trig,
                           float4* rays, float4* hits, uint64_t* total) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sun_rays, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, tris, meshes, lights, tx, ent, jobs, n, slots,
                     order, p, trig, rays, hits, total);
  return hipGetLastError();
}
hipError_t launch_sun_mid(hipStream_t s, DevScene scene, DevOcc occ, const DevLight* lights, DevEntries ent, DevTexels tx,
                          DevBounce b, const uint2* jobs, uint64_t n, uint32_t slots, uint32_t order, uint32_t colored,
                          float4* rays, const float4* hits, float* contrib) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sun_mid, dim3(grid_for(n * slots, kThreads)), dim3(kThreads), 0, s, scene, occ, lights, ent, tx, b, jobs, n,
                     slots, order, colored, rays, hits, contrib);
  return hipGetLastError();
}
hipError_t launch_sun_sum(hipStream_t s, DevTris tris, DevEntries ent, const DevLight* lights, DevTexels tx, DevBounce b,
                          const uint2* jobs, uint64_t n, uint32_t slots, uint32_t order, const float4* rays, const float* contrib) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sun_sum, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, tris, ent, lights, tx, b, jobs, n, slots, order,
                     rays, contrib);
  return hipGetLastError();
}
hipError_t launch_gather_count(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                               DevBounce b, const uint2* jobs, uint64_t njobs, BounceParams p, TrigTables trig) {
  if (njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_count, dim3(grid_for(njobs, kThreads)), dim3(kThreads), 0, s, tris, meshes, lights, tx, ent, b, jobs,
                     njobs, p, trig);
  return hipGetLastError();
}
hipError_t launch_gather_nrays(hipStream_t s, DevEntries ent, const DevLight* lights, const uint2* jobs, uint64_t n, uint64_t* out) {
  hipLaunchKernelGGL(k_gather_nrays, dim3(grid_for(n + 1, kThreads)), dim3(kThreads), 0, s, ent, lights, jobs, n, out);
  return hipGetLastError();
}
hipError_t launch_gather_split(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                               DevBounce b, const uint2* jobs, uint64_t n, const uint64_t* rs, BounceParams p, float4* rays,
                               uint32_t order, uint32_t* omask, uint32_t* tmask, uint32_t words, uint64_t* ocount,
                               uint64_t* tcount) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_split, dim3(grid_for(n, 128)), dim3(128), 0, s, tris, meshes, lights, tx, ent, b, jobs, n, rs, p, rays,
                     order, omask, tmask, words, ocount, tcount);
  return hipGetLastError();
}
hipError_t launch_gather_open(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* jobs, uint64_t n, const uint64_t* rs, BounceParams p, TrigTables trig,
                              float4* rays, float* contrib, uint32_t order, const uint32_t* list, const uint64_t* count,
                              uint64_t max_count, uint32_t* tmask, uint32_t words, uint64_t* tcount) {
  if (n == 0 || max_count == 0) return hipSuccess;
  // a grid for at most 2^20 lanes (the loop strides over the device-side count)
  hipLaunchKernelGGL(k_gather_open, dim3(grid_for(std::min<uint64_t>(max_count, 1ull << 20), 128)), dim3(128), 0, s, tris, meshes,
                     lights, tx, ent, b, jobs, n, rs, p, trig, rays, contrib, order, list, count, tmask, words,
                     reinterpret_cast<unsigned long long*>(tcount));
  return hipGetLastError();
}
hipError_t launch_test_directions(hipStream_t s, TrigTables trig, const float* a3, const float* z, const float* ang,
                                  const float* sp, uint64_t n, float* exact3, float* fast3, uint32_t* fast_ok) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_test_directions, dim3(grid_for(n, 256)), dim3(256), 0, s, trig, a3, z, ang, sp, n, exact3, fast3, fast_ok);
  return hipGetLastError();
}
hipError_t launch_gather_rays(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevTexels tx, DevEntries ent,
                              DevBounce b, const uint2* jobs, uint64_t n, const uint64_t* rs, BounceParams p, TrigTables trig,
                              float4* rays, float* contrib, uint32_t order, uint32_t* tmask, uint32_t words, uint64_t* tcount) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rays, dim3(grid_for(n, 128)), dim3(128), 0, s, tris, meshes, lights, tx, ent, b, jobs, n, rs, p, trig,
                     rays, contrib, order, tmask, words, tcount);
  return hipGetLastError();
}
template <bool CLOSEST>
static hipError_t launch_ray_stream(hipStream_t s, DevScene scene, const float4* rays, const uint64_t* total, uint64_t max_rays,
                                    const uint64_t* rs, uint64_t njobs, float* contrib, unsigned int* queue, const GatherLaunch& gl,
                                    const uint32_t* list) {
  if (max_rays >= (1ull << 31) || gl.chunk == 0 || gl.chunk > (1u << 20)) return hipErrorInvalidValue;
  if (max_rays == 0) return hipSuccess;
  const size_t lds = stack_bytes(scene, gl.block);
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  auto kernel = gl.min_waves >= 8 ? k_any_stream<8, CLOSEST> : k_any_stream<1, CLOSEST>;
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)gl.block, lds);
  if (e != hipSuccess) return e;
  if (gl.blocks_per_cu) per_cu = (int)gl.blocks_per_cu;
  uint64_t grid = std::max<uint64_t>(1, (uint64_t)std::max(1, cus) * (uint64_t)std::max(1, per_cu) / std::max(1u, gl.share));
  const uint64_t useful = (max_rays + gl.block - 1) / gl.block;
  if (grid > useful) grid = useful;
  // the 32-bit queue may pass the ray count by one run per wave before every wave has drained
  if (max_rays + (uint64_t)gl.chunk * grid * (gl.block / 64) >= (1ull << 32)) return hipErrorInvalidValue;
  if (!list && gl.order && (njobs == 0 || njobs >= (1ull << 32))) return hipErrorInvalidValue;
  if (list && !total) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(gl.block), lds, s, scene, rays, total, contrib, queue, gl.chunk,
                     std::max(1u, std::min(64u, gl.refill)), gl.order ? 1u : 0u, rs, (uint32_t)njobs, (uint32_t)max_rays, list);
  return hipGetLastError();
}
hipError_t launch_any_stream(hipStream_t s, DevScene scene, const float4* rays, const uint64_t* total, uint64_t max_rays,
                             const uint64_t* rs, uint64_t njobs, float* contrib, unsigned int* queue, const GatherLaunch& gl,
                             const uint32_t* list) {
  return launch_ray_stream<false>(s, scene, rays, total, max_rays, rs, njobs, contrib, queue, gl, list);
}
hipError_t launch_closest_stream(hipStream_t s, DevScene scene, const float4* rays, const uint64_t* total, uint64_t max_rays,
                                 uint64_t njobs, float4* hits, unsigned int* queue, const GatherLaunch& gl) {
  return launch_ray_stream<true>(s, scene, rays, total, max_rays, nullptr, njobs, reinterpret_cast<float*>(hits), queue, gl, nullptr);
}
hipError_t launch_ray_list(hipStream_t s, const float4* rays, uint64_t n, uint32_t slots, const uint64_t* rs, uint32_t order,
                           uint64_t* count, uint64_t* off, uint32_t* list, void** tmp, size_t* tmp_bytes) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_ray_list, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, rays, n, slots, rs, order, 0, count,
                     (const uint64_t*)nullptr, (uint32_t*)nullptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = scan_u64(s, count, off, n, tmp, tmp_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ray_list, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, rays, n, slots, rs, order, 1, count,
                     (const uint64_t*)off, list);
  return hipGetLastError();
}
hipError_t launch_gather_list(hipStream_t s, const uint32_t* tmask, uint32_t words, const uint64_t* toff, const uint64_t* rs,
                              uint64_t n, uint32_t order, uint32_t* list) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_list, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, tmask, words, toff, rs, n, order, list);
  return hipGetLastError();
}
hipError_t launch_gather_sum(hipStream_t s, DevEntries ent, const DevLight* lights, DevTexels tx, DevBounce b, const uint2* jobs,
                             uint64_t n, const uint64_t* rs, const float* contrib, uint32_t order, const uint32_t* tmask,
                             uint32_t words) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_sum, dim3(grid_for(n, kThreads)), dim3(kThreads), 0, s, ent, lights, tx, b, jobs, n, rs, contrib, order,
                     tmask, words);
  return hipGetLastError();
}
hipError_t launch_bounce_filter(hipStream_t s, DevBounce b, const uint32_t* win_mesh_res, const uint64_t* tile_start,
                                uint32_t nwin, uint64_t ntiles, BlurWeights weights) {
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bounce_filter, dim3((uint32_t)std::min<uint64_t>(ntiles, 1u << 20)), dim3(256), 0, s, b, win_mesh_res,
                     tile_start, nwin, ntiles, weights);
  return hipGetLastError();
}
hipError_t launch_bounce_has(hipStream_t s, DevTris tris, DevEntries ent, DevBounce b, uint8_t* has_bounce) {
  if (ent.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bounce_has, dim3(grid_for(ent.n, kThreads / 64)), dim3(kThreads), 0, s, tris, ent, b, has_bounce);
  return hipGetLastError();
}
hipError_t launch_bounce_dump(hipStream_t s, DevTris tris, DevEntries ent, DevBounce b, const uint64_t* kept_entry, uint64_t nkept,
                              const uint64_t* value_start, float* out) {
  if (nkept == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bounce_dump, dim3((uint32_t)std::min<uint64_t>(nkept, 1u << 20)), dim3(256), 0, s, tris, ent, b, kept_entry,
                     nkept, value_start, out);
  return hipGetLastError();
}

hipError_t launch_bounce_pack(hipStream_t s, DevTris tris, DevMeshes meshes, const DevLight* lights, DevEntries ent,
                              DevBounce b, const uint64_t* kept_entry, const uint64_t* word_start, uint64_t nkept,
                              uint64_t total_words, const uint64_t* kept_out, const int32_t* comp_of_slot,
                              uint32_t colored, uint64_t seed, int dithered, uint32_t* out) {
  if (nkept == 0) return hipSuccess;
  if (dithered) {
    hipLaunchKernelGGL(k_bounce_pack_seq, dim3(grid_for(nkept, 64)), dim3(64), 0, s, tris, meshes, lights, ent, b, kept_entry,
                       nkept, kept_out, word_start, comp_of_slot, colored, seed, out);
  } else {
    if (total_words == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bounce_pack_words, dim3(grid_for(total_words, kThreads)), dim3(kThreads), 0, s, tris, ent, b,
                       kept_entry, word_start, nkept, total_words, kept_out, comp_of_slot, colored, out);
  }
  return hipGetLastError();
}

}  // namespace dl

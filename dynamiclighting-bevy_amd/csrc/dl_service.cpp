// The live scene of a rebake service: the resident meshes baked with the current light set, kept
// on the device and updated edit by edit.
//
// bake() (baker.rs:27-203) couples a mesh's lights in two places only: the light slots (the
// light BVH's leaf order, baker.rs:35-50) that every triangle's light data and entry order use,
// and the mesh-wide `visited` grid that seam padding reads (mesh_bake.rs:246, seams.rs:15-39).
// Everything else of a (mesh, light) pair depends on that light and the geometry alone: its
// region's levels before padding, its own visited texels, which triangle entries optimize keeps
// (shadows.rs:54-74, on unpadded levels), its bounce payloads (the dither streams are keyed by
// light id, bounce_pack.rs:82-84) and its ray counts.  So the service keeps those per pair and an
// edit
//
//   1. traces only the changed lights against the resident geometry (an ordinary resident bake
//      of those lights, whose per-pair results are captured),
//   2. replaces the changed lights' pairs, and for every mesh whose pairs changed ORs the pairs'
//      visited bits into the mesh's grid, pads and crops every region of the mesh again,
//   3. writes every mesh's words with the new slots: changed meshes from their new payloads,
//      the others by copying their payloads (in the new entry order),
//
// which gives exactly the words, stats and lights of bake() of the edited light set.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_set>

#include "../../include/dynlight.h"
#include "dl_host.h"

using namespace dl;

namespace {

struct Slots {
  std::vector<const dl_light*> order;  // slot order: local lights in light-BVH order, then suns, then skies
  std::vector<dl_light_bvh_node> bvh;
};

// baker.rs:35-50 on a light set sorted by id.
Slots slots_of(const std::vector<dl_light>& lights) {
  std::vector<const dl_light*> local, suns, skies;
  for (const dl_light& l : lights) {
    if (l.shape == DL_SHAPE_SUN) suns.push_back(&l);
    else if (l.shape == DL_SHAPE_SKY) skies.push_back(&l);
    else local.push_back(&l);
  }
  std::vector<f3> pos;
  std::vector<float> rad;
  for (const dl_light* l : local) {
    pos.push_back(mk3(l->position[0], l->position[1], l->position[2]));
    rad.push_back(l->radius);
  }
  LightBvhOut lb;
  build_light_bvh(pos, rad, lb);
  Slots s;
  for (uint32_t i : lb.order) s.order.push_back(local[i]);
  for (const dl_light* l : suns) s.order.push_back(l);
  for (const dl_light* l : skies) s.order.push_back(l);
  s.bvh = lb.nodes;
  return s;
}

// BakedLight of a light (baker.rs:154-165): what its shape keeps, its shadow bits and bounce
// compression.
dl_baked_light baked_of(const dl_light& l, const dl_bake_options& opt) {
  const DevLight d = make_dev_light(l, opt);
  dl_baked_light b{};
  b.id = l.id;
  std::memcpy(b.origin, l.position, 12);
  b.radius = l.radius;
  b.shape = l.shape;
  const float identity[4] = {0.0f, 0.0f, 0.0f, 1.0f};
  std::memcpy(b.rotation, identity, 16);
  if (l.shape == DL_SHAPE_RECTANGLE || l.shape == DL_SHAPE_TUBE) {
    std::memcpy(b.rotation, l.rotation, 16);
    std::memcpy(b.size, l.size, 8);
  } else if (l.shape == DL_SHAPE_SUN) {
    std::memcpy(b.rotation, l.rotation, 16);
    b.angular_diameter = l.angular_diameter;
  }
  b.shadow_bits_per_texel = d.bits;
  b.bounce_compression = d.bakes_bounce ? (l.bounce_compression >= 0 ? l.bounce_compression : opt.bounce_compression) : -1;
  return b;
}

uint64_t shadow_words(const uint32_t* crop, uint32_t bits) {
  const uint64_t texels = (uint64_t)(crop[2] - crop[0] + 5) * (crop[3] - crop[1] + 5);
  const uint64_t per = 32 / bits;
  return (texels + per - 1) / per;
}

// Word ranges to copy from one source buffer: (src, dst, len) triples, merged when contiguous.
struct Ranges {
  std::vector<uint64_t> r;
  void add(uint64_t src, uint64_t dst, uint64_t len) {
    if (!len) return;
    const size_t n = r.size();
    if (n && r[n - 3] + r[n - 1] == src && r[n - 2] + r[n - 1] == dst) {
      r[n - 1] += len;
      return;
    }
    r.insert(r.end(), {src, dst, len});
  }
  void append(const Ranges& o) {
    for (size_t i = 0; i < o.r.size(); i += 3) add(o.r[i], o.r[i + 1], o.r[i + 2]);
  }
};

// Runs word copies grouped by source buffer on the context's stream.
struct CopyPlan {
  std::unordered_map<const uint32_t*, Ranges> by_src;
  const uint32_t* last = nullptr;  // the source of the previous add (runs of adds share one)
  Ranges* last_ranges = nullptr;
  void add(const uint32_t* src, uint64_t s, uint64_t d, uint64_t len) {
    if (!last_ranges || src != last) {
      last = src;
      last_ranges = &by_src[src];
    }
    last_ranges->add(s, d, len);
  }
  hipError_t run(dl_ctx* ctx, uint32_t* dst) {
    for (auto& kv : by_src) {
      const std::vector<uint64_t>& r = kv.second.r;
      const uint64_t n = r.size() / 3;
      if (!n) continue;
      std::vector<uint64_t> start(n + 1, 0);
      for (uint64_t k = 0; k < n; ++k) start[k + 1] = start[k] + r[3 * k + 2];
      DBuf<uint64_t> d_r, d_s;
      hipError_t e;
      if ((e = upload(ctx, d_r, r)) || (e = upload(ctx, d_s, start))) return e;
      if ((e = launch_copy_ranges(ctx->stream, kv.first, dst, d_r.p, d_s.p, n, start[n]))) return e;
    }
    return hipSuccess;
  }
};

struct Entry {
  uint32_t tri, slot, pair, k;
};

// Phase times of an edit on stderr when DL_SERVICE_TIMING is set (a profiling aid).
struct PhaseClock {
  const bool on = std::getenv("DL_SERVICE_TIMING") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  std::string line;
  void mark(const char* name, hipStream_t s = nullptr) {
    if (!on) return;
    if (s) (void)hipStreamSynchronize(s);
    const auto now = std::chrono::steady_clock::now();
    line += std::string(" ") + name + "=" + std::to_string(std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
  ~PhaseClock() {
    if (on) std::fprintf(stderr, "[assemble]%s\n", line.c_str());
  }
};

// Rewrites the live scene after its pairs changed: `dirty` meshes are padded and cropped again;
// every mesh gets its words in the new slot order (a mesh whose slots and pairs are unchanged is
// one copy).  The work on the host and the device grows with the meshes that change, not with
// the scene: pair levels stay in the blocks they were captured in, payload offsets are kept
// relative to their mesh, and the stats are kept up to date by merge.  Replaces svc->W.
int assemble(dl_service* svc, const std::vector<uint8_t>& dirty) {
  dl_ctx* ctx = svc->ctx;
  hipStream_t s = ctx->stream;
  ctx->cache.stream = s;
  const CacheScope cache_scope(&ctx->cache);  // stream-ordered buffers: no device-wide frees
  const dl_bake_options& opt = svc->options;
  const uint32_t nm = (uint32_t)svc->mesh.size();
  PhaseClock clock;
  const Slots sl = slots_of(svc->lights);
  std::unordered_map<uint64_t, uint32_t> slot;
  bool bounce = false;
  std::vector<dl_baked_light> baked;
  for (uint32_t i = 0; i < sl.order.size(); ++i) {
    slot[sl.order[i]->id] = i;
    baked.push_back(baked_of(*sl.order[i], opt));
    bounce |= baked.back().bounce_compression >= 0;
  }
  const uint32_t stride = bounce ? 3u : 2u;
  const bool restride = stride != svc->stride;

  // ---- per mesh, on the host: entry order, word layout, light data ------------------------
  struct Plan {
    bool whole = false;            // words unchanged (or only their slot fields): one copy of the mesh
    bool remap = false;            // whole, with the slot fields of its light data renumbered
    std::vector<uint32_t> ld;      // headers and light data (the mesh's first words)
    uint64_t len = 0;
    std::vector<Entry> ent;        // kept entries in word order
    std::vector<uint64_t> sh, bo;  // per entry: shadow and bounce payload offset in the mesh
  };
  std::vector<Plan> plan(nm);
  std::atomic<uint32_t> restaged{0}, remapped{0};
  run_parallel(nm, ctx->host_threads, [&](uint64_t mi) {
    const LiveMesh& m = svc->mesh[mi];
    Plan& p = plan[mi];
    bool same = !dirty[mi] && !restride && svc->W;
    for (const LivePair& q : m.pairs) {
      if (!same) break;
      auto it = svc->slot.find(q.light_id);
      same = it != svc->slot.end() && it->second == slot.at(q.light_id);
    }
    const uint32_t ft = svc->first_tri[mi], nt = svc->first_tri[mi + 1] - ft;
    if (same) {
      p.whole = true;
      p.len = m.w_len;
      return;
    }
    if (!dirty[mi] && !restride && svc->W) {
      // pairs unchanged, slots renumbered: when the mesh's lights keep their order, every
      // triangle's entries (slots ascending) keep theirs, so the words change in the slot fields
      // only (renumbered on the device after the copy)
      std::vector<std::pair<uint32_t, uint32_t>> os;
      os.reserve(m.pairs.size());
      bool known = true;
      for (const LivePair& q : m.pairs) {
        auto it = svc->slot.find(q.light_id);
        if (it == svc->slot.end()) {
          known = false;
          break;
        }
        os.push_back({it->second, slot.at(q.light_id)});
      }
      if (known) {
        std::sort(os.begin(), os.end());
        bool kept = true;
        for (size_t j = 1; j < os.size() && kept; ++j) kept = os[j - 1].second < os[j].second;
        if (kept) {
          p.whole = p.remap = true;
          p.len = m.w_len;
          remapped++;
          return;
        }
      }
    }
    if (!dirty[mi]) restaged++;
    // entries in (triangle, slot) order: the pairs in slot order, then a counting sort by
    // triangle (a pair keeps a triangle at most once)
    const uint32_t np = (uint32_t)m.pairs.size();
    std::vector<uint32_t> pslot(np), porder(np);
    for (uint32_t pi = 0; pi < np; ++pi) {
      pslot[pi] = slot.at(m.pairs[pi].light_id);
      porder[pi] = pi;
    }
    std::sort(porder.begin(), porder.end(), [&](uint32_t a, uint32_t b) { return pslot[a] < pslot[b]; });
    std::vector<uint32_t> count(nt, 0), at(nt, 0);
    size_t ne = 0;
    for (const LivePair& q : m.pairs) {
      for (uint32_t t : q.kept) count[t]++;
      ne += q.kept.size();
    }
    for (uint32_t t = 1; t < nt; ++t) at[t] = at[t - 1] + count[t - 1];
    p.ent.resize(ne);
    for (uint32_t pi : porder) {
      const LivePair& q = m.pairs[pi];
      for (uint32_t k = 0; k < q.kept.size(); ++k) p.ent[at[q.kept[k]]++] = Entry{q.kept[k], pslot[pi], pi, k};
    }
    std::vector<uint32_t> ld_off(nt);
    uint64_t w = 4ull * nt;
    for (uint32_t t = 0; t < nt; ++t) {
      ld_off[t] = (uint32_t)w;
      w += 1 + (uint64_t)stride * count[t];
    }
    p.ld.assign(w, 0u);
    for (uint32_t t = 0; t < nt; ++t) {  // triangle_words.rs:47-89
      const uint32_t* c = &svc->crops[4 * (size_t)(ft + t)];
      p.ld[4 * t] = ld_off[t];
      p.ld[4 * t + 1] = c[0] - 2u;
      p.ld[4 * t + 2] = c[1] - 2u;
      p.ld[4 * t + 3] = c[2] - c[0] + 1 + 4;
      p.ld[ld_off[t]] = count[t];
    }
    p.sh.resize(p.ent.size());
    p.bo.resize(p.ent.size());
    std::vector<uint32_t> filled(nt, 0);
    for (size_t j = 0; j < p.ent.size(); ++j) {
      const Entry& e = p.ent[j];
      const LivePair& q = m.pairs[e.pair];
      p.sh[j] = w;
      w += shadow_words(&svc->crops[4 * (size_t)(ft + e.tri)], q.bits);
      const uint64_t bw = q.bounce_words[e.k];
      p.bo[j] = bw ? w : 0;
      w += bw;
      uint32_t* le = &p.ld[ld_off[e.tri] + 1 + (uint64_t)stride * filled[e.tri]++];
      le[0] = e.slot;
      le[1] = (uint32_t)p.sh[j];
      if (stride == 3) le[2] = (uint32_t)p.bo[j];
    }
    p.len = w;
  });
  clock.mark("plan");
  std::vector<uint64_t> base(nm + 1, 0);
  for (uint32_t i = 0; i < nm; ++i) base[i + 1] = base[i] + plan[i].len;
  const uint64_t total = base[nm];

  // ---- the copies into the new words ------------------------------------------------------
  std::unique_ptr<DevBlock> W2(new DevBlock());
  HIP_TRY(W2->alloc(total + 1, s));
  // light data of the rewritten meshes, staged in one buffer; the copies of each rewritten mesh
  // are listed in parallel and then grouped by source
  std::vector<uint64_t> st_off(nm + 1, 0);
  for (uint32_t i = 0; i < nm; ++i) st_off[i + 1] = st_off[i] + (plan[i].whole ? 0 : plan[i].ld.size());
  std::vector<uint32_t> staged(st_off[nm]);
  struct Copy {
    const uint32_t* src;
    uint64_t s, d, len;
  };
  std::vector<std::vector<Copy>> mesh_copies(nm);
  run_parallel(nm, ctx->host_threads, [&](uint64_t i) {
    const Plan& p = plan[i];
    const LiveMesh& m = svc->mesh[i];
    if (p.whole) return;
    std::copy(p.ld.begin(), p.ld.end(), staged.begin() + (long)st_off[i]);
    std::vector<Copy>& out = mesh_copies[i];
    for (size_t j = 0; j < p.ent.size(); ++j) {
      const Entry& e = p.ent[j];
      const LivePair& q = m.pairs[e.pair];
      if (!dirty[i]) {
        const uint32_t* c = &svc->crops[4 * (size_t)(svc->first_tri[i] + e.tri)];
        out.push_back({svc->W->p, m.w_off + q.shadow_off[e.k], base[i] + p.sh[j], shadow_words(c, q.bits)});
      }
      if (q.bounce_words[e.k]) {
        if (q.src_words) out.push_back({q.src_words->p, q.bounce_off[e.k], base[i] + p.bo[j], q.bounce_words[e.k]});
        else out.push_back({svc->W->p, m.w_off + q.bounce_off[e.k], base[i] + p.bo[j], q.bounce_words[e.k]});
      }
    }
  });
  CopyPlan words;
  for (uint32_t i = 0; i < nm; ++i) {
    const Plan& p = plan[i];
    const LiveMesh& m = svc->mesh[i];
    if (p.whole) {
      words.add(svc->W->p, m.w_off, base[i], m.w_len);
      continue;
    }
    words.add(nullptr, st_off[i], base[i], p.ld.size());  // resolved to the staging buffer below
    for (const Copy& c : mesh_copies[i]) words.add(c.src, c.s, c.d, c.len);
  }
  DBuf<uint32_t> d_staged;
  HIP_TRY(upload(ctx, d_staged, staged));
  if (words.by_src.count(nullptr)) {
    words.by_src[d_staged.p] = std::move(words.by_src[nullptr]);
    words.by_src.erase(nullptr);
    words.last_ranges = nullptr;
  }
  clock.mark("word_ranges");
  if (clock.on) {
    size_t nr = 0;
    for (auto& kv : words.by_src) nr += kv.second.r.size() / 3;
    clock.line += " (ranges " + std::to_string(nr) + ", staged words " + std::to_string(staged.size()) + ")";
  }
  HIP_TRY(words.run(ctx, W2->p));
  {  // the slot fields of the meshes copied whole with renumbered slots
    std::vector<uint64_t> rbase;
    std::vector<uint32_t> rnt;
    for (uint32_t i = 0; i < nm; ++i)
      if (plan[i].remap) {
        rbase.push_back(base[i]);
        rnt.push_back(svc->first_tri[i + 1] - svc->first_tri[i]);
      }
    if (!rbase.empty()) {
      uint32_t old_n = 0;
      for (const auto& kv : svc->slot) old_n = std::max(old_n, kv.second + 1);
      std::vector<uint32_t> table(old_n, 0u);
      for (const auto& kv : svc->slot) {
        auto it = slot.find(kv.first);
        if (it != slot.end()) table[kv.second] = it->second;
      }
      DBuf<uint64_t> d_rbase;
      DBuf<uint32_t> d_rnt, d_table;
      HIP_TRY(upload(ctx, d_rbase, rbase));
      HIP_TRY(upload(ctx, d_rnt, rnt));
      HIP_TRY(upload(ctx, d_table, table));
      HIP_TRY(launch_remap_slots(s, W2->p, d_rbase.p, d_rnt.p, (uint32_t)rbase.size(), stride, d_table.p));
    }
  }
  clock.mark("word_copies", s);

  // ---- visit counts: the pending changes of merge, additions first ---------------------------
  // (so that no count passes below zero on the way)
  for (int sign : {+1, -1}) {
    std::vector<VisDelta>& list = sign > 0 ? svc->vis_add : svc->vis_sub;
    if (list.empty()) continue;
    std::vector<uint64_t> vaddr, coff, start(1, 0);
    std::vector<uint32_t> rect, vres;
    for (const VisDelta& v : list) {
      vaddr.push_back(reinterpret_cast<uint64_t>(v.block->p + v.vis_off));
      rect.insert(rect.end(), v.rect, v.rect + 4);
      vres.push_back(svc->mesh_res[v.mesh]);
      coff.push_back(svc->vc_off[v.mesh]);
      start.push_back(start.back() + v.vis_words);
    }
    DBuf<uint64_t> d_vaddr, d_coff, d_start;
    DBuf<uint32_t> d_rect, d_vres;
    HIP_TRY(upload(ctx, d_vaddr, vaddr));
    HIP_TRY(upload(ctx, d_coff, coff));
    HIP_TRY(upload(ctx, d_start, start));
    HIP_TRY(upload(ctx, d_rect, rect));
    HIP_TRY(upload(ctx, d_vres, vres));
    HIP_TRY(launch_vis_count(s, d_vaddr.p, d_rect.p, d_vres.p, d_coff.p, (uint32_t)list.size(), d_start.p, start.back(), svc->vc->p,
                             sign));
    HIP_TRY(hipStreamSynchronize(s));  // the deltas' blocks may be freed after this
    list.clear();
  }
  clock.mark("counts");

  // ---- dirty meshes: visited grids from the counts, seam padding, crop and dilation ---------
  // The levels of the dirty meshes' used pairs (those with kept entries, the only ones padded
  // and cropped) are gathered into a work buffer: padding changes them in place, and the pairs
  // keep their unpadded levels.
  {
    std::vector<uint32_t> dm;
    for (uint32_t i = 0; i < nm; ++i)
      if (dirty[i] && !plan[i].whole) dm.push_back(i);
    const uint32_t nd = (uint32_t)dm.size();
    std::vector<uint32_t> r_rect, r_bits, r_mesh, m_res32, used;
    std::vector<uint64_t> r_off, used_area(1, 0), m_voff(1, 0), m_coff;
    std::vector<uint32_t> dm_r0(nd);
    std::vector<uint64_t> dm_e0(nd + 1, 0);  // first kept entry of each dirty mesh
    std::vector<std::vector<uint32_t>> used_of(nd);  // per dirty mesh: pair -> used region
    CopyPlan levels;
    uint64_t at = 0;
    for (uint32_t d = 0; d < nd; ++d) {
      const uint32_t i = dm[d];
      const LiveMesh& m = svc->mesh[i];
      const uint32_t rs = svc->mesh_res[i];
      m_res32.push_back(rs);
      m_voff.push_back(m_voff.back() + ((uint64_t)rs * rs + 31) / 32);
      m_coff.push_back(svc->vc_off[i]);
      dm_r0[d] = (uint32_t)r_bits.size();
      used_of[d].assign(m.pairs.size(), 0);
      for (size_t pi = 0; pi < m.pairs.size(); ++pi) {
        const LivePair& q = m.pairs[pi];
        if (q.kept.empty()) continue;
        used_of[d][pi] = (uint32_t)r_bits.size() - dm_r0[d];
        used.push_back((uint32_t)r_bits.size());
        r_rect.insert(r_rect.end(), q.rect, q.rect + 4);
        r_bits.push_back(q.bits);
        r_mesh.push_back(d);
        levels.add(q.levels->p, q.grid_off, at, q.grid_words);
        r_off.push_back(at);
        at += q.grid_words;
        used_area.push_back(used_area.back() + (uint64_t)(q.rect[2] - q.rect[0] + 1) * (q.rect[3] - q.rect[1] + 1));
      }
      dm_e0[d + 1] = dm_e0[d] + plan[i].ent.size();
    }
    clock.mark("dirty_pairs");
    // the kept entries of the dirty meshes, filled mesh by mesh in parallel
    const uint64_t nkept = dm_e0.back();
    std::vector<uint32_t> k_tri(nkept), k_region(nkept);
    std::vector<uint64_t> k_words(nkept), k_out(nkept), k_start(nkept + 1, 0);
    run_parallel(nd, ctx->host_threads, [&](uint64_t d) {
      const uint32_t i = dm[d];
      const LiveMesh& m = svc->mesh[i];
      const Plan& p = plan[i];
      for (size_t j = 0; j < p.ent.size(); ++j) {
        const Entry& e = p.ent[j];
        const uint64_t k = dm_e0[d] + j;
        k_tri[k] = svc->first_tri[i] + e.tri;
        k_region[k] = dm_r0[d] + used_of[d][e.pair];
        k_words[k] = shadow_words(&svc->crops[4 * (size_t)k_tri[k]], m.pairs[e.pair].bits);
        k_out[k] = base[i] + p.sh[j];
      }
    });
    for (uint64_t k = 0; k < nkept; ++k) k_start[k + 1] = k_start[k] + k_words[k];
    const uint32_t nreg = (uint32_t)r_bits.size();
    clock.mark("dirty_host");
    if (clock.on) clock.line += " (used regions " + std::to_string(nreg) + ", kept " + std::to_string(nkept) + ")";
    if (nd) {
      DBuf<uint32_t> d_visited, d_mres32;
      DBuf<uint64_t> d_mvoff, d_mcoff;
      HIP_TRY(upload(ctx, d_mres32, m_res32));
      HIP_TRY(upload(ctx, d_mvoff, m_voff));
      HIP_TRY(upload(ctx, d_mcoff, m_coff));
      HIP_TRY(d_visited.alloc(m_voff.back() + 1));
      HIP_TRY(launch_visited_from_counts(s, d_mres32.p, d_mcoff.p, nd, d_mvoff.p, m_voff.back(), svc->vc->p, d_visited.p));
      if (nreg) {
        std::unique_ptr<DevBlock> work(new DevBlock());
        HIP_TRY(work->alloc(at + 1, s));
        HIP_TRY(levels.run(ctx, work->p));
        DBuf<uint32_t> d_rect, d_bits, d_mesh, d_used, d_ktri, d_kregion;
        DBuf<uint64_t> d_off, d_uarea, d_kstart, d_kout, d_kid;
        HIP_TRY(upload(ctx, d_rect, r_rect));
        HIP_TRY(upload(ctx, d_bits, r_bits));
        HIP_TRY(upload(ctx, d_mesh, r_mesh));
        HIP_TRY(upload(ctx, d_off, r_off));
        HIP_TRY(upload(ctx, d_used, used));
        HIP_TRY(upload(ctx, d_uarea, used_area));
        DevRegions rg{d_rect.p, d_bits.p, d_off.p, d_mesh.p, work->p};
        DevMeshes ms{};
        ms.n = nd;
        ms.res = d_mres32.p;
        ms.visited_off = d_mvoff.p;
        ms.owned_off = d_mvoff.p;
        ms.visited = d_visited.p;
        clock.mark("dirty_uploads", s);
        HIP_TRY(launch_pad_seams(s, ms, rg, d_used.p, d_uarea.p, (uint32_t)used.size(), used_area.back()));
        clock.mark("dirty_pad", s);
        if (nkept) {
          HIP_TRY(upload(ctx, d_ktri, k_tri));
          HIP_TRY(upload(ctx, d_kregion, k_region));
          HIP_TRY(upload(ctx, d_kstart, k_start));
          HIP_TRY(upload(ctx, d_kout, k_out));
          HIP_TRY(d_kid.alloc(nkept));
          HIP_TRY(launch_iota_u64(s, d_kid.p, nkept));
          DevTris tr{};
          tr.n = (uint32_t)svc->ntri;
          tr.crop = svc->d_crops->p;
          DevEntries ent{nkept, nullptr, d_kregion.p, d_ktri.p, nullptr};
          HIP_TRY(launch_crop(s, tr, ent, rg, d_kid.p, d_kstart.p, nkept, k_start.back(), d_kout.p, W2->p));
        }
        HIP_TRY(hipStreamSynchronize(s));
      }
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  clock.mark("dirty");

  // ---- the new state ------------------------------------------------------------------------
  run_parallel(nm, ctx->host_threads, [&](uint64_t i) {
    LiveMesh& m = svc->mesh[i];
    const Plan& p = plan[i];
    if (!p.whole) {
      for (size_t j = 0; j < p.ent.size(); ++j) {
        LivePair& q = m.pairs[p.ent[j].pair];
        q.shadow_off[p.ent[j].k] = p.sh[j];
        if (q.bounce_words[p.ent[j].k]) q.bounce_off[p.ent[j].k] = p.bo[j];
      }
      for (LivePair& q : m.pairs) q.src_words.reset();
    }
    m.w_off = base[i];
    m.w_len = p.len;
  });
  svc->W = std::move(W2);
  svc->slot = std::move(slot);
  svc->baked = std::move(baked);
  svc->bvh = sl.bvh;
  svc->stride = stride;
  svc->stats.triangle_words = total;
  clock.mark("state");
  if (clock.on) clock.line += " restaged_meshes=" + std::to_string(restaged.load()) + " remapped_meshes=" + std::to_string(remapped.load());
  return DL_OK;
}

// A pair's visited bits, to add to or take from its mesh's visit counts.
VisDelta vis_delta(uint32_t mesh, const LivePair& q) {
  VisDelta v;
  v.mesh = mesh;
  std::memcpy(v.rect, q.rect, 16);
  v.block = q.levels;
  v.vis_off = q.vis_off;
  v.vis_words = q.vis_words;
  return v;
}

// Adds (sign +1) or removes (-1) a pair's share of the live stats.
void count_pair(dl_bake_stats& st, const LivePair& q, int sign) {
  const uint64_t removed = q.entries - q.kept.size();
  if (sign > 0) {
    st.rays += q.rays;
    st.samples += q.samples;
    st.bounce_rays += q.bounce_rays;
    st.lights_removed += removed;
  } else {
    st.rays -= q.rays;
    st.samples -= q.samples;
    st.bounce_rays -= q.bounce_rays;
    st.lights_removed -= removed;
  }
}

// The captured pairs of a bake of `lights` replace the live pairs of those lights (and of
// `removed`); marks the meshes whose pairs changed and keeps the stats up to date.
void merge(dl_service* svc, Capture& cap, const std::unordered_set<uint64_t>& replaced, std::vector<uint8_t>& dirty) {
  // the meshes holding pairs of the replaced lights (svc->light_meshes indexes them)
  std::vector<uint32_t> hit;
  for (uint64_t id : replaced) {
    auto it = svc->light_meshes.find(id);
    if (it == svc->light_meshes.end()) continue;
    hit.insert(hit.end(), it->second.begin(), it->second.end());
    svc->light_meshes.erase(it);
  }
  std::sort(hit.begin(), hit.end());
  hit.erase(std::unique(hit.begin(), hit.end()), hit.end());
  for (uint32_t i : hit) {
    auto& pairs = svc->mesh[i].pairs;
    const size_t before = pairs.size();
    for (const LivePair& q : pairs)
      if (replaced.count(q.light_id)) {
        count_pair(svc->stats, q, -1);
        svc->vis_sub.push_back(vis_delta(i, q));
      }
    pairs.erase(std::remove_if(pairs.begin(), pairs.end(), [&](const LivePair& q) { return replaced.count(q.light_id) > 0; }),
                pairs.end());
    if (pairs.size() != before) dirty[i] = 1;
  }
  for (CapturedPair& c : cap.pairs) {
    LivePair q;
    q.light_id = c.light_id;
    std::memcpy(q.rect, c.rect, 16);
    q.bits = c.bits;
    q.grid_off = c.grid_off;
    q.grid_words = c.grid_words;
    q.vis_off = c.vis_off;
    q.vis_words = c.vis_words;
    q.rays = c.rays;
    q.samples = c.samples;
    q.bounce_rays = c.bounce_rays;
    q.entries = c.entries;
    q.kept = std::move(c.kept);
    q.bounce_off = std::move(c.bounce_off);
    q.bounce_words = std::move(c.bounce_words);
    q.shadow_off.assign(q.kept.size(), 0);
    q.levels = c.block;
    q.src_words = c.words;
    count_pair(svc->stats, q, +1);
    svc->vis_add.push_back(vis_delta(c.mesh, q));
    svc->light_meshes[c.light_id].push_back(c.mesh);
    auto& pairs = svc->mesh[c.mesh].pairs;
    pairs.insert(std::upper_bound(pairs.begin(), pairs.end(), q.light_id,
                                  [](uint64_t id, const LivePair& p) { return id < p.light_id; }),
                 std::move(q));
    dirty[c.mesh] = 1;
  }
  for (auto& kv : cap.cubes) svc->cubes[kv.first] = std::move(kv.second);
}

// Bakes `lights` (the changed ones) against the resident geometry and captures their pairs.
// Bounds of every vertex of a bake input; false when it has none.
bool bounds_of(const dl_bake_input* in, float bmin[3], float bmax[3]) {
  bool any = false;
  for (int a = 0; a < 3; ++a) {
    bmin[a] = INFINITY;
    bmax[a] = -INFINITY;
  }
  for (uint32_t m = 0; m < in->mesh_count; ++m) {
    const dl_mesh& mesh = in->meshes[m];
    if (!mesh.positions) continue;
    for (uint32_t v = 0; v < mesh.vertex_count; ++v)
      for (int a = 0; a < 3; ++a) {
        const float x = mesh.positions[3 * (size_t)v + a];
        if (!std::isfinite(x)) continue;
        bmin[a] = std::min(bmin[a], x);
        bmax[a] = std::max(bmax[a], x);
        any = true;
      }
  }
  return any;
}

// Whether the ball of a light's range (which holds its emitter, its rays and its photons) comes
// within a small margin of a part's bounds.
bool reaches(const dl_light& l, const ScenePart& part) {
  double d2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double margin = 1e-3 + 1e-5 * std::fabs((double)part.bmin[a]) + 1e-5 * std::fabs((double)part.bmax[a]);
    const double lo = (double)part.bmin[a] - margin, hi = (double)part.bmax[a] + margin, x = l.position[a];
    const double g = x < lo ? lo - x : x > hi ? x - hi : 0.0;
    d2 += g * g;
  }
  const double r = (double)l.radius * (1.0 + 1e-5) + 1e-3;
  return d2 <= r * r;
}

bool boxes_meet(const ScenePart& a, const ScenePart& b) {
  for (int k = 0; k < 3; ++k)
    if (a.bmax[k] + 1e-3f < b.bmin[k] || b.bmax[k] + 1e-3f < a.bmin[k]) return false;
  return true;
}

// The part whose bounds a light reaches (0 when it reaches none: it then meets no geometry at
// all, so any part traces it alike); -1 when it reaches two or more.
int part_of(const dl_service* svc, const dl_light& l) {
  int found = -2;
  for (size_t k = 0; k < svc->parts.size(); ++k)
    if (reaches(l, svc->parts[k])) found = found == -2 ? (int)k : -1;
  return found == -2 ? 0 : found;
}

DevScene scene_of(const Resident& r) {
  DevScene sc{};
  sc.nodes = r.d_nodes.p;
  sc.tris = reinterpret_cast<const float4*>(r.d_occ.p);
  sc.ntris = r.nocc;
  sc.stack = bvh_stack_entries(r.depth);
  sc.stack16 = r.nodes < 65536 ? 1u : 0u;
  return sc;
}

// How many of n rays (8 floats each, from device address rays8) hit some occluder of a part.
int rays_hitting(dl_ctx* ctx, const Resident& r, const float* rays8, uint64_t n, uint64_t* hits) {
  *hits = 0;
  if (n == 0 || r.nocc == 0) return DL_OK;
  ctx->cache.stream = ctx->stream;
  const CacheScope cache_scope(&ctx->cache);
  DBuf<uint32_t> flags;
  DBuf<unsigned long long> count;
  DBuf<float> unused_t;
  HIP_TRY(flags.alloc(n));
  HIP_TRY(zeroed(ctx, count, 1));
  HIP_TRY(launch_ray_queries(ctx->stream, scene_of(r), rays8, n, 0, unused_t.p, flags.p));
  HIP_TRY(launch_count_nonzero(ctx->stream, flags.p, n, count.p));
  unsigned long long c = 0;
  HIP_TRY(copy_to_host(ctx, &c, count.p, 8));
  *hits = c;
  return DL_OK;
}

// Bakes `lights` (the changed ones) against the resident geometry of the parts they reach and
// captures their pairs (mesh indices are live ones).
int trace_lights(dl_ctx* w, dl_service* svc, const std::vector<dl_light>& lights, Capture& cap) {
  if (lights.empty()) return DL_OK;
  std::vector<std::vector<dl_light>> by_part(svc->parts.size());
  for (const dl_light& l : lights) {
    const int k = part_of(svc, l);
    if (k < 0) return set_error(w, DL_E_ARGUMENT, "light " + hex_id(l.id) + " reaches two parts of the live scene (its rooms)");
    by_part[k].push_back(l);
  }
  for (size_t k = 0; k < by_part.size(); ++k) {
    if (by_part[k].empty()) continue;
    const dl_bake_input in{nullptr, 0, by_part[k].data(), (uint32_t)by_part[k].size()};
    dl_baked_scene* sc = nullptr;
    Capture part;
    const int st = dl_bake_impl(w, &in, &svc->options, nullptr, &sc, svc->parts[k].res, &part);
    delete sc;
    if (st != DL_OK) return st;
    for (CapturedPair& c : part.pairs) {
      c.mesh += svc->parts[k].mesh0;
      cap.pairs.push_back(std::move(c));
    }
    for (auto& c : part.cubes) cap.cubes.push_back(std::move(c));
  }
  return DL_OK;
}

// ---- rooms -----------------------------------------------------------------------------------
// A room joins the live scene as a part of its own.  The bake of the whole scene equals the live
// scene with the room's bake beside it when nothing that bake computes for one part reaches the
// other: no light's range reaches the other part's bounds (shadow rays, area samples, photons and
// gather rays stay inside a light's range ball) and no burial ray (unbounded) that leaves its own
// part hits the other.  join_room checks exactly these, with the burial rays that hit nothing in
// their own part recorded when each part was baked, and refuses the room otherwise.

struct RoomTrace {
  std::unique_ptr<Resident> res;
  Capture cap;
  ScenePart part;
};

// Phase 1 of a room: its bake alone (on a worker), with its pairs and escaping rays captured.
int trace_room(dl_ctx* w, dl_service* svc, const dl_bake_input* room, RoomTrace& t) {
  dl_ctx* ctx = w;  // errors are the worker's
  if (!room || room->mesh_count == 0) return set_error(w, DL_E_ARGUMENT, "a room needs meshes");
  if (!bounds_of(room, t.part.bmin, t.part.bmax)) return set_error(w, DL_E_ARGUMENT, "the room has no finite vertex");
  t.res.reset(new Resident());
  t.cap.escaping = std::make_shared<DevBlock>();
  t.cap.escaping_cap = kRoomEscapingCap;
  HIP_TRY(t.cap.escaping->alloc(8 * kRoomEscapingCap + 2, w->stream));
  dl_baked_scene* sc = nullptr;
  const int st = dl_bake_impl(w, room, &svc->options, nullptr, &sc, t.res.get(), &t.cap);
  delete sc;
  if (st != DL_OK) return st;
  HIP_TRY(hipStreamSynchronize(w->stream));
  t.part.res = t.res.get();
  return DL_OK;
}

// Phase 2 of a room: the isolation checks, then its meshes, lights and pairs join the live state
// (the caller assembles).  Marks the room's meshes dirty (`dirty` grows with the meshes).
int join_room(dl_service* svc, const dl_bake_input* room, RoomTrace& t, std::vector<uint8_t>& dirty,
              std::unordered_set<uint64_t>& added_lights) {
  dl_ctx* ctx = svc->ctx;
  const Resident& rr = *t.res;
  if (rr.meshes.empty()) return set_error(ctx, DL_E_ARGUMENT, "the room has no meshes");
  if (!svc->mesh_meta.empty() && rr.meshes.front().id <= svc->mesh_meta.back().id)
    return set_error(ctx, DL_E_ARGUMENT, "a room's mesh ids must be above every live mesh id");
  std::unordered_set<uint64_t> ids;
  for (const dl_light& l : svc->lights) ids.insert(l.id);
  for (uint32_t k = 0; k < room->light_count; ++k) {
    const dl_light& l = room->lights[k];
    if (l.shape == DL_SHAPE_SUN || l.shape == DL_SHAPE_SKY) return set_error(ctx, DL_E_ARGUMENT, "a room's lights must be local lights");
    if (!ids.insert(l.id).second) return set_error(ctx, DL_E_DUPLICATE_ID, "two lights share BakeId " + hex_id(l.id));
    for (const ScenePart& p : svc->parts)
      if (reaches(l, p)) return set_error(ctx, DL_E_ARGUMENT, "the room's light " + hex_id(l.id) + " reaches the live scene");
  }
  for (const dl_light& l : svc->lights) {
    if (l.shape == DL_SHAPE_SUN || l.shape == DL_SHAPE_SKY)
      return set_error(ctx, DL_E_ARGUMENT, "rooms join live scenes of local lights only (light " + hex_id(l.id) + ")");
    if (reaches(l, t.part)) return set_error(ctx, DL_E_ARGUMENT, "live light " + hex_id(l.id) + " reaches the room");
  }
  for (const ScenePart& p : svc->parts)
    if (boxes_meet(p, t.part)) return set_error(ctx, DL_E_ARGUMENT, "the room's bounds meet the live scene's");
  if (svc->esc_open)
    return set_error(ctx, DL_E_ARGUMENT, "too many burial rays leave the live scene to check a room's isolation (" +
                                             std::to_string(svc->esc_total) + ", at most " + std::to_string(kEscapingCap) + ")");
  if (t.cap.escaping_count > t.cap.escaping_cap)
    return set_error(ctx, DL_E_ARGUMENT, "too many burial rays leave the room to check its isolation (" +
                                             std::to_string(t.cap.escaping_count) + ", at most " +
                                             std::to_string(t.cap.escaping_cap) + ")");
  uint64_t hits = 0;
  int st = DL_OK;
  for (const auto& blk : svc->esc) {
    if ((st = rays_hitting(ctx, rr, reinterpret_cast<const float*>(blk.first->p), blk.second, &hits)) != DL_OK) return st;
    if (hits) return set_error(ctx, DL_E_ARGUMENT, "burial rays of the live scene reach the room");
  }
  for (const ScenePart& p : svc->parts) {
    st = rays_hitting(ctx, *p.res, reinterpret_cast<const float*>(t.cap.escaping->p), t.cap.escaping_count, &hits);
    if (st != DL_OK) return st;
    if (hits) return set_error(ctx, DL_E_ARGUMENT, "burial rays of the room reach the live scene");
  }

  // the room's meshes, triangles and visit counts join the live scene's
  hipStream_t s = ctx->stream;
  const uint32_t mesh0 = (uint32_t)svc->mesh.size(), nrm = (uint32_t)rr.meshes.size();
  const uint64_t tri0 = svc->ntri;
  for (uint32_t i = 0; i < nrm; ++i) {
    svc->mesh_meta.push_back(rr.meshes[i]);
    svc->mesh_res.push_back(rr.res[i]);
    svc->first_tri.push_back((uint32_t)(tri0 + rr.first_tri[i + 1]));
    LiveMesh m;
    m.texels = t.cap.texels[i];
    svc->mesh.push_back(std::move(m));
    svc->stats.texels += t.cap.texels[i];
  }
  svc->ntri += rr.ntri;
  svc->stats.meshes += nrm;
  svc->stats.triangles += rr.ntri;
  svc->crops.insert(svc->crops.end(), t.cap.crops.begin(), t.cap.crops.end());
  {
    // visit counts: the room's zeroed after the live scene's, in a buffer grown by doubling
    std::vector<uint64_t> off(svc->vc_off);
    for (uint32_t i = 0; i < nrm; ++i) off.push_back(off.back() + ((uint64_t)rr.res[i] * rr.res[i] + 1) / 2);
    if (off.back() + 1 > svc->vc->words) {
      std::unique_ptr<DevBlock> vc(new DevBlock());
      HIP_TRY(vc->alloc(std::max<uint64_t>(off.back() + 1, 2 * svc->vc->words), s));
      HIP_TRY(hipMemcpyAsync(vc->p, svc->vc->p, svc->vc_off.back() * 4, hipMemcpyDeviceToDevice, s));
      svc->vc = std::move(vc);
    }
    HIP_TRY(hipMemsetAsync(svc->vc->p + svc->vc_off.back(), 0, (off.back() + 1 - svc->vc_off.back()) * 4, s));
    std::unique_ptr<DevBlock> crops(new DevBlock());
    HIP_TRY(crops->alloc(svc->crops.size() + 1, s));
    HIP_TRY(hipMemcpyAsync(crops->p, svc->crops.data(), svc->crops.size() * 4, hipMemcpyHostToDevice, s));
    // the room's escaping rays, in a block of their own size
    if (t.cap.escaping_count) {
      std::shared_ptr<DevBlock> esc = std::make_shared<DevBlock>();
      HIP_TRY(esc->alloc(8 * t.cap.escaping_count, s));
      HIP_TRY(hipMemcpyAsync(esc->p, t.cap.escaping->p, t.cap.escaping_count * 32, hipMemcpyDeviceToDevice, s));
      svc->esc.emplace_back(std::move(esc), t.cap.escaping_count);
    }
    HIP_TRY(hipStreamSynchronize(s));
    svc->vc_off = std::move(off);
    svc->d_crops = std::move(crops);
  }
  t.part.mesh0 = mesh0;
  svc->parts.push_back(t.part);
  svc->rooms.push_back(std::move(t.res));
  // its lights and pairs
  for (uint32_t k = 0; k < room->light_count; ++k) {
    svc->lights.push_back(room->lights[k]);
    added_lights.insert(room->lights[k].id);
  }
  std::stable_sort(svc->lights.begin(), svc->lights.end(), [](const dl_light& a, const dl_light& b) { return a.id < b.id; });
  for (CapturedPair& c : t.cap.pairs) c.mesh += mesh0;
  dirty.resize(svc->mesh.size(), 0);
  for (uint32_t i = mesh0; i < svc->mesh.size(); ++i) dirty[i] = 1;
  merge(svc, t.cap, {}, dirty);
  return DL_OK;
}

// Checks an edit and returns the edited light set and the ids whose pairs it replaces.
int edited_lights(dl_service* svc, const dl_service_edit& e, std::vector<dl_light>& out, std::unordered_set<uint64_t>& replaced) {
  dl_ctx* ctx = svc->ctx;
  if ((e.light_count && !e.lights) || (e.remove_count && !e.remove)) return set_error(ctx, DL_E_ARGUMENT, "null array in an edit");
  std::unordered_set<uint64_t> current;
  for (const dl_light& l : svc->lights) current.insert(l.id);
  for (uint32_t k = 0; k < e.light_count; ++k) {
    const dl_light& l = e.lights[k];
    if (l.shape == DL_SHAPE_SUN || l.shape == DL_SHAPE_SKY)
      return set_error(ctx, DL_E_ARGUMENT, "edits move or add local lights; light " + hex_id(l.id) + " is a sun or sky");
    if (!replaced.insert(l.id).second) return set_error(ctx, DL_E_DUPLICATE_ID, "two lights share BakeId " + hex_id(l.id));
    int st = check_light(ctx, l);
    if (st) return st;
  }
  for (uint32_t k = 0; k < e.remove_count; ++k) {
    const uint64_t id = e.remove[k];
    if (!current.count(id)) return set_error(ctx, DL_E_ARGUMENT, "no light " + hex_id(id) + " to remove");
    if (!replaced.insert(id).second) return set_error(ctx, DL_E_ARGUMENT, "light " + hex_id(id) + " is edited twice");
    for (const dl_light& l : svc->lights)
      if (l.id == id && (l.shape == DL_SHAPE_SUN || l.shape == DL_SHAPE_SKY))
        return set_error(ctx, DL_E_ARGUMENT, "edits remove local lights; light " + hex_id(id) + " is a sun or sky");
  }
  for (const dl_light& l : svc->lights)
    if (!replaced.count(l.id)) out.push_back(l);
  for (uint32_t k = 0; k < e.light_count; ++k) out.push_back(e.lights[k]);
  std::stable_sort(out.begin(), out.end(), [](const dl_light& a, const dl_light& b) { return a.id < b.id; });
  return DL_OK;
}

double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

namespace dl {

int service_init_live(dl_service* svc, const dl_bake_input* scene, Capture& cap, const dl_baked_scene& first) {
  dl_ctx* ctx = svc->ctx;
  svc->lights.assign(scene->lights, scene->lights + scene->light_count);
  std::stable_sort(svc->lights.begin(), svc->lights.end(), [](const dl_light& a, const dl_light& b) { return a.id < b.id; });
  svc->mesh.assign(svc->res.meshes.size(), LiveMesh{});
  for (size_t i = 0; i < svc->mesh.size(); ++i) svc->mesh[i].texels = cap.texels[i];
  svc->crops = cap.crops;
  svc->mesh_meta = svc->res.meshes;
  svc->mesh_res = svc->res.res;
  svc->first_tri = svc->res.first_tri;
  svc->ntri = svc->res.ntri;
  svc->rooms.clear();
  svc->parts.assign(1, ScenePart{});
  svc->parts[0].res = &svc->res;
  if (!bounds_of(scene, svc->parts[0].bmin, svc->parts[0].bmax))
    for (int a = 0; a < 3; ++a) svc->parts[0].bmin[a] = svc->parts[0].bmax[a] = 0.0f;
  svc->esc.clear();
  if (cap.escaping && cap.escaping_count) svc->esc.emplace_back(cap.escaping, std::min(cap.escaping_count, cap.escaping_cap));
  svc->esc_open = !cap.escaping || cap.escaping_count > cap.escaping_cap;
  svc->esc_total = cap.escaping_count;
  svc->W.reset();
  svc->slot.clear();
  svc->light_meshes.clear();
  svc->vis_add.clear();
  svc->vis_sub.clear();
  {
    hipStream_t s = ctx->stream;
    svc->vc_off.assign(svc->mesh.size() + 1, 0);
    for (size_t i = 0; i < svc->mesh.size(); ++i)
      svc->vc_off[i + 1] = svc->vc_off[i] + ((uint64_t)svc->mesh_res[i] * svc->mesh_res[i] + 1) / 2;
    svc->vc.reset(new DevBlock());
    HIP_TRY(svc->vc->alloc(svc->vc_off.back() + 1, s));
    HIP_TRY(hipMemsetAsync(svc->vc->p, 0, (svc->vc_off.back() + 1) * 4, s));
    svc->d_crops.reset(new DevBlock());
    HIP_TRY(svc->d_crops->alloc(svc->crops.size() + 1, s));
    HIP_TRY(hipMemcpyAsync(svc->d_crops->p, svc->crops.data(), svc->crops.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  svc->stride = 0;  // forces every mesh to be written
  svc->stats = dl_bake_stats{};
  svc->stats.meshes = (uint32_t)svc->mesh.size();
  svc->stats.triangles = svc->ntri;
  for (const LiveMesh& m : svc->mesh) svc->stats.texels += m.texels;
  std::vector<uint8_t> dirty(svc->mesh.size(), 1);
  merge(svc, cap, {}, dirty);
  svc->base.reset(new dl_baked_scene());
  svc->base->settings = first.settings;
  svc->base->sun_maps = first.sun_maps;
  svc->base->probes = first.probes;
  std::memcpy(svc->base->probe_origin, first.probe_origin, sizeof first.probe_origin);
  svc->base->probe_spacing = first.probe_spacing;
  std::memcpy(svc->base->probe_counts, first.probe_counts, sizeof first.probe_counts);
  svc->base->probe_words = first.probe_words;
  int st = assemble(svc, dirty);
  if (st != DL_OK) return st;
  // the assembled words must be the bake's own, mesh by mesh
  std::vector<uint32_t> live;
  HIP_TRY(download(ctx, live, svc->W->p, svc->stats.triangle_words));
  for (size_t i = 0; i < svc->mesh.size(); ++i) {
    const dl_word_slice& w = first.meshes[i].words;
    if (w.size() != svc->mesh[i].w_len || !std::equal(w.begin(), w.end(), live.begin() + (long)svc->mesh[i].w_off))
      return set_error(ctx, DL_E_DEVICE, "the live scene's words differ from its first bake (mesh " + hex_id(first.meshes[i].id) + ")");
  }
  svc->version = 0;
  return DL_OK;
}

}  // namespace dl

extern "C" {

int dl_service_apply(dl_service* svc, const dl_service_edit* edit, dl_service_update* info) {
  if (!svc) return set_error(nullptr, DL_E_ARGUMENT, "null service");
  dl_ctx* ctx = svc->ctx;
  if (!edit) return set_error(ctx, DL_E_ARGUMENT, "null edit");
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(ctx, DL_E_DEVICE, "hipSetDevice failed");
    const auto t0 = std::chrono::steady_clock::now();
    if (edit->room) {
      if (edit->light_count || edit->remove_count) return set_error(ctx, DL_E_ARGUMENT, "an edit adds a room or edits lights, not both");
      RoomTrace t;
      int st = trace_room(ctx, svc, edit->room, t);
      if (st != DL_OK) return st;
      const auto t1 = std::chrono::steady_clock::now();
      std::vector<uint8_t> dirty(svc->mesh.size(), 0);
      std::unordered_set<uint64_t> added;
      if ((st = join_room(svc, edit->room, t, dirty, added)) != DL_OK) return st;
      if ((st = assemble(svc, dirty)) != DL_OK) return st;
      svc->version++;
      if (info) {
        *info = dl_service_update{};
        info->version = svc->version;
        for (uint8_t d : dirty) info->dirty_meshes += d;
        info->traced_pairs = t.cap.pairs.size();
        for (const CapturedPair& c : t.cap.pairs) {
          info->traced_samples += c.samples;
          info->traced_rays += c.rays;
          info->traced_bounce_rays += c.bounce_rays;
        }
        info->trace_ms = ms_between(t0, t1);
        info->assemble_ms = ms_between(t1, std::chrono::steady_clock::now());
        info->total_ms = ms_between(t0, std::chrono::steady_clock::now());
      }
      return DL_OK;
    }
    std::vector<dl_light> next;
    std::unordered_set<uint64_t> replaced;
    int st = edited_lights(svc, *edit, next, replaced);
    if (st != DL_OK) return st;
    std::vector<dl_light> changed(edit->lights, edit->lights + edit->light_count);
    Capture cap;
    if ((st = trace_lights(ctx, svc, changed, cap)) != DL_OK) return st;
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<uint8_t> dirty(svc->mesh.size(), 0);
    merge(svc, cap, replaced, dirty);
    for (uint64_t id : replaced)
      if (!std::any_of(changed.begin(), changed.end(), [&](const dl_light& l) { return l.id == id; })) svc->cubes.erase(id);
    svc->lights = std::move(next);
    if ((st = assemble(svc, dirty)) != DL_OK) return st;
    svc->version++;
    if (info) {
      *info = dl_service_update{};
      info->version = svc->version;
      for (uint8_t d : dirty) info->dirty_meshes += d;
      info->traced_pairs = cap.pairs.size();
      for (const CapturedPair& c : cap.pairs) {
        info->traced_samples += c.samples;
        info->traced_rays += c.rays;
        info->traced_bounce_rays += c.bounce_rays;
      }
      info->trace_ms = ms_between(t0, t1);
      info->assemble_ms = ms_between(t1, std::chrono::steady_clock::now());
      info->total_ms = ms_between(t0, std::chrono::steady_clock::now());
    }
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return set_error(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_error(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

int dl_service_scene(dl_service* svc, dl_baked_scene** out) {
  if (!svc) return set_error(nullptr, DL_E_ARGUMENT, "null service");
  dl_ctx* ctx = svc->ctx;
  if (!out) return set_error(ctx, DL_E_ARGUMENT, "null argument");
  *out = nullptr;
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(ctx, DL_E_DEVICE, "hipSetDevice failed");
    if (!svc->W) return set_error(ctx, DL_E_ARGUMENT, "the service has no live scene");
    std::unique_ptr<dl_baked_scene> sc(new dl_baked_scene());
    const dl_baked_scene& b = *svc->base;
    sc->settings = b.settings;
    sc->settings.bounce = svc->stride == 3;
    sc->settings.colored_bounce = svc->stride == 3 && svc->options.colored_bounce;
    sc->stats = svc->stats;
    sc->lights = svc->baked;
    sc->bvh = svc->bvh;
    auto words = std::make_shared<std::vector<uint32_t, NoInit<uint32_t>>>(svc->stats.triangle_words);
    HIP_TRY(copy_to_host(ctx, words->data(), svc->W->p, svc->stats.triangle_words * 4));
    sc->meshes.resize(svc->mesh.size());
    for (size_t i = 0; i < svc->mesh.size(); ++i) {
      dl_baked_mesh_out& mo = sc->meshes[i];
      mo.id = svc->mesh_meta[i].id;
      mo.content_hash = svc->mesh_meta[i].content_hash;
      mo.resolution = svc->mesh_res[i];
      mo.occludes = svc->mesh_meta[i].occludes;
      mo.words.share(words, words->data() + svc->mesh[i].w_off, svc->mesh[i].w_len);
    }
    if (svc->options.dynamic_geometry != DL_DYNAMIC_LIGHTING_ONLY)
      for (const dl_baked_light& l : svc->baked) {
        if (l.shape == DL_SHAPE_SUN || l.shape == DL_SHAPE_SKY) continue;
        const std::vector<uint32_t>& c = svc->cubes.at(l.id);
        sc->distance_cubes.insert(sc->distance_cubes.end(), c.begin(), c.end());
      }
    sc->sun_maps = b.sun_maps;
    sc->probes = b.probes;
    std::memcpy(sc->probe_origin, b.probe_origin, sizeof b.probe_origin);
    sc->probe_spacing = b.probe_spacing;
    std::memcpy(sc->probe_counts, b.probe_counts, sizeof b.probe_counts);
    sc->probe_words = b.probe_words;
    *out = sc.release();
    return DL_OK;
  } catch (const std::bad_alloc&) {
    return set_error(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_error(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

int dl_service_words(dl_service* svc, const uint32_t** device_words, uint64_t* count) {
  if (!svc) return set_error(nullptr, DL_E_ARGUMENT, "null service");
  if (!device_words || !count) return set_error(svc->ctx, DL_E_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lock(svc->ctx->mu);
  *device_words = svc->W ? svc->W->p : nullptr;
  *count = svc->W ? svc->stats.triangle_words : 0;
  return DL_OK;
}

}  // extern "C"

namespace {
// A fingerprint of the live scene after an edit: of its words (hashed on the device), stats and
// light order.
uint64_t live_checksum(dl_service* svc, int* st) {
  dl_ctx* ctx = svc->ctx;
  *st = DL_OK;
  unsigned long long words_hash = 0;
  if (svc->W && svc->stats.triangle_words) {
    ctx->cache.stream = ctx->stream;
    const CacheScope cache_scope(&ctx->cache);
    DBuf<unsigned long long> d;
    hipError_t e = zeroed(ctx, d, 1);
    if (e == hipSuccess) e = launch_words_hash(ctx->stream, svc->W->p, svc->stats.triangle_words, d.p);
    if (e == hipSuccess) e = hipMemcpyAsync(&words_hash, d.p, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      *st = set_error(ctx, DL_E_DEVICE, hipGetErrorString(e));
      return 0;
    }
  }
  uint64_t h = 0xcbf29ce484222325ull;
  auto put = [&](const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  };
  put(&words_hash, 8);
  put(&svc->stats, sizeof svc->stats);
  for (const dl_baked_light& l : svc->baked) put(&l.id, 8);
  return h;
}

// Device memory an edit's tracing may hold: photon cubes of its bounce lights and a share for
// its rays and regions.
uint64_t edit_bytes(const dl_service* svc, const dl_service_edit& e) {
  const uint64_t r = svc->options.photon_cube_resolution;
  uint64_t b = 0;
  for (uint32_t k = 0; k < e.light_count; ++k) {
    b += 256ull << 20;
    if (e.lights[k].illumination == DL_ILLUMINATION_SINGLE_BOUNCE && !svc->options.skip_bounce) b += 6 * r * r * 16;
  }
  return b;
}
}  // namespace

extern "C" int dl_service_replay_edits(dl_service* svc, const dl_service_edit* edits, uint32_t count, uint32_t workers,
                                       uint64_t budget_bytes, double* latency_ms, uint64_t* checksum, dl_service_update* info) {
  if (!svc) return set_error(nullptr, DL_E_ARGUMENT, "null service");
  dl_ctx* ctx = svc->ctx;
  if (count && (!edits || !latency_ms)) return set_error(ctx, DL_E_ARGUMENT, "null argument");
  if (workers == 0 || workers > 16) return set_error(ctx, DL_E_ARGUMENT, "workers must be 1 to 16");
  for (uint32_t i = 0; i < count; ++i)
    if (!(edits[i].arrival_s >= 0.0) || (i && edits[i].arrival_s < edits[i - 1].arrival_s))
      return set_error(ctx, DL_E_ARGUMENT, "arrival times must be non-negative and in order");
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    if (hipSetDevice(ctx->device) != hipSuccess) return set_error(ctx, DL_E_DEVICE, "hipSetDevice failed");
    if (!make_workers(ctx, workers)) return set_error(ctx, DL_E_DEVICE, "creating a worker stream failed");
    if (budget_bytes == 0) {
      size_t free_b = 0, total_b = 0;
      budget_bytes = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? free_b / 2 : (16ull << 30);
    }
    struct Slot {
      Capture cap;
      RoomTrace room;
      int status = DL_E_CANCELLED;
      std::string err;
      bool done = false;
      uint64_t bytes = 0;
      double trace_ms = 0;
      double queue_ms = 0;
    };
    std::vector<Slot> slots(count);
    std::mutex mu;
    std::condition_variable cv;
    uint64_t inflight = 0;
    bool stop = false;
    std::atomic<uint32_t> next{0};
    uint32_t admit = 0;  // the next edit to admit: admission follows arrival order
    const auto t0 = std::chrono::steady_clock::now();
    auto arrival = [&](uint32_t i) {
      return t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(std::chrono::duration<double>(edits[i].arrival_s));
    };
    auto work = [&](dl_ctx* w) {
      (void)hipSetDevice(w->device);
      for (;;) {
        const uint32_t i = next.fetch_add(1);
        if (i >= count) break;
        Slot& sl = slots[i];
        sl.bytes = edit_bytes(svc, edits[i]);
        {
          // Admission: edit i waits for its arrival, for every earlier edit to be admitted, and
          // for the tracing in flight to leave room for it (an edit alone is always admitted).
          // Bytes are returned when tracing ends, so an admitted edit never waits on another.
          std::unique_lock<std::mutex> l(mu);
          if (cv.wait_until(l, arrival(i), [&] { return stop; })) break;
          cv.wait(l, [&] { return stop || (admit == i && (inflight == 0 || inflight + sl.bytes <= budget_bytes)); });
          if (stop) break;
          inflight += sl.bytes;
          admit++;
          cv.notify_all();
        }
        const auto a = std::chrono::steady_clock::now();
        sl.queue_ms = std::chrono::duration<double, std::milli>(a - arrival(i)).count();
        std::vector<dl_light> changed(edits[i].lights, edits[i].lights + edits[i].light_count);
        int st;
        try {
          if (edits[i].room) {
            st = edits[i].light_count || edits[i].remove_count
                     ? set_error(w, DL_E_ARGUMENT, "an edit adds a room or edits lights, not both")
                     : trace_room(w, svc, edits[i].room, sl.room);
          } else {
            st = trace_lights(w, svc, changed, sl.cap);
          }
          if (st == DL_OK && hipStreamSynchronize(w->stream) != hipSuccess) st = set_error(w, DL_E_DEVICE, "tracing an edit failed");
        } catch (const std::bad_alloc&) {
          w->err = "host allocation failed";
          st = DL_E_OOM;
        } catch (...) {
          w->err = "unexpected failure";
          st = DL_E_DEVICE;
        }
        std::lock_guard<std::mutex> l(mu);
        inflight -= sl.bytes;
        sl.trace_ms = ms_between(a, std::chrono::steady_clock::now());
        sl.status = st;
        if (st != DL_OK) sl.err = "edit " + std::to_string(i) + ": " + w->err;
        sl.done = true;
        cv.notify_all();
      }
    };
    std::vector<std::thread> ts;
    for (uint32_t k = 0; k < workers; ++k) ts.emplace_back(work, ctx->workers[k]);
    // Edits join the live scene in arrival order.  Those whose tracing has ended by the time the
    // scene is free join together: their pairs are merged in order and the scene is assembled
    // once for all of them, which gives the scene of applying them one by one.
    constexpr uint32_t kMaxJoin = 16;
    int result = DL_OK;
    uint32_t i = 0;
    while (i < count && result == DL_OK) {
      uint32_t j = i + 1;
      {
        std::unique_lock<std::mutex> l(mu);
        cv.wait(l, [&] { return slots[i].done; });
        while (j < count && j - i < kMaxJoin && slots[j].done) ++j;
      }
      const auto a = std::chrono::steady_clock::now();
      std::vector<uint8_t> dirty(svc->mesh.size(), 0);
      uint32_t applied = i;
      int st = DL_OK;
      for (uint32_t k = i; k < j; ++k) {
        Slot& sl = slots[k];
        if (sl.status != DL_OK) {
          st = set_error(ctx, sl.status, sl.err);
          break;
        }
        if (edits[k].room) {
          std::unordered_set<uint64_t> added;
          if ((st = join_room(svc, edits[k].room, sl.room, dirty, added)) != DL_OK) break;
          applied = k + 1;
          continue;
        }
        // the parts the edit's lights reach may have grown since they were traced
        for (uint32_t q = 0; q < edits[k].light_count && st == DL_OK; ++q)
          if (svc->parts.size() > 1 && part_of(svc, edits[k].lights[q]) < 0)
            st = set_error(ctx, DL_E_ARGUMENT, "light " + hex_id(edits[k].lights[q].id) + " reaches two parts of the live scene (its rooms)");
        if (st != DL_OK) break;
        std::vector<dl_light> next_lights;
        std::unordered_set<uint64_t> replaced;
        if ((st = edited_lights(svc, edits[k], next_lights, replaced)) != DL_OK) break;
        merge(svc, sl.cap, replaced, dirty);
        for (uint64_t id : replaced)
          if (!std::any_of(edits[k].lights, edits[k].lights + edits[k].light_count, [&](const dl_light& l) { return l.id == id; }))
            svc->cubes.erase(id);
        svc->lights = std::move(next_lights);
        applied = k + 1;
      }
      const auto merged = std::chrono::steady_clock::now();
      if (applied > i) {
        const int sa = assemble(svc, dirty);
        if (sa != DL_OK) {
          result = sa;
          break;
        }
      }
      const auto done = std::chrono::steady_clock::now();
      uint32_t ndirty = 0;
      for (uint8_t d : dirty) ndirty += d;
      if (std::getenv("DL_SERVICE_TIMING"))
        std::fprintf(stderr, "[join] edits=%u merge=%f assemble=%f dirty=%u\n", applied - i, ms_between(a, merged),
                     ms_between(merged, done), ndirty);
      for (uint32_t k = i; k < applied; ++k) {
        Slot& sl = slots[k];
        latency_ms[k] = std::chrono::duration<double, std::milli>(done - arrival(k)).count();
        if (info) {
          info[k] = dl_service_update{};
          info[k].version = svc->version + (k - i) + 1;
          info[k].dirty_meshes = ndirty;
          info[k].reserved = applied - i;  // edits that joined the scene together
          const Capture& kc = edits[k].room ? sl.room.cap : sl.cap;
          info[k].traced_pairs = kc.pairs.size();
          for (const CapturedPair& c : kc.pairs) {
            info[k].traced_samples += c.samples;
            info[k].traced_rays += c.rays;
            info[k].traced_bounce_rays += c.bounce_rays;
          }
          info[k].queue_ms = sl.queue_ms;
          info[k].trace_ms = sl.trace_ms;
          info[k].assemble_ms = ms_between(a, done);
          info[k].total_ms = info[k].trace_ms + info[k].assemble_ms;
        }
        if (checksum) checksum[k] = 0;
        sl.cap = Capture{};
        sl.room = RoomTrace{};
      }
      svc->version += applied - i;
      if (checksum && applied > i) {
        int sc = DL_OK;
        checksum[applied - 1] = live_checksum(svc, &sc);
        if (sc != DL_OK) st = sc;
      }
      if (st != DL_OK) result = st;
      i = j;
    }
    {
      std::lock_guard<std::mutex> l(mu);
      stop = true;
      cv.notify_all();
    }
    next.store(count);
    for (auto& t : ts) t.join();
    return result;
  } catch (const std::bad_alloc&) {
    return set_error(ctx, DL_E_OOM, "host allocation failed");
  } catch (...) {
    return set_error(ctx, DL_E_DEVICE, "unexpected failure");
  }
}

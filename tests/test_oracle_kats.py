"""Pins the oracle against the reference's own known-answer tests and textbook values.

Every assertion below restates a test of the reference crate (cited beside it) or a closed-form
fact; a dropped term, swapped index or wrong rounding in the oracle fails one of them.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from dynlight_hip import abi, scenes

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
U8P = C.POINTER(C.c_uint8)


@pytest.fixture(scope="module")
def ol():
    return oracle.load()


def f32(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float32).reshape(-1))


def fp(a):
    return a.ctypes.data_as(F32P)


# ---- RNG: effects/rng.rs, bake/random.rs -----------------------------------------------------
def test_splitmix64_known_answer(ol):
    assert ol.dlo_splitmix64(0) == 0xE220A8397B1DCDAF
    # splitmix_spreads_neighbouring_inputs (effects/rng.rs:164-168)
    assert 16 <= bin(ol.dlo_splitmix64(1) ^ ol.dlo_splitmix64(2)).count("1") <= 48


def test_pcg32_matches_the_published_stream(ol):
    # pcg32_srandom(42, 54) reference output (PCG paper demo)
    out = np.zeros(6, np.uint32)
    ol.dlo_pcg32(42, 54, 6, out.ctypes.data_as(U32P))
    assert list(out) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_seeds_depend_on_every_part_and_order(ol):
    def seed(*parts):
        a = np.array(parts, np.uint64)
        return ol.dlo_seed(a.ctypes.data_as(U64P), len(parts))
    assert seed(1, 2, 3) == seed(1, 2, 3)
    assert seed(1, 2, 3) != seed(1, 2, 4)
    assert seed(1, 2, 3) != seed(3, 2, 1)
    assert seed(0) != seed(0, 0)
    # restated fold: acc = splitmix(acc ^ splitmix(part)) from 0x0b4ce5eed0000000
    acc = 0x0B4CE5EED0000000
    for p in (7, 8):
        acc = scenes.splitmix64(acc ^ scenes.splitmix64(p))
    assert seed(7, 8) == acc


def test_next_f32_is_uniform_in_unit_interval(ol):
    out = np.zeros(100_000, np.float32)
    key = np.array([1, 99], np.uint64)
    ol.dlo_stream_f32(key.ctypes.data_as(U64P), 2, out.size, fp(out))
    assert out.min() >= 0.0 and out.max() < 1.0
    counts = np.histogram(out, bins=10, range=(0, 1))[0] / out.size
    assert np.all(np.abs(counts - 0.1) < 0.01)
    assert np.all((out * 16777216.0) == np.floor(out * 16777216.0)), "24-bit multiples"


# ---- geometry.rs:147-213 ----------------------------------------------------------------------
@pytest.mark.parametrize("area,density,mx,want", [
    (1.0, 128.0, 2048, 128), (4.0, 128.0, 2048, 256), (2.0, 10.0, 2048, 15), (1e6, 128.0, 2048, 2048),
    (0.0, 128.0, 2048, 2), (1e-9, 128.0, 2048, 2), (float("nan"), 128.0, 2048, 2), (float("inf"), 128.0, 64, 64),
    (1.0, 128.0, 1, 2), (4.0, 10.0, 2048, 20), (360.0, 16.0, 2048, 304)])
def test_lightmap_size(ol, area, density, mx, want):
    assert ol.dlo_lightmap_size(area, density, mx) == want


def test_closest_point_covers_every_region(ol):
    a, b, c = f32(0, 0, 0), f32(1, 0, 0), f32(0, 1, 0)
    cases = [((0.25, 0.25, 1.0), (0.25, 0.25, 0.0)), ((-1, -1, 0), (0, 0, 0)), ((2, -0.5, 0), (1, 0, 0)),
             ((-0.5, 2, 0), (0, 1, 0)), ((0.5, -1, 0), (0.5, 0, 0)), ((-1, 0.5, 0), (0, 0.5, 0)),
             ((1, 1, 0), (0.5, 0.5, 0))]
    out = np.zeros(3, np.float32)
    for p, want in cases:
        ol.dlo_closest_point(fp(f32(*p)), fp(a), fp(b), fp(c), fp(out))
        assert np.allclose(out, want, atol=1e-6), (p, out)


def test_barycentrics_detect_outside_points_and_both_windings(ol):
    out = np.zeros(3, np.float32)
    assert ol.dlo_uv_barycentrics(fp(f32(0, 0, 1, 0, 0, 1)), fp(f32(0.25, 0.25)), fp(out)) == 1
    assert np.allclose(out, [0.5, 0.25, 0.25], atol=1e-6)
    assert ol.dlo_uv_barycentrics(fp(f32(0, 0, 1, 0, 0, 1)), fp(f32(0.75, 0.75)), fp(out)) == 0
    assert ol.dlo_uv_barycentrics(fp(f32(0, 0, 0, 1, 1, 0)), fp(f32(0.25, 0.25)), fp(out)) == 1
    assert np.allclose(out, [0.5, 0.25, 0.25], atol=1e-6)
    assert ol.dlo_uv_barycentrics(fp(f32(0, 0, 1, 0, 2, 0)), fp(f32(1, 0)), fp(out)) == 0


# ---- mesh_builder.rs:171-197 -----------------------------------------------------------------
def test_rects_follow_henrys_bounds_and_clamp(ol):
    out = np.zeros(4, np.uint32)
    uvs = f32(0.25, 0.25, 0.5, 0.25, 0.25, 0.5)
    ol.dlo_rect(0, fp(uvs), 16, out.ctypes.data_as(U32P))
    assert list(out) == [2, 2, 10, 10]
    ol.dlo_rect(1, fp(uvs), 16, out.ctypes.data_as(U32P))
    assert (out[0], out[2]) == (3, 8)
    ol.dlo_rect(2, fp(uvs), 16, out.ctypes.data_as(U32P))
    assert (out[0], out[2]) == (2, 9)
    ol.dlo_rect(0, fp(f32(0, 0, 0.05, 0, 0, 0.05)), 16, out.ctypes.data_as(U32P))
    assert (out[0], out[1]) == (0, 0)
    ol.dlo_rect(0, fp(f32(3, 3, 4, 4, 3, 4)), 16, out.ctypes.data_as(U32P))
    assert (out[0], out[2]) == (15, 15)


# ---- texels.rs:148-219 -----------------------------------------------------------------------
def texels(ol, v, n, uv, face, size, owned=None, touched=False):
    out = np.zeros((4096, 8), np.float32)
    if touched:
        cnt = ol.dlo_touched_texels(fp(f32(*v)), fp(f32(*n)), fp(f32(*uv)), fp(f32(*face)), size,
                                    owned.ctypes.data_as(U32P) if owned is not None else None, fp(out), 4096)
    else:
        cnt = ol.dlo_triangle_texels(fp(f32(*v)), fp(f32(*n)), fp(f32(*uv)), fp(f32(*face)), size, fp(out), 4096)
    return out[:cnt]


def test_texels_interpolate_position_and_normal(ol):
    t = texels(ol, (0, 0, 0, 4, 0, 0, 0, 0, 4), (0, 1, 0, 0, 2, 0, 0, 1, 1), (0, 0, 1, 0, 0, 1), (0, 1, 0), 4)
    assert len(t) == 10
    assert (t[0][0], t[0][1]) == (0, 0)
    assert np.allclose(t[0][2:5], [0.5, 0, 0.5], atol=1e-5)
    for r in t:
        assert abs(np.linalg.norm(r[5:8]) - 1) < 1e-5
        assert r[0] + r[1] <= 3


def test_thin_triangles_touch_texels_whose_centers_they_miss(ol):
    uv = (0.30, 0.05, 0.375, 0.05, 0.30, 0.95)
    v = (0.30, 0, 0.05, 0.375, 0, 0.05, 0.30, 0, 0.95)
    n = (0, 1, 0) * 3
    assert len(texels(ol, v, n, uv, (0, 1, 0), 4)) == 0
    t = texels(ol, v, n, uv, (0, 1, 0), 4, owned=np.zeros(1, np.uint32), touched=True)
    assert [(int(r[0]), int(r[1])) for r in t] == [(1, 0), (1, 1), (1, 2), (1, 3)]
    for r in t:
        assert 0.30 - 1e-6 <= r[2] <= 0.375 + 1e-6
        assert abs(r[4] - (r[1] + 0.5) / 4.0) < 0.13
    owned = np.zeros(1, np.uint32)
    for x in range(4):
        owned[0] |= np.uint32(1 << (x + 2 * 4))  # row y == 2
    assert len(texels(ol, v, n, uv, (0, 1, 0), 4, owned=owned, touched=True)) == 3


def test_world_origin_is_a_valid_texel_and_opposing_normals_fall_back(ol):
    out = np.zeros(3, np.float32)
    assert ol.dlo_uv_barycentrics(fp(f32(0, 0, 1, 0, 0.5, 1)), fp(f32(0.5, 1.0)), fp(out)) == 1
    t = texels(ol, (0, 0, 0, 1, 0, 0, 0, 0, 1), (0, 1, 0, 0, -1, 0, 0, 0, 0), (0, 0, 1, 0, 0, 1), (0, -1, 0), 2)
    # texel (0,0) centre (0.25, 0.25) has weights (0.5, .25, .25): normal (0, .25, 0) -> unit
    assert np.allclose(t[0][5:8], [0, 1, 0])


# ---- shadow_grid.rs / bits.rs ------------------------------------------------------------------
def test_levels_round_to_the_nearest_level(ol):
    assert ol.dlo_level(0.0, 4) == 0
    assert ol.dlo_level(0.5, 4) == 8
    assert ol.dlo_level(1.2, 4) == 15
    assert ol.dlo_level(0.49, 1) == 0
    assert ol.dlo_level(0.51, 1) == 1
    assert ol.dlo_level(float("nan"), 8) == 0


def pack(ol, levels, bits):
    lv = np.ascontiguousarray(np.array(levels, np.uint8))
    out = np.zeros(len(levels), np.uint32)
    n = ol.dlo_pack_levels(lv.ctypes.data_as(U8P), len(levels), bits, out.ctypes.data_as(U32P))
    return list(out[:n])


def test_level_and_bit_packing(ol):
    g = [0] * 9
    g[0], g[1], g[8] = 15, 1, 7
    assert pack(ol, g, 4) == [0x1F, 7]
    b = [0] * 35
    for x, y in [(0, 0), (1, 0), (2, 6)]:
        b[x + y * 5] = 1
    assert pack(ol, b, 1) == [0b11, 1]


# ---- seams.rs:135-231 --------------------------------------------------------------------------
def pad(ol, w, h, lit, visited, levels=None):
    lv = np.zeros(w * h, np.uint8)
    if levels:
        for (x, y), v in levels.items():
            lv[x + y * w] = v
    for x, y in lit:
        lv[x + y * w] = 1
    vis = np.zeros(w * h, np.uint8)
    for x, y in visited:
        vis[x + y * w] = 1
    ol.dlo_pad_seams(w, h, lv.ctypes.data_as(U8P), vis.ctypes.data_as(U8P))
    return lv.reshape(h, w)


def test_unvisited_texels_take_visited_neighbours(ol):
    g = pad(ol, 6, 1, [(2, 0)], [(2, 0)])
    assert list(g[0] != 0) == [True, True, True, True, True, False]


def test_padding_does_not_wrap_across_rows(ol):
    g = pad(ol, 4, 4, [(3, 0)], [(3, 0)])
    assert g[1, 0] == 0 and g[0, 2] and g[1, 3]
    g = pad(ol, 4, 4, [(0, 1)], [(0, 1)])
    assert g[0, 3] == 0


def test_visited_texels_are_never_changed(ol):
    g = pad(ol, 3, 3, [(1, 1)], [(1, 1), (0, 0)])
    assert g[0, 0] == 0 and g[0, 1] and g[1, 0]


def test_seams_take_the_brightest_soft_neighbour(ol):
    g = pad(ol, 3, 1, [], [(0, 0), (2, 0)], levels={(0, 0): 4, (2, 0): 11})
    assert list(g[0]) == [4, 11, 11]


def dilate(ol, w, h, levels):
    lv = np.zeros(w * h, np.uint8)
    for (x, y), v in levels.items():
        lv[x + y * w] = v
    ol.dlo_dilate(w, h, lv.ctypes.data_as(U8P))
    return lv.reshape(h, w)


def test_dilation_fills_the_two_padding_rings(ol):
    g = dilate(ol, 7, 7, {(2, 2): 1})
    for x, y in [(1, 1), (2, 1), (3, 1), (1, 2), (1, 3)]:
        assert g[y, x], (x, y)
    assert not g[1, 4]
    for x, y in [(0, 0), (1, 0), (4, 0), (0, 4)]:
        assert g[y, x], (x, y)
    assert not g[0, 5] and not g[6, 6] and not g[3, 3]


def test_dilation_spreads_the_largest_level_and_handles_tiny_grids(ol):
    g = dilate(ol, 5, 5, {(2, 2): 6, (2, 1): 3})
    assert g[1, 1] == 6 and g[0, 0] == 6 and g[2, 2] == 6
    for w, h in [(1, 1), (1, 5), (5, 1), (2, 2), (3, 3)]:
        assert dilate(ol, w, h, {(0, 0): 1})[0, 0]


def test_crop_offsets_by_two_and_dilates(ol):
    # shadows.rs:175-196: region 10x10 with (4,5) lit, crop (4,5)-(6,6) -> 7x6 grid
    lv = np.zeros(7 * 6, np.uint8)
    lv[2 + 2 * 7] = 1
    ol.dlo_dilate(7, 6, lv.ctypes.data_as(U8P))
    g = lv.reshape(6, 7)
    assert g[2, 2] == 1 and g[0, 0] == 1 and g[1, 1] == 1 and g[3, 3] == 0


# ---- triangle_words.rs:105-157 ---------------------------------------------------------------
def serialize(ol, crops, lights, bounce_layout):
    crops_a = np.array(crops, np.uint32).reshape(-1)
    counts = np.array([len(l) for l in lights], np.uint32)
    flat = [e for l in lights for e in l]
    light_ids = np.array([e[0] for e in flat] or [0], np.uint32)
    slen = np.array([0xFFFFFFFF if e[1] is None else len(e[1]) for e in flat] or [0], np.uint32)
    swords = np.array([w for e in flat if e[1] for w in e[1]] or [0], np.uint32)
    blen = np.array([0xFFFFFFFF if e[2] is None else len(e[2]) for e in flat] or [0], np.uint32)
    bwords = np.array([w for e in flat if e[2] for w in e[2]] or [0], np.uint32)
    out = np.zeros(256, np.uint32)
    p = lambda a: a.ctypes.data_as(U32P)  # noqa: E731
    n = ol.dlo_serialize(len(crops), p(crops_a), p(counts), p(light_ids), p(slen), p(swords), p(blen), p(bwords),
                         1 if bounce_layout else 0, p(out), 256)
    return [int(x) for x in out[:n]]


def test_layout_matches_henrys_builder(ol):
    words = serialize(ol, [(4, 6, 8, 9), (0, 1, 2, 2)],
                      [[(3, [0xAAAA, 0x1], None), (7, [0xFF], None)], []], False)
    assert words == [8, 2, 4, 9, 13, 0xFFFFFFFE, 0xFFFFFFFF, 7, 2, 3, 14, 7, 16, 0, 0xAAAA, 0x1, 0xFF]


def test_bounce_layout_adds_a_third_word(ol):
    recs = [[(0, [5], [9, 10]), (1, None, None)]]
    assert serialize(ol, [(2, 2, 3, 3)], recs, True) == [4, 0, 0, 6, 2, 0, 11, 12, 1, 0, 0, 5, 9, 10]
    assert serialize(ol, [(2, 2, 3, 3)], recs, False) == [4, 0, 0, 6, 2, 0, 9, 1, 0, 5]


# ---- bounce_pack.rs:130-212 ------------------------------------------------------------------
def pack_bounce(ol, values, compression, rgb=False):
    v = np.ascontiguousarray(np.array(values, np.float32).reshape(-1))
    count = len(values)
    out = np.zeros(64, np.uint32)
    n = ol.dlo_pack_bounce(fp(v), count, 1 if rgb else 0, compression, None, 0, out.ctypes.data_as(U32P), 64)
    return [int(x) for x in out[:n]]


def test_bounce_pixels_per_word_follow_henry(ol):
    counts = [len(pack_bounce(ol, [1.0] * 9, c)) for c in (0, 1, 2, 3)]
    assert counts == [3, 2, 2, 2]
    assert pack_bounce(ol, [1.0, 0.0, 1.0], 0) == [0x00FF00FF]
    assert pack_bounce(ol, [1.0] * 5, 1) == [(1 << 30) - 1]


def test_colored_bounce_pixels_pack_side_by_side(ol):
    assert pack_bounce(ol, [(1.0, 0.0, 0.25), (1, 1, 1)], 0, rgb=True) == [0xFF7F00FF, 0xFFFF]
    assert len(pack_bounce(ol, [(1.0, 0.0, 0.25), (1, 1, 1)], 2, rgb=True)) == 1
    assert len(pack_bounce(ol, [(1, 1, 1)] * 16, 1, rgb=True)) == 9


def test_black_and_invalid_bounce_values_stay_black(ol):
    for c in (0, 1, 2, 3):
        assert pack_bounce(ol, [0.0, float("nan"), -1.0], c) == [0]


# ---- light_bvh.rs:242-317 --------------------------------------------------------------------
def light_set(count):
    state = 0x12345678
    vals = []

    def nxt():
        nonlocal state
        state = scenes.splitmix64(state)
        return (state >> 40) / float(1 << 24)
    pos, rad = [], []
    for _ in range(count):
        p = [nxt() * 40.0 - 20.0 for _ in range(3)]
        pos.append(p)
        rad.append(0.5 + nxt() * 4.0)
    return np.array(pos, np.float32), np.array(rad, np.float32)


def light_bvh(ol, pos, rad):
    n = len(rad)
    nodes = (abi.dl_light_bvh_node * (2 * n + 2))()
    order = np.zeros(max(n, 1), np.uint32)
    cnt = ol.dlo_light_bvh(fp(np.ascontiguousarray(pos)), fp(np.ascontiguousarray(rad)), n, nodes, 2 * n + 2,
                           order.ctypes.data_as(U32P))
    return [nodes[i] for i in range(cnt)], order[:n]


def test_light_bvh_leaves_partition_the_lights(ol):
    pos, rad = light_set(57)
    nodes, order = light_bvh(ol, pos, rad)
    assert sorted(order) == list(range(57))
    covered = np.zeros(57, int)
    for nd in nodes:
        if nd.count > 0:
            assert nd.count <= 2
            for slot in range(nd.left_first, nd.left_first + nd.count):
                covered[slot] += 1
                li = order[slot]
                lo, hi = pos[li] - rad[li], pos[li] + rad[li]
                assert all(nd.aabb_min[a] <= lo[a] and nd.aabb_max[a] >= hi[a] for a in range(3))
    assert np.all(covered == 1)
    assert len(nodes) < 2 * 57


def test_light_bvh_small_and_empty_sets(ol):
    assert light_bvh(ol, np.zeros((0, 3), np.float32), np.zeros(0, np.float32))[0] == []
    nodes, _ = light_bvh(ol, np.ones((1, 3), np.float32), np.array([2.0], np.float32))
    assert len(nodes) == 1 and (nodes[0].left_first, nodes[0].count) == (0, 1) and list(nodes[0].aabb_min) == [-1, -1, -1]
    nodes, _ = light_bvh(ol, np.zeros((5, 3), np.float32), np.ones(5, np.float32))
    assert len(nodes) == 1 and nodes[0].count == 5
    a = light_bvh(ol, *light_set(100))
    b = light_bvh(ol, *light_set(100))
    assert [bytes(x) for x in a[0]] == [bytes(x) for x in b[0]] and np.array_equal(a[1], b[1])


# ---- mesh_texels.rs:170-199 ------------------------------------------------------------------
def buried(ol, normal, turn, answers):
    ans = np.ascontiguousarray(np.array(answers + [0] * (8 - len(answers)), np.uint8))
    rays = C.c_uint32()
    dirs = np.zeros(24, np.float32)
    b = ol.dlo_is_buried(fp(f32(*normal)), turn, ans.ctypes.data_as(U8P), C.byref(rays), fp(dirs))
    return b, rays.value, dirs.reshape(8, 3)[: rays.value]


def test_burial_rays_cover_the_hemisphere_above_the_face(ol):
    b, n, d = buried(ol, (0, 1, 0), 0.3, [1] * 8)
    assert b == 1 and n == 5
    b2, n2, d2 = buried(ol, (0, 1, 0), 0.3, [0] * 8)
    assert b2 == 0 and n2 == 4
    for v in np.concatenate([d, d2]):
        assert abs(np.linalg.norm(v) - 1) < 1e-5 and v[1] > 0


def test_a_majority_of_back_faces_buries_the_texel(ol):
    assert buried(ol, (0, 0, 1), 0.0, [1, 0, 1, 1, 1, 1, 1, 1])[0] == 1
    assert buried(ol, (0, 0, 1), 0.0, [0, 1, 0, 1, 0, 1, 0, 1])[0] == 0


# ---- spread.rs:74-127 ------------------------------------------------------------------------
def spread(ol, i, off, rect):
    out = np.zeros(2, np.float32)
    ol.dlo_spread(i, fp(f32(*off)), 1 if rect else 0, fp(out))
    return out


def test_powers_of_two_put_one_point_in_each_strip(ol):
    for count in (1, 2, 4, 16, 64):
        for rect, axis in ((True, 0), (True, 1), (False, 0)):
            strips = sorted(int(spread(ol, i, (0, 0), rect)[axis] * count) for i in range(count))
            assert strips == list(range(count))


def test_rectangles_put_one_point_in_each_cell(ol):
    cells = sorted((int(p[0] * 4), int(p[1] * 4)) for p in (spread(ol, i, (0, 0), True) for i in range(16)))
    assert cells == [(x, y) for x in range(4) for y in range(4)]


def test_shifted_points_leave_no_gap_wider_than_two_strips(ol):
    for axis in (0, 1):
        vals = sorted(float(spread(ol, i, (0.371, 0.913), True)[axis]) for i in range(16))
        gaps = [vals[0] + 1 - vals[-1]] + [b - a for a, b in zip(vals, vals[1:])]
        assert max(gaps) <= 2 / 16 + 1e-6
    assert spread(ol, 5, (0.371, 0.913), False)[1] == np.float32(0.913)


def disk(ol, sq):
    out = np.zeros(2, np.float32)
    ol.dlo_concentric_disk(fp(f32(*sq)), fp(out))
    return out


def test_the_disk_map_keeps_areas_and_rings(ol):
    assert list(disk(ol, (0.5, 0.5))) == [0, 0]
    for c in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        assert abs(np.linalg.norm(disk(ol, c)) - 1) < 1e-6
    assert np.allclose(disk(ol, (1.0, 0.5)), [1, 0], atol=1e-6)
    assert abs(np.linalg.norm(disk(ol, (0.75, 0.6))) - 0.5) < 1e-6
    inside = sum(np.linalg.norm(disk(ol, spread(ol, i, (0, 0), True))) < 0.5 for i in range(256))
    assert 60 <= inside <= 68


# ---- area.rs:207-298, sun.rs:89-143, sky.rs:54-86 ---------------------------------------------
def panel_light(rect=True, length=2.0):
    if rect:
        return scenes.area_light(1, (0, 0, 0), 10.0, scenes.quat_from_rotation_x(-math.pi / 2), (2.0, 1.0))
    return scenes.area_light(1, (0, 0, 0), 10.0, (0, 0, 0, 1), (length, 0.0))


def area_vis(ol, light, pos, normal, minimum, kind, blocker=(0, 0), samples=32, reach=float("nan")):
    inp = abi.BakeInput([], [light])
    frac = C.c_float()
    rays = C.c_uint32()
    key = np.array([1, 2, 3], np.uint64)
    ok = ol.dlo_area_visibility(C.cast(inp._light_arr.ctypes.data, C.POINTER(abi.dl_light)), reach, fp(f32(*pos)),
                                fp(f32(*normal)), samples, minimum, key.ctypes.data_as(U64P), 3, kind,
                                fp(f32(*blocker)), C.byref(frac), C.byref(rays))
    return (frac.value, rays.value) if ok else None


def test_open_texels_stop_after_the_probes(ol):
    below = ((0.3, -2.0, 0.1), (0, 1, 0))
    assert area_vis(ol, panel_light(), *below, 0, 0) == (1.0, 5)
    assert area_vis(ol, panel_light(), *below, 0, 1) == (0.0, 5)
    assert area_vis(ol, panel_light(False), *below, 0, 0) == (1.0, 3)
    assert area_vis(ol, panel_light(), *below, 16, 0) == (1.0, 21)
    assert area_vis(ol, panel_light(False), *below, 16, 1) == (0.0, 19)
    assert area_vis(ol, panel_light(), *below, 100, 0) == (1.0, 37)


def test_minimum_samples_catch_occluders_between_the_probes(ol):
    below = ((0.0, -2.0, 0.0), (0, 1, 0))
    assert area_vis(ol, panel_light(), *below, 0, 2, (-0.7, -0.4)) == (1.0, 5)
    frac, rays = area_vis(ol, panel_light(), *below, 16, 2, (-0.7, -0.4))
    assert rays == 37 and 0.7 < frac < 0.95


def test_half_covered_panels_give_a_penumbra(ol):
    frac, rays = area_vis(ol, panel_light(), (0, -2, 0), (0, 1, 0), 0, 3)
    assert rays == 37 and abs(frac - 0.5) < 0.08
    near, _ = area_vis(ol, panel_light(), (1.5, -1.0, 0), (0, 1, 0), 0, 3)
    assert near > 0.6


def test_texels_behind_panels_or_out_of_reach_see_nothing(ol):
    assert area_vis(ol, panel_light(), (0, 1, 0), (0, -1, 0), 0, 0) is None
    assert area_vis(ol, panel_light(), (0, -2, 0), (0, -1, 0), 0, 0) is None
    tube = panel_light(False, 4.0)
    assert area_vis(ol, tube, (1.9, 1.0, 0), (0, -1, 0), 0, 0) is not None
    assert area_vis(ol, tube, (3.5, 0, 0), (-1, 0, 0), 0, 0, samples=8, reach=1.6) is not None
    assert area_vis(ol, tube, (3.5, 0, 0), (-1, 0, 0), 0, 0, samples=8, reach=1.4) is None
    assert area_vis(ol, tube, (3.5, 0, 0), (-1, 0, 0), 0, 0, samples=8, reach=-2.0) is None


def sun_vis(ol, degrees, normal, minimum, kind=0, blocker=(0,)):
    sun = scenes.LightData(1, (0, 0, 0), 0.5, shape=abi.SHAPE_SUN, rotation=scenes.quat_from_rotation_x(-math.pi / 2),
                           angular_diameter=degrees)
    inp = abi.BakeInput([], [sun])
    frac = C.c_float()
    rays = C.c_uint32()
    key = np.array([4, 5, 6], np.uint64)
    ok = ol.dlo_sun_visibility(C.cast(inp._light_arr.ctypes.data, C.POINTER(abi.dl_light)), fp(f32(*normal)), 64,
                               minimum, key.ctypes.data_as(U64P), 3, kind, fp(f32(*blocker)), C.byref(frac),
                               C.byref(rays))
    return (frac.value, rays.value) if ok else None


def test_suns_without_a_disk_take_one_hard_ray(ol):
    assert sun_vis(ol, 0.0, (0, 1, 0), 16) == (1.0, 1)
    assert sun_vis(ol, 0.0, (0, 1, 0), 16, 1) == (0.0, 1)
    assert sun_vis(ol, 0.0, (0, -1, 0), 16) is None
    g = np.array([1.0, -0.05, 0.0])
    assert sun_vis(ol, 0.0, tuple(g / np.linalg.norm(g)), 0) is not None


def test_soft_suns_stop_after_the_probes_and_edges_give_penumbras(ol):
    assert sun_vis(ol, 0.53, (0, 1, 0), 0) == (1.0, 5)
    assert sun_vis(ol, 0.53, (0, 1, 0), 16) == (1.0, 21)
    frac, rays = sun_vis(ol, 4.0, (0, 1, 0), 16, 2, (0.0,))
    assert rays == 69 and abs(frac - 0.5) < 0.06
    edge = math.tan(-0.5 * math.radians(2.0))
    frac, _ = sun_vis(ol, 4.0, (0, 1, 0), 16, 4, (edge,))
    assert 0.75 < frac < 0.95


def sky_vis(ol, normal, samples, kind, blocker=(0,)):
    key = np.array([1, 2, 3], np.uint64)
    return ol.dlo_sky_visibility(fp(f32(*normal)), samples, key.ctypes.data_as(U64P), 3, kind, fp(f32(*blocker)))


def test_sky_open_covered_and_half(ol):
    assert sky_vis(ol, (0, 1, 0), 64, 0) == 1.0
    assert sky_vis(ol, (0, 1, 0), 64, 1) == 0.0
    assert abs(sky_vis(ol, (0, 1, 0), 64, 2) - 0.5) < 0.04
    low = math.sin(math.radians(30.0))
    assert abs(sky_vis(ol, (0, 1, 0), 256, 5, (low,)) - 0.75) < 0.02


# ---- distance_cubes.rs (half 2.7.1 round to nearest even) -------------------------------------
def test_f16_pairs_round_to_nearest_even(ol):
    assert ol.dlo_f16_pair(1.0, 2.0) == (0x4000 << 16) | 0x3C00
    assert ol.dlo_f16_pair(65504.0, 1e6) == (0x7C00 << 16) | 0x7BFF
    assert ol.dlo_f16_pair(1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11) == (0x3C02 << 16) | 0x3C00
    assert ol.dlo_f16_pair(2 ** -24, -0.0) == (0x8000 << 16) | 0x0001
    rng = np.random.default_rng(1)
    vals = rng.uniform(0, 100, 2000).astype(np.float32)
    for a, b in vals.reshape(-1, 2):
        w = ol.dlo_f16_pair(float(a), float(b))
        assert (w & 0xFFFF) == np.float16(a).view(np.uint16) and (w >> 16) == np.float16(b).view(np.uint16)


# ---- occluder queries: src/bake/tests/occluder.rs ---------------------------------------------
def any_hit(ol, tris, o, d, mx, brute=1):
    t = np.ascontiguousarray(np.array(tris, np.float32).reshape(-1))
    return ol.dlo_ray_any(fp(t), len(t) // 9, fp(f32(*o)), fp(f32(*d)), mx, brute)


def closest(ol, tris, o, d, mx, brute=1):
    t = np.ascontiguousarray(np.array(tris, np.float32).reshape(-1))
    tt = C.c_float()
    idx = C.c_uint32()
    ok = ol.dlo_ray_closest(fp(t), len(t) // 9, fp(f32(*o)), fp(f32(*d)), mx, brute, C.byref(tt), C.byref(idx))
    return (tt.value, idx.value) if ok else None


QUAD = [0, 0, 0, 1, 0, 0, 0, 0, 1]


def test_hits_are_two_sided_and_include_the_maximum(ol):
    o = (0.25, 1.0, 0.25)
    assert closest(ol, QUAD, o, (0, -1, 0), 10.0)[0] == 1.0
    assert any_hit(ol, QUAD, (0.25, -1.0, 0.25), (0, 1, 0), 10.0)
    assert any_hit(ol, QUAD, o, (0, -1, 0), 1.0)
    assert not any_hit(ol, QUAD, o, (0, -1, 0), 0.999)
    assert not any_hit(ol, QUAD, o, (0, 1, 0), 10.0)
    assert not any_hit(ol, QUAD, o, (1, 0, 0), 10.0)


def test_negative_zero_directions_still_hit(ol):
    fl = [-2, 0, -2, 2, 0, 2, 2, 0, -2, -2, 0, -2, -2, 0, 2, 2, 0, 2]
    for brute in (0, 1):
        assert any_hit(ol, fl, (0.2, 1.0, -0.3), (-0.0, -1.0, -0.0), 2.0, brute)
        assert abs(closest(ol, fl, (0.2, 1.0, -0.3), (-0.0, -1.0, -0.0), 2.0, brute)[0] - 1.0) < 1e-6


def test_closest_hits_take_the_nearer_surface_and_lowest_index_on_ties(ol):
    lower = [-1, 0, -1, 1, 0, 1, 1, 0, -1]
    upper = [-1, 1, -1, 1, 1, 1, 1, 1, -1]
    assert closest(ol, lower + upper, (0.3, 2.0, -0.4), (0, -1, 0), 10.0)[1] == 1
    twin = lower + lower
    assert closest(ol, twin, (0.3, 2.0, -0.4), (0, -1, 0), 10.0) == (2.0, 0)


def test_oracle_bvh_equals_a_scan_of_every_occluder(ol):
    rng = np.random.default_rng(11)
    n = 500
    c = rng.uniform(-5, 5, (n, 1, 3))
    tris = (c + rng.normal(scale=0.7, size=(n, 3, 3))).astype(np.float32).reshape(-1, 9)
    tris[:50] = np.round(tris[:50])  # shared vertices and edges
    tris = np.ascontiguousarray(tris)
    m = 20000
    rays = np.zeros((m, 7), np.float32)
    rays[:, :3] = rng.uniform(-6, 6, (m, 3))
    d = rng.normal(size=(m, 3))
    d[: m // 8, 2] = 0.0
    rays[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 6] = rng.choice([np.inf, 2.0, 8.0], m)
    rays = np.ascontiguousarray(rays)
    out = {}
    for brute in (0, 1):
        for cl in (0, 1):
            t = np.zeros(m, np.float32)
            i = np.zeros(m, np.uint32)
            ol.dlo_rays_batch(fp(tris), n, fp(rays), m, cl, brute, fp(t), i.ctypes.data_as(U32P))
            out[(brute, cl)] = (t, i)
    assert np.array_equal(out[(0, 0)][1], out[(1, 0)][1])
    assert np.array_equal(out[(0, 1)][1], out[(1, 1)][1])
    assert np.array_equal(out[(0, 1)][0].view(np.uint32), out[(1, 1)][0].view(np.uint32))
    assert out[(1, 0)][1].sum() > m // 10

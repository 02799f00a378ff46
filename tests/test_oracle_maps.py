"""Sun maps and sky probes of the oracle, pinned by the reference's own tests of sun_map.rs,
sky_probes.rs and src/bake/tests/sky.rs (restated with their tolerances) plus closed forms."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from dynlight_hip import abi, scenes
from dynlight_hip.abi import BakeInput, LightData

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
U8P = C.POINTER(C.c_uint8)


@pytest.fixture(scope="module")
def ol():
    return oracle.load()


def opts(**kw):
    o = abi.default_options()
    o.dynamic_geometry = abi.DYNAMIC_ANGULAR
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def sun(rotation, degrees, light_id=10):
    return LightData(light_id, (500.0, -300.0, 70.0), 0.5, shape=abi.SHAPE_SUN, rotation=rotation,
                     angular_diameter=degrees)


def depths(m, world):
    """The stored depth and the depth of `world` in map space (sun_map.rs:163-171)."""
    rows = np.float32(m["rows"]).reshape(3, 4)
    w = np.float32(world)
    mm = [float(np.dot(r[:3], w) + r[3]) for r in rows]
    size = float(m["resolution"])
    tx, ty = math.floor(mm[0] * size), math.floor(mm[1] * size)
    index = int(ty * size + tx)
    word = int(m["texels"][index // 2]) >> (16 * (index % 2))
    return (word & 0xFFFF) / 65535.0, mm[2]


def test_sun_depth_packing(ol):
    ol.dlo_sun_depth_pair.restype = C.c_uint32
    ol.dlo_sun_depth_pair.argtypes = [C.c_float, C.c_float]
    assert ol.dlo_sun_depth_pair(C.c_float(0.0), C.c_float(1.0)) == 0xFFFF0000
    assert ol.dlo_sun_depth_pair(C.c_float(0.5), C.c_float(-3.0)) == 32768  # round half away; clamp
    assert ol.dlo_sun_depth_pair(C.c_float(2.0), C.c_float(float("nan"))) == 0xFFFF


def test_maps_store_the_first_surface_along_the_rays():
    """sun_map.rs:186-220"""
    inp = BakeInput([scenes.floor(1, (0.0, 0.0, 0.0), 5.0), scenes.cuboid(2, (-1.0, 0.0, -1.0), (1.0, 2.0, 1.0))],
                    [sun(scenes.quat_from_rotation_x(-math.pi / 2), 0.53)])
    s = oracle.bake(inp, opts(sun_map_resolution=64, pixel_density=2.0))
    assert len(s.sun_maps) == 1
    m = s.sun_maps[0]
    assert (m["resolution"], len(m["texels"])) == (64, 64 * 64 // 2)
    for receiver, shaded in [((0.2, 1.0, -0.3), True), ((0.2, 2.5, -0.3), False), ((3.0, 0.5, 3.0), False),
                             ((3.0, -0.5, 3.0), True)]:
        stored, depth = depths(m, receiver)
        assert (depth > stored) == shaded, receiver
    top, _ = depths(m, (0.0, 0.0, 0.0))
    ground, _ = depths(m, (4.0, 0.0, -4.0))
    per_meter = float(np.linalg.norm(np.float32(m["rows"][8:11])))
    assert abs((ground - top) / per_meter - 2.0) < 0.01, "box height"


def test_maps_cover_the_scene_across_slanted_rays():
    """sun_map.rs:222-243"""
    rot = scenes.quat_from_rotation_arc([0, 0, 1.0], [1.0, 1.0, 0.5])
    inp = BakeInput([scenes.floor(1, (10.0, 0.0, 0.0), 2.0)], [sun(rot, 0.0)])
    m = oracle.bake(inp, opts(sun_map_resolution=32, pixel_density=2.0)).sun_maps[0]
    rows = np.float32(m["rows"]).reshape(3, 4)
    for cx, cz in [(8.0, -2.0), (12.0, 2.0), (8.0, 2.0), (12.0, -2.0)]:
        w = np.float32([cx, 0.0, cz])
        mm = [float(np.dot(r[:3], w) + r[3]) for r in rows]
        assert all(0.0 <= c <= 1.0 for c in mm)
        stored, depth = depths(m, (cx, 0.5, cz))
        assert depth < stored, "above the floor is lit"
    stored, depth = depths(m, (10.0, -0.5, 0.0))
    assert depth > stored, "below the floor is shaded"


def test_empty_scenes_give_a_map_nothing_is_behind():
    """sun_map.rs:245-255"""
    s = oracle.bake(BakeInput([], [sun(scenes.quat_from_rotation_x(-1.0), 0.53)]), opts(sun_map_resolution=256))
    m = s.sun_maps[0]
    assert m["resolution"] == 1
    stored, depth = depths(m, (3.0, -40.0, 2.0))
    assert depth <= stored


def test_lighting_only_bakes_no_maps():
    """src/bake/tests/sun.rs:167-171 and sky.rs:152-153"""
    o = opts(sun_map_resolution=64)
    o.dynamic_geometry = abi.DYNAMIC_LIGHTING_ONLY
    inp = BakeInput([scenes.floor(1, (0.0, 0.0, 0.0), 5.0)], [sun(scenes.quat_from_rotation_x(-1.0), 0.53),
                                                             scenes.sky_light(11)])
    s = oracle.bake(inp, o)
    assert s.sun_maps == [] and s.sky_probes is None and len(s.distance_cubes) == 0


def test_sun_maps_follow_sun_slot_order():
    """determinism.rs:110: one map per sun, in the suns' slot (id) order."""
    inp = BakeInput([scenes.floor(1, (0.0, 0.0, 0.0), 5.0), scenes.cuboid(2, (-1.0, 0.0, -1.0), (1.0, 2.0, 1.0))],
                    [sun(scenes.quat_from_rotation_x(-1.2), 0.5, 30), sun(scenes.quat_from_rotation_x(-0.4), 0.5, 20)])
    a = oracle.bake(inp, opts(sun_map_resolution=16, pixel_density=1.0))
    b = oracle.bake(BakeInput(inp.meshes, inp.lights[::-1]), opts(sun_map_resolution=16, pixel_density=1.0))
    assert len(a.sun_maps) == 2
    assert all(np.array_equal(x["texels"], y["texels"]) for x, y in zip(a.sun_maps, b.sun_maps))
    assert not np.array_equal(a.sun_maps[0]["texels"], a.sun_maps[1]["texels"])


def grid(ol, mn, mx, spacing):
    origin = np.zeros(3, np.float32)
    sp = C.c_float()
    counts = np.zeros(3, np.uint32)
    ol.dlo_probe_grid(np.float32(mn).ctypes.data_as(F32P), np.float32(mx).ctypes.data_as(F32P), C.c_float(spacing),
                      origin.ctypes.data_as(F32P), C.byref(sp), counts.ctypes.data_as(U32P))
    return origin, sp.value, counts


def test_grids_cover_the_bounds_and_stay_within_the_cap(ol):
    """sky_probes.rs:196-213"""
    origin, spacing, counts = grid(ol, (0, 0, 0), (4.0, 2.0, 0.0), 1.0)
    assert list(counts) == [6, 4, 2]
    assert list(origin) == [-0.5, -0.5, -0.5]
    assert list(origin + (counts - 1) * np.float32(spacing)) == [4.5, 2.5, 0.5]
    origin, spacing, counts = grid(ol, (0, 0, 0), (500.0, 20.0, 300.0), 1.0)
    assert int(np.prod(counts.astype(np.int64))) <= 65536
    assert spacing > 1.0
    reach = (counts - 1).astype(np.float32) * np.float32(spacing)
    assert np.all(reach >= np.float32([500.0, 20.0, 300.0]))


def test_inside_probes_take_their_neighbours_values(ol):
    """sky_probes.rs:215-229"""
    counts = np.uint32([4, 1, 1])
    op = np.zeros(24, np.float32)
    op[0:6] = 0.5
    op[18:24] = 1.0
    valid = np.uint8([1, 0, 0, 1])
    ol.dlo_fill_inside(counts.ctypes.data_as(U32P), op.ctypes.data_as(F32P), valid.ctypes.data_as(U8P))
    assert list(valid) == [1, 1, 1, 1]
    assert np.all(op[6:12] == 0.5) and np.all(op[12:18] == 1.0)
    lost = np.zeros(4, np.uint8)
    op = np.zeros(24, np.float32)
    ol.dlo_fill_inside(counts.ctypes.data_as(U32P), op.ctypes.data_as(F32P), lost.ctypes.data_as(U8P))
    assert not lost.any()


def test_probe_packing(ol):
    words = np.zeros(2, np.uint32)
    op = np.float32([1.0, 0.0, 0.5, 2.0, -1.0, 0.2])
    ol.dlo_pack_probe(op.ctypes.data_as(F32P), words.ctypes.data_as(U32P))
    # axes 0-3 in bytes 0-3 of word 0: 255, 0, round(127.5) = 128, 255; axes 4-5 in bytes 0-1 of
    # word 1: 0, round(51.0) = 51
    assert words[0] == 255 | (0 << 8) | (128 << 16) | (255 << 24)
    assert words[1] == 51 << 8


def openness(p, x, y, z):
    cx, cy = p["counts"][0], p["counts"][1]
    i = (z * cy + y) * cx + x
    w = p["words"][2 * i:2 * i + 2]
    return [((int(w[a // 4]) >> (8 * (a % 4))) & 0xFF) / 255.0 for a in range(6)]


def test_sky_probes_see_the_roof_and_fill_from_above_under_the_floor():
    """src/bake/tests/sky.rs:122-156"""
    meshes = [scenes.floor(1, (0.0, 0.0, 0.0), 3.0), scenes.cuboid(2, (-1.0, 1.4, -1.0), (1.0, 1.6, 1.0))]
    o = opts(pixel_density=8.0, sky_samples=1024, sky_probe_spacing=1.0)
    s = oracle.bake(BakeInput(meshes, [scenes.sky_light(10)]), o)
    p = s.sky_probes
    assert p is not None
    assert p["counts"] == [8, 4, 8]
    assert np.float32(p["origin"]).tolist() == np.float32([-3.5, -0.7, -3.5]).tolist()
    for x in range(8):
        for z in range(8):
            assert openness(p, x, 3, z)[2] == 1.0, "above everything"
    _, _, up, down, _, _ = openness(p, 3, 1, 3)
    # 0.3 m over the floor, 1.1 m under the roof reaching 0.5 and 1.5 m past the probe: the
    # roof's view factor is 0.411
    assert abs(up - 0.589) < 0.01, up
    assert down < 0.05, down
    assert openness(p, 3, 0, 3) == openness(p, 3, 1, 3), "under the floor takes the probe above"
    empty = oracle.bake(BakeInput([], [scenes.sky_light(10)]), o)
    assert empty.sky_probes is None, "nothing to shade"

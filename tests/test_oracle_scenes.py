"""Whole-bake properties of the oracle: its BVH equals a scan of every occluder, analytic shadow
footprints (src/bake/tests/footprint.rs), thread-count and input-order determinism
(src/bake/tests/determinism.rs:57-81), error cases and empty scenes
(src/bake/tests/determinism.rs:125-169)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import scene_cases
from dynlight_hip import abi, mirror, scenes
from dynlight_hip.abi import BakeInput
from helpers import assert_scenes_equal


@pytest.mark.parametrize("name", sorted(scene_cases.CASES))
def test_oracle_bvh_bake_equals_brute_force_bake(name):
    inp, opts = scene_cases.CASES[name]()
    a = oracle.bake(inp, opts, threads=4)
    b = oracle.bake(inp, opts, threads=4, brute_force=True)
    assert_scenes_equal(a, b)
    assert a.stats["samples"] > 0


@pytest.mark.parametrize("name", ["mixed", "panel"])
def test_bytes_do_not_depend_on_thread_count_or_input_order(name):
    inp, opts = scene_cases.CASES[name]()
    base = oracle.bake(inp, opts, threads=1)
    for t in (4, 7):
        assert_scenes_equal(oracle.bake(inp, opts, threads=t), base)
    rev = BakeInput(list(reversed(inp.meshes)), list(reversed(inp.lights)))
    assert_scenes_equal(oracle.bake(rev, opts, threads=3), base)


# ---- footprint.rs:101-157 ----------------------------------------------------------------------
BOX_MIN = np.array([-0.5, 0.5, -0.5], np.float32)
BOX_MAX = np.array([0.5, 1.0, 0.5], np.float32)


def analytic_shadow(p, light):
    d = light - p
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if abs(d[a]) < 1e-9:
            if p[a] < BOX_MIN[a] or p[a] > BOX_MAX[a]:
                return False
            continue
        x, y = (BOX_MIN[a] - p[a]) / d[a], (BOX_MAX[a] - p[a]) / d[a]
        t0, t1 = max(t0, min(x, y)), min(t1, max(x, y))
    return t0 <= t1


def footprint(light):
    fl = scenes.floor(1, (0, 0, 0), 2.0)
    inp = BakeInput([fl, scenes.cuboid(2, BOX_MIN, BOX_MAX)], [scenes.point_light(10, light, 10.0)])
    o = abi.default_options()
    o.pixel_density = 16.0
    o.dynamic_geometry = abi.DYNAMIC_LIGHTING_ONLY
    baked = oracle.bake(inp, o)
    mesh = next(m for m in baked.meshes if m.id == 1)
    size = mesh.resolution
    texel = 4.0 / size
    ol = oracle.load()
    light = np.array(light, np.float32)
    amb = lit = shadowed = 0
    for tri in range(2):
        idx = fl.indices[tri * 3: tri * 3 + 3]
        v = np.ascontiguousarray(fl.positions[idx].reshape(-1))
        uv = np.ascontiguousarray(fl.uv1[idx].reshape(-1))
        n = np.ascontiguousarray(np.tile(np.array([0, 1, 0], np.float32), 3))
        out = np.zeros((8192, 8), np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
        cnt = ol.dlo_triangle_texels(fp(v), fp(n), fp(uv), fp(np.array([0, 1, 0], np.float32)), size, fp(out), 8192)
        view = mirror.TriangleView(mesh.triangles, tri, bool(baked.settings["bounce"]))
        entry = next((e for e in view.entries() if e.light_index == 0), None)
        for r in out[:cnt]:
            x, y, p = int(r[0]), int(r[1]), r[2:5]
            expected = not analytic_shadow(p, light)
            agree = all((not analytic_shadow(p + np.array([dx, 0, dz], np.float32) * texel, light)) == expected
                        for dx in (-1, 0, 1) for dz in (-1, 0, 1))
            if not agree:
                amb += 1
                continue
            actual = entry is not None and view.shadow_bit(entry, x, y) is True
            assert actual == expected, (x, y, p)
            lit += expected
            shadowed += not expected
    return size, amb, lit, shadowed


def test_shadow_under_a_box_matches_the_projection():
    size, amb, lit, shadowed = footprint((0.0, 3.0, 0.0))
    assert size == 64
    side = 1.5 * 16.0
    assert (side - 2) ** 2 < shadowed <= side * side
    assert lit > 3000 and amb < 4 * 26 * 3


def test_off_center_light_casts_a_skewed_shadow():
    _, _, lit, shadowed = footprint((1.2, 2.5, -0.7))
    assert shadowed > 300 and lit > 3000


# ---- errors and empty scenes (determinism.rs:125-169, input.rs:233-373) ------------------------
def error_of(inp, opts=None):
    with pytest.raises(abi.BakeError) as e:
        oracle.bake(inp, opts or abi.default_options())
    return e.value.code, str(e.value)


def test_duplicate_ids_are_rejected():
    a, b = scenes.floor(5, (0, 0, 0), 1.0), scenes.floor(5, (0, 1, 0), 1.0)
    assert error_of(BakeInput([a, b], [])) == (abi.DL_E_DUPLICATE_ID, "two meshes share BakeId 0x0000000000000005")
    l1, l2 = scenes.point_light(7, (0, 1, 0), 2.0), scenes.point_light(7, (0, 2, 0), 2.0)
    assert error_of(BakeInput([], [l1, l2])) == (abi.DL_E_DUPLICATE_ID, "two lights share BakeId 0x0000000000000007")


def test_invalid_meshes_lights_and_options_are_rejected():
    m = scenes.floor(1, (0, 0, 0), 1.0)
    m.indices = m.indices[:5]
    assert error_of(BakeInput([m], [])) == (abi.DL_E_INVALID_MESH, "mesh 0x0000000000000001: 5 indices is not a multiple of 3")
    m = scenes.floor(1, (0, 0, 0), 1.0)
    m.indices = np.array([0, 1, 9], np.uint32)
    assert error_of(BakeInput([m], []))[1] == "mesh 0x0000000000000001: index 9 is out of range for 4 vertices"
    m = scenes.floor(1, (0, 0, 0), 1.0)
    m.positions = m.positions.copy()
    m.positions[2, 1] = np.inf
    assert error_of(BakeInput([m], []))[1] == "mesh 0x0000000000000001: non-finite vertex position"
    l = scenes.point_light(3, (0, 1, 0), -1.0)
    assert error_of(BakeInput([], [l])) == (abi.DL_E_INVALID_LIGHT, "light 0x0000000000000003: radius must be positive and finite")
    o = abi.default_options()
    o.pixel_density = 0.0
    assert error_of(BakeInput([], []), o) == (abi.DL_E_INVALID_OPTIONS, "invalid bake options: pixel density must be positive")
    o = abi.default_options()
    o.sky_samples = 0
    assert error_of(BakeInput([], []), o)[1] == "invalid bake options: skies need at least one sample"


def test_empty_scenes_bake():
    s = oracle.bake(BakeInput([], []), abi.default_options())
    assert s.meshes == [] and s.stats["triangles"] == 0 and len(s.light_bvh) == 0
    s = oracle.bake(BakeInput([scenes.floor(1, (0, 0, 0), 1.0)], []), abi.default_options())
    assert len(s.meshes) == 1 and mirror.triangle_count(s.meshes[0].triangles) == 2


def test_alpha_tested_lights_pass_transparent_texels():
    """occluder.rs:99-121 at the level of a bake: under a screen whose x < 0 half is transparent,
    the floor is lit on that side and shadowed on the other by an alpha-tested light straight
    above, while a light without transparency is blocked by the whole screen."""
    screen = scene_cases.half_transparent(2, (0.0, 1.0, 0.0), 1.0)
    ol = oracle.load()
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    for transparency, lit_left in ((abi.TRANSPARENCY_ENABLED, True), (abi.TRANSPARENCY_DISABLED, False)):
        light = scenes.point_light(10, (0.0, 3.0, 0.0), 6.0)
        light.transparency = transparency
        fl = scenes.floor(1, (0, 0, 0), 2.0)
        o = scene_cases.opts(8.0)
        o.dynamic_geometry = abi.DYNAMIC_LIGHTING_ONLY
        baked = oracle.bake(BakeInput([fl, screen], [light]), o)
        mesh = next(m for m in baked.meshes if m.id == 1)
        left, right = [], []
        for tri in range(2):
            idx = fl.indices[tri * 3: tri * 3 + 3]
            v = np.ascontiguousarray(fl.positions[idx].reshape(-1))
            uv = np.ascontiguousarray(fl.uv1[idx].reshape(-1))
            n = np.ascontiguousarray(np.tile(np.array([0, 1, 0], np.float32), 3))
            out = np.zeros((8192, 8), np.float32)
            cnt = ol.dlo_triangle_texels(fp(v), fp(n), fp(uv), fp(np.array([0, 1, 0], np.float32)), mesh.resolution,
                                         fp(out), 8192)
            view = mirror.TriangleView(mesh.triangles, tri, False)
            entry = next((e for e in view.entries() if e.light_index == 0), None)
            for r in out[:cnt]:
                x, y, p = int(r[0]), int(r[1]), r[2:5]
                lit = entry is not None and view.shadow_bit(entry, x, y) is True
                if abs(p[2]) < 0.5 and -0.8 < p[0] < -0.2:
                    left.append(lit)
                if abs(p[2]) < 0.5 and 0.2 < p[0] < 0.8:
                    right.append(lit)
        assert len(left) > 20 and len(right) > 20
        assert all(v == lit_left for v in left), transparency
        assert not any(right)

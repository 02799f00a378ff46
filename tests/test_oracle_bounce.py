"""Pins the oracle's bounce half: the reference's own bounce-side tests, restated against the
oracle, plus closed forms and an FP64 restatement of the gather.

* bounce_map.rs:223-281 (windowed filter = full-map filter, Gaussian weights, padded crops);
* shared/cube.rs:113-202 (face table, axes, texel round trips, index order, degenerate input);
* src/bake/tests/colored_bounce.rs:85-184 and src/bake/tests/bounce.rs:66-126 (view factor,
  colored sun and photon bounce, white bounce, Preview, thread-count byte equality);
* closed forms: the photon of a light above a floor (bounce.rs:87-115), glam's `acos_approx`
  against acos within Abramowitz & Stegun 4.4.46's published bound, `slerp` against the float64
  closed form (sin((1-s)t) a + sin(st) b) / sin t;
* the f32 bounce crops against the same samples gathered and filtered in double.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from dynlight_hip import abi, scenes
from dynlight_hip.abi import BakeInput
from dynlight_hip.mirror import TriangleView
from helpers import assert_scenes_equal

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)


@pytest.fixture(scope="module")
def ol():
    return oracle.load()


def fp(a):
    return a.ctypes.data_as(F32P)


def up(a):
    return a.ctypes.data_as(U32P)


def f32(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float32).reshape(-1))


# ---- bounce_map.rs:223-281 -------------------------------------------------------------------
def bounce_map(ol, size, crops, texels, ops=3):
    t = np.ascontiguousarray(np.array(texels, np.float32).reshape(-1, 5))
    out = np.zeros(size * size * 3, np.float32)
    ol.dlo_bounce_map(size, up(np.array(crops, np.uint32)), fp(t), len(t), ops, fp(out))
    return out.reshape(size, size, 3)


def full_map(size, texels):
    """The filters on a plain full-size map, out of place (the reference test's own model)."""
    m = np.zeros((size, size), np.float32)
    for x, y, v in texels:
        m[y, x] = v

    def at(a, x, y):
        return a[y, x] if 0 <= x < size and 0 <= y < size else np.float32(0)

    order = [(0, 0), (-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]
    dil = np.zeros_like(m)
    for y in range(size):
        for x in range(size):
            for dx, dy in order:
                v = at(m, x + dx, y + dy)
                if v > 0:
                    dil[y, x] = v
                    break
    w = np.exp(-np.arange(-3, 4, dtype=np.float64) ** 2 / 50.0)
    w = (w / w.sum()).astype(np.float32)
    cur = dil
    for vertical in (False, True):
        nxt = np.zeros_like(cur)
        for y in range(size):
            for x in range(size):
                s = np.float32(0)
                for k in range(-3, 4):
                    sx, sy = (x, min(max(y + k, 0), size - 1)) if vertical else (min(max(x + k, 0), size - 1), y)
                    s += at(cur, sx, sy) * w[k + 3]
                nxt[y, x] = s / w.sum()
        cur = nxt
    return cur


def test_windowed_filters_match_a_full_map(ol):
    size = 24
    texels = [(1, 1, 0.5), (2, 1, 0.25), (12, 13, 1.0), (13, 13, 0.75), (0, 3, 0.1)]
    got = bounce_map(ol, size, (0, 0, 14, 14), [(x, y, v, v * 0.5, v * 0.25) for x, y, v in texels])
    full = full_map(size, texels)
    for y in range(15):
        for x in range(15):
            e = full[y, x]
            assert np.allclose(got[y, x], [e, e * 0.5, e * 0.25], atol=1e-6), (x, y, got[y, x], e)
    # outside the window (crops grown by 4) nothing is stored
    assert np.all(got[19:, :] == 0) and np.all(got[:, 19:] == 0)


def test_dilation_takes_the_first_positive_neighbour_in_henrys_order(ol):
    # (0, 0) first, then the row above left to right, then the sides, then the row below
    t = [(5, 4, 0.1, 0, 0), (6, 6, 0.2, 0, 0), (4, 5, 0.3, 0, 0), (6, 5, 0.0, 0.4, 0)]
    got = bounce_map(ol, 16, (4, 4, 8, 8), t, ops=1)
    assert got[5, 5, 0] == np.float32(0.1)          # above first
    assert got[5, 7].tolist() == [0.0, np.float32(0.4), 0.0]  # a value with any positive channel counts
    assert got[7, 7, 0] == np.float32(0.2)           # (-1, -1) of (7, 7) is (6, 6)
    assert got[5, 4, 0] == np.float32(0.3)           # a set texel keeps its own value


def test_weights_are_normalized_and_symmetric(ol):
    w = np.zeros(7, np.float32)
    s = np.zeros(1, np.float32)
    ol.dlo_gaussian_weights(fp(w), fp(s))
    assert abs(float(w.sum()) - 1.0) < 1e-6 and abs(float(s[0]) - 1.0) < 1e-6
    assert w[0] == w[6] and w[1] == w[5] and w[2] == w[4]
    assert w[3] > w[2] > w[1] > w[0]
    closed = np.exp(-np.arange(-3, 4) ** 2 / 50.0)
    assert np.allclose(w, closed / closed.sum(), rtol=2e-7, atol=0)


def test_crops_are_padded_and_dropped_when_dark(ol):
    t = f32(5, 4, 0.0, 0.5, 0.0)
    crops = np.array([4, 4, 6, 5], np.uint32)
    out = np.zeros(7 * 6 * 3, np.float32)
    lit = ol.dlo_bounce_crop(16, up(crops), fp(t), 1, up(np.array([4, 4, 6, 5], np.uint32)), fp(out))
    grid = out.reshape(6, 7, 3)
    assert lit == 1
    assert grid[2, 3].tolist() == [0.0, 0.5, 0.0], "texel (5, 4) lands at (3, 2)"
    assert int((grid.max(axis=2) > 0).sum()) == 1, "padding stays zero"
    dark = np.zeros(6 * 6 * 3, np.float32)
    assert ol.dlo_bounce_crop(16, up(crops), fp(t), 1, up(np.array([10, 10, 11, 11], np.uint32)), fp(dark)) == 0


# ---- shared/cube.rs:113-202 ------------------------------------------------------------------
def face_uv(ol, d):
    face = np.zeros(1, np.uint32)
    uv = np.zeros(2, np.float32)
    ol.dlo_cube_face_uv(fp(f32(*d)), up(face), fp(uv))
    return int(face[0]), uv


def direction(ol, face, uv):
    out = np.zeros(3, np.float32)
    ol.dlo_cube_direction(face, fp(f32(*uv)), fp(out))
    return out


def texel_of(ol, d, res):
    out = np.zeros(3, np.uint32)
    ol.dlo_cube_texel(fp(f32(*d)), res, up(out))
    return tuple(int(v) for v in out)


AXES = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def test_axes_land_in_face_centers(ol):
    for face, axis in enumerate(AXES):
        f, uv = face_uv(ol, axis)
        assert f == face
        assert np.allclose(uv, 0.5, atol=1e-6)
        assert np.allclose(direction(ol, face, (0.5, 0.5)), axis, atol=1e-6)


@pytest.mark.parametrize("d,face", [((1.0, -0.5, -0.5), 0), ((0.5, 1.0, 0.5), 2), ((0.5, -1.0, -0.5), 3),
                                    ((-0.5, -0.5, -1.0), 5)])
def test_matches_the_vulkan_face_table(ol, d, face):
    f, uv = face_uv(ol, d)
    assert f == face
    assert np.allclose(uv, [0.75, 0.75], atol=1e-6)


def test_every_texel_center_round_trips(ol):
    res = 16
    out = np.zeros(3, np.float32)
    for face in range(6):
        for y in range(res):
            for x in range(res):
                ol.dlo_cube_texel_direction(face, x, y, res, fp(out))
                assert abs(float(np.linalg.norm(out)) - 1.0) < 1e-5
                assert texel_of(ol, out, res) == (face, x, y)
                assert texel_of(ol, out * np.float32(7.5), res) == (face, x, y)


def test_uv_round_trips_inside_faces(ol):
    for face in range(6):
        for uv in [(0.1, 0.9), (0.3, 0.2), (0.99, 0.01)]:
            f, back = face_uv(ol, direction(ol, face, uv))
            assert f == face
            assert np.allclose(back, uv, atol=1e-5), (face, uv, back)


def test_indices_are_face_major_then_rows(ol):
    ix = ol.dlo_cube_texel_index
    assert ix(0, 0, 0, 64) == 0 and ix(0, 1, 0, 64) == 1 and ix(0, 0, 1, 64) == 64
    assert ix(1, 0, 0, 64) == 4096 and ix(5, 63, 63, 64) == 6 * 4096 - 1


def test_degenerate_directions_do_not_panic(ol):
    f, uv = face_uv(ol, (0, 0, 0))
    assert uv.tolist() == [0.5, 0.5]
    _, x, y = texel_of(ol, (0, 0, 0), 64)
    assert x < 64 and y < 64


def test_ties_go_to_the_later_axis(ol):
    # direction_to_face_uv (cube.rs:31-45) compares strictly: X only when |x| beats both, Y when
    # |y| beats |z|, else Z
    assert face_uv(ol, (1, 1, 1))[0] == 4
    assert face_uv(ol, (1, 0.5, 1))[0] == 4
    assert face_uv(ol, (1, 1, 0.5))[0] == 2
    assert face_uv(ol, (-1, -1, 0.5))[0] == 3
    assert face_uv(ol, (1, 0.5, 0.5))[0] == 0


# ---- closed forms: acos_approx, slerp, photons -----------------------------------------------
def test_acos_approx_is_within_the_published_bound(ol):
    # Abramowitz & Stegun 4.4.46: |acos(x) - sqrt(1-x) P7(x)| <= 2e-8 on [0, 1]; the f32
    # evaluation adds a few ulps of pi/2.  A dropped or wrong coefficient errs by > 1e-5.
    xs = np.linspace(-1.0, 1.0, 20001, dtype=np.float32)
    err = max(abs(ol.dlo_acos_approx(float(x)) - math.acos(float(x))) for x in xs)
    assert err < 6e-7, err
    assert ol.dlo_acos_approx(1.0) == 0.0
    assert ol.dlo_acos_approx(-1.0) == np.float32(math.pi)
    assert ol.dlo_acos_approx(1.5) == 0.0, "inputs beyond 1 clamp"


def slerp(ol, a, b, s):
    out = np.zeros(3, np.float32)
    ol.dlo_slerp(fp(f32(*a)), fp(f32(*b)), s, fp(out))
    return out


def test_slerp_follows_the_closed_form(ol):
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(3000):
        a = rng.normal(size=3)
        b = rng.normal(size=3)
        la, lb = rng.uniform(0.5, 2.0, 2)
        a = (a / np.linalg.norm(a) * la).astype(np.float32)
        b = (b / np.linalg.norm(b) * lb).astype(np.float32)
        s = np.float32(rng.uniform(0, 1))
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        na, nb = np.linalg.norm(a64), np.linalg.norm(b64)
        th = math.acos(np.clip(np.dot(a64, b64) / (na * nb), -1, 1))
        if math.sin(th) < 1e-3:
            continue
        want = (math.sin((1 - s) * th) * a64 / na + math.sin(s * th) * b64 / nb) / math.sin(th) * (na + (nb - na) * s)
        got = slerp(ol, a, b, float(s)).astype(np.float64)
        worst = max(worst, float(np.max(np.abs(got - want)) / max(na, nb)))
    assert worst < 2e-5, worst
    # endpoints and the parallel fallback (lerp)
    assert np.allclose(slerp(ol, (1, 0, 0), (0, 2, 0), 0.0), (1, 0, 0), atol=1e-6)
    assert np.allclose(slerp(ol, (1, 0, 0), (0, 2, 0), 1.0), (0, 2, 0), atol=1e-6)
    assert np.allclose(slerp(ol, (1, 0, 0), (2, 0, 0), 0.5), (1.5, 0, 0), atol=1e-6)
    half = slerp(ol, (1, 0, 0), (-1, 0, 0), 0.5)  # opposite: a quarter turn about an orthogonal axis
    assert abs(np.linalg.norm(half) - 1) < 1e-6 and abs(half[0]) < 1e-6


def test_on_unit_sphere_draws_uniform_unit_vectors(ol):
    n = 200_000
    out = np.zeros(n * 3, np.float32)
    key = np.array([3, 1, 4, 1, 5], np.uint64)
    ol.dlo_on_unit_sphere(key.ctypes.data_as(U64P), len(key), n, fp(out))
    v = out.reshape(n, 3).astype(np.float64)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=2e-6)
    assert np.all(np.abs(v.mean(axis=0)) < 0.01)
    assert np.allclose(np.cov(v.T), np.eye(3) / 3, atol=0.01)
    # z is the first draw mapped to [-1, 1]: its histogram is flat (Archimedes)
    h = np.histogram(v[:, 2], bins=10, range=(-1, 1))[0] / n
    assert np.all(np.abs(h - 0.1) < 0.005)


def _opts(**kw):
    o = abi.default_options()
    o.dynamic_geometry = abi.DYNAMIC_LIGHTING_ONLY
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_photons_of_a_light_above_a_floor(ol):
    """bounce.rs:87-115: the straight-down photon of a light 1.5 m above a floor, a light below
    the floor (unlit side), and misses."""
    res = 8
    ground = scenes.floor(1, (0, 0, 0), 3.0)
    ground.albedo = (0.8, 0.1, 0.0)
    above = scenes.point_light(10, (0.0, 1.5, 0.0), 5.0, bounce=True)
    below = scenes.point_light(11, (0.0, -1.0, 0.0), 4.0, bounce=True)
    _, _, cubes = oracle.bounce_crops(BakeInput([ground], [above, below]), _opts(photon_cube_resolution=res),
                                      photon_cubes=True)
    def expected(light_y, face):
        """distance of the photon through texel (res/2, res/2) of `face`: the floor hit, moved
        back along the ray and off the floor (normal +Y) by max(0.001 t, 0.001)."""
        d = np.zeros(3, np.float32)
        ol.dlo_cube_texel_direction(face, res // 2, res // 2, res, fp(d))
        d = d.astype(np.float64)
        t = -light_y / d[1]
        light = np.array([0.0, light_y, 0.0])
        bias = max(t * 0.001, 0.001)
        nudged = light + d * t - d * bias + np.array([0.0, 1.0, 0.0]) * bias
        return float(np.linalg.norm(nudged - light))

    dist, nrm, alb = cubes[10]
    assert abs(float(dist[3, res // 2, res // 2]) - expected(1.5, 3)) < 2e-6
    assert nrm[3, res // 2, res // 2].tolist() == [0, 127, 0]
    assert alb[3, res // 2, res // 2].tolist() == [204, 26, 0], "round(c * 255)"
    # upward directions miss: the radius and no normal
    assert np.all(dist[2] == np.float32(5.0)) and np.all(nrm[2] == 0) and np.all(alb[2] == 0)
    # the light below sees the floor's back: a distance and albedo but no normal
    dist_b, nrm_b, alb_b = cubes[11]
    assert np.all(nrm_b == 0)
    assert abs(float(dist_b[2, res // 2, res // 2]) - expected(-1.0, 2)) < 2e-6
    assert alb_b[2, res // 2, res // 2].tolist() == [204, 26, 0]


# ---- src/bake/tests/colored_bounce.rs:85-184 ---------------------------------------------------
RED = (0.8, 0.1, 0.0)


def sun_input(floor_albedo):
    sun = scenes.sun_light(10, (0.0, 1.0, 0.0), 0.0)
    sun.illumination = abi.ILLUMINATION_SINGLE_BOUNCE
    sun.bounce_samples = 64
    ground = scenes.floor(1, (0, 0, 0), 10.0)
    ground.albedo = floor_albedo
    return BakeInput([ground, scenes.cuboid(2, (-0.02, 0.0, -0.5), (0.0, 1.0, 0.5))], [sun])


def colored_options(colored):
    return _opts(pixel_density=8.0, photon_cube_resolution=48, bounce_compression=abi.BOUNCE_BITS8,
                 colored_bounce=1 if colored else 0)


def triangle_texels(ol, mesh, t, size):
    idx = mesh.indices[3 * t:3 * t + 3]
    v = np.ascontiguousarray(mesh.positions[idx].reshape(-1), np.float32)
    n = np.ascontiguousarray(mesh.normals[idx].reshape(-1), np.float32)
    uv = np.ascontiguousarray(mesh.uv1[idx].reshape(-1), np.float32)
    e1, e2 = v[3:6] - v[0:3], v[6:9] - v[0:3]
    c = np.cross(e1, e2).astype(np.float32)
    face = (c * (np.float32(1) / np.sqrt(np.float32(np.dot(c, c))))).astype(np.float32)
    out = np.zeros((65536, 8), np.float32)
    cnt = ol.dlo_triangle_texels(fp(v), fp(n), fp(uv), fp(face), size, fp(out), 65536)
    return out[:cnt]


def bounce_of(ol, baked, mesh, t, slot, bits=8):
    """Decoded bounce (squared back) over the texels of triangle t for light slot `slot`."""
    m = next(x for x in baked.meshes if x.id == mesh.id)
    view = TriangleView(m.triangles, t, baked.settings["bounce"])
    entry = next((e for e in view.entries() if e.light_index == slot), None)
    assert entry is not None, "the light reaches the triangle"
    tx = triangle_texels(ol, mesh, t, m.resolution)
    assert len(tx)
    if entry.bounce_offset == 0:
        return np.zeros((len(tx), 3))
    vals = []
    for r in tx:
        x, y = int(r[0]), int(r[1])
        if baked.settings["colored_bounce"]:
            s = np.array(view.bounce_texel_rgb(entry, x, y, bits))
        else:
            s = np.repeat(view.bounce_texel(entry, x, y, bits), 3)
        vals.append(s * s)
    return np.array(vals)


def test_a_sunlit_floor_bounces_onto_a_wall_by_its_view_factor(ol):
    inp = sun_input((1.0, 1.0, 1.0))
    baked = oracle.bake(inp, colored_options(False))
    assert baked.settings["bounce"] and not baked.settings["colored_bounce"]
    assert baked.lights[0]["bounce_compression"] == abi.BOUNCE_BITS8
    wall = inp.meshes[1]
    values = np.concatenate([bounce_of(ol, baked, wall, t, 0) for t in range(2)])
    # a floor reaching 10 m out fills 1/2 (1 - h / sqrt(h^2 + L^2)) ~ 0.48 of the view at h = 0.5
    average = values[:, 0].mean()
    assert 0.40 <= average < 0.50, average
    assert np.all(values[:, 0] > 0.3), "every texel sees the floor"
    ground = bounce_of(ol, baked, inp.meshes[0], 0, 0)[:, 0].mean()
    assert ground < 0.01, ground


def test_colored_sun_bounce_takes_the_floor_color(ol):
    inp = sun_input(RED)
    baked = oracle.bake(inp, colored_options(True))
    assert baked.settings["colored_bounce"]
    colored = np.concatenate([bounce_of(ol, baked, inp.meshes[1], t, 0) for t in range(2)]).mean(axis=0)
    white_in = sun_input((1.0, 1.0, 1.0))
    white_b = oracle.bake(white_in, colored_options(False))
    white = np.concatenate([bounce_of(ol, white_b, white_in.meshes[1], t, 0) for t in range(2)]).mean(axis=0)
    assert np.allclose(colored, np.array(RED) * white[0], atol=0.01), (colored, white)


def photon_color_input():
    light = scenes.point_light(10, (0.0, 1.5, 0.0), 5.0, bounce=True)
    light.bounce_samples = 12
    ground = scenes.floor(1, (0, 0, 0), 3.0)
    ground.albedo = RED
    return BakeInput([ground, scenes.cuboid(2, (1.0, 0.0, -0.5), (1.5, 1.0, 0.5))], [light])


def test_photons_carry_the_color_of_what_they_hit(ol):
    inp = photon_color_input()
    baked = oracle.bake(inp, colored_options(True))
    values = np.concatenate([bounce_of(ol, baked, inp.meshes[1], t, 0) for t in range(2)])
    lit = values[values[:, 0] > 0]
    assert len(lit) * 2 > len(values)
    assert np.all(lit[:, 2] == 0), "blue stays 0"
    assert np.all(np.abs(lit[:, 1] - lit[:, 0] / 8) < 0.01), "green follows red at an eighth"


def test_white_bounce_ignores_colors():
    red = oracle.bake(sun_input(RED), colored_options(False))
    white = oracle.bake(sun_input((1.0, 1.0, 1.0)), colored_options(False))
    assert_scenes_equal(red, white)


@pytest.mark.parametrize("colored", [True, False])
def test_bounce_bakes_are_identical_on_any_thread_count(colored):
    inp = sun_input(RED) if colored else bounce_input()
    o = colored_options(True) if colored else bounce_options(False)
    single = oracle.bake(inp, o, threads=1)
    for threads in (3, 8):
        assert_scenes_equal(oracle.bake(inp, o, threads=threads), single)


# ---- src/bake/tests/bounce.rs:66-126 -------------------------------------------------------
def bounce_input():
    light = scenes.point_light(10, (0.0, 1.5, 0.0), 5.0, bounce=True)
    light.bounce_samples = 12
    return BakeInput([scenes.floor(1, (0, 0, 0), 3.0), scenes.cuboid(2, (1.0, 0.0, -0.5), (1.5, 1.0, 0.5))], [light])


def bounce_options(skip):
    return _opts(pixel_density=8.0, photon_cube_resolution=48, skip_bounce=1 if skip else 0,
                 bounce_compression=abi.BOUNCE_BITS6)


def test_the_floor_bounces_onto_the_far_side_of_the_box(ol):
    inp = bounce_input()
    baked = oracle.bake(inp, bounce_options(False))
    assert baked.settings["bounce"]
    assert baked.lights[0]["bounce_compression"] == abi.BOUNCE_BITS6
    assert baked.stats["bounce_rays"] > 0
    box = inp.meshes[1]
    for t in range(2):  # +X faces away from the light: it keeps the light only for its bounce
        v = bounce_of(ol, baked, box, t, 0, bits=6)[:, 0]
        assert (v > 0).sum() * 2 > len(v)
        assert np.all((v >= 0) & (v <= 1))
    assert np.any(bounce_of(ol, baked, box, 2, 0, bits=6)[:, 0] > 0), "the lit -X face too"
    floor_vals = [bounce_of(ol, baked, inp.meshes[0], t, 0, bits=6)[:, 0] for t in range(2)]
    assert np.any(np.concatenate(floor_vals) > 0)


def test_preview_skips_bounce():
    baked = oracle.bake(bounce_input(), bounce_options(True))
    assert not baked.settings["bounce"]
    assert baked.lights[0]["bounce_compression"] == -1
    assert baked.stats["bounce_rays"] == 0
    cube = next(m for m in baked.meshes if m.id == 2)
    for t in range(2):
        assert TriangleView(cube.triangles, t, False).light_count == 0, "no direct light reaches the far side"


# ---- f32 bounce against the FP64 restatement -------------------------------------------------
def fp64_error(inp, o):
    _, f = oracle.bounce_crops(inp, o)
    _, d = oracle.bounce_crops(inp, o, fp64=True)
    assert f.keys() == d.keys() and len(f)
    rel, absd = [], 0.0
    for k in f:
        x, y = f[k].ravel(), d[k].ravel()
        m = (x != 0) | (y != 0)
        rel.append(np.abs(x[m] - y[m]) / np.maximum(np.abs(y[m]), 1e-30))
        absd = max(absd, float(np.abs(x - y).max()))
    return np.concatenate(rel), absd


@pytest.mark.parametrize("case", ["bounce_point_case", "bounce_sun_case"])
def test_f32_bounce_stays_within_the_fp64_gather(case):
    """Median error is f32 rounding (~1e-7); what exceeds 1e-4 are discrete flips, a sample whose
    cube texel or visibility is decided the other way, a few tenths of a percent of the values.
    A dropped term, a wrong sample index or a swapped operand moves most values by far more."""
    import scene_cases
    rel, absd = fp64_error(*getattr(scene_cases, case)())
    assert np.median(rel) < 1e-6, np.median(rel)
    assert np.percentile(rel, 99) < 1e-4, np.percentile(rel, 99)
    assert np.mean(rel <= 1e-4) > 0.99, np.mean(rel <= 1e-4)
    assert absd < 2e-3, absd

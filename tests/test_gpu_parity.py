"""Parity of the HIP bake with the oracle (bit-exact words, stats and distance cubes)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from dynlight_hip import abi, scenes
from helpers import assert_scenes_equal

pytestmark = pytest.mark.gpu

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)


def _fp(a):
    return a.ctypes.data_as(F32P)


def test_trig_matches_host_libm(ctx):
    """Device f32 sin/cos equal the host's libm on the bake's argument domain."""
    rng = np.random.default_rng(7)
    pos = rng.integers(0x37800000, 0x41000000, 4_000_000, dtype=np.uint32)
    neg = rng.integers(0xB7800000, 0xBF4CCCCD, 2_000_000, dtype=np.uint32)
    turn = np.arange(0, 1 << 24, 13, dtype=np.uint32).astype(np.float32) * np.float32(1.0 / 16777216.0)
    burial = [((np.float32(i) + turn) * np.float32(6.2831855)).astype(np.float32) / np.float32(8.0) for i in range(8)]
    x = np.ascontiguousarray(np.concatenate([pos.view(np.float32), neg.view(np.float32)] + burial).astype(np.float32))
    ds, dc = np.zeros_like(x), np.zeros_like(x)
    hs, hc = np.zeros_like(x), np.zeros_like(x)
    assert ctx.lib.dl_test_trig(ctx.handle, _fp(x), x.size, _fp(ds), _fp(dc)) == 0, ctx.error()
    assert ctx.lib.dl_test_trig_host(_fp(x), x.size, _fp(hs), _fp(hc)) == 0
    assert np.array_equal(ds.view(np.uint32), hs.view(np.uint32))
    assert np.array_equal(dc.view(np.uint32), hc.view(np.uint32))
    # every exception input (where the host libm is not the nearest float)
    for which in (0, 1):
        n = ctx.lib.dl_test_trig_keys(which, None, 0)
        keys = np.zeros(n, np.uint32)
        ctx.lib.dl_test_trig_keys(which, keys.ctypes.data_as(U32P), n)
        xe = np.ascontiguousarray(keys.view(np.float32))
        es, ec, hs2, hc2 = (np.zeros_like(xe) for _ in range(4))
        assert ctx.lib.dl_test_trig(ctx.handle, _fp(xe), xe.size, _fp(es), _fp(ec)) == 0, ctx.error()
        assert ctx.lib.dl_test_trig_host(_fp(xe), xe.size, _fp(hs2), _fp(hc2)) == 0
        assert np.array_equal(es.view(np.uint32), hs2.view(np.uint32))
        assert np.array_equal(ec.view(np.uint32), hc2.view(np.uint32))
    ol = oracle.load()
    pick = rng.integers(0, x.size, 20000)
    gs = np.array([ol.dlo_sinf(float(v)) for v in x[pick]], dtype=np.float32)
    gc = np.array([ol.dlo_cosf(float(v)) for v in x[pick]], dtype=np.float32)
    assert np.array_equal(gs.view(np.uint32), ds[pick].view(np.uint32))
    assert np.array_equal(gc.view(np.uint32), dc[pick].view(np.uint32))


def _room_triangles():
    inp, _ = scenes.box_room(1)
    tris = []
    for m in inp.meshes:
        idx = np.asarray(m.indices).reshape(-1, 3)
        tris.append(np.asarray(m.positions, np.float32)[idx].reshape(-1, 9))
    return np.ascontiguousarray(np.concatenate(tris).astype(np.float32))


@pytest.mark.parametrize("closest", [0, 1])
def test_ray_queries_equal_a_scan_of_every_occluder(ctx, closest):
    tris = _room_triangles()
    rng = np.random.default_rng(3 + closest)
    n = 200_000
    o = rng.uniform([-5.5, -0.5, -5.5], [5.5, 4.5, 5.5], (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[: n // 10, 1] = 0.0  # axis-parallel components
    d[n // 10: n // 5, 0] = -0.0
    mx = rng.choice(np.array([np.inf, 3.0, 0.5, 12.0], np.float32), n)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 3:6], rays[:, 6] = o, d, mx
    rays = np.ascontiguousarray(rays)
    t_gpu = np.zeros(n, np.float32)
    i_gpu = np.zeros(n, np.uint32)
    st = ctx.lib.dl_test_ray_queries(ctx.handle, _fp(tris), tris.shape[0], _fp(rays), n, closest, _fp(t_gpu),
                                     i_gpu.ctypes.data_as(U32P))
    assert st == 0, ctx.error()
    ol = oracle.load()
    r7 = np.ascontiguousarray(rays[:, :7])
    t_ref = np.zeros(n, np.float32)
    i_ref = np.zeros(n, np.uint32)
    ol.dlo_rays_batch(_fp(tris), tris.shape[0], _fp(r7), n, closest, 1, _fp(t_ref), i_ref.ctypes.data_as(U32P))
    assert np.array_equal(i_gpu, i_ref)
    if closest:
        assert np.array_equal(t_gpu.view(np.uint32), t_ref.view(np.uint32))


def test_box_room_bake_matches_oracle(ctx):
    inp, opts = scenes.box_room(1)
    got = ctx.bake(inp, opts)
    want = oracle.bake(inp, opts, threads=0)
    assert_scenes_equal(got, want)
    assert got.stats["samples"] > 100_000


import scene_cases  # noqa: E402


@pytest.mark.parametrize("name", sorted(scene_cases.CASES))
def test_light_shapes_and_edge_cases_match_oracle(ctx, name):
    inp, opts = scene_cases.CASES[name]()
    assert_scenes_equal(ctx.bake(inp, opts), oracle.bake(inp, opts))


def _errors(bake, inp, opts):
    try:
        bake(inp, opts)
    except abi.BakeError as e:
        return e.code, str(e)
    return None


def test_errors_match_oracle(ctx):
    from dynlight_hip.abi import BakeInput
    a, b = scenes.floor(5, (0, 0, 0), 1.0), scenes.floor(5, (0, 1, 0), 1.0)
    bad_mesh = scenes.floor(1, (0, 0, 0), 1.0)
    bad_mesh.indices = np.array([0, 1, 9], np.uint32)
    cases = [BakeInput([a, b], []), BakeInput([bad_mesh], []),
             BakeInput([], [scenes.point_light(3, (0, 1, 0), -1.0)])]
    for inp in cases:
        want = _errors(oracle.bake, inp, abi.default_options())
        assert want is not None
        assert _errors(ctx.bake, inp, abi.default_options()) == want
    o = abi.default_options()
    o.max_lightmap_size = 1
    assert _errors(ctx.bake, BakeInput([], []), o) == _errors(oracle.bake, BakeInput([], []), o)


def test_empty_scenes_match_oracle(ctx):
    from dynlight_hip.abi import BakeInput
    for inp in (BakeInput([], []), BakeInput([scenes.floor(1, (0, 0, 0), 1.0)], []),
                BakeInput([], [scenes.point_light(1, (0, 0, 0), 1.0)])):
        assert_scenes_equal(ctx.bake(inp, abi.default_options()), oracle.bake(inp, abi.default_options()))


@pytest.mark.parametrize("mode,order", [(0, 0), (1, 1), (1, 2)])
def test_bounce_gather_modes_match_oracle(ctx, mode, order):
    """Both point-light gathers (one wave per triangle, light and 64 texels; streamed per-lane
    jobs with light-major job order, with job-major and sample-major ray slots) against the
    oracle on a 2x2-room interior with 12 bounce lights, and the bounce edge cases."""
    try:
        ctx.set_gather(mode, order=order)
        inp, opts = scenes.interior(seed=9, target_triangles=3000, lights=12, grid=2, density=12.0)
        got = ctx.bake(inp, opts)
        assert got.stats["bounce_rays"] > 1_000_000
        assert_scenes_equal(got, oracle.bake(inp, opts))
        for name in sorted(scene_cases.CASES):
            if name.startswith("bounce"):
                inp, opts = scene_cases.CASES[name]()
                assert_scenes_equal(ctx.bake(inp, opts), oracle.bake(inp, opts))
    finally:
        ctx.set_gather(1, order=2)


def test_outdoor_matches_oracle(ctx):
    """A 48x32 heightfield with 12 buildings: soft sun with one bounce, rectangle and tube area
    lights (16 samples), a baked sky, sun map and sky probes."""
    inp, opts = scenes.outdoor(nx=48, nz=32, buildings=12, area_lights=6, density=2.0)
    opts.sun_map_resolution = 128
    opts.sky_samples = 16
    opts.sky_probe_spacing = 4.0
    got = ctx.bake(inp, opts)
    assert got.stats["bounce_rays"] > 100_000 and len(got.sun_maps) == 1 and got.sky_probes is not None
    assert_scenes_equal(got, oracle.bake(inp, opts))


def test_per_lane_multi_ray_path_matches_oracle(ctx):
    """Area lights, soft suns and skies traced per lane in the shadow-ray kernel (the streamed
    any-hit path switched off) against the oracle."""
    try:
        ctx.set_trace_stream(False)
        for name in ("panel", "sun", "sky", "mixed", "maps"):
            inp, opts = scene_cases.CASES[name]()
            assert_scenes_equal(ctx.bake(inp, opts), oracle.bake(inp, opts))
        inp, opts = scenes.outdoor(nx=48, nz=32, buildings=12, area_lights=6, density=2.0)
        opts.sun_map_resolution = 128
        opts.sky_samples = 16
        opts.sky_probe_spacing = 4.0
        assert_scenes_equal(ctx.bake(inp, opts), oracle.bake(inp, opts))
    finally:
        ctx.set_trace_stream(True)

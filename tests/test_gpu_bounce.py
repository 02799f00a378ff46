"""Bounce RGB before packing: the f32 crops the device's bounce packing reads (its filtered bounce
maps cropped to each triangle) against the oracle's, entry by entry.  The path's tolerance is 1e-4
relative to the reference's f32 result; the device follows the reference's operation order, so the
crops are also required to be bit-identical.  Photon cubes are compared at 48^2/64^2 and at the
default 512^2."""
import numpy as np
import pytest

import oracle
import scene_cases
from dynlight_hip import abi, scenes
from helpers import assert_scenes_equal
from test_oracle_bounce import colored_options, photon_color_input, sun_input, RED

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4  # bounce RGB within 1e-4 relative of the reference's f32 result


def _bounce_512():
    inp, o = scene_cases.bounce_point_case()
    o.photon_cube_resolution = 512
    return inp, o


CASES = {
    "point": scene_cases.bounce_point_case,
    "sun": scene_cases.bounce_sun_case,
    "dither": scene_cases.bounce_dither_case,
    "white_dither": scene_cases.bounce_white_dither_case,
    "colored_sun_floor": lambda: (sun_input(RED), colored_options(True)),
    "photon_color": lambda: (photon_color_input(), colored_options(True)),
    "point_512": _bounce_512,
}


def compare_crops(got, want):
    assert got.keys() == want.keys(), (sorted(set(got) ^ set(want)))[:8]
    worst = 0.0
    for k in want:
        g = got[k].astype(np.float64)
        w = want[k]
        assert g.shape == w.shape, k
        scale = np.maximum(np.abs(w), np.float64(np.finfo(np.float32).tiny))
        worst = max(worst, float(np.max(np.abs(g - w) / scale)))
    assert worst <= REL_TOL, worst
    for k in want:
        assert np.array_equal(got[k].astype(np.float64), want[k]), f"entry {k} is within tolerance but not bit-identical"
    return worst


@pytest.mark.parametrize("name", list(CASES))
def test_pre_pack_bounce_equals_the_oracle(ctx, name):
    inp, o = CASES[name]()
    print(f"bounce case {name}: cubes {o.photon_cube_resolution}", flush=True)
    scene, got = ctx.bake_bounce_crops(inp, o)
    want_scene, want = oracle.bounce_crops(inp, o)
    assert len(want) > 0
    compare_crops(got, want)
    assert_scenes_equal(scene, want_scene)
    # the error of the f32 bounce against the same samples in double (stated, not a parity gate)
    _, d = oracle.bounce_crops(inp, o, fp64=True)
    rel = np.concatenate([np.abs(got[k].ravel() - d[k].ravel()) / np.maximum(np.abs(d[k].ravel()), 1e-30)
                          for k in d if np.any(d[k] != 0)])
    print(f"  f32 vs fp64: median {np.median(rel):.2e}, p99 {np.percentile(rel, 99):.2e}, "
          f"share above 1e-4 {np.mean(rel > 1e-4):.4f}", flush=True)


def test_capture_is_off_unless_asked(ctx):
    inp, o = scene_cases.bounce_point_case()
    ctx.bake(inp, o)
    assert ctx.lib.dl_test_bounce_capture(ctx.handle, 0, None, None) == 0


@pytest.mark.parametrize("res", [5, 37, 512])
def test_photon_tiles_change_no_result(ctx, res):
    """The gather skips the photon reads its tile summaries decide; bakes with and without the
    tiles are identical (words, bounce crops, stats) and equal the oracle, at cube sizes that are
    not multiples of the 8-texel tile too.  Counting mode shows that fewer photons were read."""
    inp, o = scene_cases.bounce_point_case()
    o.photon_cube_resolution = res
    print(f"photon tiles at cube resolution {res}", flush=True)
    ctx.set_counting(True)
    try:
        ctx.set_photon_tiles(False)
        try:
            plain, plain_crops = ctx.bake_bounce_crops(inp, o)
            plain_t = ctx.timings()
        finally:
            ctx.set_photon_tiles(True)
        tiled, tiled_crops = ctx.bake_bounce_crops(inp, o)
        tiled_t = ctx.timings()
    finally:
        ctx.set_counting(False)
    assert_scenes_equal(tiled, plain)
    assert tiled_crops.keys() == plain_crops.keys()
    for k in plain_crops:
        assert np.array_equal(tiled_crops[k].view(np.uint32), plain_crops[k].view(np.uint32)), k
    assert plain_t["gather_samples"] > 0
    assert plain_t["gather_photon_reads"] == plain_t["gather_samples"]
    assert tiled_t["gather_samples"] == plain_t["gather_samples"]
    assert tiled_t["gather_photon_reads"] < tiled_t["gather_samples"]
    print(f"  photon reads {tiled_t['gather_photon_reads']} of {tiled_t['gather_samples']} samples", flush=True)
    want_scene, want = oracle.bounce_crops(inp, o)
    assert_scenes_equal(tiled, want_scene)
    compare_crops(tiled_crops, want)


@pytest.mark.parametrize("res", [5, 37, 512])
def test_lazy_photon_cubes_change_no_result(ctx, res):
    """A bake whose gathers set up few samples per cube texel traces only the texels its samples can
    read: the result equals that of full cubes (words, bounce crops, stats) and the oracle; at 512^2
    fewer texels are traced, at 5^2 (more samples than texels) every texel."""
    inp, o = scene_cases.bounce_point_case()
    o.photon_cube_resolution = res
    print(f"lazy photon cubes at cube resolution {res}", flush=True)
    ctx.set_lazy_photons(False)
    try:
        full, full_crops = ctx.bake_bounce_crops(inp, o)
        full_t = ctx.timings()
    finally:
        ctx.set_lazy_photons(True)
    lazy, lazy_crops = ctx.bake_bounce_crops(inp, o)
    lazy_t = ctx.timings()
    assert_scenes_equal(lazy, full)
    assert lazy_crops.keys() == full_crops.keys()
    for k in full_crops:
        assert np.array_equal(lazy_crops[k].view(np.uint32), full_crops[k].view(np.uint32)), k
    total = full_t["photon_texels"]
    assert total > 0 and total % (6 * res * res) == 0
    print(f"  traced {lazy_t['photon_texels']} of {total} cube texels", flush=True)
    assert 0 < lazy_t["photon_texels"] <= total
    if res == 512:
        assert lazy_t["photon_texels"] < total // 4
    if res == 5:
        assert lazy_t["photon_texels"] == total
    want_scene, want = oracle.bounce_crops(inp, o)
    assert_scenes_equal(lazy, want_scene)
    compare_crops(lazy_crops, want)


SPLIT_DIR_ERR = 1.0e-4  # kSplitDirErr: the bound the split setup's margins assume on each component


def test_fast_directions_stay_far_within_the_split_bound(ctx):
    """The split setup's fast pass (hardware sin/cos/rcp/sqrt) against the exact chain on 16M samples
    drawn as the bake draws them (z and azimuth from 24-bit uniforms, the 32 spreads of a 32-sample
    light, random unit directions to the texel): every component within a tenth of the bound."""
    rng = np.random.default_rng(7)
    n = 1 << 24
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    u1 = (rng.integers(0, 1 << 24, n) * np.float32(1.0 / 16777216.0)).astype(np.float32)
    u2 = (rng.integers(0, 1 << 24, n) * np.float32(1.0 / 16777216.0)).astype(np.float32)
    z = u1 * np.float32(2.0) - np.float32(1.0)
    ang = u2 * np.float32(6.2831855)
    step = np.float32(0.9) / np.float32(31)
    sp = (np.float32(0.1) + rng.integers(0, 32, n).astype(np.float32) * step).astype(np.float32)
    exact, fast, ok = ctx.gather_directions(d.astype(np.float32), z, ang, sp)
    assert ok.mean() > 0.98
    err = np.abs(exact[ok].astype(np.float64) - fast[ok]).max()
    print(f"fast pass: {ok.mean():.4f} of samples, largest component error {err:.3e} (bound {SPLIT_DIR_ERR:.1e})", flush=True)
    assert err < SPLIT_DIR_ERR / 10
    assert np.all(np.abs(np.linalg.norm(exact.astype(np.float64), axis=1) - 1.0) < 1e-5)


@pytest.mark.parametrize("res", [5, 37, 512, 2048])
def test_split_setup_changes_no_result(ctx, res):
    """Bakes with the split sample setup (default) and with the exact chain for every sample are
    identical (words, bounce crops, stats) and equal the oracle; at 2048^2 cubes the fast pass's
    candidate span can exceed 2x2 texels, and those samples go to the exact pass."""
    inp, o = scene_cases.bounce_point_case()
    o.photon_cube_resolution = res
    print(f"split setup at cube resolution {res}", flush=True)
    ctx.set_gather_split(False)
    try:
        plain, plain_crops = ctx.bake_bounce_crops(inp, o)
    finally:
        ctx.set_gather_split(True)
    split, split_crops = ctx.bake_bounce_crops(inp, o)
    assert_scenes_equal(split, plain)
    assert split_crops.keys() == plain_crops.keys()
    for k in plain_crops:
        assert np.array_equal(split_crops[k].view(np.uint32), plain_crops[k].view(np.uint32)), k
    want_scene, want = oracle.bounce_crops(inp, o)
    assert_scenes_equal(split, want_scene)
    compare_crops(split_crops, want)

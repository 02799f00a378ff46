"""Parity at full size: the whole interior (200k triangles, 256 bounce lights, 512^2 photon cubes as
benchmarked) and the whole outdoor scene (1.07 M triangles, sun, 64 area lights, sky) baked on the
GPU, against the oracle baking a deterministic sample of receiver meshes (splitmix64(id) % m == 0)
against every occluder and light; the oracle builds the photon cubes of the lights that can reach
a sampled mesh.  Every sampled mesh's words, the lights, the light BVH, the distance cubes and
(outdoor) the sun map and sky probes must be bit-identical.  The outdoor terrain (one mesh of
1.05 M triangles holding most of the texels) is compared on a texel window: the oracle bakes the
terrain triangles whose crops meet the window and the per-triangle payloads of those 8 texels
inside it must equal the GPU's."""
import os

import numpy as np
import pytest

import oracle
from dynlight_hip import scenes
from helpers import first_diff, terrain_arrays, triangle_payloads, window_triangles

pytestmark = pytest.mark.gpu

THREADS = len(os.sched_getaffinity(0))


def compare_sample(got, want):
    assert got.lights == want.lights
    assert np.array_equal(got.light_bvh, want.light_bvh)
    by_id = {m.id: m for m in got.meshes}
    assert want.meshes, "the sample holds meshes"
    for w in want.meshes:
        g = by_id[w.id]
        assert (g.content_hash, g.resolution, g.occludes) == (w.content_hash, w.resolution, w.occludes)
        d = first_diff(g.triangles, w.triangles)
        assert d is None, f"mesh {w.id:#x}: {d}"
    assert first_diff(got.distance_cubes, want.distance_cubes) is None


def test_full_interior_matches_the_oracle_on_a_mesh_sample(ctx):
    """Meshes spread over every room (64^2 cubes, so the oracle builds every light's cube fast)."""
    inp, opts = scenes.interior(2)
    opts.photon_cube_resolution = 64
    got = ctx.bake(inp, opts)
    mask = scenes.mesh_sample_mask(inp.mesh_array["id"], 512)
    print(f"interior: {int(mask.sum())} sampled meshes, 64^2 cubes", flush=True)
    want = oracle.bake_receivers(inp, opts, mask, threads=THREADS)
    compare_sample(got, want)
    assert want.stats["samples"] > 100_000 and want.stats["bounce_rays"] > 1_000_000


def test_full_interior_at_512_cubes_matches_the_oracle_in_a_corner_room(ctx):
    """The benchmarked configuration (512^2 photon cubes): the room shell at (0, 0) and a quarter
    of the boxes and pillars inside it; the oracle builds the cubes of the lights that reach it."""
    inp, opts = scenes.interior(2)
    assert opts.photon_cube_resolution == 512
    got = ctx.bake(inp, opts)
    vo, pos = inp.vertex_offsets, inp._positions
    centre = np.array([pos[vo[i]:vo[i + 1]].mean(axis=0) for i in range(len(vo) - 1)])
    ids = inp.mesh_array["id"]
    mask = (centre[:, 0] < 6.0) & (centre[:, 2] < 6.0) & (scenes.mesh_sample_mask(ids, 4) | (ids == 1))
    print(f"interior corner room: {int(mask.sum())} sampled meshes, 512^2 cubes", flush=True)
    want = oracle.bake_receivers(inp, opts, mask, threads=THREADS)
    compare_sample(got, want)
    assert want.stats["samples"] > 1_000_000 and want.stats["bounce_rays"] > 10_000_000


@pytest.fixture(scope="module")
def outdoor(ctx):
    inp, opts = scenes.outdoor(3)
    return inp, opts, ctx.bake(inp, opts)


def test_full_outdoor_matches_the_oracle_on_a_mesh_sample(outdoor):
    inp, opts, got = outdoor
    mask = scenes.mesh_sample_mask(inp.mesh_array["id"], 256)
    print(f"outdoor: {int(mask.sum())} sampled meshes", flush=True)
    want = oracle.bake_receivers(inp, opts, mask, threads=THREADS)
    compare_sample(got, want)
    assert len(got.sun_maps) == len(want.sun_maps) == 1
    assert first_diff(got.sun_maps[0]["texels"], want.sun_maps[0]["texels"]) is None
    assert first_diff(got.sky_probes["words"], want.sky_probes["words"]) is None


def test_full_outdoor_terrain_matches_the_oracle_on_a_texel_window(outdoor):
    inp, opts, got = outdoor
    terrain = next(m for m in got.meshes if m.id == 1)
    res = terrain.resolution
    bounce = got.settings["bounce"]
    idx, uv = terrain_arrays(inp)
    ntri = len(idx) // 3
    # a window around a triangle that an area light reaches besides the sun and the sky
    probe = triangle_payloads(terrain.triangles, bounce, range(0, ntri, 97))
    t0 = next(t for t, (_, entries) in probe.items() if len(entries) >= 3)
    cx, cy = probe[t0][0][0] + 2, probe[t0][0][1] + 2
    rect = (max(0, cx - 40), max(0, cy - 40), min(res - 1, cx + 40), min(res - 1, cy + 40))
    tris = window_triangles(idx, uv, res, rect)
    print(f"terrain window {rect} of {res}^2: {len(tris)} triangles", flush=True)
    assert len(tris) > 100
    want = oracle.bake_window(inp, opts, 1, rect, threads=THREADS)
    g = triangle_payloads(terrain.triangles, bounce, tris)
    w = triangle_payloads(want.meshes[0].triangles, bounce, tris)
    bad = [t for t in tris if g[t] != w[t]]
    assert not bad, f"{len(bad)} of {len(tris)} terrain triangles differ, first {bad[0]}"
    assert sum(len(e) for _, e in w.values()) > 2 * len(tris)

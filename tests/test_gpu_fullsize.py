"""Parity at full size: the whole interior (200k triangles, 256 bounce lights) and the whole
outdoor scene (1.07 M triangles, sun, 64 area lights, sky) baked on the GPU, against the oracle
baking a deterministic sample of receiver meshes (splitmix64(id) % m == 0) against every occluder
and light.  Every sampled mesh's words, the lights, the light BVH, the distance cubes and (outdoor)
the sun map and sky probes must be bit-identical.  The interior uses 64^2 photon cubes so the
oracle's cubes take seconds; the gather itself does not depend on the cube size."""
import os

import numpy as np
import pytest

import oracle
from dynlight_hip import scenes
from helpers import first_diff

pytestmark = pytest.mark.gpu


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
    inp, opts = scenes.interior(2)
    opts.photon_cube_resolution = 64
    got = ctx.bake(inp, opts)
    mask = scenes.mesh_sample_mask(inp.mesh_array["id"], 512)
    print(f"interior: {int(mask.sum())} sampled meshes", flush=True)
    want = oracle.bake_receivers(inp, opts, mask, threads=len(os.sched_getaffinity(0)))
    compare_sample(got, want)
    assert want.stats["samples"] > 100_000 and want.stats["bounce_rays"] > 1_000_000


def test_full_outdoor_matches_the_oracle_on_a_mesh_sample(ctx):
    inp, opts = scenes.outdoor(3)
    got = ctx.bake(inp, opts)
    mask = scenes.mesh_sample_mask(inp.mesh_array["id"], 256)
    print(f"outdoor: {int(mask.sum())} sampled meshes (the terrain is not among them)", flush=True)
    want = oracle.bake_receivers(inp, opts, mask, threads=len(os.sched_getaffinity(0)))
    compare_sample(got, want)
    assert len(got.sun_maps) == len(want.sun_maps) == 1
    assert first_diff(got.sun_maps[0]["texels"], want.sun_maps[0]["texels"]) is None
    assert first_diff(got.sky_probes["words"], want.sky_probes["words"]) is None

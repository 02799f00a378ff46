"""Comparison helpers shared by the parity tests."""
import numpy as np


def first_diff(a: np.ndarray, b: np.ndarray):
    if a.shape != b.shape:
        return f"length {a.shape} != {b.shape}"
    idx = np.nonzero(a != b)[0]
    if idx.size == 0:
        return None
    i = int(idx[0])
    return f"{idx.size} words differ, first at {i}: {a[i]:#x} != {b[i]:#x}"


def assert_scenes_equal(got, want, check_stats=True):
    """Bit-exact equality of two BakedScene read-backs (product vs oracle)."""
    assert got.settings == want.settings
    assert got.lights == want.lights
    assert np.array_equal(got.light_bvh, want.light_bvh)
    assert len(got.meshes) == len(want.meshes)
    bad = []
    for g, w in zip(got.meshes, want.meshes):
        assert (g.id, g.content_hash, g.resolution, g.occludes) == (w.id, w.content_hash, w.resolution, w.occludes)
        d = first_diff(g.triangles, w.triangles)
        if d:
            bad.append(f"mesh {g.id:#x}: {d}")
    assert not bad, "; ".join(bad[:8])
    d = first_diff(got.distance_cubes, want.distance_cubes)
    assert d is None, f"distance cubes: {d}"
    assert len(got.sun_maps) == len(want.sun_maps)
    for k, (g, w) in enumerate(zip(got.sun_maps, want.sun_maps)):
        assert np.array_equal(np.float32(g["rows"]).view(np.uint32), np.float32(w["rows"]).view(np.uint32)), f"sun map {k} rows"
        assert g["resolution"] == w["resolution"]
        d = first_diff(g["texels"], w["texels"])
        assert d is None, f"sun map {k}: {d}"
    assert (got.sky_probes is None) == (want.sky_probes is None)
    if got.sky_probes is not None:
        g, w = got.sky_probes, want.sky_probes
        assert np.array_equal(np.float32(g["origin"] + [g["spacing"]]).view(np.uint32),
                              np.float32(w["origin"] + [w["spacing"]]).view(np.uint32))
        assert g["counts"] == w["counts"]
        d = first_diff(g["words"], w["words"])
        assert d is None, f"sky probes: {d}"
    if check_stats:
        assert got.stats == want.stats

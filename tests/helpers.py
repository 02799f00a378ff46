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


def triangle_payloads(words: np.ndarray, bounce_layout: bool, triangles):
    """Per triangle of `triangles`: its header bounds and, per light entry, (light index, shadow
    words, bounce words), read through the runtime's TriangleView (mirror.rs:38-50).  Payloads
    are contiguous in serialization order, so one ends where the next begins."""
    import bisect
    from dynlight_hip.mirror import TriangleView, triangle_count
    n = triangle_count(words)
    offs = set()
    views = [TriangleView(words, t, bounce_layout) for t in range(n)]
    for v in views:
        for e in v.entries():
            offs.update(o for o in (e.shadow_offset, e.bounce_offset) if o)
    offs = sorted(offs) + [len(words)]

    def part(o):
        if o == 0:
            return None
        return tuple(int(w) for w in words[o:offs[bisect.bisect_right(offs, o)]])

    return {t: (views[t].bounds, [(e.light_index, part(e.shadow_offset), part(e.bounce_offset)) for e in views[t].entries()])
            for t in triangles}


def window_triangles(indices, uv1, resolution: int, rect, inset: int = 8):
    """Triangles (of a mesh's indices and UV1) whose crop rectangle (mesh_builder.rs:103-113)
    lies `inset` texels inside rect = (min_x, min_y, max_x, max_y)."""
    idx = np.asarray(indices).reshape(-1, 3)
    uv = np.asarray(uv1, np.float32)[idx]  # (T, 3, 2)
    lo = np.floor(uv.min(axis=1) * np.float32(resolution)) - 2
    hi = np.ceil(uv.max(axis=1) * np.float32(resolution)) + 2
    ok = (lo[:, 0] >= rect[0] + inset) & (lo[:, 1] >= rect[1] + inset) & (hi[:, 0] <= rect[2] - inset) & \
         (hi[:, 1] <= rect[3] - inset)
    return [int(t) for t in np.nonzero(ok)[0]]


def terrain_arrays(inp):
    """Indices (mesh-local) and UV1 of the first mesh of a BakeInput built from arrays."""
    vo, io = inp.vertex_offsets, inp.index_offsets
    return inp._indices[io[0]:io[1]], inp._uv1[vo[0]:vo[1]]

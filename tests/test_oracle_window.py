"""The oracle's texel-window mode (dlo_set_window): baking only the triangles of one mesh whose
crops meet a window gives, for the triangles 8 texels inside it, exactly the words of the whole
bake.  This is what lets the full-size terrain be compared with the GPU on a window of it."""
from dynlight_hip import scenes
import oracle
from helpers import terrain_arrays, triangle_payloads, window_triangles


def test_a_texel_window_bakes_like_the_whole_mesh():
    inp, opts = scenes.outdoor(nx=48, nz=32, buildings=12, area_lights=6, density=2.0)
    full = oracle.bake(inp, opts)
    terrain = next(m for m in full.meshes if m.id == 1)
    res = terrain.resolution
    bounce = full.settings["bounce"]
    idx, uv = terrain_arrays(inp)
    # centre the window on the first triangle that two area lights, the sun and the sky reach
    every = triangle_payloads(terrain.triangles, bounce, range(len(idx) // 3))
    t0 = next(t for t, (_, entries) in every.items() if len(entries) >= 4)
    cx, cy = every[t0][0][0] + 2, every[t0][0][1] + 2
    rect = (max(0, cx - 24), max(0, cy - 20), min(res - 1, cx + 24), min(res - 1, cy + 20))
    part = oracle.bake_window(inp, opts, 1, rect)
    assert [m.id for m in part.meshes] == [1]
    tris = window_triangles(idx, uv, res, rect)
    assert len(tris) > 40
    want = {t: every[t] for t in tris}
    got = triangle_payloads(part.meshes[0].triangles, bounce, tris)
    assert any(len(entries) >= 3 for _, entries in want.values()), "sun, sky and an area light reach the window"
    assert got == want
    # the windowed bake did less: triangles outside it have no lights
    assert part.stats["samples"] < full.stats["samples"]

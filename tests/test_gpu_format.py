"""Baked scenes through the .dynlight codec on the GPU path: the file of a bake holds exactly the
payload the format documents (src/bake/format/mod.rs:1-21, written independently in
test_format.py), and decoding it gives the bake back (decode(encode(s)) == s)."""
import io

import numpy as np
import pytest

import dynlight_hip
import scene_cases
from dynlight_hip import abi, scenes
from test_format import file_of, payload, unpacked

pytestmark = pytest.mark.gpu

BOUNCE_NAME = {-1: None, abi.BOUNCE_BITS8: "Bits8", abi.BOUNCE_BITS6: "Bits6", abi.BOUNCE_BITS5: "Bits5",
               abi.BOUNCE_BITS4: "Bits4"}
SHAPE_NAME = {0: "point", 1: "rect", 2: "tube", 3: "sun", 4: "sky"}


def as_layout(s):
    """A BakedScene in the test layout's terms (what encode must write)."""
    st = s.settings
    dyn = {abi.DYNAMIC_LIGHTING_ONLY: 0, abi.DYNAMIC_DISTANCE_CUBES: 1, abi.DYNAMIC_ANGULAR: 2}[st["dynamic_geometry"]]
    lights = []
    for l in s.lights:
        shape = SHAPE_NAME[l["shape"]]
        rot, size = (0.0, 0.0, 0.0, 1.0), (0.0, 0.0)
        if shape in ("rect", "tube"):
            rot, size = tuple(l["rotation"]), tuple(l["size"])
        elif shape == "sun":
            rot, size = tuple(l["rotation"]), (l["angular_diameter"], 0.0)
        lights.append(dict(id=l["id"], origin=tuple(l["origin"]), radius=l["radius"], shape=shape, rotation=rot,
                           size=size, bounce=BOUNCE_NAME[l["bounce_compression"]], bits=l["shadow_bits_per_texel"]))
    return dict(
        flags=int(st["bounce"]) | int(st["colored_bounce"]) << 1, dynamic=dyn, density=st["pixel_density"],
        max_size=st["max_lightmap_size"], cube_res=st["distance_cube_resolution"], lights=lights,
        bvh=[tuple(int(w) for w in node) for node in np.asarray(s.light_bvh)],
        meshes=[dict(id=m.id, hash=m.content_hash, res=m.resolution, words=[int(w) for w in m.triangles],
                     occludes=int(m.occludes)) for m in s.meshes],
        cubes=[int(w) for w in s.distance_cubes],
        sun_maps=[dict(res=m["resolution"], rows=m["rows"], texels=[int(w) for w in m["texels"]]) for m in s.sun_maps],
        probes=None if s.sky_probes is None else dict(origin=tuple(s.sky_probes["origin"]), spacing=s.sky_probes["spacing"],
                                                      counts=tuple(s.sky_probes["counts"]),
                                                      words=[int(w) for w in s.sky_probes["words"]]),
        stats=tuple(s.stats[k] for k in ("meshes", "triangles", "texels", "rays", "lights_removed", "bounce_rays",
                                         "triangle_words")))


def same_scene(got, want):
    assert got.settings == want.settings and got.lights == want.lights
    assert np.array_equal(got.light_bvh, want.light_bvh)
    assert [(m.id, m.content_hash, m.resolution, m.occludes) for m in got.meshes] == \
        [(m.id, m.content_hash, m.resolution, m.occludes) for m in want.meshes]
    for a, b in zip(got.meshes, want.meshes):
        assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(got.distance_cubes, want.distance_cubes)
    assert len(got.sun_maps) == len(want.sun_maps)
    for a, b in zip(got.sun_maps, want.sun_maps):
        assert a["resolution"] == b["resolution"] and a["rows"] == b["rows"] and np.array_equal(a["texels"], b["texels"])
    assert (got.sky_probes is None) == (want.sky_probes is None)
    if got.sky_probes is not None:
        assert np.array_equal(got.sky_probes["words"], want.sky_probes["words"])
    assert {k: v for k, v in got.stats.items() if k != "samples"} == {k: v for k, v in want.stats.items() if k != "samples"}


@pytest.mark.parametrize("name", ["box_room", "interior_bounce", "outdoor"])
def test_baked_scenes_round_trip_through_the_file(ctx, name):
    if name == "box_room":
        inp, opts = scenes.box_room(1)
    elif name == "interior_bounce":
        inp, opts = scenes.interior(seed=9, target_triangles=3000, lights=12, grid=2, density=12.0)
        inp = inp.with_lights(inp.lights + [scenes.area_light(77, (3.0, 2.5, 3.0), 6.0, scenes.quat_from_rotation_x(-1.2),
                                                             (1.0, 0.5))])
    else:
        inp, opts = scenes.outdoor(nx=48, nz=32, buildings=12, area_lights=6, density=2.0)
        opts.sun_map_resolution = 128
        opts.sky_samples = 16
        opts.sky_probe_spacing = 4.0
    scene, data = ctx.bake_encoded(inp, opts)
    assert data[:8] == b"DYNLIGHT"
    assert unpacked(data) == payload(as_layout(scene)), "the file holds the documented payload"
    same_scene(dynlight_hip.decode(data), scene)
    assert dynlight_hip.reencode(data) == data, "decode then encode gives the same bytes"
    # a file written by another gzip implementation reads the same
    same_scene(dynlight_hip.decode(file_of(unpacked(data))), scene)


def py_buffers(s):
    """BufferBuilder (src/baked_lighting/layout.rs:148-214) over a BakedScene, written here."""
    tris, recs = [], []
    for m in s.meshes:
        words = [int(w) for w in m.triangles]
        recs += [len(tris), m.resolution, words[0] // 4 if words else 0, 0]
        tris += words
    tris = tris or [0]
    aux = [0] * 12
    aux[0], aux[1] = len(aux), len(recs) // 4
    aux += recs
    bvh = [int(w) for w in np.asarray(s.light_bvh).reshape(-1)]
    aux[2], aux[3] = len(aux), len(bvh) // 8
    aux += bvh
    local = len(s.lights)
    while local and s.lights[local - 1]["shape"] in (3, 4):
        local -= 1
    aux[4], aux[5] = len(aux), (local if len(s.distance_cubes) else 0)
    aux += [int(w) for w in s.distance_cubes]
    st = s.settings
    aux[6] = (1 if st["bounce"] else 0) | {abi.DYNAMIC_LIGHTING_ONLY: 0, abi.DYNAMIC_DISTANCE_CUBES: 2,
                                            abi.DYNAMIC_ANGULAR: 4}[st["dynamic_geometry"]]
    aux[7] = len(s.lights)
    aux[8], aux[9] = len(aux), len(s.sun_maps)
    texels = len(aux) + 16 * len(s.sun_maps)
    for m in s.sun_maps:
        aux += list(np.asarray(m["rows"], np.float32).view(np.uint32)) + [m["resolution"], texels, 0, 0]
        texels += len(m["texels"])
    for m in s.sun_maps:
        aux += [int(w) for w in m["texels"]]
    if s.sky_probes is not None:
        p = s.sky_probes
        aux[10] = len(aux)
        aux += list(np.asarray(list(p["origin"]) + [p["spacing"]], np.float32).view(np.uint32)) + list(p["counts"])
        aux.append(len(aux) + 1)
        aux += [int(w) for w in p["words"]]
    return tris, aux


@pytest.mark.parametrize("name", ["box_room", "outdoor"])
def test_scene_buffers_of_a_bake(ctx, name):
    if name == "box_room":
        inp, opts = scenes.box_room(1)
    else:
        inp, opts = scenes.outdoor(nx=48, nz=32, buildings=12, area_lights=6, density=2.0)
        opts.sun_map_resolution = 128
        opts.sky_samples = 16
        opts.sky_probe_spacing = 4.0
    scene, b = ctx.bake_buffers(inp, opts)
    tris, aux = py_buffers(scene)
    assert [int(w) for w in b["triangles"]] == tris
    assert [int(w) for w in b["aux"]] == aux
    assert [r["id"] for r in b["receivers"]] == [m.id for m in scene.meshes]
    assert [r["tag"] for r in b["receivers"]] == list(range(1, len(scene.meshes) + 1))


def test_a_full_outdoor_bake_loads_with_one_copy_of_its_words(ctx, tmp_path):
    """The loader path on a real bake: the full outdoor scene's file (hundreds of MB of words) is
    stream-decoded in a fresh process and its words are taken into the buffers mesh by mesh; the
    peak host memory grows by one copy of the words plus the mesh being taken, and the buffers
    equal scene_buffers of the baked scene."""
    from test_loader_memory import _peak
    inp, opts = scenes.outdoor(3)
    scene, data = ctx.bake_encoded(inp, opts)
    path = str(tmp_path / "outdoor.dynlight")
    with open(path, "wb") as f:
        f.write(data)
    words = sum(len(m.triangles) for m in scene.meshes)
    largest = max(len(m.triangles) for m in scene.meshes)
    take = _peak(path, "take")
    assert take["words"] == words
    grow = take["peak"] - take["base"]
    print(f"outdoor file: {len(data) / 2**20:.0f} MiB packed, {words * 4 / 2**20:.0f} MiB of words (largest mesh "
          f"{largest * 4 / 2**20:.0f} MiB); load peak growth {grow / 2**20:.0f} MiB", flush=True)
    # meshes are taken one at a time: the buffers plus the mesh being copied (the terrain is one
    # mesh holding most of the words)
    assert grow < 1.05 * (words + largest) * 4 + (64 << 20)
    loaded, buffers = dynlight_hip.load_bake(io.BytesIO(data))
    assert loaded.triangles_taken
    tris = np.concatenate([m.triangles for m in scene.meshes])
    assert np.array_equal(buffers["triangles"], tris)

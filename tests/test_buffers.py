"""BakedBuffers (dl_scene_buffers) against the reference's layout tests
(src/baked_lighting/layout.rs:217-330): scenes come in as .dynlight files written by the
independent payload writer of test_format.py."""
import struct

import numpy as np

import dynlight_hip
from test_format import empty, file_of, payload

AUX_HEADER_WORDS, RECEIVER_WORDS, SUN_MAP_RECORD_WORDS = 12, 4, 16
RECEIVERS, RECEIVER_COUNT, BVH, BVH_NODE_COUNT, CUBES, CUBE_COUNT, FLAGS, BAKED_LIGHTS = range(8)
SUN_MAPS, SUN_MAP_COUNT, SKY_PROBES = 8, 9, 10
SUN_MAP_RESOLUTION, SUN_MAP_TEXELS = 12, 13
SKY_PROBE_SPACING, SKY_PROBE_COUNTS, SKY_PROBE_WORDS, SKY_PROBE_RECORD_WORDS = 3, 4, 7, 8
BOUNCE_LAYOUT, CUBE_BIAS, CUBE_ANGULAR = 1, 2, 4


def light(i, shape):
    return dict(id=i, origin=(0, 0, 0), radius=1.0, shape=shape, rotation=(0, 0, 0, 1), size=(0.5, 0.0), bounce=None,
                bits={"point": 1, "sun": 4, "sky": 8}[shape])


def record(buffers, tag):
    aux = buffers["aux"]
    index = tag - 1
    if index < 0 or index >= aux[RECEIVER_COUNT]:
        return None
    start = aux[RECEIVERS] + index * RECEIVER_WORDS
    return list(aux[start:start + RECEIVER_WORDS])


def f32(word):
    return struct.unpack("<f", struct.pack("<I", int(word)))[0]


def test_empty_buffers_are_valid_storage():
    b = dynlight_hip.scene_buffers_of_file(file_of(payload(empty())))
    assert list(b["triangles"]) == [0]
    assert len(b["aux"]) == AUX_HEADER_WORDS
    assert b["aux"][RECEIVER_COUNT] == 0 and b["aux"][SKY_PROBES] == 0
    assert record(b, 1) is None and record(b, 0) is None
    assert b["receivers"] == []


def test_meshes_are_concatenated_with_their_base():
    a = [8, 0, 0, 5, 4, 0, 0, 7, 0, 0]
    bw = [4, 1, 1, 6, 1, 3, 0]
    maps = [dict(res=r, rows=[float(r)] * 12, texels=[r * 10] * ((r * r + 1) // 2)) for r in (1, 2)]
    s = empty()
    s.update(flags=1, dynamic=0,
             lights=[light(1, "point"), light(2, "sun"), light(3, "sun"), light(4, "sky")],
             bvh=[(0,) * 8],
             meshes=[dict(id=10, hash=1, res=64, words=a, occludes=1), dict(id=20, hash=2, res=32, words=bw, occludes=1)],
             cubes=[7, 8], sun_maps=maps,
             probes=dict(origin=(1.0, 2.0, 3.0), spacing=0.5, counts=(1, 2, 1), words=[11, 12, 13, 14]))
    b = dynlight_hip.scene_buffers_of_file(file_of(payload(s)))
    assert len(b["triangles"]) == len(a) + len(bw)
    assert list(b["triangles"][len(a):]) == bw
    first, second = b["receivers"]
    assert (first["id"], first["tag"], first["triangle_count"]) == (10, 1, 2)
    assert (second["id"], second["tag"], second["content_hash"]) == (20, 2, 2)
    assert record(b, 1) == [0, 64, 2, 0]
    assert record(b, 2) == [len(a), 32, 1, 0]
    assert record(b, 3) is None
    aux = b["aux"]
    assert aux[BVH] == AUX_HEADER_WORDS + 2 * RECEIVER_WORDS
    assert aux[BVH_NODE_COUNT] == 1
    assert aux[CUBES] == aux[BVH] + 8
    assert list(aux[aux[CUBES]:aux[CUBES] + 2]) == [7, 8]
    assert aux[CUBE_COUNT] == 1, "one light before the suns and the sky"
    assert aux[FLAGS] == BOUNCE_LAYOUT and aux[BAKED_LIGHTS] == 4
    assert aux[SUN_MAPS] == aux[CUBES] + 2 and aux[SUN_MAP_COUNT] == 2
    for k, m in enumerate(maps):
        rec = aux[SUN_MAPS] + k * SUN_MAP_RECORD_WORDS
        assert f32(aux[rec + 11]) == float(m["res"])
        assert aux[rec + SUN_MAP_RESOLUTION] == m["res"]
        tex = aux[rec + SUN_MAP_TEXELS]
        assert list(aux[tex:tex + len(m["texels"])]) == m["texels"]
    last = aux[aux[SUN_MAPS] + SUN_MAP_RECORD_WORDS + SUN_MAP_TEXELS]
    rec = aux[SKY_PROBES]
    assert last + 2 == rec, "the 2x2 map's 2 words end the maps"
    assert f32(aux[rec + 2]) == 3.0 and f32(aux[rec + SKY_PROBE_SPACING]) == 0.5
    assert list(aux[rec + SKY_PROBE_COUNTS:rec + SKY_PROBE_COUNTS + 3]) == [1, 2, 1]
    first_probe = aux[rec + SKY_PROBE_WORDS]
    assert first_probe == rec + SKY_PROBE_RECORD_WORDS
    assert list(aux[first_probe:]) == [11, 12, 13, 14]


def test_cube_flags_follow_the_dynamic_geometry_mode():
    for dynamic, flag in ((0, 0), (1, CUBE_BIAS), (2, CUBE_ANGULAR)):
        s = empty()
        s["dynamic"] = dynamic
        assert dynlight_hip.scene_buffers_of_file(file_of(payload(s)))["aux"][FLAGS] == flag

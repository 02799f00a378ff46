"""The .dynlight codec (dl_encode, dl_decode, dl_decode_stream) against the file layout of
src/bake/format/mod.rs:1-21, lights.rs, maps.rs and codec.rs, written here independently, and the
reference's format tests (src/bake/format/tests.rs:125-359): round trips, deterministic bytes,
light order and sun map rules, sky probe rules, versions, corruption, invalid values, streamed
reads."""
import gzip
import io
import math
import struct
import zlib

import numpy as np
import pytest

import dynlight_hip
from dynlight_hip import abi

NOT_DYNLIGHT, UNSUPPORTED, TRUNCATED, DECOMPRESS, CHECKSUM, INVALID = 1, 2, 3, 4, 5, 6
BITS = {None: 0, "Bits8": 8, "Bits6": 6, "Bits5": 5, "Bits4": 4}
SHAPES = {"point": 0, "rect": 1, "tube": 2, "sun": 3, "sky": 4}


def quat_x(a):
    return (math.sin(a / 2), 0.0, 0.0, math.cos(a / 2))


def sun(i):
    return dict(id=i, origin=(0, 0, 0), radius=10.0, shape="sun", rotation=quat_x(-1.0), size=(0.53, 0.0),
                bounce=None, bits=4)


def sky(i):
    return dict(id=i, origin=(0, 0, 0), radius=10.0, shape="sky", rotation=(0, 0, 0, 1), size=(0, 0), bounce=None, bits=8)


def sample():
    """format/tests.rs:31-122"""
    return dict(
        flags=3, dynamic=1, density=64.0, max_size=1024, cube_res=64,
        lights=[dict(id=42, origin=(1.0, -2.5, 3.25), radius=6.0, shape="rect", rotation=quat_x(-0.5), size=(1.6, 0.4),
                     bounce="Bits5", bits=4),
                dict(id=7, origin=(0, 0, 0), radius=0.5, shape="point", rotation=(0, 0, 0, 1), size=(0, 0), bounce=None, bits=1),
                dict(id=9, origin=(0, 1, 0), radius=3.0, shape="tube", rotation=(0, 0, 0, 1), size=(2.0, 0.0), bounce=None, bits=2),
                sun(3), sky(11)],
        bvh=[struct.unpack("<8I", struct.pack("<3fI3fI", -1.0, -2.0, -3.0, 0, 4.0, 5.0, 6.0, 2))],
        meshes=[dict(id=1, hash=0xABCD, res=128, words=[4, 0, 0, 5, 0], occludes=1),
                dict(id=2, hash=3, res=2, words=[], occludes=0)],
        cubes=[0x3C003C00, 7],
        sun_maps=[dict(res=3, rows=[1, 0, 0, 0, 0, 0.5, 0, 0.25, -0.125, -0.125, -0.125, -0.125],
                       texels=[0xFFFF0000, 1, 2, 3, 0x8000])],
        probes=dict(origin=(-0.5, 0.25, 2.0), spacing=1.5, counts=(2, 1, 3), words=[0x01010101 * i for i in range(12)]),
        stats=(2, 1, 10, 20, 1, 0, 5))


def empty():
    return dict(flags=0, dynamic=0, density=0.0, max_size=0, cube_res=0, lights=[], bvh=[], meshes=[], cubes=[],
                sun_maps=[], probes=None, stats=(0, 0, 0, 0, 0, 0, 0))


def payload(s):
    """The payload as format/mod.rs:173-212 writes it."""
    b = bytearray()
    b += struct.pack("<IIfII", s["flags"], s["dynamic"], s["density"], s["max_size"], s["cube_res"])
    b += struct.pack("<Q", len(s["lights"]))
    for l in s["lights"]:
        b += struct.pack("<Q4fII4f2fI", l["id"], *l["origin"], l["radius"], BITS[l["bounce"]], SHAPES[l["shape"]],
                         *l["rotation"], *l["size"], l["bits"])
    words = [w for node in s["bvh"] for w in node]
    b += struct.pack("<Q", len(words)) + struct.pack(f"<{len(words)}I", *words)
    b += struct.pack("<Q", len(s["meshes"]))
    for m in s["meshes"]:
        b += struct.pack("<QQII", m["id"], m["hash"], m["res"], m["occludes"])
        b += struct.pack("<Q", len(m["words"])) + struct.pack(f"<{len(m['words'])}I", *m["words"])
    b += struct.pack("<Q", len(s["cubes"])) + struct.pack(f"<{len(s['cubes'])}I", *s["cubes"])
    b += struct.pack("<Q", len(s["sun_maps"]))
    for m in s["sun_maps"]:
        b += struct.pack("<I12f", m["res"], *m["rows"])
        b += struct.pack("<Q", len(m["texels"])) + struct.pack(f"<{len(m['texels'])}I", *m["texels"])
    if s["probes"] is None:
        b += struct.pack("<I", 0)
    else:
        p = s["probes"]
        b += struct.pack("<I3ff3I", 1, *p["origin"], p["spacing"], *p["counts"])
        b += struct.pack("<Q", len(p["words"])) + struct.pack(f"<{len(p['words'])}I", *p["words"])
    b += struct.pack("<I6Q", *s["stats"])
    return bytes(b)


def file_of(p, version=6):
    """Header (magic, version, length, CRC-32) and a gzip member of the payload."""
    return b"DYNLIGHT" + struct.pack("<IQI", version, len(p), zlib.crc32(p)) + gzip.compress(p, mtime=0)


def unpacked(data):
    return gzip.decompress(data[24:])


def check_decoded(got, s):
    st = got.settings
    assert (st["bounce"], st["colored_bounce"]) == (bool(s["flags"] & 1), bool(s["flags"] & 2))
    assert st["dynamic_geometry"] == {0: abi.DYNAMIC_LIGHTING_ONLY, 1: abi.DYNAMIC_DISTANCE_CUBES, 2: abi.DYNAMIC_ANGULAR}[s["dynamic"]]
    assert st["pixel_density"] == np.float32(s["density"]) and st["max_lightmap_size"] == s["max_size"]
    assert [l["id"] for l in got.lights] == [l["id"] for l in s["lights"]]
    for g, w in zip(got.lights, s["lights"]):
        assert g["shape"] == SHAPES[w["shape"]] and g["shadow_bits_per_texel"] == w["bits"]
        assert g["radius"] == np.float32(w["radius"])
        if w["shape"] in ("rect", "tube"):
            assert tuple(g["rotation"]) == tuple(np.float32(v) for v in w["rotation"])
            assert tuple(g["size"]) == tuple(np.float32(v) for v in w["size"])
        if w["shape"] == "sun":
            assert g["angular_diameter"] == np.float32(w["size"][0])
    assert [(m.id, m.content_hash, m.resolution, m.occludes, list(m.triangles)) for m in got.meshes] == \
        [(m["id"], m["hash"], m["res"], bool(m["occludes"]), m["words"]) for m in s["meshes"]]
    assert list(got.distance_cubes) == s["cubes"]
    assert len(got.sun_maps) == len(s["sun_maps"])
    assert (got.sky_probes is None) == (s["probes"] is None)
    assert tuple(got.stats[k] for k in ("meshes", "triangles", "texels", "rays", "lights_removed", "bounce_rays",
                                         "triangle_words")) == s["stats"]


def format_error(data, stream=False):
    with pytest.raises(abi.BakeError) as e:
        (dynlight_hip.decode_stream(io.BytesIO(data)) if stream else dynlight_hip.decode(data))
    assert e.value.code == abi.DL_E_FORMAT
    return e.value.format_kind, str(e.value)


def test_scenes_round_trip():
    for s in (sample(), empty()):
        data = file_of(payload(s))
        check_decoded(dynlight_hip.decode(data), s)
        again = dynlight_hip.reencode(data)
        assert again[:8] == b"DYNLIGHT" and again[:24] == data[:24]
        assert unpacked(again) == payload(s), "encode writes the documented payload"
        assert dynlight_hip.reencode(again) == again, "encoding is deterministic"
        assert again[24:34] == bytes([0x1F, 0x8B, 8, 0, 0, 0, 0, 0, 0, 255]), "gzip header: mtime 0, OS 255"


def test_suns_come_last_with_one_sun_map_each_or_none():
    early = sample()
    early["lights"].insert(0, sun(4))
    assert format_error(file_of(payload(early))) == (INVALID, ".dynlight file has an invalid light order")
    after_sky = sample()
    after_sky["lights"].append(sun(5))
    assert format_error(file_of(payload(after_sky)))[1].endswith("light order")
    two = sample()
    two["lights"].insert(4, sun(5))
    assert format_error(file_of(payload(two)))[1].endswith("sun map count")
    two["sun_maps"] = []
    dynlight_hip.decode(file_of(payload(two)))
    short = sample()
    short["sun_maps"][0]["texels"].pop()
    assert format_error(file_of(payload(short)))[1].endswith("invalid sun map")


def test_sky_probes_need_a_sky_and_whole_probes():
    no_sky = sample()
    no_sky["lights"].pop()
    assert format_error(file_of(payload(no_sky)))[1].endswith("sky probes")
    no_sky["probes"] = None
    dynlight_hip.decode(file_of(payload(no_sky)))
    short, zero, huge = sample(), sample(), sample()
    short["probes"]["words"].pop()
    zero["probes"]["counts"] = (2, 0, 3)
    huge["probes"]["counts"] = (1 << 16, 1 << 16, 1 << 16)
    for broken in (short, zero, huge):
        assert format_error(file_of(payload(broken))) == (INVALID, ".dynlight file has an invalid sky probes")


def test_unknown_versions_are_rejected_clearly():
    data = bytearray(file_of(payload(sample())))
    for v in (7, 1, 4, 5):
        data[8:12] = struct.pack("<I", v)
        kind, msg = format_error(bytes(data))
        assert kind == UNSUPPORTED and f"version {v} is not supported" in msg and "re-bake" in msg


def test_corruption_is_detected():
    data = file_of(payload(sample()))
    assert format_error(b"PNG\0\0\0\0\0garbage-garbage-garbage") == (NOT_DYNLIGHT, "not a .dynlight file (bad magic)")
    assert format_error(data[:10])[0] == TRUNCATED
    format_error(data[:-3])
    wrong = bytearray(data)
    wrong[20] ^= 1
    assert format_error(bytes(wrong)) == (CHECKSUM, ".dynlight payload checksum mismatch")
    wrong = bytearray(data)
    wrong[12] ^= 1
    format_error(bytes(wrong))
    flipped = bytearray(data)
    flipped[len(flipped) - 12] ^= 0x55
    format_error(bytes(flipped))


def test_invalid_payload_values_are_rejected():
    s = sample()
    s["max_size"] = 1
    p = bytearray(payload(s))
    p[4] = 9
    assert format_error(file_of(bytes(p)))[1].endswith("dynamic geometry mode")
    shape = 20 + 8 + 28  # settings, light count, then id, origin, radius, bounce precision
    p = bytearray(payload(sample()))
    p[shape] = 7
    assert format_error(file_of(bytes(p)))[1].endswith("light shape")
    for bits in (0, 3, 16):
        p = bytearray(payload(sample()))
        p[shape + 28] = bits
        assert format_error(file_of(bytes(p)))[1].endswith("shadow bits per texel")
    p = payload(sample()) + b"\0"
    assert format_error(file_of(p)) == (INVALID, ".dynlight file has an invalid trailing payload")
    p = bytearray(payload(sample()))
    p[0] = 4
    assert format_error(file_of(bytes(p)))[1].endswith("settings flag")


class Dribble(io.RawIOBase):
    """At most 7 bytes per read (format/tests.rs:297-308)."""

    def __init__(self, data):
        self.data = data

    def read(self, n=-1):
        k = min(7, len(self.data) if n is None or n < 0 else n)
        out, self.data = self.data[:k], self.data[k:]
        return out


def test_streamed_and_whole_buffer_reads_agree():
    big = sample()
    big["meshes"] = [dict(id=i + 1, hash=i, res=8, words=[(w * 2654435761) & 0xFFFFFFFF for w in range(10000 + i)],
                          occludes=int(i % 2 == 0)) for i in range(5)]
    for s in (sample(), empty(), big):
        data = file_of(payload(s))
        whole = dynlight_hip.decode(data)
        check_decoded(whole, s)
        for src in (io.BytesIO(data), Dribble(data)):
            streamed = dynlight_hip.decode_stream(src)
            assert [list(m.triangles) for m in streamed.meshes] == [list(m.triangles) for m in whole.meshes]
            assert streamed.stats == whole.stats and streamed.lights == whole.lights


def test_streamed_reads_report_the_same_errors():
    data = file_of(payload(sample()))
    wrong_sum, wrong_len, flipped, old = (bytearray(data) for _ in range(4))
    wrong_sum[20] ^= 1
    wrong_len[12] ^= 1
    flipped[len(flipped) - 12] ^= 0x55
    old[8:12] = struct.pack("<I", 5)
    for broken in (wrong_sum, wrong_len, flipped, old, data[:10], data[:-3], b"PNG\0\0\0\0\0garbage-garbage-garbage"):
        broken = bytes(broken)
        assert format_error(broken, stream=True)[0] == format_error(broken)[0]

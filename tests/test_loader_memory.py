"""The loader's memory path (baked_lighting/loader.rs:47-68, examples/bake_load_memory.rs): a
.dynlight file is inflated as it is parsed (decode_stream) and its triangle words are moved into
the GPU buffers one mesh at a time (take_scene_buffers, apply.rs:45-61), so one copy of the words
is held at a time.  The file here is synthetic, in the shape bake_load_memory's `gen` writes:
meshes of 2^20 words of small values, 200 MB unpacked."""
import gzip
import io
import json
import os
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest

import dynlight_hip
from dynlight_hip import BakeError, abi
from test_format import file_of, payload, sample

HERE = os.path.dirname(os.path.abspath(__file__))
MESH_WORDS = 1 << 20


def test_taking_buffers_equals_scene_buffers_and_marks_the_scene():
    data = file_of(payload(sample()))
    want = dynlight_hip.scene_buffers_of_file(data)
    scene, got = dynlight_hip.load_bake(io.BytesIO(data))
    assert scene.triangles_taken
    for k in ("triangles", "aux"):
        assert np.array_equal(got[k], want[k]), k
    assert got["receivers"] == want["receivers"]
    assert [m.id for m in scene.meshes] == [1, 2] and all(len(m.triangles) == 0 for m in scene.meshes)
    assert [m.resolution for m in scene.meshes] == [128, 2], "meshes keep their ids and resolutions"


def test_a_taken_scene_is_refused_by_encode_and_scene_buffers():
    import ctypes as C
    lib = dynlight_hip.load()
    h = dynlight_hip._decode_handle(lib, file_of(payload(sample())))
    try:
        b = C.c_void_p()
        assert lib.dl_take_scene_buffers(h, C.byref(b)) == 0
        lib.dl_baked_buffers_free(b)
        assert lib.dl_scene_triangles_taken(h) == 1
        buf, n = C.POINTER(C.c_uint8)(), C.c_uint64()
        assert lib.dl_encode(h, C.byref(buf), C.byref(n)) == abi.DL_E_TRIANGLES_TAKEN
        assert lib.dl_last_error(None).decode() == \
            "the scene's triangle words were moved into its GPU buffers, so it can't be written"
        assert lib.dl_scene_buffers(h, C.byref(b)) == abi.DL_E_TRIANGLES_TAKEN
        assert lib.dl_take_scene_buffers(h, C.byref(b)) == abi.DL_E_TRIANGLES_TAKEN
        assert lib.dl_last_error(None).decode() == "the scene's triangle words were already taken"
    finally:
        lib.dl_baked_scene_free(h)


def big_file(path, megabytes):
    """A synthetic bake of `megabytes` unpacked MB (bake_load_memory.rs `gen`): xorshift words
    masked to small values so the gzip ratio is realistic."""
    meshes = (megabytes * 1024 * 1024 // 4 + MESH_WORDS - 1) // MESH_WORDS
    s = np.uint64(0x9E3779B97F4A7C15)
    state = np.arange(1, MESH_WORDS + 1, dtype=np.uint64) * s
    head = bytearray(struct.pack("<IIfII", 0, 0, 16.0, 2048, 0) + struct.pack("<Q", 0) + struct.pack("<Q", 0))
    head += struct.pack("<Q", meshes)
    crc = zlib.crc32(head)
    total = len(head)
    comp = zlib.compressobj(6, zlib.DEFLATED, 31)
    with open(path, "wb") as f:
        f.write(b"\0" * 24)  # header, written at the end
        f.write(comp.compress(bytes(head)))
        for i in range(meshes):
            state ^= state << np.uint64(13)
            state ^= state >> np.uint64(7)
            state ^= state << np.uint64(17)
            words = ((state >> np.uint64(40)).astype(np.uint32) & np.uint32(0x0F0F))
            words[0] = 4 * 1000
            rec = struct.pack("<QQII", i + 1, i, 16, 1) + struct.pack("<Q", MESH_WORDS) + words.tobytes()
            crc = zlib.crc32(rec, crc)
            total += len(rec)
            f.write(comp.compress(rec))
        tail = struct.pack("<Q", 0) + struct.pack("<Q", 0) + struct.pack("<I", 0) + struct.pack("<I6Q", meshes, 0, 0, 0, 0, 0,
                                                                                            meshes * MESH_WORDS)
        crc = zlib.crc32(tail, crc)
        total += len(tail)
        f.write(comp.compress(tail))
        f.write(comp.flush())
        f.seek(0)
        f.write(b"DYNLIGHT" + struct.pack("<IQI", 6, total, crc & 0xFFFFFFFF))
    return meshes * MESH_WORDS * 4


CHILD = r"""
import io, json, sys
sys.path[:0] = sys.argv[3].split(":")
import dynlight_hip

def hwm():
    for line in open("/proc/self/status"):
        if line.startswith("VmHWM:"):
            return int(line.split()[1]) * 1024
def rss():
    for line in open("/proc/self/status"):
        if line.startswith("VmRSS:"):
            return int(line.split()[1]) * 1024

dynlight_hip.load()
packed = open(sys.argv[1], "rb").read()  # the asset reader reads the packed file whole
base = rss()
if sys.argv[2] == "take":
    scene, buffers = dynlight_hip.load_bake(io.BytesIO(packed), copy=False)
else:  # decode, then scene_buffers that copies the words while the scene keeps its own
    import ctypes as C
    lib = dynlight_hip.load()
    h = dynlight_hip._decode_stream_handle(lib, io.BytesIO(packed))
    buffers = dynlight_hip._buffers_of_handle(lib, h, take=False, copy=False)
    lib.dl_baked_scene_free(h)
print(json.dumps({"base": base, "peak": hwm(), "packed": len(packed), "words": buffers["triangles"]}))
"""


def _peak(path, mode):
    pkg = os.path.join(os.path.dirname(HERE), "dynamiclighting-bevy_amd")
    out = subprocess.run([sys.executable, "-c", CHILD, path, mode, pkg], capture_output=True, text=True, timeout=600,
                         check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_loading_a_large_bake_holds_one_copy_of_the_words(tmp_path):
    path = str(tmp_path / "big.dynlight")
    unpacked = big_file(path, 200)
    take = _peak(path, "take")
    copy = _peak(path, "copy")
    assert take["words"] * 4 == unpacked
    grow_take = take["peak"] - take["base"]
    grow_copy = copy["peak"] - copy["base"]
    print(f"unpacked {unpacked / 2**20:.0f} MiB, packed {take['packed'] / 2**20:.0f} MiB: peak growth "
          f"{grow_take / 2**20:.0f} MiB with take_scene_buffers, {grow_copy / 2**20:.0f} MiB with scene_buffers", flush=True)
    assert grow_take < 1.15 * unpacked, "one copy of the words, plus one mesh and the inflate window"
    assert grow_copy > 1.8 * unpacked, "the copying path holds both"

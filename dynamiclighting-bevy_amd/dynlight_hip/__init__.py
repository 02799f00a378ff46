"""MI355X bake backend: Python face of libdynlight_hip.so.

Mirrors the reference's pure bake API (src/bake/baker.rs:27-31):

    scene = dynlight_hip.bake(inp, options)          # BakedScene, or raises BakeError

`inp` is an abi.BakeInput (BakeMesh/BakeLight data), `options` an abi.dl_bake_options
(BakeOptions; `abi.default_options()` is BakeOptions::default()).  The work runs in the HIP
library on the GPU; there is no CPU fallback, and loading fails loudly when the library or a
device is missing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import os

from . import abi
from .abi import BakeError, BakeInput, LightData, MeshData, default_options  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdynlight_hip.so")
PKG_DIR = os.path.dirname(HERE)

_lib = None


def build() -> None:
    """Compiles libdynlight_hip.so for gfx950 with the package Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", PKG_DIR, "-j8"], check=True)


def load() -> C.CDLL:
    """Loads the HIP library (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}`")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.dl_ctx_create.restype = C.c_int
    lib.dl_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
    lib.dl_ctx_destroy.restype = None
    lib.dl_ctx_destroy.argtypes = [P]
    lib.dl_last_error.restype = C.c_char_p
    lib.dl_last_error.argtypes = [P]
    lib.dl_default_options.restype = None
    lib.dl_default_options.argtypes = [C.POINTER(abi.dl_bake_options)]
    lib.dl_bake.restype = C.c_int
    lib.dl_bake.argtypes = [P, C.POINTER(abi.dl_bake_input), C.POINTER(abi.dl_bake_options),
                            C.POINTER(abi.dl_progress), C.POINTER(P)]
    lib.dl_bake_batch.restype = C.c_int
    lib.dl_bake_batch.argtypes = [P, C.POINTER(abi.dl_bake_input), C.POINTER(abi.dl_bake_options), C.c_uint32,
                                  C.c_uint32, C.POINTER(P), C.POINTER(C.c_int)]
    lib.dl_service_create.restype = C.c_int
    lib.dl_service_create.argtypes = [P, C.POINTER(abi.dl_bake_input), C.POINTER(abi.dl_bake_options), C.POINTER(P),
                                      C.POINTER(P)]
    lib.dl_service_bake.restype = C.c_int
    lib.dl_service_bake.argtypes = [P, C.c_void_p, C.c_uint32, C.POINTER(P)]
    lib.dl_service_replay.restype = C.c_int
    lib.dl_service_replay.argtypes = [P, C.POINTER(abi.dl_service_request), C.c_uint32, C.c_uint32,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(P)]
    lib.dl_service_ray_queries_noalloc.restype = C.c_int
    lib.dl_service_ray_queries_noalloc.argtypes = [P, C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p,
                                                   C.POINTER(C.c_uint64)]
    lib.dl_service_apply.restype = C.c_int
    lib.dl_service_apply.argtypes = [P, C.POINTER(abi.dl_service_edit), C.POINTER(abi.dl_service_update)]
    lib.dl_service_replay_edits.restype = C.c_int
    lib.dl_service_replay_edits.argtypes = [P, C.POINTER(abi.dl_service_edit), C.c_uint32, C.c_uint32, C.c_uint64,
                                            C.c_void_p, C.c_void_p, C.POINTER(abi.dl_service_update)]
    lib.dl_service_scene.restype = C.c_int
    lib.dl_service_scene.argtypes = [P, C.POINTER(P)]
    lib.dl_service_words.restype = C.c_int
    lib.dl_service_words.argtypes = [P, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
    lib.dl_service_destroy.restype = None
    lib.dl_service_destroy.argtypes = [P]
    lib.dl_encode.restype = C.c_int
    lib.dl_encode.argtypes = [P, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint64)]
    lib.dl_bytes_free.restype = None
    lib.dl_bytes_free.argtypes = [C.POINTER(C.c_uint8)]
    lib.dl_decode.restype = C.c_int
    lib.dl_decode.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(P)]
    lib.dl_decode_stream.restype = C.c_int
    lib.dl_decode_stream.argtypes = [READ_FN, C.c_void_p, C.POINTER(P)]
    lib.dl_format_error_kind.restype = C.c_int
    lib.dl_format_error_kind.argtypes = []
    lib.dl_scene_buffers.restype = C.c_int
    lib.dl_scene_buffers.argtypes = [P, C.POINTER(P)]
    f32p_, u32p_ = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    lib.dl_mesh_content_hash.restype = C.c_uint64
    lib.dl_mesh_content_hash.argtypes = [f32p_, C.c_uint32, u32p_, C.c_uint64, f32p_, C.c_uint32]
    lib.dl_bake_id_from_path.restype = C.c_uint64
    lib.dl_bake_id_from_path.argtypes = [C.POINTER(C.c_char_p), C.POINTER(C.c_uint64), C.c_uint32]
    lib.dl_assign_ids.restype = C.c_int
    lib.dl_assign_ids.argtypes = [C.POINTER(abi.dl_id_candidate), C.c_uint32, C.POINTER(C.c_uint64)]
    lib.dl_mesh_to_world.restype = C.c_int
    lib.dl_mesh_to_world.argtypes = [f32p_, f32p_, C.c_uint32, f32p_, u32p_, C.c_uint64, f32p_, C.c_uint32, f32p_, f32p_,
                                     u32p_, C.POINTER(C.c_uint64)]
    lib.dl_take_scene_buffers.restype = C.c_int
    lib.dl_take_scene_buffers.argtypes = [P, C.POINTER(P)]
    lib.dl_scene_triangles_taken.restype = C.c_int
    lib.dl_scene_triangles_taken.argtypes = [P]
    for name in ("dl_buffers_triangles", "dl_buffers_aux"):
        getattr(lib, name).restype = C.c_uint64
        getattr(lib, name).argtypes = [P, C.POINTER(C.POINTER(C.c_uint32))]
    lib.dl_buffers_receiver_count.restype = C.c_uint32
    lib.dl_buffers_receiver_count.argtypes = [P]
    lib.dl_buffers_receiver.restype = C.c_int
    lib.dl_buffers_receiver.argtypes = [P, C.c_uint32, C.POINTER(abi.dl_receiver)]
    lib.dl_baked_buffers_free.restype = None
    lib.dl_baked_buffers_free.argtypes = [P]
    lib.dl_baked_scene_free.restype = None
    lib.dl_baked_scene_free.argtypes = [P]
    lib.dl_last_timings.restype = C.c_int
    lib.dl_last_timings.argtypes = [P, C.POINTER(abi.dl_timings)]
    abi.bind_scene_api(lib, "dl_")
    f32p, u32p = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
    lib.dl_test_ray_queries.restype = C.c_int
    lib.dl_test_ray_queries.argtypes = [P, f32p, C.c_uint32, f32p, C.c_uint64, C.c_int, f32p, u32p]
    lib.dl_test_trig.restype = C.c_int
    lib.dl_test_trig.argtypes = [P, f32p, C.c_uint64, f32p, f32p]
    lib.dl_test_trig_host.restype = C.c_int
    lib.dl_test_trig_host.argtypes = [f32p, C.c_uint64, f32p, f32p]
    lib.dl_test_trig_exhaustive.restype = C.c_int
    lib.dl_test_trig_exhaustive.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
    lib.dl_test_bvh_structure.restype = C.c_int
    lib.dl_test_bvh_structure.argtypes = [f32p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
    lib.dl_test_trig_keys.restype = C.c_uint64
    lib.dl_test_trig_keys.argtypes = [C.c_int, C.POINTER(C.c_uint32), C.c_uint64]
    lib.dl_test_trig_table_sizes.restype = None
    lib.dl_test_trig_table_sizes.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.dl_set_burial.restype = C.c_int
    lib.dl_set_burial.argtypes = [P, C.c_uint32]
    lib.dl_set_counting.restype = C.c_int
    lib.dl_set_counting.argtypes = [P, C.c_int]
    lib.dl_set_bounce_capture.restype = C.c_int
    lib.dl_set_bounce_capture.argtypes = [P, C.c_int]
    lib.dl_test_bounce_capture.restype = C.c_uint64
    lib.dl_test_bounce_capture.argtypes = [P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(f32p)]
    lib.dl_set_launch.restype = C.c_int
    lib.dl_set_launch.argtypes = [P, C.c_uint32, C.c_uint32]
    lib.dl_test_batch_trace.restype = C.c_uint32
    lib.dl_test_batch_trace.argtypes = [P, C.POINTER(abi.dl_timings), C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                        C.c_uint32]
    lib.dl_set_trace_stream.restype = C.c_int
    lib.dl_set_trace_stream.argtypes = [P, C.c_int]
    lib.dl_autotune_batch.restype = C.c_int
    lib.dl_autotune_batch.argtypes = [P, C.POINTER(abi.dl_bake_input), C.POINTER(abi.dl_bake_options), C.c_uint32,
                                      C.c_uint32, C.c_char_p, C.POINTER(abi.dl_tuned_launch)]
    lib.dl_set_gather_split.restype = C.c_int
    lib.dl_set_gather_split.argtypes = [P, C.c_int]
    f32p = C.POINTER(C.c_float)
    lib.dl_test_gather_directions.restype = C.c_int
    lib.dl_test_gather_directions.argtypes = [P, f32p, f32p, f32p, f32p, C.c_uint64, f32p, f32p, C.POINTER(C.c_uint32)]
    lib.dl_set_photon_tiles.restype = C.c_int
    lib.dl_set_photon_tiles.argtypes = [P, C.c_int]
    lib.dl_set_lazy_photons.restype = C.c_int
    lib.dl_set_lazy_photons.argtypes = [P, C.c_int]
    lib.dl_test_light_bvh.restype = C.c_uint32
    lib.dl_test_light_bvh.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32,
                                      C.POINTER(abi.dl_light_bvh_node), C.c_uint32, C.POINTER(C.c_uint32)]
    lib.dl_set_gather.restype = C.c_int
    lib.dl_set_gather.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32]
    _lib = lib
    return lib


READ_FN = C.CFUNCTYPE(C.c_int64, C.c_void_p, C.POINTER(C.c_uint8), C.c_uint64)


def _format_error(lib) -> BakeError:
    e = BakeError(abi.DL_E_FORMAT, lib.dl_last_error(None).decode())
    e.format_kind = lib.dl_format_error_kind()
    return e


def _encode_handle(lib, h) -> bytes:
    buf, n = C.POINTER(C.c_uint8)(), C.c_uint64()
    st = lib.dl_encode(h, C.byref(buf), C.byref(n))
    if st != 0:
        raise BakeError(st, lib.dl_last_error(None).decode())
    try:
        return C.string_at(buf, n.value)
    finally:
        lib.dl_bytes_free(buf)


def _decode_handle(lib, data: bytes):
    h = C.c_void_p()
    st = lib.dl_decode(data, len(data), C.byref(h))
    if st == abi.DL_E_FORMAT:
        raise _format_error(lib)
    if st != 0:
        raise BakeError(st, lib.dl_last_error(None).decode())
    return h


def decode(data: bytes) -> abi.BakedScene:
    """decode (src/bake/format/mod.rs:152-169): a .dynlight file held in memory."""
    lib = load()
    h = _decode_handle(lib, bytes(data))
    try:
        return abi.read_scene(lib, "dl_", h)
    finally:
        lib.dl_baked_scene_free(h)


def _decode_stream_handle(lib, source):
    failed = []

    def read(_user, buf, cap):
        try:
            chunk = source.read(int(cap))
        except Exception as exc:  # the source's own failure
            failed.append(exc)
            return -1
        C.memmove(buf, chunk, len(chunk))
        return len(chunk)

    cb = READ_FN(read)
    h = C.c_void_p()
    st = lib.dl_decode_stream(cb, None, C.byref(h))
    if st == abi.DL_E_FORMAT:
        raise _format_error(lib)
    if st != 0:
        raise BakeError(st, lib.dl_last_error(None).decode())
    return h


def decode_stream(source) -> abi.BakedScene:
    """decode_stream (src/bake/format/stream.rs:34-71): a .dynlight file read from `source`
    (anything with read(n)) as it is parsed."""
    lib = load()
    h = _decode_stream_handle(lib, source)
    try:
        return abi.read_scene(lib, "dl_", h)
    finally:
        lib.dl_baked_scene_free(h)


def load_bake(source, copy: bool = True):
    """The loader path (baked_lighting/loader.rs:47-68): decode_stream of `source`, then
    take_scene_buffers, which releases each mesh's words as it copies them into the buffers.
    Returns (scene, buffers): the scene's meshes keep their ids, hashes and resolutions but no
    words (its `triangles_taken` is set).  copy=False returns only the buffer sizes (no Python
    copies of the words), for memory measurements."""
    lib = load()
    h = _decode_stream_handle(lib, source)
    try:
        buffers = _buffers_of_handle(lib, h, take=True, copy=copy)
        scene = abi.read_scene(lib, "dl_", h)
        scene.triangles_taken = bool(lib.dl_scene_triangles_taken(h))
        return scene, buffers
    finally:
        lib.dl_baked_scene_free(h)


def _buffers_of_handle(lib, h, take: bool = False, copy: bool = True) -> dict:
    """scene_buffers (src/baked_lighting/apply.rs:27-38), or with take=True take_scene_buffers
    (apply.rs:45-61), of a library scene: the triangles and aux word arrays and the receivers in
    mesh order (tag = index + 1)."""
    b = C.c_void_p()
    st = (lib.dl_take_scene_buffers if take else lib.dl_scene_buffers)(h, C.byref(b))
    if st != 0:
        raise BakeError(st, lib.dl_last_error(None).decode())
    try:
        out = {}
        for name in ("triangles", "aux"):
            p = C.POINTER(C.c_uint32)()
            n = getattr(lib, f"dl_buffers_{name}")(b, C.byref(p))
            if not copy:
                out[name] = int(n)
                continue
            out[name] = np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
        recs = []
        for i in range(lib.dl_buffers_receiver_count(b)):
            r = abi.dl_receiver()
            lib.dl_buffers_receiver(b, i, C.byref(r))
            recs.append({"id": r.id, "content_hash": r.content_hash, "tag": r.tag, "triangle_count": r.triangle_count,
                         "occludes": bool(r.occludes)})
        out["receivers"] = recs
        return out
    finally:
        lib.dl_baked_buffers_free(b)


def scene_buffers_of_file(data: bytes) -> dict:
    """The BakedBuffers of a decoded .dynlight file (decode, then scene_buffers)."""
    lib = load()
    h = _decode_handle(lib, bytes(data))
    try:
        return _buffers_of_handle(lib, h)
    finally:
        lib.dl_baked_scene_free(h)


def reencode(data: bytes) -> bytes:
    """decode followed by encode (src/bake/format/mod.rs:101-119) of the decoded scene."""
    lib = load()
    h = _decode_handle(lib, bytes(data))
    try:
        return _encode_handle(lib, h)
    finally:
        lib.dl_baked_scene_free(h)


class Context:
    """One HIP device with its stream and resident sin/cos tables (dl_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.dl_ctx_create(device, C.byref(h))
        if st != 0:
            raise BakeError(st, self.lib.dl_last_error(None).decode())
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.lib.dl_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def error(self) -> str:
        return self.lib.dl_last_error(self.handle).decode()

    def bake(self, inp: BakeInput, options: abi.dl_bake_options | None = None,
             progress: abi.dl_progress | None = None) -> abi.BakedScene:
        """bake() (src/bake/baker.rs:27-203) on the GPU."""
        options = options if options is not None else default_options()
        h = C.c_void_p()
        st = self.lib.dl_bake(self.handle, C.byref(inp.struct), C.byref(options),
                              C.byref(progress) if progress is not None else None, C.byref(h))
        if st != 0:
            raise BakeError(st, self.error())
        try:
            return abi.read_scene(self.lib, "dl_", h)
        finally:
            self.lib.dl_baked_scene_free(h)

    def bake_batch(self, inputs: list[BakeInput], options: list[abi.dl_bake_options] | abi.dl_bake_options | None = None,
                   streams: int = 4, read: bool | str = True) -> list | None:
        """Bakes independent jobs on `streams` HIP streams at once (dl_bake_batch); each result
        equals bake() of the same job alone.  read=True returns BakedScenes, read="stats" only
        each job's stats, read=False nothing (the results are freed unread)."""
        n = len(inputs)
        if options is None or isinstance(options, abi.dl_bake_options):
            options = [options if options is not None else default_options()] * n
        ins = (abi.dl_bake_input * max(1, n))(*[i.struct for i in inputs])
        ops = (abi.dl_bake_options * max(1, n))(*options)
        outs = (C.c_void_p * max(1, n))()
        status = (C.c_int * max(1, n))()
        st = self.lib.dl_bake_batch(self.handle, ins, ops, n, streams, outs, status)
        try:
            if st != 0:
                raise BakeError(st, self.error())
            if read == "stats":
                out = []
                for i in range(n):
                    st = abi.dl_bake_stats()
                    self.lib.dl_scene_stats(outs[i], C.byref(st))
                    out.append({name: getattr(st, name) for name, _ in st._fields_})
                return out
            return [abi.read_scene(self.lib, "dl_", outs[i]) for i in range(n)] if read else None
        finally:
            for i in range(n):
                if outs[i]:
                    self.lib.dl_baked_scene_free(outs[i])

    def autotune_batch(self, inputs: list[BakeInput], options: list[abi.dl_bake_options] | abi.dl_bake_options | None = None,
                       max_streams: int = 16, cache_path: str | None = None) -> dict:
        """dl_autotune_batch: times launch shapes on a sample of the jobs, applies the fastest to
        this context and returns them (`streams` is for bake_batch); cache_path keeps them per
        batch signature."""
        n = len(inputs)
        if options is None or isinstance(options, abi.dl_bake_options):
            options = [options if options is not None else default_options()] * n
        ins = (abi.dl_bake_input * max(1, n))(*[i.struct for i in inputs])
        ops = (abi.dl_bake_options * max(1, n))(*options)
        t = abi.dl_tuned_launch()
        st = self.lib.dl_autotune_batch(self.handle, ins, ops, n, max_streams,
                                        cache_path.encode() if cache_path else None, C.byref(t))
        if st != 0:
            raise BakeError(st, self.error())
        return {name: getattr(t, name) for name, _ in t._fields_}

    def batch_trace(self) -> list[dict]:
        """Per job of the last bake_batch: worker, start/end ms from the batch start, timings."""
        n = self.lib.dl_test_batch_trace(self.handle, None, None, None, 0)
        tm = (abi.dl_timings * max(1, n))()
        wk = (C.c_uint32 * max(1, n))()
        sp = (C.c_double * max(2, 2 * n))()
        self.lib.dl_test_batch_trace(self.handle, tm, wk, sp, n)
        return [{"worker": wk[i], "start_ms": sp[2 * i], "end_ms": sp[2 * i + 1],
                 **{name: getattr(tm[i], name) for name, _ in tm[i]._fields_}} for i in range(n)]

    def bake_encoded(self, inp: BakeInput, options: abi.dl_bake_options | None = None) -> tuple[abi.BakedScene, bytes]:
        """bake() and encode() of its result: the scene and its .dynlight bytes."""
        options = options if options is not None else default_options()
        h = C.c_void_p()
        st = self.lib.dl_bake(self.handle, C.byref(inp.struct), C.byref(options), None, C.byref(h))
        if st != 0:
            raise BakeError(st, self.error())
        try:
            return abi.read_scene(self.lib, "dl_", h), _encode_handle(self.lib, h)
        finally:
            self.lib.dl_baked_scene_free(h)

    def bake_buffers(self, inp: BakeInput, options: abi.dl_bake_options | None = None) -> tuple[abi.BakedScene, dict]:
        """bake() and scene_buffers() of its result."""
        options = options if options is not None else default_options()
        h = C.c_void_p()
        st = self.lib.dl_bake(self.handle, C.byref(inp.struct), C.byref(options), None, C.byref(h))
        if st != 0:
            raise BakeError(st, self.error())
        try:
            return abi.read_scene(self.lib, "dl_", h), _buffers_of_handle(self.lib, h)
        finally:
            self.lib.dl_baked_scene_free(h)

    def timings(self) -> dict:
        t = abi.dl_timings()
        self.lib.dl_last_timings(self.handle, C.byref(t))
        return {name: getattr(t, name) for name, _ in t._fields_}

    def set_counting(self, enable: bool) -> None:
        """Counts BVH node fetches and triangle tests of the shadow-ray kernel in later bakes."""
        self.lib.dl_set_counting(self.handle, 1 if enable else 0)

    def bake_bounce_crops(self, inp: BakeInput, options: abi.dl_bake_options | None = None):
        """bake() with the bounce capture on: (scene, {(mesh id, triangle, light id): (h, w, 3)
        float32}), the pre-pack f32 crop of every kept bounce entry."""
        self.lib.dl_set_bounce_capture(self.handle, 1)
        try:
            scene = self.bake(inp, options)
            n = self.lib.dl_test_bounce_capture(self.handle, 0, None, None)
            meta = (C.c_uint64 * 5)()
            vals = C.POINTER(C.c_float)()
            out = {}
            for i in range(n):
                self.lib.dl_test_bounce_capture(self.handle, i, meta, C.byref(vals))
                w, h = int(meta[3]), int(meta[4])
                out[(int(meta[0]), int(meta[1]), int(meta[2]))] = \
                    np.ctypeslib.as_array(vals, shape=(h * w * 3,)).copy().reshape(h, w, 3)
        finally:
            self.lib.dl_set_bounce_capture(self.handle, 0)
        return scene, out

    def set_burial(self, lanes_block: int = 0) -> None:
        """Burial test form: 0 streams its rays (closest-hit ray stream); otherwise one lane per
        texel, lanes_block threads per block."""
        st = self.lib.dl_set_burial(self.handle, lanes_block)
        if st != 0:
            raise BakeError(st, self.error())

    def set_launch(self, trace_block: int = 0, burial_block: int = 0) -> None:
        st = self.lib.dl_set_launch(self.handle, trace_block, burial_block)
        if st != 0:
            raise BakeError(st, self.error())

    def set_trace_stream(self, enable: bool) -> None:
        """Area lights, soft suns and skies through the streamed any-hit path (default) or per lane."""
        self.lib.dl_set_trace_stream(self.handle, 1 if enable else 0)

    def set_gather_split(self, enable: bool) -> None:
        """The gather's sample setup as a fast pass plus an exact pass for the samples it leaves open
        (default), or the exact chain for every sample; the results are the same."""
        self.lib.dl_set_gather_split(self.handle, 1 if enable else 0)

    def gather_directions(self, dirs, z, azimuth, spread):
        """dl_test_gather_directions: (exact (n, 3), fast (n, 3), fast_ok (n,)) float32/bool arrays."""
        dirs = np.ascontiguousarray(dirs, np.float32)
        z, azimuth, spread = (np.ascontiguousarray(v, np.float32) for v in (z, azimuth, spread))
        n = len(z)
        exact = np.zeros((n, 3), np.float32)
        fast = np.zeros((n, 3), np.float32)
        ok = np.zeros(n, np.uint32)
        f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        st = self.lib.dl_test_gather_directions(self.handle, f(dirs), f(z), f(azimuth), f(spread), n, f(exact), f(fast),
                                                ok.ctypes.data_as(C.POINTER(C.c_uint32)))
        if st != 0:
            raise BakeError(st, self.error())
        return exact, fast, ok.astype(bool)

    def set_photon_tiles(self, enable: bool) -> None:
        """The point-light gather skips the photon reads its cube's tile summaries decide (default)
        or reads every sample's photon; the results are the same."""
        self.lib.dl_set_photon_tiles(self.handle, 1 if enable else 0)

    def set_lazy_photons(self, enable: bool) -> None:
        """Bakes with few gather samples per photon-cube texel trace only the texels their samples
        can read (default), or every texel of every cube; the results are the same."""
        self.lib.dl_set_lazy_photons(self.handle, 1 if enable else 0)

    def set_gather(self, mode: int = 1, block: int = 0, chunk: int = 0, chunk_rays: int = 0, blocks_per_cu: int = 0,
                   min_waves: int = 0, refill: int = 0, order: int = 0) -> None:
        """Point-light bounce gather: 0 = one wave per (triangle, light, 64 texels), 1 = split
        gather (rays, persistent any-hit traversal, ordered sums).  order: 1 job-major ray slots,
        2 sample-major, 0 keep."""
        st = self.lib.dl_set_gather(self.handle, mode, block, chunk, chunk_rays, blocks_per_cu, min_waves, refill, order)
        if st != 0:
            raise BakeError(st, self.error())


class Service:
    """A rebake service (dl_service): the scene's geometry stays resident on the device and later
    bakes of other lights reuse it.  `initial` is the first bake of the scene with its lights."""

    def __init__(self, ctx: Context, inp: BakeInput, options: abi.dl_bake_options | None = None):
        self.ctx, self.lib = ctx, ctx.lib
        self.options = options if options is not None else default_options()
        h, first = C.c_void_p(), C.c_void_p()
        st = self.lib.dl_service_create(ctx.handle, C.byref(inp.struct), C.byref(self.options), C.byref(h), C.byref(first))
        if st != 0:
            raise BakeError(st, ctx.error())
        self.handle = h
        try:
            self.initial = abi.read_scene(self.lib, "dl_", first)
        finally:
            self.lib.dl_baked_scene_free(first)

    def bake(self, lights) -> abi.BakedScene:
        """The scene's meshes baked with `lights` (a list of LightData), reusing the resident geometry."""
        packed = BakeInput([], list(lights))
        h = C.c_void_p()
        st = self.lib.dl_service_bake(self.handle, packed.struct.lights, packed.struct.light_count, C.byref(h))
        if st != 0:
            raise BakeError(st, self.ctx.error())
        try:
            return abi.read_scene(self.lib, "dl_", h)
        finally:
            self.lib.dl_baked_scene_free(h)

    def apply(self, lights=(), remove=()) -> dict:
        """Edits the live scene (dl_service_apply): `lights` are added or replace the lights of the same
        id, `remove` lists ids to drop.  Returns the update record."""
        packed = BakeInput([], list(lights))
        rm = (C.c_uint64 * max(1, len(remove)))(*remove)
        e = abi.dl_service_edit(packed.struct.lights, packed.struct.light_count, rm, len(remove), 0.0, None)
        u = abi.dl_service_update()
        st = self.lib.dl_service_apply(self.handle, C.byref(e), C.byref(u))
        if st != 0:
            raise BakeError(st, self.ctx.error())
        return {name: getattr(u, name) for name, _ in u._fields_}

    def add_room(self, room: BakeInput) -> dict:
        """Adds a room to the live scene (dl_service_apply with a room): a scene of its own that
        joins as a part of the live scene when isolated from it.  Returns the update record."""
        e = abi.dl_service_edit(None, 0, None, 0, 0.0, C.cast(C.pointer(room.struct), C.c_void_p))
        u = abi.dl_service_update()
        st = self.lib.dl_service_apply(self.handle, C.byref(e), C.byref(u))
        if st != 0:
            raise BakeError(st, self.ctx.error())
        return {name: getattr(u, name) for name, _ in u._fields_}

    def replay_edits(self, edits, workers: int = 4, budget_bytes: int = 0, checksums: bool = False):
        """Serves edits as they arrive (dl_service_replay_edits): (lights, remove, arrival_s) light
        edits, or (room BakeInput, None, arrival_s) room additions; tracing on `workers` streams
        under a device-memory budget, joining the live scene in arrival order.  Returns
        (latencies in ms, update records, checksums or None)."""
        n = len(edits)
        rooms = [e[0] if isinstance(e[0], BakeInput) else None for e in edits]
        packed = [BakeInput([], []) if rm is not None else BakeInput([], list(e[0])) for rm, e in zip(rooms, edits)]
        rms = [(C.c_uint64 * max(1, len(e[1] or ())))(*(e[1] or ())) for e in edits]
        arr = (abi.dl_service_edit * max(1, n))()
        for i, (p, rm, e) in enumerate(zip(packed, rms, edits)):
            room = C.cast(C.pointer(rooms[i].struct), C.c_void_p) if rooms[i] is not None else None
            arr[i] = abi.dl_service_edit(p.struct.lights, p.struct.light_count, rm, len(e[1] or ()), float(e[2]), room)
        lat = (C.c_double * max(1, n))()
        sums = (C.c_uint64 * max(1, n))()
        info = (abi.dl_service_update * max(1, n))()
        st = self.lib.dl_service_replay_edits(self.handle, arr, n, workers, budget_bytes, lat, sums if checksums else None, info)
        if st != 0:
            raise BakeError(st, self.ctx.error())
        ups = [{name: getattr(info[i], name) for name, _ in info[i]._fields_} for i in range(n)]
        return list(lat[:n]), ups, (list(sums[:n]) if checksums else None)

    def scene(self) -> abi.BakedScene:
        """The live scene (dl_service_scene), read back to the host."""
        h = C.c_void_p()
        st = self.lib.dl_service_scene(self.handle, C.byref(h))
        if st != 0:
            raise BakeError(st, self.ctx.error())
        try:
            return abi.read_scene(self.lib, "dl_", h)
        finally:
            self.lib.dl_baked_scene_free(h)

    def words(self) -> tuple[int, int]:
        """(device address, word count) of the live scene's triangle words (dl_service_words)."""
        p, n = C.c_void_p(), C.c_uint64()
        st = self.lib.dl_service_words(self.handle, C.byref(p), C.byref(n))
        if st != 0:
            raise BakeError(st, self.ctx.error())
        return p.value or 0, n.value

    def replay(self, requests, workers: int = 4, read: bool | str = False, service_times: list | None = None):
        """Serves (lights, arrival_s) requests as they arrive on `workers` streams.  Returns the
        latencies in ms (arrival to result on the host) and, with read=True, the results
        (read="stats": only their stats).  service_times, if a list, receives each request's time
        from admission to result."""
        n = len(requests)
        packed = [BakeInput([], list(l)) for l, _ in requests]
        reqs = (abi.dl_service_request * max(1, n))()
        for i, (p, (_, t)) in enumerate(zip(packed, requests)):
            reqs[i].lights = p.struct.lights
            reqs[i].light_count = p.struct.light_count
            reqs[i].arrival_s = float(t)
        lat = (C.c_double * max(1, n))()
        svc_ms = (C.c_double * max(1, n))()
        status = (C.c_int * max(1, n))()
        outs = (C.c_void_p * max(1, n))()
        st = self.lib.dl_service_replay(self.handle, reqs, n, workers, lat, svc_ms, status, outs if read else None)
        if service_times is not None:
            service_times[:] = list(svc_ms[:n])
        try:
            if st != 0:
                raise BakeError(st, self.ctx.error())
            results = None
            if read == "stats":
                results = []
                for i in range(n):
                    sst = abi.dl_bake_stats()
                    self.lib.dl_scene_stats(outs[i], C.byref(sst))
                    results.append({name: getattr(sst, name) for name, _ in sst._fields_})
            elif read:
                results = [abi.read_scene(self.lib, "dl_", outs[i]) for i in range(n)]
            return list(lat[:n]), results
        finally:
            if read:
                for i in range(n):
                    if outs[i]:
                        self.lib.dl_baked_scene_free(outs[i])

    def close(self):
        if getattr(self, "handle", None):
            self.lib.dl_service_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


_default_ctx = None


def bake(inp: BakeInput, options: abi.dl_bake_options | None = None,
         progress: abi.dl_progress | None = None) -> abi.BakedScene:
    """bake() on device 0 with a process-wide context."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx.bake(inp, options, progress)


# ---- the ECS caller's conversions (src/bake/snapshot.rs, identity.rs, shared/bake_id.rs) -------
def _f32(a, cols):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, cols))


def mesh_content_hash(positions, indices, uv) -> int:
    """mesh_content_hash (src/shared/mesh_hash.rs:16-33)."""
    lib = load()
    p, u = _f32(positions, 3), _f32(uv, 2)
    i = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1))
    return int(lib.dl_mesh_content_hash(p.ctypes.data_as(C.POINTER(C.c_float)), len(p), i.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        len(i), u.ctypes.data_as(C.POINTER(C.c_float)), len(u)))


def _segments(segments):
    raw = [s.encode() if isinstance(s, str) else bytes(s) for s in segments]
    ptrs = (C.c_char_p * max(1, len(raw)))(*raw)
    lens = (C.c_uint64 * max(1, len(raw)))(*[len(r) for r in raw])
    return raw, ptrs, lens


def bake_id_from_path(segments) -> int:
    """BakeId::from_path (src/shared/bake_id.rs:15-23)."""
    raw, ptrs, lens = _segments(segments)
    return int(load().dl_bake_id_from_path(ptrs, lens, len(raw)))


def assign_ids(candidates) -> list[int]:
    """assign_ids (src/bake/identity.rs:64-101) of (explicit id or None, name path, tiebreak hash,
    position) candidates."""
    lib = load()
    keep, arr = [], (abi.dl_id_candidate * max(1, len(candidates)))()
    for k, (explicit, segments, tiebreak, position) in enumerate(candidates):
        raw, ptrs, lens = _segments(segments)
        keep.append((raw, ptrs, lens))
        c = arr[k]
        c.has_explicit = 0 if explicit is None else 1
        c.explicit_id = 0 if explicit is None else explicit
        c.segment_count = len(raw)
        c.segments = ptrs
        c.segment_lengths = lens
        c.tiebreak_hash = tiebreak
        c.position = tuple(position)
    out = (C.c_uint64 * max(1, len(candidates)))()
    st = lib.dl_assign_ids(arr, len(candidates), out)
    if st != 0:
        raise BakeError(st, lib.dl_last_error(None).decode())
    return [int(out[k]) for k in range(len(candidates))]


def mesh_to_world(mesh_id: int, affine, positions, indices, uv1, normals=None) -> MeshData:
    """MeshData::to_world (src/bake/snapshot.rs:229-272): a BakeMesh in world space from local
    data and a 3x4 affine transform (columns of the linear part, then the translation)."""
    lib = load()
    fp = C.POINTER(C.c_float)
    a = np.ascontiguousarray(np.asarray(affine, np.float32).reshape(12))
    p, u = _f32(positions, 3), _f32(uv1, 2)
    n = None if normals is None else _f32(normals, 3)
    i = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1))
    po, no, io_ = np.zeros_like(p), (None if n is None else np.zeros_like(n)), np.zeros_like(i)
    h = C.c_uint64()
    st = lib.dl_mesh_to_world(a.ctypes.data_as(fp), p.ctypes.data_as(fp), len(p), None if n is None else n.ctypes.data_as(fp),
                              i.ctypes.data_as(C.POINTER(C.c_uint32)), len(i), u.ctypes.data_as(fp), len(u),
                              po.ctypes.data_as(fp), None if no is None else no.ctypes.data_as(fp),
                              io_.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(h))
    if st < 0:
        raise BakeError(st, lib.dl_last_error(None).decode())
    return MeshData(mesh_id, po, io_, u.copy(), normals=no if st == 1 else None, content_hash=int(h.value))

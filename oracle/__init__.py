"""ORACLE loader -- test infrastructure only.

Loads oracle/_build/libdl_oracle.so (the CPU restatement of the reference bake, built by
oracle/Makefile).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libdl_oracle.so")
_PKG = os.path.join(os.path.dirname(HERE), "dynamiclighting-bevy_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from dynlight_hip import abi  # noqa: E402

_lib = None


def build() -> None:
    """Compiles the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.dlo_bake.restype = C.c_int
    lib.dlo_bake.argtypes = [C.POINTER(abi.dl_bake_input), C.POINTER(abi.dl_bake_options), C.c_uint32,
                             C.c_uint32, C.POINTER(P)]
    lib.dlo_last_error.restype = C.c_char_p
    lib.dlo_bake_receivers.restype = C.c_int
    lib.dlo_bake_receivers.argtypes = [C.POINTER(abi.dl_bake_input), C.POINTER(abi.dl_bake_options), C.c_uint32,
                                       C.POINTER(C.c_uint8), C.POINTER(P)]
    abi.bind_scene_api(lib, "dlo_")
    f32p = C.POINTER(C.c_float)
    u32p = C.POINTER(C.c_uint32)
    u64p = C.POINTER(C.c_uint64)
    sigs = {
        "dlo_splitmix64": (C.c_uint64, [C.c_uint64]),
        "dlo_seed": (C.c_uint64, [u64p, C.c_uint32]),
        "dlo_pcg32": (None, [C.c_uint64, C.c_uint64, C.c_uint32, u32p]),
        "dlo_stream_f32": (None, [u64p, C.c_uint32, C.c_uint32, f32p]),
        "dlo_lightmap_size": (C.c_uint32, [C.c_float, C.c_float, C.c_uint32]),
        "dlo_rect": (None, [C.c_uint32, f32p, C.c_uint32, u32p]),
        "dlo_closest_point": (None, [f32p, f32p, f32p, f32p, f32p]),
        "dlo_uv_barycentrics": (C.c_int, [f32p, f32p, f32p]),
        "dlo_triangle_texels": (C.c_uint32, [f32p, f32p, f32p, f32p, C.c_uint32, f32p, C.c_uint32]),
        "dlo_touched_texels": (C.c_uint32, [f32p, f32p, f32p, f32p, C.c_uint32, u32p, f32p, C.c_uint32]),
        "dlo_level": (C.c_uint8, [C.c_float, C.c_uint32]),
        "dlo_spread": (None, [C.c_uint32, f32p, C.c_uint32, f32p]),
        "dlo_concentric_disk": (None, [f32p, f32p]),
        "dlo_is_buried": (C.c_int, [f32p, C.c_float, C.POINTER(C.c_uint8), u32p, f32p]),
        "dlo_area_visibility": (C.c_int, [C.POINTER(abi.dl_light), C.c_float, f32p, f32p, C.c_uint32,
                                          C.c_uint32, u64p, C.c_uint32, C.c_int32, f32p, f32p, u32p]),
        "dlo_sun_visibility": (C.c_int, [C.POINTER(abi.dl_light), f32p, C.c_uint32, C.c_uint32, u64p,
                                         C.c_uint32, C.c_int32, f32p, f32p, u32p]),
        "dlo_sky_visibility": (C.c_float, [f32p, C.c_uint32, u64p, C.c_uint32, C.c_int32, f32p]),
        "dlo_light_bvh": (C.c_uint32, [f32p, f32p, C.c_uint32, P, C.c_uint32, u32p]),
        "dlo_pad_seams": (None, [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
        "dlo_dilate": (None, [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8)]),
        "dlo_pack_levels": (C.c_uint32, [C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32, u32p]),
        "dlo_serialize": (C.c_uint32, [C.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p, u32p, C.c_uint32,
                                       u32p, C.c_uint32]),
        "dlo_pack_bounce": (C.c_uint32, [f32p, C.c_uint32, C.c_uint32, C.c_int32, u64p, C.c_uint32, u32p,
                                         C.c_uint32]),
        "dlo_f16_pair": (C.c_uint32, [C.c_float, C.c_float]),
        "dlo_ray_any": (C.c_int, [f32p, C.c_uint32, f32p, f32p, C.c_float, C.c_uint32]),
        "dlo_ray_closest": (C.c_int, [f32p, C.c_uint32, f32p, f32p, C.c_float, C.c_uint32, f32p, u32p]),
        "dlo_rays_batch": (None, [f32p, C.c_uint32, f32p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, u32p]),
        "dlo_sinf": (C.c_float, [C.c_float]),
        "dlo_cosf": (C.c_float, [C.c_float]),
        "dlo_cube_face_uv": (None, [f32p, u32p, f32p]),
        "dlo_cube_direction": (None, [C.c_uint32, f32p, f32p]),
        "dlo_cube_texel_direction": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, f32p]),
        "dlo_cube_texel": (None, [f32p, C.c_uint32, u32p]),
        "dlo_cube_texel_index": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
        "dlo_gaussian_weights": (None, [f32p, f32p]),
        "dlo_bounce_map": (None, [C.c_uint32, u32p, f32p, C.c_uint32, C.c_uint32, f32p]),
        "dlo_bounce_crop": (C.c_int, [C.c_uint32, u32p, f32p, C.c_uint32, u32p, f32p]),
        "dlo_slerp": (None, [f32p, f32p, C.c_float, f32p]),
        "dlo_acos_approx": (C.c_float, [C.c_float]),
        "dlo_on_unit_sphere": (None, [u64p, C.c_uint32, C.c_uint32, f32p]),
        "dlo_set_bounce_capture": (None, [C.c_uint32]),
        "dlo_set_window": (None, [C.c_uint64, u32p]),
        "dlo_bounce_capture_count": (C.c_uint64, []),
        "dlo_bounce_capture_get": (C.c_int, [C.c_uint64, u64p, C.POINTER(C.c_double), C.c_uint64]),
        "dlo_photon_cube": (C.c_int, [C.c_uint64, u32p, C.POINTER(f32p), C.POINTER(C.POINTER(C.c_int8)),
                                      C.POINTER(C.POINTER(C.c_uint8))]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_timings() -> dict:
    """Stage wall times (ms) of the last oracle bake on this thread."""
    lib = load()
    out = (C.c_double * 3)()
    lib.dlo_last_timings(out)
    return {"photon_cubes_ms": out[0], "meshes_ms": out[1], "distance_cubes_ms": out[2]}


def bake_receivers(inp: "abi.BakeInput", options: "abi.dl_bake_options", receivers, threads: int = 0):
    """Bakes only the meshes where `receivers` is true; every mesh still occludes."""
    import numpy as np
    lib = load()
    mask = np.ascontiguousarray(np.asarray(receivers, dtype=np.uint8))
    h = C.c_void_p()
    st = lib.dlo_bake_receivers(C.byref(inp.struct), C.byref(options), threads,
                                mask.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(h))
    if st != 0:
        raise abi.BakeError(st, lib.dlo_last_error().decode())
    try:
        return abi.read_scene(lib, "dlo_", h)
    finally:
        lib.dlo_scene_free(h)


def bake(inp: "abi.BakeInput", options: "abi.dl_bake_options", threads: int = 0,
         brute_force: bool = False) -> "abi.BakedScene":
    """The oracle's bake; raises abi.BakeError like the product."""
    lib = load()
    h = C.c_void_p()
    st = lib.dlo_bake(C.byref(inp.struct), C.byref(options), threads, 1 if brute_force else 0, C.byref(h))
    if st != 0:
        raise abi.BakeError(st, lib.dlo_last_error().decode())
    try:
        return abi.read_scene(lib, "dlo_", h)
    finally:
        lib.dlo_scene_free(h)


def bounce_crops(inp: "abi.BakeInput", options: "abi.dl_bake_options", fp64: bool = False, threads: int = 0,
                 photon_cubes: bool = False):
    """Bakes with the bounce capture on and returns {(mesh id, triangle, light id): (h, w, 3) float64}
    of every kept bounce entry: the f32 crops the bounce packing reads, or (fp64=True) the same
    entries gathered and filtered in double.  With photon_cubes=True also returns the bake's photon
    cubes {light id: (dist (6, r, r), normal (6, r, r, 3) int8, albedo (6, r, r, 3) uint8)}."""
    import numpy as np
    lib = load()
    lib.dlo_set_bounce_capture(2 if fp64 else 1)
    try:
        scene = bake(inp, options, threads=threads)
        n = lib.dlo_bounce_capture_count()
        out = {}
        meta = (C.c_uint64 * 5)()
        for i in range(n):
            lib.dlo_bounce_capture_get(i, meta, None, 0)
            w, h = int(meta[3]), int(meta[4])
            buf = np.zeros(w * h * 3, np.float64)
            lib.dlo_bounce_capture_get(i, meta, buf.ctypes.data_as(C.POINTER(C.c_double)), buf.size)
            out[(int(meta[0]), int(meta[1]), int(meta[2]))] = buf.reshape(h, w, 3)
        cubes = {}
        if photon_cubes:
            for light in inp.lights:
                res = C.c_uint32()
                d, nr, al = C.POINTER(C.c_float)(), C.POINTER(C.c_int8)(), C.POINTER(C.c_uint8)()
                if lib.dlo_photon_cube(light.id, C.byref(res), C.byref(d), C.byref(nr), C.byref(al)) != 0:
                    continue
                r = res.value
                k = 6 * r * r
                cubes[light.id] = (np.ctypeslib.as_array(d, (k,)).copy().reshape(6, r, r),
                                   np.ctypeslib.as_array(nr, (k * 3,)).copy().reshape(6, r, r, 3),
                                   np.ctypeslib.as_array(al, (k * 3,)).copy().reshape(6, r, r, 3))
    finally:
        lib.dlo_set_bounce_capture(0)
    return (scene, out, cubes) if photon_cubes else (scene, out)


def bake_window(inp: "abi.BakeInput", options: "abi.dl_bake_options", mesh_id: int, rect, threads: int = 0):
    """Bakes only mesh `mesh_id`, and in it only the triangles whose crop meets rect = (min_x,
    min_y, max_x, max_y); every mesh still occludes.  Triangles whose crop lies 8 texels inside
    rect get the words of the whole bake."""
    import numpy as np
    lib = load()
    r = np.array(rect, np.uint32)
    lib.dlo_set_window(mesh_id, r.ctypes.data_as(C.POINTER(C.c_uint32)))
    try:
        ids = inp.mesh_array["id"]
        return bake_receivers(inp, options, ids == mesh_id, threads=threads)
    finally:
        lib.dlo_set_window(0, None)

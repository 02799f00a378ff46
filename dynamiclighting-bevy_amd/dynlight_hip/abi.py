"""ctypes view of include/dynlight.h.

Input structs are built as numpy structured arrays whose layout equals the C structs (checked
against ctypes at import), so scenes with tens of thousands of meshes are marshalled without
per-mesh Python work.  The same arrays feed both the HIP library (prefix ``dl_``) and any other
library exporting the same accessors under another prefix.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

# ---- enums (include/dynlight.h) --------------------------------------------------------------
DL_OK = 0
DL_E_CANCELLED = -1
DL_E_DUPLICATE_ID = -2
DL_E_INVALID_MESH = -3
DL_E_INVALID_LIGHT = -4
DL_E_INVALID_OPTIONS = -5
DL_E_DEVICE = -16
DL_E_OOM = -17
DL_E_ARGUMENT = -18
DL_E_FORMAT = -19
DL_E_TRIANGLES_TAKEN = -20
DL_NOALLOC_MAX_RAYS = 1 << 30  # include/dynlight.h

ILLUMINATION_DIRECT, ILLUMINATION_SINGLE_BOUNCE = 0, 1
TRANSPARENCY_DISABLED, TRANSPARENCY_ENABLED, TRANSPARENCY_ENABLED_MAX = 0, 1, 2
BOUNCE_BITS8, BOUNCE_BITS6, BOUNCE_BITS5, BOUNCE_BITS4, BOUNCE_DEFAULT = 0, 1, 2, 3, -1
SHAPE_POINT, SHAPE_RECTANGLE, SHAPE_TUBE, SHAPE_SUN, SHAPE_SKY = 0, 1, 2, 3, 4
DYNAMIC_LIGHTING_ONLY, DYNAMIC_ANGULAR, DYNAMIC_DISTANCE_CUBES = 0, 1, 2

STATUS_NAMES = {
    DL_E_CANCELLED: "Cancelled",
    DL_E_DUPLICATE_ID: "DuplicateId",
    DL_E_INVALID_MESH: "InvalidMesh",
    DL_E_INVALID_LIGHT: "InvalidLight",
    DL_E_INVALID_OPTIONS: "InvalidOptions",
    DL_E_DEVICE: "Device",
    DL_E_OOM: "OutOfMemory",
    DL_E_ARGUMENT: "Argument",
    DL_E_FORMAT: "Format",
    DL_E_TRIANGLES_TAKEN: "TrianglesTaken",
}


class BakeError(Exception):
    """BakeError (src/bake/input.rs:198-225): `kind` names the variant, `code` the ABI status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code
        self.kind = STATUS_NAMES.get(code, f"status {code}")


# ---- ctypes structs --------------------------------------------------------------------------
class dl_alpha_texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba8", C.c_void_p),
                ("rgba8_len", C.c_uint64), ("cutoff", C.c_float)]


class dl_mesh(C.Structure):
    _fields_ = [("id", C.c_uint64), ("content_hash", C.c_uint64), ("vertex_count", C.c_uint32),
                ("positions", C.c_void_p), ("normals", C.c_void_p), ("index_count", C.c_uint32),
                ("indices", C.c_void_p), ("uv1", C.c_void_p), ("uv0", C.c_void_p),
                ("alpha", C.c_void_p), ("occludes", C.c_uint32), ("albedo", C.c_float * 3),
                ("uv1_count", C.c_uint32), ("uv0_count", C.c_uint32),
                ("normal_count", C.c_uint32)]


class dl_light(C.Structure):
    _fields_ = [("id", C.c_uint64), ("position", C.c_float * 3), ("radius", C.c_float),
                ("illumination", C.c_int32), ("transparency", C.c_int32),
                ("bounce_samples", C.c_uint32), ("bounce_intensity", C.c_float),
                ("bounce_compression", C.c_int32), ("shape", C.c_int32),
                ("rotation", C.c_float * 4), ("size", C.c_float * 2),
                ("angular_diameter", C.c_float)]


class dl_bake_input(C.Structure):
    _fields_ = [("meshes", C.c_void_p), ("mesh_count", C.c_uint32), ("lights", C.c_void_p),
                ("light_count", C.c_uint32)]


class dl_receiver(C.Structure):
    _fields_ = [("id", C.c_uint64), ("content_hash", C.c_uint64), ("tag", C.c_uint32), ("triangle_count", C.c_uint32),
                ("occludes", C.c_uint32), ("reserved", C.c_uint32)]


class dl_id_candidate(C.Structure):
    _fields_ = [("has_explicit", C.c_uint32), ("segment_count", C.c_uint32), ("explicit_id", C.c_uint64),
                ("segments", C.POINTER(C.c_char_p)), ("segment_lengths", C.POINTER(C.c_uint64)),
                ("tiebreak_hash", C.c_uint64), ("position", C.c_float * 3), ("reserved", C.c_uint32)]


class dl_service_edit(C.Structure):
    _fields_ = [("lights", C.c_void_p), ("light_count", C.c_uint32), ("remove", C.POINTER(C.c_uint64)),
                ("remove_count", C.c_uint32), ("arrival_s", C.c_double), ("room", C.c_void_p)]


class dl_service_update(C.Structure):
    _fields_ = [("version", C.c_uint64), ("dirty_meshes", C.c_uint32), ("reserved", C.c_uint32),
                ("traced_pairs", C.c_uint64), ("trace_ms", C.c_double), ("assemble_ms", C.c_double),
                ("total_ms", C.c_double), ("traced_samples", C.c_uint64), ("traced_rays", C.c_uint64),
                ("traced_bounce_rays", C.c_uint64), ("queue_ms", C.c_double)]


class dl_service_request(C.Structure):
    _fields_ = [("lights", C.c_void_p), ("light_count", C.c_uint32), ("arrival_s", C.c_double)]


class dl_bake_options(C.Structure):
    _fields_ = [("pixel_density", C.c_float), ("max_lightmap_size", C.c_uint32),
                ("bounce_compression", C.c_int32), ("dynamic_geometry", C.c_int32),
                ("skip_bounce", C.c_uint32), ("photon_cube_resolution", C.c_uint32),
                ("area_light_samples", C.c_uint32), ("area_light_min_samples", C.c_uint32),
                ("sun_map_resolution", C.c_uint32), ("sky_samples", C.c_uint32),
                ("sky_probe_spacing", C.c_float), ("colored_bounce", C.c_uint32),
                ("seed", C.c_uint64)]


class dl_progress(C.Structure):
    _fields_ = [("stage", C.c_uint32), ("done", C.c_uint64), ("total", C.c_uint64),
                ("cancel", C.c_uint32)]


class dl_bake_stats(C.Structure):
    _fields_ = [("meshes", C.c_uint32), ("triangles", C.c_uint64), ("texels", C.c_uint64),
                ("rays", C.c_uint64), ("lights_removed", C.c_uint64),
                ("bounce_rays", C.c_uint64), ("triangle_words", C.c_uint64),
                ("samples", C.c_uint64)]


class dl_light_bvh_node(C.Structure):
    _fields_ = [("aabb_min", C.c_float * 3), ("left_first", C.c_uint32),
                ("aabb_max", C.c_float * 3), ("count", C.c_uint32)]


class dl_baked_light(C.Structure):
    _fields_ = [("id", C.c_uint64), ("origin", C.c_float * 3), ("radius", C.c_float),
                ("shape", C.c_int32), ("rotation", C.c_float * 4), ("size", C.c_float * 2),
                ("angular_diameter", C.c_float), ("shadow_bits_per_texel", C.c_uint32),
                ("bounce_compression", C.c_int32)]


class dl_baked_mesh_view(C.Structure):
    _fields_ = [("id", C.c_uint64), ("content_hash", C.c_uint64), ("resolution", C.c_uint32),
                ("occludes", C.c_uint32), ("triangles", C.c_void_p),
                ("triangle_word_count", C.c_uint64)]


class dl_sun_map_view(C.Structure):
    _fields_ = [("rows", C.c_float * 12), ("resolution", C.c_uint32), ("texels", C.c_void_p),
                ("texel_word_count", C.c_uint64)]


class dl_sky_probes_view(C.Structure):
    _fields_ = [("present", C.c_uint32), ("origin", C.c_float * 3), ("spacing", C.c_float),
                ("counts", C.c_uint32 * 3), ("words", C.c_void_p), ("word_count", C.c_uint64)]


class dl_baked_settings(C.Structure):
    _fields_ = [("bounce", C.c_uint32), ("colored_bounce", C.c_uint32),
                ("dynamic_geometry", C.c_int32), ("pixel_density", C.c_float),
                ("max_lightmap_size", C.c_uint32), ("distance_cube_resolution", C.c_uint32)]


class dl_tuned_launch(C.Structure):
    _fields_ = [("streams", C.c_uint32), ("trace_block", C.c_uint32), ("stream_block", C.c_uint32),
                ("packet", C.c_uint32), ("burial_block", C.c_uint32), ("burial_lanes", C.c_uint32),
                ("trace_stream", C.c_uint32), ("trials", C.c_uint32), ("ms", C.c_double),
                ("from_cache", C.c_uint32)]


class dl_timings(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("upload_ms", C.c_double), ("texels_ms", C.c_double),
                ("trace_ms", C.c_double), ("bounce_ms", C.c_double), ("post_ms", C.c_double),
                ("maps_ms", C.c_double), ("download_ms", C.c_double),
                ("trace_kernel_ms", C.c_double), ("trace_launches", C.c_uint64),
                ("trace_rays", C.c_uint64), ("bvh_nodes", C.c_uint64),
                ("occluder_triangles", C.c_uint64), ("trace_node_visits", C.c_uint64),
                ("trace_tri_tests", C.c_uint64), ("bounce_kernel_ms", C.c_double),
                ("bounce_node_visits", C.c_uint64), ("bounce_tri_tests", C.c_uint64),
                ("bounce_launches", C.c_uint64), ("gather_trace_ms", C.c_double),
                ("gather_trace_launches", C.c_uint64), ("prepare_ms", C.c_double), ("bvh_ms", C.c_double),
                ("gather_rays_ms", C.c_double), ("gather_rays_launches", C.c_uint64), ("gather_jobs", C.c_uint64),
                ("gather_samples", C.c_uint64), ("gather_traced", C.c_uint64), ("burial_ms", C.c_double),
                ("burial_rays", C.c_uint64), ("burial_node_visits", C.c_uint64), ("burial_tri_tests", C.c_uint64),
                ("gather_photon_reads", C.c_uint64), ("gather_open", C.c_uint64),
                ("photon_texels", C.c_uint64)]


def _np_dtype(struct) -> np.dtype:
    names, formats, offsets = [], [], []
    for name, ctype in struct._fields_:
        names.append(name)
        offsets.append(getattr(struct, name).offset)
        if isinstance(ctype, type) and issubclass(ctype, C.Array):
            base = np.dtype(ctype._type_)
            formats.append((base, (ctype._length_,)))
        elif ctype is C.c_void_p:
            formats.append(np.uint64)
        else:
            formats.append(np.dtype(ctype))
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": C.sizeof(struct)})


MESH_DTYPE = _np_dtype(dl_mesh)
LIGHT_DTYPE = _np_dtype(dl_light)
ALPHA_DTYPE = _np_dtype(dl_alpha_texture)
assert MESH_DTYPE.itemsize == C.sizeof(dl_mesh)
assert LIGHT_DTYPE.itemsize == C.sizeof(dl_light)


def default_options() -> dl_bake_options:
    """BakeOptions::default() (src/bake/input.rs:166-171, src/bake/settings.rs:67-81)."""
    return dl_bake_options(
        pixel_density=128.0, max_lightmap_size=2048, bounce_compression=BOUNCE_BITS8,
        dynamic_geometry=DYNAMIC_ANGULAR, skip_bounce=0, photon_cube_resolution=512,
        area_light_samples=64, area_light_min_samples=16, sun_map_resolution=1024,
        sky_samples=64, sky_probe_spacing=1.0, colored_bounce=0, seed=0)


# ---- bake input: numpy-backed, pointer-stable ------------------------------------------------
@dataclass
class MeshData:
    """One BakeMesh (src/bake/input.rs:24-45) as numpy arrays."""
    id: int
    positions: np.ndarray            # (V, 3) float32
    indices: np.ndarray              # (3T,) uint32
    uv1: np.ndarray                  # (V, 2) float32
    normals: np.ndarray | None = None
    uv0: np.ndarray | None = None
    alpha: tuple | None = None       # (width, height, rgba8 uint8 array, cutoff)
    occludes: bool = True
    albedo: tuple = (1.0, 1.0, 1.0)
    content_hash: int = 0


@dataclass
class LightData:
    """One BakeLight (src/bake/input.rs:101-132)."""
    id: int
    position: tuple
    radius: float
    illumination: int = ILLUMINATION_DIRECT
    transparency: int = TRANSPARENCY_DISABLED
    bounce_samples: int = 32
    bounce_intensity: float = 1.0
    bounce_compression: int = BOUNCE_DEFAULT
    shape: int = SHAPE_POINT
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    size: tuple = (0.0, 0.0)
    angular_diameter: float = 0.0


class BakeInput:
    """BakeInput (src/bake/input.rs:17-20) packed into C-layout arrays.

    Meshes share concatenated vertex/index buffers; the struct array points into them. Keep the
    object alive while a bake that borrows it runs.
    """

    def __init__(self, meshes: list[MeshData] | None = None, lights: list[LightData] | None = None):
        self.meshes = list(meshes or [])
        self.lights = list(lights or [])
        self._pack()

    @classmethod
    def from_arrays(cls, *, ids, positions, indices, uv1, vertex_offsets, index_offsets,
                    normals=None, albedo=None, occludes=None, content_hash=None,
                    lights: list[LightData] | None = None):
        """Builds an input from already concatenated buffers (large procedural scenes)."""
        self = cls.__new__(cls)
        self.meshes = None
        self.lights = list(lights or [])
        n = len(ids)
        self._positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self._indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        self._uv1 = np.ascontiguousarray(uv1, dtype=np.float32).reshape(-1, 2)
        self._normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        vo = np.asarray(vertex_offsets, dtype=np.int64)
        io = np.asarray(index_offsets, dtype=np.int64)
        arr = np.zeros(n, dtype=MESH_DTYPE)
        arr["id"] = np.asarray(ids, dtype=np.uint64)
        arr["content_hash"] = 0 if content_hash is None else np.asarray(content_hash, dtype=np.uint64)
        vc = (vo[1:] - vo[:-1]).astype(np.uint32)
        ic = (io[1:] - io[:-1]).astype(np.uint32)
        arr["vertex_count"] = vc
        arr["index_count"] = ic
        arr["uv1_count"] = vc
        pbase = self._positions.ctypes.data
        arr["positions"] = pbase + vo[:-1] * 12
        arr["indices"] = self._indices.ctypes.data + io[:-1] * 4
        arr["uv1"] = self._uv1.ctypes.data + vo[:-1] * 8
        if self._normals is not None:
            arr["normals"] = self._normals.ctypes.data + vo[:-1] * 12
            arr["normal_count"] = vc
        arr["occludes"] = 1 if occludes is None else np.asarray(occludes, dtype=np.uint32)
        arr["albedo"] = 1.0 if albedo is None else np.asarray(albedo, dtype=np.float32).reshape(n, 3)
        self._mesh_arr = arr
        self._alphas = []
        self._keep = []
        self._pack_lights()
        self._make_struct()
        self.vertex_offsets = vo
        self.index_offsets = io
        return self

    def _pack(self):
        ms = self.meshes
        n = len(ms)
        arr = np.zeros(n, dtype=MESH_DTYPE)
        self._keep = []
        self._alphas = []
        for i, m in enumerate(ms):
            pos = np.ascontiguousarray(m.positions, dtype=np.float32).reshape(-1, 3)
            idx = np.ascontiguousarray(m.indices, dtype=np.uint32).reshape(-1)
            uv1 = np.ascontiguousarray(m.uv1, dtype=np.float32).reshape(-1, 2)
            self._keep += [pos, idx, uv1]
            arr[i]["id"] = m.id
            arr[i]["content_hash"] = m.content_hash
            arr[i]["vertex_count"] = pos.shape[0]
            arr[i]["positions"] = pos.ctypes.data
            arr[i]["index_count"] = idx.shape[0]
            arr[i]["indices"] = idx.ctypes.data if idx.size else 0
            arr[i]["uv1"] = uv1.ctypes.data if uv1.size else 0
            arr[i]["uv1_count"] = uv1.shape[0]
            if m.normals is not None and len(m.normals):
                nrm = np.ascontiguousarray(m.normals, dtype=np.float32).reshape(-1, 3)
                self._keep.append(nrm)
                arr[i]["normals"] = nrm.ctypes.data
                arr[i]["normal_count"] = nrm.shape[0]
            if m.uv0 is not None:
                uv0 = np.ascontiguousarray(m.uv0, dtype=np.float32).reshape(-1, 2)
                self._keep.append(uv0)
                arr[i]["uv0"] = uv0.ctypes.data if uv0.size else 1
                arr[i]["uv0_count"] = uv0.shape[0]
            if m.alpha is not None:
                w, h, rgba, cutoff = m.alpha
                rgba = np.ascontiguousarray(rgba, dtype=np.uint8).reshape(-1)
                a = dl_alpha_texture(w, h, rgba.ctypes.data, rgba.size, cutoff)
                self._keep.append(rgba)
                self._alphas.append(a)
                arr[i]["alpha"] = C.addressof(a)
            arr[i]["occludes"] = 1 if m.occludes else 0
            arr[i]["albedo"] = m.albedo
        self._mesh_arr = arr
        self._pack_lights()
        self._make_struct()

    def _pack_lights(self):
        la = np.zeros(len(self.lights), dtype=LIGHT_DTYPE)
        for i, l in enumerate(self.lights):
            la[i]["id"] = l.id
            la[i]["position"] = l.position
            la[i]["radius"] = l.radius
            la[i]["illumination"] = l.illumination
            la[i]["transparency"] = l.transparency
            la[i]["bounce_samples"] = l.bounce_samples
            la[i]["bounce_intensity"] = l.bounce_intensity
            la[i]["bounce_compression"] = l.bounce_compression
            la[i]["shape"] = l.shape
            la[i]["rotation"] = l.rotation
            la[i]["size"] = l.size
            la[i]["angular_diameter"] = l.angular_diameter
        self._light_arr = la

    def _make_struct(self):
        self.struct = dl_bake_input(
            self._mesh_arr.ctypes.data if len(self._mesh_arr) else None, len(self._mesh_arr),
            self._light_arr.ctypes.data if len(self._light_arr) else None, len(self._light_arr))

    @property
    def mesh_count(self) -> int:
        return len(self._mesh_arr)

    @property
    def mesh_array(self) -> np.ndarray:
        return self._mesh_arr

    @property
    def triangle_count(self) -> int:
        return int(self._mesh_arr["index_count"].sum() // 3)

    def with_lights(self, lights: list[LightData]) -> "BakeInput":
        """The same meshes (buffers are shared) with other lights."""
        other = BakeInput.__new__(BakeInput)
        other.meshes = self.meshes
        other.lights = list(lights)
        other._keep = [self]
        other._alphas = []
        other._mesh_arr = self._mesh_arr
        other._pack_lights()
        other._make_struct()
        return other

    def joined(self, other: "BakeInput") -> "BakeInput":
        """This input's meshes and lights followed by `other`'s (buffers are shared)."""
        j = BakeInput.__new__(BakeInput)
        j.meshes = None
        j.lights = list(self.lights) + list(other.lights)
        j._keep = [self, other]
        j._alphas = []
        j._mesh_arr = np.zeros(len(self._mesh_arr) + len(other._mesh_arr), dtype=self._mesh_arr.dtype)  # the C layout
        j._mesh_arr[:len(self._mesh_arr)] = self._mesh_arr
        j._mesh_arr[len(self._mesh_arr):] = other._mesh_arr
        j._pack_lights()
        j._make_struct()
        return j

    def subset(self, keep: np.ndarray) -> "BakeInput":
        """A view with only the meshes where `keep` is true (buffers are shared)."""
        sub = BakeInput.__new__(BakeInput)
        sub.meshes = None
        sub.lights = self.lights
        sub._keep = [self]
        sub._alphas = []
        sub._mesh_arr = np.ascontiguousarray(self._mesh_arr[keep])
        sub._light_arr = self._light_arr
        sub._make_struct()
        return sub


# ---- baked scene read-back -------------------------------------------------------------------
@dataclass
class BakedMesh:
    id: int
    content_hash: int
    resolution: int
    occludes: bool
    triangles: np.ndarray


@dataclass
class BakedScene:
    """BakedScene (src/bake/format/types.rs:13-37) copied out of a library-owned handle."""
    settings: dict
    stats: dict
    lights: list
    light_bvh: np.ndarray
    meshes: list = field(default_factory=list)
    distance_cubes: np.ndarray | None = None
    sun_maps: list = field(default_factory=list)
    sky_probes: dict | None = None
    triangles_taken: bool = False


def _struct_dict(s) -> dict:
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if isinstance(v, C.Array):
            v = list(v)
        out[name] = v
    return out


def bind_scene_api(lib: C.CDLL, prefix: str):
    """Declares the BakedScene accessors `<prefix>scene_*` on `lib`."""
    H = C.c_void_p
    sig = {
        "scene_settings": (None, [H, C.POINTER(dl_baked_settings)]),
        "scene_stats": (None, [H, C.POINTER(dl_bake_stats)]),
        "scene_light_count": (C.c_uint32, [H]),
        "scene_light": (C.c_int, [H, C.c_uint32, C.POINTER(dl_baked_light)]),
        "scene_light_bvh": (C.c_uint32, [H, C.POINTER(C.c_void_p)]),
        "scene_mesh_count": (C.c_uint32, [H]),
        "scene_mesh": (C.c_int, [H, C.c_uint32, C.POINTER(dl_baked_mesh_view)]),
        "scene_meshes": (C.c_uint32, [H, C.c_uint32, C.c_uint32, C.POINTER(dl_baked_mesh_view)]),
        "scene_distance_cubes": (C.c_uint64, [H, C.POINTER(C.c_void_p)]),
        "scene_sun_map_count": (C.c_uint32, [H]),
        "scene_sun_map": (C.c_int, [H, C.c_uint32, C.POINTER(dl_sun_map_view)]),
        "scene_sky_probes": (None, [H, C.POINTER(dl_sky_probes_view)]),
        "scene_free": (None, [H]),
    }
    for name, (res, args) in sig.items():
        if not hasattr(lib, prefix + name):
            continue
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args


def _words(ptr, count) -> np.ndarray:
    if count == 0 or not ptr:
        return np.zeros(0, dtype=np.uint32)
    buf = (C.c_uint32 * count).from_address(ptr)
    return np.frombuffer(buf, dtype=np.uint32).copy()


def read_scene(lib: C.CDLL, prefix: str, handle) -> BakedScene:
    """Copies everything out of a scene handle into Python objects."""
    g = lambda n: getattr(lib, prefix + n)  # noqa: E731
    st = dl_baked_settings()
    g("scene_settings")(handle, C.byref(st))
    stats = dl_bake_stats()
    g("scene_stats")(handle, C.byref(stats))
    lights = []
    for i in range(g("scene_light_count")(handle)):
        bl = dl_baked_light()
        g("scene_light")(handle, i, C.byref(bl))
        lights.append(_struct_dict(bl))
    p = C.c_void_p()
    n = g("scene_light_bvh")(handle, C.byref(p))
    bvh = np.zeros((n, 8), dtype=np.uint32)
    if n:
        bvh = np.frombuffer((C.c_uint32 * (8 * n)).from_address(p.value), dtype=np.uint32).reshape(n, 8).copy()
    meshes = []
    nm = g("scene_mesh_count")(handle)
    if hasattr(lib, prefix + "scene_meshes") and nm:
        # every view in one call, read as a numpy record array; words copied once per run of
        # meshes whose words lie end to end
        views = (dl_baked_mesh_view * nm)()
        got = g("scene_meshes")(handle, 0, nm, views)
        assert got == nm
        rec = np.frombuffer(views, dtype=np.dtype({"names": ["id", "hash", "res", "occ", "ptr", "n"],
                                                   "formats": ["<u8", "<u8", "<u4", "<u4", "<u8", "<u8"],
                                                   "offsets": [0, 8, 16, 20, 24, 32], "itemsize": C.sizeof(dl_baked_mesh_view)}))
        ptr, cnt = rec["ptr"].astype(np.int64), rec["n"].astype(np.int64)
        brk = np.ones(nm, bool)
        brk[1:] = (ptr[1:] != ptr[:-1] + 4 * cnt[:-1]) | (ptr[1:] == 0) | (ptr[:-1] == 0)
        starts = np.flatnonzero(brk).tolist() + [nm]
        ids, hashes, res, occ = rec["id"].tolist(), rec["hash"].tolist(), rec["res"].tolist(), rec["occ"].tolist()
        off = np.concatenate([[0], np.cumsum(cnt)]).tolist()
        for a, b in zip(starts[:-1], starts[1:]):
            block = _words(int(ptr[a]), off[b] - off[a])
            for k in range(a, b):
                meshes.append(BakedMesh(ids[k], hashes[k], res[k], bool(occ[k]), block[off[k] - off[a]:off[k + 1] - off[a]]))
    else:
        for i in range(nm):
            mv = dl_baked_mesh_view()
            g("scene_mesh")(handle, i, C.byref(mv))
            meshes.append(BakedMesh(mv.id, mv.content_hash, mv.resolution, bool(mv.occludes),
                                    _words(mv.triangles, mv.triangle_word_count)))
    p = C.c_void_p()
    n = g("scene_distance_cubes")(handle, C.byref(p))
    cubes = _words(p.value, n)
    suns = []
    for i in range(g("scene_sun_map_count")(handle)):
        sv = dl_sun_map_view()
        g("scene_sun_map")(handle, i, C.byref(sv))
        suns.append({"rows": list(sv.rows), "resolution": sv.resolution,
                     "texels": _words(sv.texels, sv.texel_word_count)})
    sp = dl_sky_probes_view()
    g("scene_sky_probes")(handle, C.byref(sp))
    probes = None
    if sp.present:
        probes = {"origin": list(sp.origin), "spacing": sp.spacing, "counts": list(sp.counts),
                  "words": _words(sp.words, sp.word_count)}
    return BakedScene(_struct_dict(st), _struct_dict(stats), lights, bvh, meshes, cubes, suns, probes)

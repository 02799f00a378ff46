"""dl_service_ray_queries_noalloc: ray queries against a service's resident BVH that allocate no
device memory (a fixed per-GPU scratch holds the queue head, hit count and overflow flag).  Its
results must equal dl_test_ray_queries over the same triangles, on two calls in a row, the second
smaller (the scratch keeps the first call's contents and must be cleared)."""
import ctypes as C

import numpy as np
import pytest

from dynlight_hip import BakeError, Service, abi, scenes

pytestmark = pytest.mark.gpu

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)


class DeviceBuffer:
    """Device memory through the HIP runtime the library itself uses (test infrastructure)."""
    hip = None

    def __init__(self, nbytes):
        if DeviceBuffer.hip is None:
            DeviceBuffer.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
            DeviceBuffer.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            DeviceBuffer.hip.hipFree.argtypes = [C.c_void_p]
            DeviceBuffer.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            DeviceBuffer.hip.hipDeviceSynchronize.argtypes = []
        self.ptr = C.c_void_p()
        assert DeviceBuffer.hip.hipMalloc(C.byref(self.ptr), max(4, nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        assert a.nbytes <= self.nbytes
        assert DeviceBuffer.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def download(self, a):
        assert DeviceBuffer.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        if self.ptr:
            DeviceBuffer.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _scene():
    inp, o = scenes.box_room(1)
    o.dynamic_geometry = abi.DYNAMIC_LIGHTING_ONLY
    order = sorted(range(len(inp.meshes)), key=lambda i: inp.meshes[i].id)
    tris = []
    for i in order:  # occluder order: meshes by id, triangles in order (none degenerate here)
        m = inp.meshes[i]
        tris.append(np.asarray(m.positions, np.float32)[np.asarray(m.indices).reshape(-1, 3)].reshape(-1, 9))
    return inp, o, np.ascontiguousarray(np.concatenate(tris).astype(np.float32))


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform([-4.5, 0.1, -4.5], [4.5, 3.9, 4.5], (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 6] = rng.uniform(0.1, 12.0, n)
    r[: n // 8, 3:6] = np.array([1.0, -0.0, 0.0], np.float32)  # axis-parallel with a -0 component
    return np.ascontiguousarray(r)


def _reference(ctx, tris, rays, closest):
    n = len(rays)
    t = np.zeros(n, np.float32)
    idx = np.zeros(n, np.uint32)
    st = ctx.lib.dl_test_ray_queries(ctx.handle, tris.ctypes.data_as(F32P), len(tris), rays.ctypes.data_as(F32P), n, closest,
                                     t.ctypes.data_as(F32P), idx.ctypes.data_as(U32P))
    assert st == 0, ctx.error()
    return t, idx


def _noalloc(svc, rays, closest):
    n = len(rays)
    d_rays, d_t, d_idx = DeviceBuffer(rays.nbytes), DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    try:
        d_rays.upload(rays)
        hits = C.c_uint64(12345)
        st = svc.lib.dl_service_ray_queries_noalloc(svc.handle, d_rays.ptr, n, closest, d_t.ptr, d_idx.ptr, C.byref(hits))
        if st != 0:
            raise BakeError(st, svc.ctx.error())
        return d_t.download(np.zeros(n, np.float32)), d_idx.download(np.zeros(n, np.uint32)), hits.value
    finally:
        for b in (d_rays, d_t, d_idx):
            b.free()


@pytest.mark.parametrize("closest", [0, 1])
def test_noalloc_queries_equal_the_allocating_entry_point(ctx, closest):
    inp, o, tris = _scene()
    svc = Service(ctx, inp, o)
    try:
        for n, seed in ((200_000, 1), (5_000, 2)):  # the second call is smaller
            print(f"noalloc closest={closest} rays={n}", flush=True)
            rays = _rays(n, seed)
            want_t, want_idx = _reference(ctx, tris, rays, closest)
            got_t, got_idx, hits = _noalloc(svc, rays, closest)
            assert np.array_equal(got_idx, want_idx)
            if closest:
                assert np.array_equal(got_t.view(np.uint32), want_t.view(np.uint32))
                assert hits == int((want_idx != 0xFFFFFFFF).sum())
            else:
                assert hits == int(want_idx.sum())
            assert 0 < hits < n
    finally:
        svc.close()


def test_noalloc_queries_reject_bad_arguments(ctx):
    inp, o, _ = _scene()
    svc = Service(ctx, inp, o)
    try:
        n = abi.DL_NOALLOC_MAX_RAYS + 1
        assert svc.lib.dl_service_ray_queries_noalloc(svc.handle, None, n, 0, None, None, None) == abi.DL_E_ARGUMENT
        assert ctx.error() == f"at most {abi.DL_NOALLOC_MAX_RAYS} rays per call"
        assert svc.lib.dl_service_ray_queries_noalloc(svc.handle, None, 10, 0, None, None, None) == abi.DL_E_ARGUMENT
        # zero rays: nothing to do, no hits
        hits = C.c_uint64(7)
        assert svc.lib.dl_service_ray_queries_noalloc(svc.handle, None, 0, 1, None, None, C.byref(hits)) == 0
        assert hits.value == 0
    finally:
        svc.close()

"""The light BVH the bake builds on the host (dl_test_light_bvh, no device needed) against the
oracle's restatement of LightBvh::build (light_bvh.rs:59-189): nodes and slot order bit for bit on
light sets of many sizes, with coordinates that tie along an axis, coincide or spread."""
import ctypes as C

import numpy as np
import pytest

import dynlight_hip
import oracle
from dynlight_hip import abi

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)


@pytest.fixture(scope="module")
def lib():
    try:
        return dynlight_hip.load()
    except OSError as e:
        pytest.skip(str(e))


@pytest.fixture(scope="module")
def ol():
    return oracle.load()


def build(fn, pos, rad):
    n = len(rad)
    pos = np.ascontiguousarray(pos, np.float32)
    rad = np.ascontiguousarray(rad, np.float32)
    nodes = (abi.dl_light_bvh_node * (2 * n + 2))()
    order = np.zeros(max(n, 1), np.uint32)
    cnt = fn(pos.ctypes.data_as(F32P), rad.ctypes.data_as(F32P), n, nodes, 2 * n + 2, order.ctypes.data_as(U32P))
    raw = bytes(memoryview(nodes).cast("B")[: cnt * C.sizeof(abi.dl_light_bvh_node)])
    return cnt, raw, order[:n].copy()


def light_sets():
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 5, 17, 64, 257, 1000, 2500):
        yield f"spread-{n}", rng.uniform(-50, 50, (n, 3)), rng.uniform(0.5, 12.0, n)
        grid = rng.integers(0, 8, (n, 3)).astype(np.float64)
        grid[:, 1] = 1.5
        yield f"ties-{n}", grid, np.full(n, 4.0)
        yield f"coincident-{n}", np.tile(rng.uniform(-5, 5, (1, 3)), (n, 1)), rng.uniform(0.5, 2.0, n)
        room = np.stack([rng.integers(0, 600, n) * 0.1, rng.integers(0, 30, n) * 0.1, rng.integers(0, 600, n) * 0.1], 1)
        yield f"rooms-{n}", room, rng.uniform(4.0, 12.0, n)


@pytest.mark.parametrize("name,pos,rad", list(light_sets()), ids=lambda v: v if isinstance(v, str) else "")
def test_light_bvh_equals_the_oracle(lib, ol, name, pos, rad):
    got = build(lib.dl_test_light_bvh, pos, rad)
    want = build(ol.dlo_light_bvh, pos, rad)
    assert got[0] == want[0] > 0
    assert got[1] == want[1], "nodes differ"
    assert np.array_equal(got[2], want[2]), "slot order differs"
    assert sorted(got[2]) == list(range(len(rad)))


def test_light_bvh_of_no_lights(lib):
    assert build(lib.dl_test_light_bvh, np.zeros((0, 3)), np.zeros(0))[0] == 0

"""Structure of the occluder BVH built on the host (no device needed).

The device traversal is checked against brute force in tests/test_gpu_parity.py; here the host
builder's output is checked on its own: every node reachable once, the leaves' reference ranges
partitioning the references, every leaf record holding its triangle, and the same tree for every
thread count.  The scenes reach the builder's paths: references cut from large triangles (and
the over-budget fallback that cuts the largest first), ranges large enough to be split by several
threads, coincident centres and flat scenes.
"""
import ctypes as C
import os

import numpy as np
import pytest

import dynlight_hip


@pytest.fixture(scope="module")
def lib():
    try:
        return dynlight_hip.load()
    except OSError as e:
        pytest.skip(str(e))


def build(lib, tris, threads):
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    out = (C.c_uint64 * 4)()
    st = lib.dl_test_bvh_structure(tris.ctypes.data_as(C.POINTER(C.c_float)), tris.shape[0], threads, out)
    assert st == 0, lib.dl_last_error(None).decode()
    return tuple(out)


def small_tris(rng, n, extent=100.0, size=0.5):
    c = rng.uniform(0.0, extent, (n, 1, 3))
    return (c + rng.uniform(-size, size, (n, 3, 3))).reshape(n, 9)


def large_tris(rng, n, extent=100.0):
    return rng.uniform(0.0, extent, (n, 9))


def test_tiny_inputs(lib):
    rng = np.random.default_rng(1)
    assert build(lib, np.zeros((0, 9)), 4)[:2] == (0, 0)
    for n in (1, 3, 4, 5, 17):
        nodes, refs, depth, _ = build(lib, small_tris(rng, n), 4)
        assert nodes >= 1 and refs == n and depth >= 1


@pytest.mark.parametrize("n", [3000, 150_000])
def test_same_tree_for_every_thread_count(lib, n):
    rng = np.random.default_rng(n)
    tris = np.concatenate([small_tris(rng, n), large_tris(rng, n // 100)])
    got = {t: build(lib, tris, t) for t in (1, 3, 8)}
    assert len(set(got.values())) == 1, got
    nodes, refs, depth, _ = got[1]
    assert refs > tris.shape[0]  # the large triangles were cut into several references
    assert depth >= 4


def test_over_budget_cuts_the_largest_first(lib, monkeypatch):
    rng = np.random.default_rng(7)
    tris = np.concatenate([small_tris(rng, 5000), large_tris(rng, 500)])
    full = build(lib, tris, 4)
    monkeypatch.setenv("DL_BVH_SPLIT_BUDGET", "0.01")  # at most 55 extra references
    capped = build(lib, tris, 4)
    assert capped[1] == tris.shape[0] + 55 < full[1]
    assert build(lib, tris, 1) == capped


def test_coincident_and_flat(lib):
    rng = np.random.default_rng(3)
    same = np.tile(small_tris(rng, 1), (2000, 1))
    nodes, refs, _, _ = build(lib, same, 4)
    assert refs == 2000 and nodes >= 2000 // 16
    flat = small_tris(rng, 20_000)
    flat[:, 1::3] = 0.0  # every vertex on the plane y = 0
    flat = flat[np.abs(np.cross(flat[:, 3:6] - flat[:, 0:3], flat[:, 6:9] - flat[:, 0:3])).sum(axis=1) > 0]
    assert build(lib, flat, 8) == build(lib, flat, 1)

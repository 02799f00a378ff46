"""The public bake API through the C ABI: cancellation (src/bake/tests/determinism.rs:150-157) and
the baked words read back the way the runtime shader reads them (tests/bake_api.rs:86-130)."""
import ctypes as C
import threading
import time

import numpy as np
import pytest

import oracle
from dynlight_hip import BakeError, abi, scenes
from dynlight_hip.abi import BakeInput, LightData, MeshData
from dynlight_hip.mirror import DISTANCE_CUBE_WORDS, TriangleView, distance_cube_value, triangle_count
from helpers import assert_scenes_equal

pytestmark = pytest.mark.gpu


def test_cancelled_bakes_stop_and_the_context_bakes_again(ctx):
    inp, o = scenes.box_room(1)
    progress = abi.dl_progress()
    progress.cancel = 1
    with pytest.raises(BakeError) as e:
        ctx.bake(inp, o, progress)
    assert e.value.code == abi.DL_E_CANCELLED
    assert str(e.value) == "the bake was cancelled"
    # the same context bakes afterwards, bit-exact against the oracle
    done = abi.dl_progress()
    assert_scenes_equal(ctx.bake(inp, o, done), oracle.bake(inp, o))
    assert done.stage == 4, "BakeStage::Done"


def test_a_cancel_raised_mid_bake_stops_it(ctx):
    inp, o = scenes.interior(2, target_triangles=20_000, lights=48)
    progress = abi.dl_progress()
    result = {}

    def run():
        try:
            result["scene"] = ctx.bake(inp, o, progress)
        except BakeError as err:
            result["error"] = err

    t = threading.Thread(target=run)
    t.start()
    deadline = time.time() + 60
    while progress.stage < 2 and t.is_alive() and time.time() < deadline:  # into the mesh passes
        time.sleep(0.0005)
    progress.cancel = 1
    t.join(120)
    assert not t.is_alive()
    stage = progress.stage
    print(f"cancel raised at stage {stage}: {'cancelled' if 'error' in result else 'finished first'}", flush=True)
    if "error" in result:
        assert result["error"].code == abi.DL_E_CANCELLED
    # whatever was in flight, the context's next bakes are whole and equal
    a = ctx.bake(inp, o)
    b = ctx.bake(inp, o)
    assert_scenes_equal(a, b)
    small, so = scenes.box_room(1)
    assert_scenes_equal(ctx.bake(small, so), oracle.bake(small, so))


def _floor():
    pos = np.array([[-2, 0, -2], [2, 0, -2], [2, 0, 2], [-2, 0, 2]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return MeshData(7, pos, np.array([0, 2, 1, 0, 3, 2], np.uint32), uv,
                    normals=np.tile(np.array([0, 1, 0], np.float32), (4, 1)))


def test_baked_scene_reads_like_the_shader(ctx):
    inp = BakeInput([_floor()], [LightData(2, (1.0, 1.0, 0.0), 3.0), LightData(1, (-1.0, 1.0, 0.0), 3.0)])
    o = abi.default_options()
    o.pixel_density = 4.0
    o.dynamic_geometry = abi.DYNAMIC_ANGULAR
    scene = ctx.bake(inp, o)
    assert len(scene.light_bvh) == 1, "two lights fit one leaf"
    assert scene.light_bvh[0][7] == 2, "count"
    mesh = scene.meshes[0]
    assert mesh.resolution == 16 and triangle_count(mesh.triangles) == 2
    for t in range(2):
        view = TriangleView(mesh.triangles, t, False)
        assert view.light_count == 2
        for k in range(2):
            e = view.light(k)
            assert e.light_index == k, "entries in BVH slot order"
            assert view.shadow_bit(e, 8, 8) is True, "nothing blocks the floor"
    assert len(scene.distance_cubes) == 2 * DISTANCE_CUBE_WORDS
    ol = oracle.load()
    out = np.zeros(3, np.uint32)
    for slot in range(2):
        for d, want in (((0, -1, 0), 1.0), ((0, 1, 0), 3.0)):
            ol.dlo_cube_texel(np.array(d, np.float32).ctypes.data_as(C.POINTER(C.c_float)), 64,
                              out.ctypes.data_as(C.POINTER(C.c_uint32)))
            v = distance_cube_value(scene.distance_cubes, slot, ol.dlo_cube_texel_index(int(out[0]), int(out[1]),
                                                                                      int(out[2]), 64))
            if want == 3.0:
                assert v == 3.0, "misses store the radius"
            else:
                assert abs(v - 1.0) < 0.01, v

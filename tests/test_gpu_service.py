"""The rebake service (dl_service_*): a resident scene baked with other lights equals dl_bake of
the same meshes with those lights, bit-exact, and every request's result is the same served alone
and in a mix of overlapping requests."""
import math

import numpy as np
import pytest

import dynlight_hip
import oracle
from dynlight_hip import BakeError, abi, scenes
from helpers import assert_scenes_equal

pytestmark = pytest.mark.gpu


def _scene():
    return scenes.interior(seed=9, target_triangles=3000, lights=12, grid=2, density=12.0)


def _requests(inp, count, seed=5):
    """Light requests on the 2x2-room scene: moves of 1-4 existing lights (same ids, displaced
    0.5-3 m), additions of 1-4 new point, area or sun lights."""
    rng = scenes.Pcg32.seeded(seed)
    out = []
    for r in range(count):
        k = 1 + rng.below(4)
        lights = []
        if rng.uniform() < 0.6:
            for idx in sorted({rng.below(len(inp.lights)) for _ in range(k)}):
                base = inp.lights[idx]
                d = rng.uniform(0.5, 3.0)
                a, b = rng.uniform(0.0, 2 * math.pi), rng.uniform(-0.3, 0.3)
                p = (base.position[0] + d * math.cos(a), min(2.9, max(0.1, base.position[1] + b)), base.position[2] + d * math.sin(a))
                lights.append(scenes.point_light(base.id, p, base.radius, bounce=True))
        else:
            for j in range(k):
                lid = 50_000 + 10 * r + j
                p = (rng.uniform(0.1, 11.9), rng.uniform(0.1, 2.9), rng.uniform(0.1, 11.9))
                kind = rng.uniform()
                if kind < 0.6:
                    lights.append(scenes.point_light(lid, p, rng.uniform(4.0, 12.0), bounce=True))
                elif kind < 0.9:
                    lights.append(scenes.area_light(lid, p, rng.uniform(4.0, 12.0), scenes.quat_from_rotation_x(-math.pi / 2),
                                                    (rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0))))
                else:
                    lights.append(scenes.sun_light(lid, (0.3, 0.9, 0.2), 0.53))
        out.append(lights)
    return out


@pytest.fixture(scope="module")
def service(ctx):
    inp, opts = _scene()
    svc = dynlight_hip.Service(ctx, inp, opts)
    yield inp, opts, svc
    svc.close()


def test_first_bake_equals_a_plain_bake(ctx, service):
    inp, opts, svc = service
    assert_scenes_equal(svc.initial, ctx.bake(inp, opts))


def test_rebakes_equal_plain_bakes_of_the_same_lights(ctx, service):
    inp, opts, svc = service
    for i, lights in enumerate(_requests(inp, 6)):
        print(f"request {i}: {len(lights)} lights", flush=True)
        got = svc.bake(lights)
        assert_scenes_equal(got, ctx.bake(inp.with_lights(lights), opts))
        assert got.stats["samples"] > 0
    # and one against the oracle
    lights = _requests(inp, 1, seed=11)[0]
    assert_scenes_equal(svc.bake(lights), oracle.bake(inp.with_lights(lights), opts))


def test_requests_in_a_mix_equal_requests_alone(service):
    inp, opts, svc = service
    reqs = _requests(inp, 12, seed=7)
    alone = [svc.bake(l) for l in reqs]
    # arrivals 5 ms apart: requests overlap on three worker streams
    lat, mixed = svc.replay([(l, 0.005 * i) for i, l in enumerate(reqs)], workers=3, read=True)
    assert len(lat) == len(reqs) and all(x > 0 for x in lat)
    for a, m in zip(alone, mixed):
        assert_scenes_equal(m, a)


def test_service_errors(ctx, service):
    inp, opts, svc = service
    with pytest.raises(BakeError) as e:
        svc.bake([scenes.point_light(3, (1, 1, 1), -1.0)])
    assert e.value.code == abi.DL_E_INVALID_LIGHT
    with pytest.raises(BakeError) as e:
        svc.replay([([], 0.0)], workers=0)
    assert e.value.code == abi.DL_E_ARGUMENT
    # still serving afterwards
    lights = [scenes.point_light(7, (3.0, 1.5, 3.0), 6.0, bounce=True)]
    assert_scenes_equal(svc.bake(lights), ctx.bake(inp.with_lights(lights), opts))

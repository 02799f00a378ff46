"""Batched bakes (dl_bake_batch): every job's result equals the same job baked alone, and the
cheapest jobs equal the oracle; failing jobs and argument errors are reported."""
import numpy as np
import pytest

import oracle
from dynlight_hip import BakeError, abi, scenes
from helpers import assert_scenes_equal

pytestmark = pytest.mark.gpu


def _cheap_jobs(count, pool=96):
    """The `count` cheapest jobs of the first `pool` batch jobs, by triangles x lights x density^2,
    half with bounce (odd j) and half without."""
    jobs = [scenes.batch_job(j) for j in range(pool)]
    cost = [i.triangle_count * len(i.lights) * o.pixel_density ** 2 for i, o in jobs]
    order = np.argsort(cost)
    odd = [int(j) for j in order if j % 2][: count // 2]
    even = [int(j) for j in order if j % 2 == 0][: count - len(odd)]
    pick = sorted(odd + even)
    return pick, [jobs[j][0] for j in pick], [jobs[j][1] for j in pick]


def test_batched_jobs_equal_single_bakes(ctx):
    pick, ins, ops = _cheap_jobs(12)
    print(f"batch jobs {pick}", flush=True)
    got = ctx.bake_batch(ins, ops, streams=3)
    assert len(got) == len(ins)
    for j, inp, o, g in zip(pick, ins, ops, got):
        alone = ctx.bake(inp, o)
        assert_scenes_equal(g, alone)
        assert g.stats["samples"] > 0, j
    # the same batch again on more streams than jobs left per stream
    again = ctx.bake_batch(ins[:5], ops[:5], streams=8)
    for a, g in zip(again, got[:5]):
        assert_scenes_equal(a, g)


def test_cheapest_batched_jobs_equal_the_oracle(ctx):
    pick, ins, ops = _cheap_jobs(4)
    got = ctx.bake_batch(ins, ops, streams=2)
    for j, inp, o, g in zip(pick, ins, ops, got):
        print(f"oracle job {j}", flush=True)
        assert_scenes_equal(g, oracle.bake(inp, o, threads=8))


def test_batch_errors(ctx):
    assert ctx.bake_batch([], [], streams=2) == []
    good, o = scenes.box_room(1)
    a, b = scenes.floor(5, (0, 0, 0), 1.0), scenes.floor(5, (0, 1, 0), 1.0)
    bad = abi.BakeInput([a, b], [])
    with pytest.raises(BakeError) as e:
        ctx.bake_batch([good, bad, good], [o, abi.default_options(), o], streams=2)
    assert e.value.code == abi.DL_E_DUPLICATE_ID
    assert str(e.value) == "job 1: two meshes share BakeId 0x0000000000000005"
    with pytest.raises(BakeError) as e:
        ctx.bake_batch([good], [o], streams=0)
    assert e.value.code == abi.DL_E_ARGUMENT
    # the context still bakes afterwards
    assert_scenes_equal(ctx.bake_batch([good], [o], streams=1)[0], ctx.bake(good, o))

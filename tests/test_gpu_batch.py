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


def _spanning_jobs(count=32, pool=128):
    """`count` jobs spread over the cost order of the first `pool` batch jobs (every pool/count-th
    by triangles x lights x density^2): small and large, with and without bounce."""
    jobs = [scenes.batch_job(j) for j in range(pool)]
    cost = [i.triangle_count * len(i.lights) * o.pixel_density ** 2 for i, o in jobs]
    pick = sorted(int(j) for j in np.argsort(cost)[::pool // count])
    return pick, [jobs[j][0] for j in pick], [jobs[j][1] for j in pick]


def test_batched_jobs_equal_single_bakes(ctx):
    pick, ins, ops = _spanning_jobs(32)
    print(f"batch jobs {pick}", flush=True)
    assert sum(j % 2 for j in pick) >= 8 and sum(1 - j % 2 for j in pick) >= 8, "with and without bounce"
    got = ctx.bake_batch(ins, ops, streams=4)
    assert len(got) == len(ins)
    for j, inp, o, g in zip(pick, ins, ops, got):
        alone = ctx.bake(inp, o)
        assert_scenes_equal(g, alone)
        assert g.stats["samples"] > 0, j
    # the same batch again on more streams than jobs left per stream
    again = ctx.bake_batch(ins[:5], ops[:5], streams=8)
    for a, g in zip(again, got[:5]):
        assert_scenes_equal(a, g)


def test_lazy_photon_cubes_keep_batched_results(ctx):
    """The bounce jobs of the batch workload with lazy photon cubes (the default: jobs that set up
    few gather samples per cube texel trace only the texels their samples can read) equal the same
    jobs with every texel traced."""
    pick, ins, ops = _spanning_jobs(32)
    sel = [i for i, j in enumerate(pick) if j % 2]
    pick, ins, ops = [pick[i] for i in sel], [ins[i] for i in sel], [ops[i] for i in sel]
    print(f"bounce jobs {pick}", flush=True)
    ctx.set_lazy_photons(False)
    try:
        full = ctx.bake_batch(ins, ops, streams=4)
        full_t = ctx.batch_trace()
    finally:
        ctx.set_lazy_photons(True)
    lazy = ctx.bake_batch(ins, ops, streams=4)
    lazy_t = ctx.batch_trace()
    for j, a, b in zip(pick, lazy, full):
        assert_scenes_equal(a, b)
    traced = sum(t["photon_texels"] for t in lazy_t)
    total = sum(t["photon_texels"] for t in full_t)
    print(f"  traced {traced} of {total} cube texels", flush=True)
    assert 0 < traced < total
    assert any(a["photon_texels"] < b["photon_texels"] for a, b in zip(lazy_t, full_t))


def test_cheapest_batched_jobs_equal_the_oracle(ctx):
    pick, ins, ops = _cheap_jobs(6)
    got = ctx.bake_batch(ins, ops, streams=2)
    for j, inp, o, g in zip(pick, ins, ops, got):
        print(f"oracle job {j}", flush=True)
        assert_scenes_equal(g, oracle.bake(inp, o, threads=8))


def test_spanning_batched_jobs_equal_the_oracle_on_a_mesh_sample(ctx):
    """Jobs across the size distribution against the oracle on an eighth of their meshes."""
    pick, ins, ops = _spanning_jobs(32)
    pick, ins, ops = pick[::4], ins[::4], ops[::4]
    got = ctx.bake_batch(ins, ops, streams=4)
    for j, inp, o, g in zip(pick, ins, ops, got):
        mask = scenes.mesh_sample_mask(inp.mesh_array["id"], 8)
        mask[0] = True  # the first room shell
        print(f"oracle job {j}: {int(mask.sum())} of {len(mask)} meshes", flush=True)
        want = oracle.bake_receivers(inp, o, mask, threads=8)
        by_id = {m.id: m for m in g.meshes}
        assert g.lights == want.lights and np.array_equal(g.light_bvh, want.light_bvh)
        for w in want.meshes:
            assert np.array_equal(by_id[w.id].triangles, w.triangles), f"job {j} mesh {w.id:#x}"
        assert np.array_equal(g.distance_cubes, want.distance_cubes)


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


def test_autotuned_launches_keep_results_and_are_cached(ctx, tmp_path):
    """dl_autotune_batch picks launch shapes on a sample of the jobs; a batch baked with them equals
    single bakes, and the second call reads the same settings from the cache file."""
    pick, ins, ops = _spanning_jobs(16, pool=64)
    print(f"autotune on jobs {pick}", flush=True)
    cache = str(tmp_path / "tune.txt")
    try:
        t = ctx.autotune_batch(ins, ops, max_streams=8, cache_path=cache)
        print(f"  tuned {t}", flush=True)
        assert t["from_cache"] == 0 and t["trials"] >= 5
        assert 1 <= t["streams"] <= 8
        assert t["trace_block"] in (64, 128, 256) and t["stream_block"] in (128, 256, 512)
        assert t["packet"] in (64, 128, 256) and t["burial_block"] in (64, 128, 256)
        assert t["burial_lanes"] in (0, 64, 128) and t["trace_stream"] in (0, 1)
        again = ctx.autotune_batch(ins, ops, max_streams=8, cache_path=cache)
        assert again["from_cache"] == 1
        keys = ("streams", "trace_block", "stream_block", "packet", "burial_block", "burial_lanes", "trace_stream")
        assert {k: again[k] for k in keys} == {k: t[k] for k in keys}
        got = ctx.bake_batch(ins, ops, streams=t["streams"])
    finally:
        ctx.set_launch(128, 256)
        ctx.set_gather(1, block=256, chunk=128)
        ctx.set_burial(0)
        ctx.set_trace_stream(True)
    for j, inp, o, g in zip(pick, ins, ops, got):
        assert_scenes_equal(g, ctx.bake(inp, o))
    with pytest.raises(BakeError):
        ctx.autotune_batch([], None)

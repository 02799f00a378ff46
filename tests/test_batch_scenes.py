"""The batch workload generator: PCG32 block draws equal the scalar recurrence, jobs are
deterministic, and their sizes stay in the stated ranges (500-5e4 triangles, 1-64 lights with
40/30/30 point/spot/area kinds, bounce on odd jobs)."""
import numpy as np

from dynlight_hip import abi, scenes


def test_block_draws_equal_scalar_draws():
    for seed in (0, 7, 0xDEADBEEF):
        a, b = scenes.Pcg32.seeded(seed), scenes.Pcg32.seeded(seed)
        for n in (1, 2, 3, 17, 64, 1000):
            assert [a.next_u32() for _ in range(n)] == list(b.u32s(n))
            assert a.state == b.state
        assert a.uniform(2.0, 5.0) == b.uniforms(1, 2.0, 5.0)[0]


def test_pcg32_known_outputs():
    """PCG32 XSH-RR with the reference's seeding recurrence (src/effects/rng.rs:48-71): the published
    pcg32 demo generator (initstate 42, initseq 54) starts 0xa15c02b7, 0x7b47f409, 0xba1d3330."""
    r = scenes.Pcg32(42, 54)
    assert [r.next_u32() for _ in range(3)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330]
    r = scenes.Pcg32(42, 54)
    assert list(r.u32s(3)) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330]


def test_jobs_are_deterministic():
    a, oa = scenes.batch_job(5)
    b, ob = scenes.batch_job(5)
    assert np.array_equal(a._positions, b._positions) and np.array_equal(a._indices, b._indices)
    assert oa.pixel_density == ob.pixel_density
    assert [l.position for l in a.lights] == [l.position for l in b.lights]
    c, _ = scenes.batch_job(6)
    assert a.triangle_count != c.triangle_count or not np.array_equal(a._positions, c._positions)


def test_job_sizes_follow_the_distributions():
    ins, ops = scenes.batch(512)
    tris = np.array([i.triangle_count for i in ins])
    lights = np.array([len(i.lights) for i in ins])
    assert tris.min() >= 500 - 12 and tris.max() <= 50_000 + 12
    assert lights.min() >= 1 and lights.max() <= 64
    # log-uniform: the median sits near the geometric mean of the range
    assert 3000 < np.median(tris) < 8000
    kinds = [l.shape for i in ins for l in i.lights]
    area = np.mean([k in (abi.SHAPE_RECTANGLE, abi.SHAPE_TUBE) for k in kinds])
    assert 0.25 < area < 0.35
    for j, i in enumerate(ins):
        bounce = {l.illumination for l in i.lights}
        assert bounce == {abi.ILLUMINATION_SINGLE_BOUNCE if j % 2 else abi.ILLUMINATION_DIRECT}
    assert all(o.pixel_density > 0 for o in ops)

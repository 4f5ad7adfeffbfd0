"""The served form on the GPU: jobs of both families arriving over time, admitted under a tight HBM
budget so that the resident set changes from tick to tick; every job's final state must be
bit-identical to the same job run alone."""
import pytest

from braingrid_amd import serve as S
from braingrid_amd import sweep as SW

pytestmark = pytest.mark.gpu


def test_served_jobs_equal_alone(hip_ctx):
    jobs = S.served_requests(10, rate=50.0, seed=11, side_lo=10, side_hi=28, frac_lif=0.5, dur_lo=1, dur_hi=3)
    jobs.append(S.Job(id=100, spec=dict(id=100, side=0, kind="izh_stdp", sim_s=1.0, seed=1), arrival=0.01))
    print(f"\n[served {len(jobs)} jobs]", flush=True)
    eng = S.DeviceEngine(hip_ctx, fingerprint=True)
    fps = [SW.footprint(SW.plan(j.spec)) for j in jobs if j.spec["side"] > 0]
    budget = 3 * max(fps)
    srv = S.Server(engine=eng, budget=budget)
    srv.run(jobs)
    eng.close()
    counts = srv.counts()
    assert counts.get(S.DONE) == 10 and counts.get(S.FAILED) == 1, counts
    assert max(u for _, u in srv.occupancy) <= budget
    assert srv.stats()["latency_p99_s"] > 0
    for j in jobs:
        if j.state != S.DONE:
            continue
        alone = SW.SweepNet(hip_ctx, j.spec)
        for _ in range(alone.epochs):
            alone.net.advance(alone.steps_per_epoch)
            alone.end_epoch()
        fp = alone.fingerprint()
        summ = alone.summary()
        alone.close()
        assert fp == j.result["fingerprint"], (j.id, j.spec)
        for k in ("burstinessHist", "spikesHistory"):
            assert (summ[k] == j.result[k]).all(), (j.id, k)
        assert summ[k].sum() > 0


def test_evicted_jobs_resume_bit_for_bit(hip_ctx):
    """Long residents of both families fill a tight budget; a large job arriving later is passed
    over until residents are checkpointed and evicted for it. Every job, evicted or not, ends in the
    state of the same job run alone."""
    mk = lambda i, side, kind, sim_s, t: S.Job(id=i, arrival=t, spec=dict(
        id=i, side=side, kind=kind, tR=1.0, fE=0.9, seed=SW.net_seed(5, i), sim_s=sim_s))
    jobs = [mk(0, 16, "lif_growth", 4.0, 0.0), mk(1, 14, "izh_stdp", 4.0, 0.0), mk(2, 18, "lif_growth", 4.0, 0.0),
            mk(3, 18, "izh_stdp", 2.0, 0.05)]
    fps = {j.id: SW.footprint(SW.plan(j.spec)) for j in jobs}
    budget = fps[0] + fps[1] + fps[2]
    assert fps[2] < fps[3] <= budget - fps[0]   # the late job fits only when both newer residents leave
    print(f"\n[served eviction budget={budget}]", flush=True)
    eng = S.DeviceEngine(hip_ctx, fingerprint=True)
    srv = S.Server(engine=eng, budget=budget, max_skips=0, evict=True)
    try:
        srv.run(jobs)
    finally:
        eng.close()
        srv.close()
    assert srv.counts() == {S.DONE: 4} and srv.evicted >= 1 and srv.restored == srv.evicted
    assert any(j.evictions for j in jobs)
    for j in jobs:
        alone = SW.SweepNet(hip_ctx, j.spec)
        for _ in range(alone.epochs):
            alone.net.advance(alone.steps_per_epoch)
            alone.end_epoch()
        fp = alone.fingerprint()
        summ = alone.summary()
        alone.close()
        assert fp == j.result["fingerprint"], (j.id, j.evictions)
        for k in ("burstinessHist", "spikesHistory"):
            assert (summ[k] == j.result[k]).all(), (j.id, k)

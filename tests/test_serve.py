"""Scheduling of the served form (braingrid_amd/serve.py) on the CPU, with an engine that only
keeps accounts: admission by footprint, rejection, failures, cancellation, fairness, statistics."""
import json
import os

import numpy as np

from braingrid_amd import serve as S


class FakeEngine:
    """Footprint = 1000 * side; a spec with side 13 fails at creation, one with kind "boom" raises
    a device error at its second epoch. Tracks resident bytes to check the budget is never passed."""

    def __init__(self, clock):
        self.clock = clock
        self.resident = {}
        self.max_resident = 0
        self.released = []
        self.advanced = []
        self.saved = []

    def total_bytes(self):
        return 10**9

    def plan(self, spec):
        if spec.get("side", 0) <= 0:
            raise ValueError("bad side")
        return {"bytes": 1000 * spec["side"]}

    def footprint(self, plan):
        return plan["bytes"]

    def create(self, spec, plan):
        if spec["side"] == 13:
            raise RuntimeError("out of device memory")
        h = {"spec": spec, "epochs": 0}
        self.resident[id(h)] = plan["bytes"]
        self.max_resident = max(self.max_resident, sum(self.resident.values()))
        return h

    def advance(self, handles, steps):
        self.advanced.append(len(handles))
        self.clock.t += 0.01 * len(handles)

    def end_epoch(self, h):
        h["epochs"] += 1

    def check(self, h):
        if h["spec"].get("kind") == "boom" and h["epochs"] == 2:
            raise RuntimeError("device error bits 0x2")

    def result(self, h):
        return {"epochs": h["epochs"], "side": h["spec"]["side"]}

    def release(self, h):
        self.released.append(h["spec"]["id"])
        del self.resident[id(h)]

    def save(self, h, path):
        with open(path, "w") as f:
            json.dump({"id": h["spec"]["id"], "epochs": h["epochs"]}, f)
        self.saved.append(h["spec"]["id"])

    def restore(self, spec, plan, path):
        with open(path) as f:
            st = json.load(f)
        assert st["id"] == spec["id"]
        h = self.create(spec, plan)
        h["epochs"] = st["epochs"]
        return h


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


def job(i, side, sim_s=2.0, arrival=0.0, kind="lif_growth"):
    return S.Job(id=i, spec=dict(id=i, side=side, sim_s=sim_s, kind=kind), arrival=arrival)


def make(budget=10_000, **kw):
    clock = Clock()
    eng = FakeEngine(clock)
    return S.Server(engine=eng, budget=budget, clock=clock, **kw), eng, clock


def test_all_jobs_complete_within_budget():
    srv, eng, clock = make(budget=10_000)
    jobs = [job(i, side=1 + (i * 7) % 9, sim_s=float(1 + i % 3), arrival=0.05 * i) for i in range(40)]
    srv.run(jobs, sleep=clock.sleep)
    assert srv.counts() == {S.DONE: 40}
    assert eng.max_resident <= 10_000
    for j in jobs:
        assert j.result["epochs"] == int(j.spec["sim_s"])
        assert j.finished >= srv.t0 + j.arrival
    assert not eng.resident and sorted(eng.released) == list(range(40))
    st = srv.stats(total_bytes=eng.total_bytes())
    assert st["completed"] == 40 and st["latency_p99_s"] >= st["latency_p50_s"] > 0
    assert 0 < st["hbm_occupancy_max"] <= 10_000 / 10**9


def test_oversized_job_is_rejected_and_others_run():
    srv, eng, clock = make(budget=5_000)
    jobs = [job(0, 3), job(1, 6, arrival=0.01), job(2, 4, arrival=0.02)]
    srv.run(jobs, sleep=clock.sleep)
    assert jobs[1].state == S.REJECTED and "exceeds the budget" in jobs[1].error
    assert jobs[0].state == S.DONE and jobs[2].state == S.DONE
    assert jobs[1].handle is None and 1 not in eng.released


def test_poisoned_requests_fail_alone():
    srv, eng, clock = make(budget=20_000)
    jobs = [job(0, 2), job(1, 0), job(2, 13), job(3, 3, kind="boom", sim_s=4.0), job(4, 2, sim_s=3.0)]
    srv.run(jobs, sleep=clock.sleep)
    assert jobs[1].state == S.FAILED and "bad side" in jobs[1].error          # invalid specification
    assert jobs[2].state == S.FAILED and "out of device memory" in jobs[2].error   # creation failed
    assert jobs[3].state == S.FAILED and "device error" in jobs[3].error      # failed while running
    assert 3 in eng.released and jobs[3].epochs_done == 1
    assert jobs[0].state == S.DONE and jobs[4].state == S.DONE
    assert srv.used == 0 and not eng.resident


def test_cancel_queued_and_active():
    srv, eng, clock = make(budget=3_000)
    a, b, c = job(0, 3, sim_s=5.0), job(1, 3, sim_s=5.0), job(2, 1, sim_s=5.0)
    for j in (a, b, c):
        srv.submit(j)
    srv.tick()                     # only a fits (3000 of 3000): b and c wait
    assert a.state == S.ACTIVE and b.state == S.QUEUED and c.state == S.QUEUED
    assert srv.cancel(b.id) and b.state == S.CANCELLED
    assert srv.cancel(a.id) and a.state == S.CANCELLED and 0 in eng.released and srv.used == 0
    assert not srv.cancel(a.id)
    srv.tick()
    assert c.state == S.ACTIVE and b.handle is None
    while srv.active:
        srv.tick()
    assert c.state == S.DONE and srv.counts() == {S.CANCELLED: 2, S.DONE: 1}


def test_small_jobs_pass_a_waiting_large_one_but_not_forever():
    srv, eng, clock = make(budget=10_000, max_skips=3)
    big = job(0, 8, sim_s=2.0, arrival=0.001)
    smalls = [job(i, 3, sim_s=4.0, arrival=0.002 * i) for i in range(1, 30)]
    runner = job(99, 5, sim_s=6.0, arrival=0.0)   # holds half the budget at the start
    srv.run([runner, big] + smalls, sleep=clock.sleep)
    assert srv.counts() == {S.DONE: 31}
    overtook = [s for s in smalls if s.admitted < big.admitted]
    assert big.skipped > 3 and 0 < len(overtook) < len(smalls)


def test_request_stream_shape():
    jobs = S.served_requests(2000, rate=20.0, seed=7)
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0) and abs(arr[-1] / 2000 - 1 / 20) < 0.005
    sides = np.array([j.spec["side"] for j in jobs])
    assert sides.min() >= 30 and sides.max() <= 200
    assert abs(np.mean(np.log(sides)) - (np.log(30) + np.log(200)) / 2) < 0.05
    frac = np.mean([j.spec["kind"] == "lif_growth" for j in jobs])
    assert abs(frac - 0.7) < 0.04
    d = np.array([j.spec["sim_s"] for j in jobs])
    assert d.min() == 1 and d.max() == 20 and np.all(d == np.round(d))


def test_eviction_makes_room_for_a_blocked_job():
    """Long residents fill the budget; a large job passed over max_skips times gets room at an epoch
    boundary by checkpointing the newest residents, which later continue from their checkpoints."""
    srv, eng, clock = make(budget=10_000, max_skips=2, evict=True)
    residents = [job(i, 3, sim_s=30.0, arrival=0.001 * i) for i in range(3)]      # 9000 B for 30 epochs
    big = job(10, 6, sim_s=2.0, arrival=0.01)
    try:
        srv.run(residents + [big], sleep=clock.sleep)
        assert srv.counts() == {S.DONE: 4}
        assert srv.evicted >= 1 and srv.restored == srv.evicted and sorted(eng.saved) == [1, 2]
        assert big.finished < min(r.finished for r in residents[1:])
        for r in residents:
            assert r.result["epochs"] == 30 and r.epochs_done == 30
        assert residents[0].evictions == 0 and residents[2].evictions == 1
        assert eng.max_resident <= 10_000 and srv.used == 0
        assert all(r.checkpoint is None for r in residents) and not os.listdir(srv.spill_dir)
        st = srv.stats()
        assert st["evicted"] == srv.evicted and st["spill_bytes"] > 0
    finally:
        srv.close()


def test_no_eviction_when_it_cannot_free_enough():
    srv, eng, clock = make(budget=10_000, max_skips=1, evict=True, max_evictions=0)
    jobs = [job(0, 5, sim_s=6.0), job(1, 4, sim_s=6.0, arrival=0.001), job(2, 6, sim_s=1.0, arrival=0.002)]
    try:
        srv.run(jobs, sleep=clock.sleep)
        assert srv.evicted == 0 and not eng.saved and srv.counts() == {S.DONE: 3}
        assert jobs[2].admitted >= jobs[0].finished and jobs[2].skipped > 1
    finally:
        srv.close()


def test_cancel_of_an_evicted_job_removes_its_checkpoint():
    srv, eng, clock = make(budget=6_000, max_skips=0, evict=True)
    a, b = job(0, 3, sim_s=10.0), job(1, 3, sim_s=10.0, arrival=0.001)
    big = job(2, 5, sim_s=3.0, arrival=0.002)
    try:
        for j in (a, b, big):
            srv.submit(j)
        srv.tick()            # a and b admitted; big blocked at once (max_skips 0)
        srv.tick()            # epoch boundary: both residents evicted to make room for big
        assert a.state == S.QUEUED and b.state == S.QUEUED and a.checkpoint and b.checkpoint
        path = b.checkpoint
        assert os.path.exists(path) and srv.cancel(b.id)
        assert not os.path.exists(path) and b.state == S.CANCELLED
        while srv.active or srv.queue:
            srv.tick()
        assert big.state == S.DONE and a.state == S.DONE and a.result["epochs"] == 10
    finally:
        srv.close()


class _FakeBatch:
    def __init__(self):
        self.sets = []
        self.steps = 0

    def set(self, nets):
        self.sets.append(list(nets))

    def advance(self, steps):
        self.steps += steps


class _FakeHandle:
    def __init__(self, tag):
        self.net = tag
        self.closed = False

    def close(self):
        self.closed = True


def test_device_engine_sets_batch_again_after_a_member_is_released():
    """The served engine caches its batch membership between epochs. A job released after the
    queue drained and a new job admitted later must form a new membership even when the new
    handle object reuses the released one's address (the cache compares the objects themselves)."""
    eng = S.DeviceEngine.__new__(S.DeviceEngine)
    eng.batch = _FakeBatch()
    eng._members = None
    eng.fingerprint = False
    h1 = _FakeHandle("net1")
    eng.advance([h1], 10)
    eng.advance([h1], 10)
    assert len(eng.batch.sets) == 1
    eng.release(h1)
    assert h1.closed
    del h1
    # many new handles: CPython reuses freed addresses, so some of these share the old id()
    for k in range(50):
        h = _FakeHandle(f"net{k + 2}")
        eng.advance([h], 10)
        assert eng.batch.sets[-1] == [f"net{k + 2}"]
        eng.release(h)
        del h
    assert len(eng.batch.sets) == 51
    # the same members in the same order are not set again; another order or set is
    a, b = _FakeHandle("a"), _FakeHandle("b")
    eng.advance([a, b], 1)
    eng.advance([a, b], 1)
    eng.advance([b, a], 1)
    eng.advance([b], 1)
    assert [x for x in eng.batch.sets[51:]] == [["a", "b"], ["b", "a"], ["b"]]

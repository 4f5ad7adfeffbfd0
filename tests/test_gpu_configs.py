"""The benchmark configurations at their own workloads on the GPU, against the oracle (deterministic
noise mode, bit for bit) and, for the batched and served forms, against the same networks run alone.

- 10 x 10 LIF + DS + ConnGrowth, 10 epochs of 1 s, no input, as configured (startRadius 0.4: no
  synapse can form) and with startRadius 0.6 (4-neighbour synapses from the first growth update);
- 1,000 Izhikevich neurons with 100 random DynamicSTDP synapses each, delays of 11..201 steps on a
  256-bit ring, and 10 Hz Poisson input into every neuron, for 1 s (10,000 steps) through the
  per-step STDP kernels;
- the 256-network sweep (seed 2024) batched for 0.3 s, every network against its alone run, and a
  stratified sample against the oracle;
- a served stream of 32 jobs of both families, sides 30..120, under a budget that forces turnover,
  every job against its alone run and the smallest jobs of each family against the oracle.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from braingrid_amd import bghip as B
from braingrid_amd import networks as NW
from braingrid_amd import serve as S
from braingrid_amd import sweep as SW
from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("start_radius", [0.4, 0.6])
def test_plain_method_10x10_ten_epochs(hip_ctx, start_radius):
    """The plain method as configured: 100 LIF neurons (10% inhibitory and 10% starters placed by
    DynamicLayout, seed 1), DS synapses, growth epsilon 0.6, beta 0.1, rho 1e-4, targetRate 1.9,
    minRadius 0.1, ten 1 s epochs with the growth update and history step after each."""
    print(f"\n[config0] 10x10, 10 x 1 s, startRadius {start_radius}", flush=True)
    cfg = O.default_config()
    cfg.update(width=10, height=10, layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1, seed=1,
               epoch_duration=1.0, num_epochs=10, start_radius=start_radius, epsilon=0.6, beta=0.1, rho=1e-4,
               target_rate=1.9, min_radius=0.1, noise_mode=1, math_mode=1, sum_mode=1)
    ora = O.OracleNet(cfg)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    try:
        steps = H.steps_per_epoch(cfg)
        assert steps == 10000
        bad = []
        for ep in range(10):
            ora.advance(steps)
            net.advance(steps)
            ora.update_connections()
            net.update_connections()
            ora.update_history()
            net.update_history()
            net.sync()
            bad += H.compare_state(ora, net, cfg, f"epoch{ep + 1}")
        total = net.total_synapses
        print(f"[config0] {total} synapses, {len(bad)} differences", flush=True)
        assert not bad, bad[:10]
        assert (total == 0) if start_radius == 0.4 else (total > 300)
        assert net.error_bits() == 0
    finally:
        net.close()
        ora.close()


def test_izh_dynstdp_poisson_one_second(hip_ctx):
    """1,000 Izhikevich neurons (800 E / 200 I, r-correlated a, b, c, d), 100 random outgoing
    DynamicSTDP synapses per neuron, delays U{1..20} ms (11..201 steps, 256-bit ring), STDP
    tau+ = tau- = 20 ms, A+ 0.005, A- -0.00525, and 10 Hz Poisson input of weight 25 into every
    neuron (the product path: braingrid_amd.networks builder, input initialised on the device), for
    10,000 steps. The input state, every neuron and every synapse equal the oracle's after each
    quarter."""
    N, k, seed = 1000, 100, 2
    print(f"\n[config2] N={N} k={k} Poisson 10 Hz, 10000 steps", flush=True)
    neu, syn, meta = NW.izh_stdp_arrays(N, seed, k=k)
    assert meta["ring_bits"] == 256
    ora, cfg = H.izh_static_oracle(neu, syn, meta, seed, input=dict(kind="poisson", fr_mean=10.0, weight=25.0))
    net, _ = NW.izh_stdp(hip_ctx, N, seed, k=k)
    net.upload_input(B.INPUT_POISSON, neuron_type=meta["ntype"], fr_mean=10.0, weight=25.0)
    try:
        bad = H.compare_input(ora, net, cfg, "upload") + H.compare_state(ora, net, cfg, "upload")
        assert not bad, bad
        for q in range(4):
            ora.advance(2500)
            net.advance(2500)
            bad = H.compare_input(ora, net, cfg, f"step{ora.step}") + H.compare_state(ora, net, cfg, f"step{ora.step}")
            assert not bad, bad
        spikes = int(ora.array("spikeCount").sum())
        inputs = int((ora.array("in_lastSpike") != np.uint64(0xFFFFFFFFFFFFFFFF)).sum())
        print(f"[config2] {spikes} spikes, {inputs} neurons received input", flush=True)
        assert spikes > 0 and inputs > 900
        assert net.error_bits() == 0
    finally:
        net.close()
        ora.close()


SWEEP_STEPS = 3000   # 0.3 s simulated


def _sweep_sample(specs):
    """The stratified sample of the sweep compared with the oracle: the smallest network, the
    smallest and largest growth networks, the largest and the median Izhikevich networks (for seed
    2024: the Izhikevich network of side 10, growth networks of sides 11 and 116, Izhikevich networks
    of sides 112 (12,544 neurons, 1.25M synapses) and 35)."""
    lif = [s for s in specs if s["kind"] == "lif_growth"]
    izh = sorted([s for s in specs if s["kind"] == "izh_stdp"], key=lambda s: s["side"])
    pick = [min(specs, key=lambda s: s["side"]), min(lif, key=lambda s: s["side"]),
            max(lif, key=lambda s: s["side"]), izh[-1], izh[len(izh) // 2]]
    out, seen = [], set()
    for s in pick:
        if s["id"] not in seen:
            seen.add(s["id"])
            out.append(s)
    return out


def _oracle_state(spec, steps):
    ora, cfg = H.oracle_for_spec(spec)
    ora.advance(steps)
    return ora, cfg


def test_sweep_256_networks_batched_equal_alone_and_oracle(hip_ctx):
    """All 256 networks of the sweep advanced together by shared launches (launch-shape autotuner
    on) for 3,000 steps give each network's alone result bit for bit (full-state digest); a
    stratified sample equals the oracle in every state array."""
    specs = SW.sweep_specs(256, 2024, sim_s=1.0)
    sample = _sweep_sample(specs)
    print(f"\n[config3] 256 networks, {SWEEP_STEPS} steps; oracle sample ids {[s['id'] for s in sample]}", flush=True)
    # the oracle runs of the sample, on host threads while the device works
    pool = ThreadPoolExecutor(max_workers=len(sample))
    futs = {s["id"]: pool.submit(_oracle_state, s, SWEEP_STEPS) for s in sample}
    try:
        nets = SW.build(hip_ctx, specs)
        for x in nets:
            x.net.advance(SWEEP_STEPS)
            x.net.sync()
        alone = {x.spec["id"]: x.fingerprint() for x in nets}
        for x in nets:
            x.close()
        print("[config3] alone done", flush=True)
        nets = SW.build(hip_ctx, specs)
        batch = B.Batch(hip_ctx)
        batch.autotune(SW.TUNE_STEPS)
        batch.set([x.net for x in nets])
        batch.advance(SWEEP_STEPS)
        batch.sync()
        shapes = batch.shapes()
        diff = [x.spec["id"] for x in nets if x.fingerprint() != alone[x.spec["id"]]]
        errs = [x.spec["id"] for x in nets if x.net.error_bits()]
        print(f"[config3] batched: {len(diff)} networks differ from alone; groups "
              f"{[(g['mode'], g['networks'], g['tuned']) for g in shapes]}", flush=True)
        assert not diff, diff[:10]
        assert not errs, errs[:10]
        byid = {x.spec["id"]: x for x in nets}
        bad = []
        for s in sample:
            ora, cfg = futs[s["id"]].result()
            where = f"net{s['id']}-side{s['side']}"
            bad += H.compare_state(ora, byid[s["id"]].net, cfg, where) + H.compare_input(ora, byid[s["id"]].net, cfg, where)
            ora.close()
        print(f"[config3] oracle sample: {len(bad)} differences", flush=True)
        batch.close()
        for x in nets:
            x.close()
        assert not bad, bad[:10]
    finally:
        pool.shutdown(wait=True)


def test_served_stream_equal_alone_and_oracle(hip_ctx):
    """32 jobs (70% growth / 30% Izhikevich, sides log-uniform 30..120, 1..3 s each) arriving at
    40/s into a budget of about six median jobs, so jobs queue, join and leave between ticks. Every
    completed job's result (spike-count histograms, final radii) equals its alone run's, the whole
    final state digest included; the two smallest jobs of each family are also run start to end in
    the oracle and their results are equal to the served ones."""
    jobs = S.served_requests(32, rate=40.0, seed=23, side_lo=30, side_hi=120, frac_lif=0.7, dur_lo=1, dur_hi=3)
    fps = sorted(SW.footprint(SW.plan(j.spec)) for j in jobs)
    budget = max(fps[-1], 6 * fps[len(fps) // 2])
    kinds = {k: sorted([j for j in jobs if j.spec["kind"] == k], key=lambda j: (j.spec["side"], j.spec["sim_s"]))
             for k in ("lif_growth", "izh_stdp")}
    check = kinds["lif_growth"][:2] + kinds["izh_stdp"][:2]
    print(f"\n[config4] {len(jobs)} jobs, budget {budget / 2**20:.0f} MiB; oracle jobs "
          f"{[(j.id, j.spec['kind'], j.spec['side'], j.spec['sim_s']) for j in check]}", flush=True)
    pool = ThreadPoolExecutor(max_workers=len(check))
    futs = {j.id: pool.submit(H.run_spec_in_oracle, j.spec) for j in check}
    eng = S.DeviceEngine(hip_ctx, fingerprint=True)
    srv = S.Server(engine=eng, budget=budget)
    try:
        srv.run(jobs)
    finally:
        eng.close()
    try:
        counts = srv.counts()
        assert counts.get(S.DONE) == len(jobs), counts
        assert max(u for _, u in srv.occupancy) <= budget
        # turnover: some job was admitted only after another had finished
        assert any(j.admitted > min(i.finished for i in jobs) for j in jobs)
        for j in jobs:
            alone = SW.SweepNet(hip_ctx, j.spec)
            for _ in range(alone.epochs):
                alone.net.advance(alone.steps_per_epoch)
                alone.end_epoch()
            fp = alone.fingerprint()
            alone.close()
            assert fp == j.result["fingerprint"], (j.id, j.spec)
        print("[config4] every job equal to its alone run", flush=True)
        bad = []
        for j in check:
            ref = futs[j.id].result()
            for k, v in ref.items():
                got = np.asarray(j.result[k])
                if got.shape != v.shape or not np.array_equal(got.view(np.uint32), v.astype(np.float32).view(np.uint32)):
                    bad.append((j.id, k))
            assert ref["burstinessHist"].sum() > 0
        print(f"[config4] oracle jobs: {len(bad)} differences", flush=True)
        assert not bad, bad
    finally:
        pool.shutdown(wait=True)

"""Long-run firing rates of a network whose synapses grow during the run, against the compiled
reference's CPU path with its own noise stream.

Workload: the reference's configfiles/test-medium.xml (tests/golden/ref/configfiles): 30 x 30
AllLIFNeurons with FixedLayout (9 starter and 18 inhibitory neurons from its NList files),
AllDSSynapses, ConnGrowth from startRadius 0.4, twenty 100 s epochs at dt = 1e-4 s (2e7 steps).
The 9 starter neurons fire from the start (the others stay below threshold throughout); every
neurite grows until neighbours overlap after epoch 10 and synapses form, through which the
starters' spikes then reach their neighbours. The reference side
is the committed fixture tests/golden/ref_rates_medium.npz: oracle/_ref/growth (the reference's
simulator built from its sources) once per noise seed, written by
tests/golden/make_medium_rates_fixture.py. The GPU side runs the host model for the same seeds,
all eight networks advanced together by one batch (bghip_batch_*), with its counter-based noise.
The two noise streams differ, so the comparison is statistical."""
import os

import numpy as np
import pytest

from braingrid_amd import bghip as B
from braingrid_amd import host

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "ref")


def _run_medium(ctx, seeds, epochs):
    cfg = host.read_config(os.path.join(REF, "configfiles", "test-medium.xml"), REF)
    assert cfg["width"] == 30 and cfg["num_epochs"] == epochs
    models = [host.Model(ctx, dict(cfg, seed=int(s))) for s in seeds]
    steps = int(round(cfg["epoch_duration"] / cfg["dt"]))
    batch = B.Batch(ctx)
    try:
        batch.set([m.net for m in models])
        for e in range(epochs):
            batch.advance(steps)
            for m in models:
                m.update_connections(stats=False)
                m.update_history()
            if e % 5 == 4:
                print(f"[medium] epoch {e + 1}: {[m.net.total_synapses for m in models]} synapses", flush=True)
        batch.sync()
        N = cfg["width"] * cfg["height"]
        rates = np.array([m.array("ratesHistory").reshape(epochs + 1, N) for m in models], np.float64)
        radii = np.array([m.array("radiiHistory").reshape(epochs + 1, N) for m in models], np.float64)
        burst = np.array([m.array("burstinessHist") for m in models], np.float64)
        synapses = [m.net.total_synapses for m in models]
        errs = [m.net.error_bits() for m in models]
    finally:
        batch.close()
        for m in models:
            m.close()
    counts = np.rint(rates[:, 1:] * cfg["epoch_duration"])
    return counts, radii, burst, synapses, errs


def test_firing_rates_of_the_growing_test_medium_network(hip_ctx):
    """Tolerances, stated here. The starters' firing is close to a Poisson process (an epoch's
    eight-seed spike sum varies by about 1.1% between independent noise streams, near the Poisson
    value), so the difference of two independent runs has a standard deviation of about 0.34% for
    the whole run's total, 0.5% for a ten-epoch half and 1.5% for one epoch's eight-seed sum:
      - the mean spike total over the eight seeds within 1% of the reference's, and so the mean
        firing rate;
      - the spike totals of the first ten epochs (no synapses) and of the last ten (with
        synapses) each within 2%;
      - burstinessHist: the burst totals equal the spike counts on each side, and the eight-run
        sums agree over every 100 s epoch within 6% (4 standard deviations: a missing or
        mis-binned part of an epoch breaks it, the noise of 20 epochs does not);
      - per neuron, over all epochs and seeds: the same neurons fire, and the mean of
        z = (g - r) / sqrt(g + r) over the 9 firing neurons is within 1.2 of 0 (the mean of 9
        such values has a standard error of about 1/3) with a standard deviation below 2.5;
      - final radii: the mean over neurons and seeds within 1e-3 relative of the reference's
        (seed-to-seed 1e-5: the silent neurons' radii follow the same deterministic outgrowth),
        and every neuron's eight-seed mean within 5% (a starter's radius follows its rate)."""
    fx = np.load(os.path.join(HERE, "golden", "ref_rates_medium.npz"))
    ref = fx["counts"].astype(np.float64)
    seeds = [int(s) for s in fx["seeds"]]
    S, E, N = ref.shape
    gpu, g_radii, g_burst, synapses, errs = _run_medium(hip_ctx, seeds, E)
    assert errs == [0] * S
    print(f"[medium] synapses at the end: {synapses}", flush=True)
    assert min(synapses) > 0

    g_tot, r_tot = gpu.sum(axis=(1, 2)), ref.sum(axis=(1, 2))
    rel = (g_tot.mean() - r_tot.mean()) / r_tot.mean()
    spread = np.std(r_tot / r_tot.mean(), ddof=1)
    print(f"[medium] spikes per seed: gpu {g_tot.tolist()} reference {r_tot.tolist()}; mean difference "
          f"{rel * 100:.3f}%, reference seed-to-seed spread {spread * 100:.3f}%", flush=True)
    assert abs(rel) < 0.01
    half = E // 2
    for name, sl in (("epochs 1-10", slice(0, half)), ("epochs 11-20", slice(half, E))):
        g, r = gpu[:, sl].sum(), ref[:, sl].sum()
        print(f"[medium] {name}: gpu {g:.0f} reference {r:.0f} ({(g - r) / r * 100:.3f}%)", flush=True)
        assert abs(g - r) / r < 0.02

    r_burst = fx["burst"].astype(np.float64)
    assert np.array_equal(g_burst.sum(axis=1), g_tot)
    assert np.array_equal(r_burst.sum(axis=1), r_tot)
    per_epoch_g = g_burst.sum(axis=0).reshape(E, -1).sum(axis=1)
    per_epoch_r = r_burst.sum(axis=0).reshape(E, -1).sum(axis=1)
    dev = (per_epoch_g - per_epoch_r) / per_epoch_r
    print(f"[medium] burstinessHist per epoch: {np.round(dev * 100, 2).tolist()}%", flush=True)
    assert np.all(np.abs(dev) < 0.06)

    cg, cr = gpu.sum(axis=(0, 1)), ref.sum(axis=(0, 1))
    assert np.array_equal(cg > 0, cr > 0)
    f = cr > 0
    z = (cg[f] - cr[f]) / np.sqrt(cg[f] + cr[f])
    print(f"[medium] {int(f.sum())} firing neurons: z mean {z.mean():.4f}, sd {z.std(ddof=1):.4f}", flush=True)
    assert abs(z.mean()) < 1.2
    assert z.std(ddof=1) < 2.5

    r_radii = fx["radii"].astype(np.float64)
    gr, rr = g_radii[:, -1].mean(axis=0), r_radii[:, -1].mean(axis=0)
    rel_r = (gr.mean() - rr.mean()) / rr.mean()
    worst = np.max(np.abs(gr - rr) / rr)
    print(f"[medium] final radius: mean {gr.mean():.5f} vs reference {rr.mean():.5f} ({rel_r * 100:.3f}%), "
          f"worst neuron {worst * 100:.2f}%", flush=True)
    assert abs(rel_r) < 1e-3
    assert worst < 0.05

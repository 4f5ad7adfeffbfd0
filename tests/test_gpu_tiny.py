"""The smallest networks, against the oracle bit for bit: networks of fewer neurons than a
workgroup's four destinations or a streaming wavefront's four rows, so the kernels run with empty
destinations in their workgroups and wavefronts.

- LIF + DS + ConnGrowth on lines of 1, 2, 3 and 5 neurons, neuron 0 a starter, startRadius 2.0 (each
  neuron reaches the next ones at the first growth update; one neuron alone forms none), three
  0.1 s epochs with the growth update and history step after each: after the first update the
  synapses advance in the implicit-ring fused kernel;
- Izhikevich + DynamicSTDP + 10 Hz Poisson networks of 4 and 9 neurons (the sweep's family at
  its smallest sides), 3,000 steps through the per-step STDP kernels;
- a batch of such networks of both families against each network run alone."""
import pytest

from braingrid_amd import bghip as B
from braingrid_amd import sweep as SW
from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", [1, 2, 3, 5])
def test_tiny_growth_line(hip_ctx, width):
    print(f"\n[tiny] LIF line of {width}", flush=True)
    cfg = O.default_config()
    cfg.update(width=width, height=1, active_list=[0], start_radius=2.0, epoch_duration=0.1, num_epochs=3,
               noise_mode=1, math_mode=1, sum_mode=1)
    ora = O.OracleNet(cfg)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    try:
        steps = H.steps_per_epoch(cfg)
        bad = []
        for ep in range(3):
            ora.advance(steps)
            net.advance(steps)
            ora.update_connections()
            net.update_connections()
            ora.update_history()
            net.update_history()
            net.sync()
            bad += H.compare_state(ora, net, cfg, f"width{width}-epoch{ep + 1}")
        total = net.total_synapses
        spikes = int(ora.array("spikeCount").sum() + ora.array("spikeCountOffset").sum())
        print(f"[tiny] {total} synapses, {len(bad)} differences", flush=True)
        assert not bad, bad[:10]
        assert (total == 0) if width == 1 else (total > 0)
        assert spikes > 0
        assert net.error_bits() == 0
    finally:
        net.close()
        ora.close()


def _izh_spec(side, k):
    return dict(id=k, side=side, kind="izh_stdp", tR=0.5, fE=0.9, seed=SW.net_seed(31, k), sim_s=1.0)


@pytest.mark.parametrize("side", [2, 3])
def test_tiny_izh_stdp_poisson(hip_ctx, side):
    print(f"\n[tiny] Izhikevich + DynamicSTDP + Poisson, {side}x{side}", flush=True)
    spec = _izh_spec(side, side)
    ora, cfg = H.oracle_for_spec(spec)
    x = SW.SweepNet(hip_ctx, spec)
    try:
        bad = []
        for c in (1000, 2000):
            ora.advance(c)
            x.net.advance(c)
            x.net.sync()
            where = f"side{side}-step{ora.step}"
            bad += H.compare_state(ora, x.net, cfg, where) + H.compare_input(ora, x.net, cfg, where)
        assert not bad, bad[:10]
        assert x.net.error_bits() == 0
    finally:
        x.close()
        ora.close()


def test_tiny_networks_batched_equal_alone(hip_ctx):
    specs = [_izh_spec(s, k) for k, s in enumerate((2, 3, 2, 4))]
    specs += [dict(id=10 + k, side=s, kind="lif_growth", tR=0.5, fE=0.9, seed=SW.net_seed(31, 10 + k), sim_s=1.0,
                   r0=2.0) for k, s in enumerate((2, 3))]
    print(f"\n[tiny] batch of {len(specs)} tiny networks", flush=True)
    steps = 2000
    alone = {}
    for s in specs:
        x = SW.SweepNet(hip_ctx, s)
        x.net.advance(steps)
        x.net.sync()
        alone[s["id"]] = x.fingerprint()
        x.close()
    nets = SW.build(hip_ctx, specs)
    batch = B.Batch(hip_ctx)
    try:
        batch.set([x.net for x in nets])
        batch.advance(steps)
        batch.sync()
        diff = [x.spec["id"] for x in nets if x.fingerprint() != alone[x.spec["id"]]]
        errs = [x.spec["id"] for x in nets if x.net.error_bits()]
        print(f"[tiny] groups {[(g['mode'], g['networks']) for g in batch.shapes()]}", flush=True)
        assert not diff, diff
        assert not errs, errs
    finally:
        batch.close()
        for x in nets:
            x.close()

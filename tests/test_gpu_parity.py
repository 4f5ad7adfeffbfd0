"""GPU parity: the HIP path against the CPU oracle in deterministic mode, bit for bit.

Every config in tests/golden/configs runs on both sides with the Philox noise stream, the
project-owned transcendental functions and the wave64 summation order; after each epoch's steps,
after each growth update and after each history update every state array must be identical:
membrane state, refractory counters, spike histories and counts, synapse slots (source, type, W,
psr, delay ring contents and phase, DS r/u/lastSpike, STDP post rings), index maps, rates, radii.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _run(hip_ctx, name, block_steps=0, epochs=None):
    cfg = H.load_config(name)
    ora = O.OracleNet(cfg)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg, block_steps=block_steps)
    steps = H.steps_per_epoch(cfg)
    bad = H.compare_state(ora, net, cfg, "setup")
    n_ep = cfg["num_epochs"] if epochs is None else epochs
    spikes = 0
    for ep in range(1, n_ep + 1):
        ora.advance(steps)
        net.advance(steps)
        net.sync()
        spikes += int(ora.array("spikeCount").sum())
        bad += H.compare_state(ora, net, cfg, f"ep{ep}-advanced")
        ora.update_connections()
        if cfg["conn"] == "ConnGrowth":   # Model::updateConnections changes nothing for ConnStatic
            net.update_connections()
        bad += H.compare_state(ora, net, cfg, f"ep{ep}-grown")
        ora.update_history()
        net.update_history()
        net.sync()
        bad += H.compare_state(ora, net, cfg, f"ep{ep}-history", check_maps=False)
        if bad:
            break
    net.close()
    return bad, spikes, ora.total_synapses


@pytest.mark.parametrize("name", H.GROWTH_CONFIGS)
def test_growth_configs_bit_exact(hip_ctx, name):
    print(f"[parity] {name}", flush=True)
    bad, spikes, total = _run(hip_ctx, name)
    assert not bad, bad[:10]
    assert spikes > 0
    assert total > 0


@pytest.mark.parametrize("name", H.STATIC_CONFIGS)
def test_static_configs_bit_exact(hip_ctx, name):
    """ConnStatic networks (nearest-k, Spiking and DS synapses, up to all-to-all): the oracle's
    synapses uploaded, every epoch bit for bit."""
    print(f"[parity] {name}", flush=True)
    bad, spikes, total = _run(hip_ctx, name)
    assert not bad, bad[:10]
    assert spikes > 0 and total > 0


@pytest.mark.parametrize("block_steps", [1, 2, 5])
def test_block_size_does_not_change_results(hip_ctx, block_steps):
    print(f"[parity] conn-lif-ds-dense block_steps={block_steps}", flush=True)
    bad, spikes, total = _run(hip_ctx, "conn-lif-ds-dense", block_steps=block_steps)
    assert not bad, bad[:10]


@pytest.mark.parametrize("M", [30, 8200])
def test_dense_weight_update_with_erasures_matches_oracle(hip_ctx, M):
    """bghip_update_weights (ConnGrowth::updateSynapsesWeights with a caller's W matrix): existing
    synapses erased (their slots reused by later sources' new synapses), reweighted or kept at zero
    weight, new ones created; slot state and index maps equal the oracle's, and so does the network
    over the steps that follow and a second update. Rows of 30 slots go through the workgroup path
    for rows with holes or erasures, rows of 8,200 slots (beyond its LDS slot bitmaps) through the
    one-lane replay."""
    print(f"[parity] weight update with erasures, M={M}", flush=True)
    cfg = O.default_config()
    cfg.update(width=6, height=6, max_synapses_per_neuron=M, layout="DynamicLayout", frac_exc=0.8,
               frac_starter=0.1, noise_mode=1, math_mode=1, sum_mode=1, epoch_duration=0.05, num_epochs=2)
    ora = O.OracleNet(cfg)
    ora.set_radii(0.75)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    ora.update_connections()
    net.update_connections()
    N = 36
    rng = np.random.default_rng(11)
    bad = []
    for rnd in range(2):
        src = ora.array("sourceNeuronIndex")
        existing = np.zeros((N, N), bool)
        for k in np.flatnonzero(ora.array("in_use")):
            existing[src[k], k // M] = True
        Wm = rng.uniform(-1.0, 1.0, (N, N)).astype(np.float32)
        Wm[rng.random((N, N)) < 0.75] = 0.0
        Wm[existing] = rng.choice(np.array([-0.5, 0.0, 0.25, 3.0], np.float32), existing.sum())
        ora.update_weights(Wm)
        st = net.update_weights(Wm)
        print(f"[weights] round {rnd}: {st.created} created, {st.erased} erased, {st.total_synapses} synapses",
              flush=True)
        assert st.erased == int(((Wm < 0) & existing).sum()) > 0
        assert st.created == int(((Wm > 0) & ~existing).sum()) > 0
        assert st.total_synapses == ora.total_synapses
        bad += H.compare_state(ora, net, cfg, f"weights-{rnd}")
        ora.advance(300)
        net.advance(300)
        net.sync()
        bad += H.compare_state(ora, net, cfg, f"weights-{rnd}-advanced")
    net.close()
    assert not bad, bad[:10]


def test_spike_download_and_clear_match_oracle(hip_ctx):
    """bghip_download_spikes returns the oracle's spike counts, offsets and history rings, and
    bghip_clear_spike_counts rotates the offsets and zeroes the counts as clearSpikeCounts does
    (the oracle clears inside its history update); the network then runs on bit for bit."""
    cfg = H.load_config("static-lif-ds")
    ora = O.OracleNet(cfg)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    bad = []
    for rnd in range(3):
        ora.advance(1700)
        net.advance(1700)
        sp = net.download_spikes()
        for k in ("spikeCount", "spikeCountOffset", "spike_history"):
            if not np.array_equal(sp[k], ora.array(k).reshape(-1).astype(sp[k].dtype)):
                bad.append(f"round {rnd}: {k}")
        assert int(sp["spikeCount"].sum()) > 0
        ora.update_history()
        net.clear_spike_counts()
        sp = net.download_spikes()
        assert int(sp["spikeCount"].sum()) == 0
        if not np.array_equal(sp["spikeCountOffset"], ora.array("spikeCountOffset")):
            bad.append(f"round {rnd}: offsets after clear")
        bad += H.compare_state(ora, net, cfg, f"cleared-{rnd}")
    net.close()
    assert not bad, bad[:10]


FULL_SIZE = {
    # side, M, radius, seed, synapses after the first growth update, minimum spikes in 270 steps
    "10k-benchmark": (100, 1000, 8.9, 1, 8440432, 50000),
    "40k-largest-served": (200, 100, 2.6, 4, 3441176, 200000),
}


@pytest.mark.parametrize("case", list(FULL_SIZE))
def test_benchmark_network_bit_exact_at_full_size(hip_ctx, case):
    """Full sizes: the measured workload (100 x 100 LIF + DS growth network, test-small
    distributions, DynamicLayout 0.9 / 0.1, M = 1000, max_firing_rate 1000, every radius set to 8.9:
    8,440,432 synapses) and the largest network of the served workload (200 x 200, 40,000 neurons,
    M = 100, radius 2.6: 3,441,176 synapses). One growth update on each side creates the synapses on
    the device and in the oracle; membranes start within 10% of the reset-to-threshold range below
    threshold (seeded) so that nearly every neuron fires; 270 steps (thirty 9-step fused launches,
    80,000 / 255,000 spikes); then a second growth update over the existing synapses. Every state
    array equals the oracle's after each stage."""
    side, M, radius, seed, want, min_spikes = FULL_SIZE[case]
    cfg = O.default_config()
    cfg.update(width=side, height=side, max_synapses_per_neuron=M, max_firing_rate=1000, num_epochs=1,
               epoch_duration=10.0, seed=seed, layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1,
               noise_mode=1, math_mode=1, sum_mode=1)
    ora = O.OracleNet(cfg)
    ora.set_radii(radius)
    vm, vth, vr = ora.view("Vm"), ora.array("Vthresh"), ora.array("Vreset")
    u = np.random.default_rng(5).random(vm.size).astype(np.float32)
    vm[:] = (vth - u * np.float32(0.1) * (vth - vr)).astype(np.float32)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    try:
        print(f"[parity] {case}: growth update", flush=True)
        ora.update_connections()
        st = net.update_connections()
        assert st.total_synapses == ora.total_synapses == want
        bad = H.compare_state(ora, net, cfg, "grown")
        print(f"[parity] {case}: 270 steps", flush=True)
        ora.advance(270)
        net.advance(270)
        net.sync()
        spikes = int(ora.array("spikeCount").sum())
        bad += H.compare_state(ora, net, cfg, "advanced")
        ora.update_connections()
        net.update_connections()
        bad += H.compare_state(ora, net, cfg, "grown-again")
        print(f"[parity] {case}: {spikes} spikes, {len(bad)} differences", flush=True)
        assert spikes > min_spikes
        assert not bad, bad[:10]
    finally:
        net.close()
        ora.close()


def test_weight_update_with_erasures_at_full_size(hip_ctx):
    """The 10,000-neuron benchmark network (8,440,432 synapses after the first growth update) given
    a caller's matrix that erases about 5% of the existing synapses, re-weights the rest and creates
    about 10,000 new ones in the freed and the free slots; then 270 steps over the holed rows and a
    growth update (overlap areas) over them. Every state array equals the oracle's after each stage."""
    side, M, radius, seed = 100, 1000, 8.9, 1
    N = side * side
    cfg = O.default_config()
    cfg.update(width=side, height=side, max_synapses_per_neuron=M, max_firing_rate=1000, num_epochs=1,
               epoch_duration=10.0, seed=seed, layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1,
               noise_mode=1, math_mode=1, sum_mode=1)
    ora = O.OracleNet(cfg)
    ora.set_radii(radius)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    try:
        print("[parity] 10k erasures: growth update", flush=True)
        ora.update_connections()
        net.update_connections()
        use = ora.array("in_use").astype(bool)
        slots = np.flatnonzero(use)
        src = ora.array("sourceNeuronIndex")[slots].astype(np.int64)
        dst = slots // M
        rng = np.random.default_rng(17)
        Wm = np.zeros((N, N), np.float32)
        new = rng.random((N, N), dtype=np.float32) < np.float32(1e-4)
        Wm[new] = rng.uniform(0.5, 2.0, int(new.sum())).astype(np.float32)
        Wm[src, dst] = rng.uniform(0.1, 3.0, slots.size).astype(np.float32)
        gone = rng.random(slots.size) < 0.05
        Wm[src[gone], dst[gone]] = -1.0
        existing = np.zeros((N, N), bool)
        existing[src, dst] = True
        print("[parity] 10k erasures: weight update", flush=True)
        ora.update_weights(Wm)
        st = net.update_weights(Wm)
        print(f"[parity] 10k erasures: {st.created} created, {st.erased} erased, {st.total_synapses} synapses",
              flush=True)
        assert st.erased == int(gone.sum()) > 0
        assert st.created == int(((Wm > 0) & ~existing).sum()) > 0
        assert st.total_synapses == ora.total_synapses
        del Wm, new, existing
        bad = H.compare_state(ora, net, cfg, "weights")
        print("[parity] 10k erasures: 270 steps", flush=True)
        ora.advance(270)
        net.advance(270)
        net.sync()
        bad += H.compare_state(ora, net, cfg, "advanced")
        ora.update_connections()
        net.update_connections()
        bad += H.compare_state(ora, net, cfg, "grown-again")
        print(f"[parity] 10k erasures: {len(bad)} differences", flush=True)
        assert not bad, bad[:10]
    finally:
        net.close()
        ora.close()

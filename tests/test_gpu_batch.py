"""Batched form: networks advanced together by the multi-network kernels give exactly the results
each gives alone (here: exactly the oracle's), whatever their sizes, models and current steps."""
import numpy as np
import pytest

from braingrid_amd import bghip as B
from braingrid_amd import sweep as SW
from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

MIX = ["conn-lif-ds", "conn-lif-ds-dense", "conn-izh-ds", "conn-lif-stdp", "conn-izh-stdp", "conn-lif-dynstdp",
       "conn-lif-ds-dynamic", "conn-lif-spiking"]


def test_batch_results_equal_alone(hip_ctx):
    print("[batch] mixed batch of 8 networks", flush=True)
    cfgs = [H.load_config(n) for n in MIX]
    oras = [O.OracleNet(c) for c in cfgs]
    nets = [H.gpu_net_from_oracle(hip_ctx, o, c) for o, c in zip(oras, cfgs)]
    # desynchronise the members: network k first advances k*7 steps alone
    for k, (o, n) in enumerate(zip(oras, nets)):
        o.advance(7 * k)
        n.advance(7 * k)
    batch = B.Batch(hip_ctx)
    batch.set(nets)
    bad = []
    for rnd in range(3):
        # one growth epoch worth of steps for everyone, in uneven chunks
        chunk = [1000, 2345, 1655][rnd]
        for o in oras:
            o.advance(chunk)
        batch.advance(chunk)
        for k, (o, n, c) in enumerate(zip(oras, nets, cfgs)):
            n.sync()
            bad += H.compare_state(o, n, c, f"{MIX[k]}-round{rnd}", check_maps=False)
        # a growth update on every member, then the next round
        for o, n in zip(oras, nets):
            o.update_connections()
            n.update_connections()
            o.update_history()
            n.update_history()
    batch.close()
    for n in nets:
        n.close()
    assert not bad, bad[:10]


SHAPES = [
    dict(block=64, dests_per_wave=1, block_stream=64, syn_blocks_per_shard=1, sum_blocks_per_shard=1),
    dict(block=128, dests_per_wave=2, block_stream=128, syn_blocks_per_shard=16, sum_blocks_per_shard=32),
    dict(block=256, dests_per_wave=4, block_stream=64, syn_blocks_per_shard=3, sum_blocks_per_shard=5),
]


def _grown_mix(hip_ctx):
    """The mixed networks after one growth update (so the fused and per-step groups exist)."""
    cfgs = [H.load_config(n) for n in MIX]
    oras = [O.OracleNet(c) for c in cfgs]
    nets = [H.gpu_net_from_oracle(hip_ctx, o, c) for o, c in zip(oras, cfgs)]
    for o, n in zip(oras, nets):
        o.advance(400)
        n.advance(400)
        o.update_connections()
        n.update_connections()
    return cfgs, oras, nets


@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_batch_launch_shapes_give_identical_results(hip_ctx, shape):
    """Every launch shape (block sizes, destinations per streaming wavefront, event-pass grids)
    gives the oracle's results bit for bit, through direct launches and captured graph replays."""
    print(f"[batch] launch shape {SHAPES[shape]}", flush=True)
    cfgs, oras, nets = _grown_mix(hip_ctx)
    batch = B.Batch(hip_ctx)
    batch.set(nets)
    groups = batch.shapes()
    assert {g["mode"] for g in groups} >= {1, 2}, groups
    for i in range(len(groups)):
        batch.set_shape(i, **SHAPES[shape])
    bad = []
    for rnd, chunk in enumerate((300, 611)):   # 300: one graph replay + 44 direct steps
        for o in oras:
            o.advance(chunk)
        batch.advance(chunk)
        for k, (o, n, c) in enumerate(zip(oras, nets, cfgs)):
            n.sync()
            bad += H.compare_state(o, n, c, f"{MIX[k]}-shape{shape}-round{rnd}", check_maps=False)
    for g in batch.shapes():
        assert g["block"] == SHAPES[shape]["block"], g
    batch.close()
    for n in nets:
        n.close()
    assert not bad, bad[:10]


def test_batch_shape_is_validated(hip_ctx):
    cfgs, oras, nets = _grown_mix(hip_ctx)
    batch = B.Batch(hip_ctx)
    batch.set(nets[:2])
    for bad in (dict(block=96), dict(dests_per_wave=3), dict(block_stream=512), dict(syn_blocks_per_shard=0),
                dict(sum_blocks_per_shard=65)):
        with pytest.raises(B.BgHipError):
            batch.set_shape(0, **bad)
    with pytest.raises(B.BgHipError):
        batch.set_shape(len(batch.shapes()), block=64)
    batch.close()
    for n in nets:
        n.close()


def test_autotuner_tunes_every_group_and_keeps_results(hip_ctx):
    """The autotuner's trials run on the members' own steps: afterwards every group has a measured
    shape, the results are still the oracle's, and a new batch of the same kinds and sizes starts
    from the remembered shapes."""
    print("[batch] autotuner", flush=True)
    cfgs = [H.load_config(n) for n in MIX]
    oras = [O.OracleNet(c) for c in cfgs]
    nets = [H.gpu_net_from_oracle(hip_ctx, o, c) for o, c in zip(oras, cfgs)]
    batch = B.Batch(hip_ctx)
    batch.autotune(4)
    batch.set(nets)
    bad = []
    modes = set()
    for rnd, chunk in enumerate((1000, 1500)):   # first chunk: no synapses yet (neuron-only groups)
        for o in oras:
            o.advance(chunk)
        batch.advance(chunk)
        for k, (o, n, c) in enumerate(zip(oras, nets, cfgs)):
            n.sync()
            bad += H.compare_state(o, n, c, f"{MIX[k]}-tuned-round{rnd}", check_maps=False)
        for g in batch.shapes():
            modes.add(g["mode"])
            assert g["tuned"] == 1 and g["step_us"] > 0, g
        for o, n in zip(oras, nets):
            o.update_connections()
            n.update_connections()
    assert modes == {0, 1, 2}, modes
    remembered = {(g["mode"], g["neuron_model"], g["synapse_model"], g["ring_words"]): g for g in batch.shapes()}
    batch.close()
    again = B.Batch(hip_ctx)
    again.set(nets)
    for g in again.shapes():
        key = (g["mode"], g["neuron_model"], g["synapse_model"], g["ring_words"])
        if key in remembered and remembered[key]["work_items"] == g["work_items"]:
            assert g["tuned"] == 1, g
            for f in ("block", "dests_per_wave", "block_stream", "syn_blocks_per_shard", "sum_blocks_per_shard"):
                assert g[f] == remembered[key][f], (f, g, remembered[key])
    again.close()
    for n in nets:
        n.close()
    assert not bad, bad[:10]


def test_batch_membership_changes(hip_ctx):
    cfgs = [H.load_config(n) for n in ("conn-lif-ds", "conn-izh-stdp", "conn-lif-ds-dense")]
    oras = [O.OracleNet(c) for c in cfgs]
    nets = [H.gpu_net_from_oracle(hip_ctx, o, c) for o, c in zip(oras, cfgs)]
    batch = B.Batch(hip_ctx)
    batch.set(nets[:2])
    batch.advance(500)
    for o in oras[:2]:
        o.advance(500)
    batch.set(nets[1:])
    batch.advance(300)
    for o in oras[1:]:
        o.advance(300)
    batch.set([nets[0]])
    batch.advance(200)
    oras[0].advance(200)
    bad = []
    for o, n, c in zip(oras, nets, cfgs):
        n.sync()
        bad += H.compare_state(o, n, c, c["synapse"], check_maps=False)
    batch.close()
    for n in nets:
        n.close()
    assert not bad, bad[:10]


def test_stdp_rows_growing_after_the_group_formed(hip_ctx):
    """Weight updates that add sources to existing destinations (bghip_update_weights) between batch
    advances lengthen rows without changing any member's mode: the per-step STDP group, its tables
    and its captured graph stay, and every member still equals the oracle."""
    names = ["conn-lif-stdp", "conn-izh-stdp", "conn-lif-dynstdp"]
    rng = np.random.default_rng(3)
    cfgs = [H.load_config(n) for n in names]
    oras = [O.OracleNet(c) for c in cfgs]
    nets = [H.gpu_net_from_oracle(hip_ctx, o, c) for o, c in zip(oras, cfgs)]
    for o, n in zip(oras, nets):
        o.advance(400)
        n.advance(400)
        o.update_connections()
        n.update_connections()
    batch = B.Batch(hip_ctx)
    batch.set(nets)
    bad = []
    grew = 0
    for rnd in range(3):
        for o in oras:
            o.advance(700)
        batch.advance(700)
        assert all(g["mode"] == 2 for g in batch.shapes()), batch.shapes()
        for k, (o, n, c) in enumerate(zip(oras, nets, cfgs)):
            n.sync()
            bad += H.compare_state(o, n, c, f"{names[k]}-grow-round{rnd}", check_maps=False)
        before = [o.total_synapses for o in oras]
        for o, n, c in zip(oras, nets, cfgs):
            N = c["width"] * c["height"]
            Wm = np.where(rng.random((N, N)) < 0.25, 0.5, 0.0).astype(np.float32)
            np.fill_diagonal(Wm, 0.0)
            M = c["max_synapses_per_neuron"]
            src = o.array("sourceNeuronIndex")
            for k in np.flatnonzero(o.array("in_use")):
                Wm[src[k], k // M] = 0.75    # existing synapses stay (reweighted)
            o.update_weights(Wm)
            n.update_weights(Wm)
        grew += sum(o.total_synapses > b for o, b in zip(oras, before))
    print(f"[batch] {grew} growth updates lengthened rows inside one group", flush=True)
    assert grew >= 3
    batch.close()
    for n in nets:
        n.close()
    assert not bad, bad[:10]


def test_batch_follows_a_member_whose_blocks_shorten(hip_ctx):
    """A fused (Spiking/DS) member gains a synapse with a 3-step delay between batch advances
    (bghip_add_synapses_ex); its block length drops from 9 to 3 steps and the batch re-forms its
    groups, so no spike is delivered inside the block it was fired in: the member equals the
    oracle, which gets the same synapse through addSynapse and the same total delay."""
    cfg = H.load_config("conn-izh-ds")   # active: every neuron fires several times in 0.5 s
    ora = O.OracleNet(cfg)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    other_cfg = H.load_config("conn-lif-ds")
    other_ora = O.OracleNet(other_cfg)
    other = H.gpu_net_from_oracle(hip_ctx, other_ora, other_cfg)
    batch = B.Batch(hip_ctx)
    try:
        for o, n in ((ora, net), (other_ora, other)):
            o.advance(400)
            n.advance(400)
            o.update_connections()
            n.update_connections()
        batch.set([net, other])
        batch.advance(500)
        for o in (ora, other_ora):
            o.advance(500)
        assert any(g["mode"] == 1 for g in batch.shapes())
        M = cfg["max_synapses_per_neuron"]
        src = ora.array("sourceNeuronIndex")
        have = {int(src[k]) for k in range(M) if ora.array("in_use")[k]}   # sources of neuron 0
        fired = ora.array("spikeCount")
        s = max((i for i in range(1, cfg["width"] * cfg["height"]) if i not in have), key=lambda i: fired[i])
        assert fired[s] > 0
        w = np.float32(2.0e-7)
        net.add_synapses(np.array([s]), np.array([0]), np.array([w]), total_delay=np.array([3]))
        slot = ora.add_synapse(s, 0, float(w))
        ora.view("total_delay")[slot] = 3
        ora.rebuild_index_map()
        bad = []
        for rnd, chunk in enumerate((600, 1300)):
            for o in (ora, other_ora):
                o.advance(chunk)
            batch.advance(chunk)
            net.sync()
            other.sync()
            bad += H.compare_state(ora, net, cfg, f"short-delay-{rnd}", check_maps=False)
            bad += H.compare_state(other_ora, other, other_cfg, f"other-{rnd}", check_maps=False)
        assert ora.array("spikeCount").sum() > 0
        assert not bad, bad[:10]
    finally:
        batch.close()
        net.close()
        other.close()


def test_destroyed_member_leaves_the_batch(hip_ctx):
    """A network destroyed while it is a batch member leaves the batch: later batch calls advance
    the remaining members (exactly as the oracle) and never touch the destroyed one, also when a
    new network is created in its place before the next call."""
    print("[batch] member destroyed between batch calls", flush=True)
    cfgs, oras, nets = _grown_mix(hip_ctx)
    batch = B.Batch(hip_ctx)
    batch.set(nets)
    for o in oras:
        o.advance(300)
    batch.advance(300)
    gone = [1, 4]
    for k in gone:
        nets[k].close()
    keep = [k for k in range(len(nets)) if k not in gone]
    # a new network of another size takes the freed memory before the batch runs again
    cnew = H.load_config("conn-lif-ds")
    onew = O.OracleNet(cnew)
    nnew = H.gpu_net_from_oracle(hip_ctx, onew, cnew)
    assert len(batch.shapes()) >= 1
    for k in keep:
        oras[k].advance(500)
    batch.advance(500)
    bad = []
    for k in keep:
        nets[k].sync()
        bad += H.compare_state(oras[k], nets[k], cfgs[k], f"{MIX[k]}-after-destroy", check_maps=False)
    onew.advance(20)
    nnew.advance(20)
    nnew.sync()
    bad += H.compare_state(onew, nnew, cnew, "new-net", check_maps=False)
    batch.close()
    for k in keep:
        nets[k].close()
    nnew.close()
    assert not bad, bad[:10]


def test_more_stdp_members_than_event_shards(hip_ctx):
    """1,040 Izhikevich + DynamicSTDP + Poisson networks of 36 neurons in one batch: more members
    than the batch's 1,024 event-list shards, so the members' runs go round all shards and a shard
    lists several networks' events (with fewer members each has shards of its own). After 2,000
    batched steps every sampled network equals its alone run bit for bit (full-state digest), no
    member has a device error, and the networks spiked (so the event lists were used)."""
    n, steps = 1040, 2000
    specs = [dict(id=k, side=6, kind="izh_stdp", tR=0.5, fE=0.9, seed=SW.net_seed(77, k), sim_s=1.0)
             for k in range(n)]
    sample = list(range(0, n, 97)) + [n - 1]
    print(f"\n[batch] {n} STDP members, {steps} steps; alone sample {sample}", flush=True)
    alone = {}
    for k in sample:
        x = SW.SweepNet(hip_ctx, specs[k])
        x.net.advance(steps)
        x.net.sync()
        alone[k] = x.fingerprint()
        x.close()
    nets = SW.build(hip_ctx, specs)
    batch = B.Batch(hip_ctx)
    try:
        batch.set([x.net for x in nets])
        batch.advance(steps)
        batch.sync()
        groups = batch.shapes()
        print(f"[batch] groups {[(g['mode'], g['networks']) for g in groups]}", flush=True)
        assert any(g["networks"] == n for g in groups)
        errs = [x.spec["id"] for x in nets if x.net.error_bits()]
        assert not errs, errs[:10]
        spikes = sum(int(x.net.download_spikes()["spikeCount"].sum()) for x in nets[::13])
        print(f"[batch] spikes in every 13th network: {spikes}", flush=True)
        assert spikes > 0
        diff = [k for k in sample if nets[k].fingerprint() != alone[k]]
        assert not diff, diff
    finally:
        batch.close()
        for x in nets:
            x.close()

"""Batched form: networks advanced together by the multi-network kernels give exactly the results
each gives alone (here: exactly the oracle's), whatever their sizes, models and current steps."""
import numpy as np
import pytest

from braingrid_amd import bghip as B
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

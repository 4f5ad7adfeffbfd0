"""Checkpoints (bghip_net_save / bghip_net_load): a network saved mid-run and loaded into a new
network continues bit for bit like the original, for the fused DS path with growth and regular
input, the per-step DynamicSTDP path with Poisson input, and a fused block left pending; damaged
files are refused."""
import os

import numpy as np
import pytest

from braingrid_amd import bghip as B
from braingrid_amd import networks as NW
from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _same(a: B.Net, b: B.Net, max_spikes: int) -> list[str]:
    bad = []
    na, nb = a.download_neurons(max_spikes), b.download_neurons(max_spikes)
    for k in na:
        if na[k].tobytes() != nb[k].tobytes():
            bad.append(k)
    sa, sb = a.download_synapses(), b.download_synapses()
    for k in sa:
        if sa[k].tobytes() != sb[k].tobytes():
            bad.append(k)
    ia, ib = a.download_input(), b.download_input()
    for k in ia:
        if np.asarray(ia[k]).tobytes() != np.asarray(ib[k]).tobytes():
            bad.append("input:" + k)
    if a.total_synapses != b.total_synapses:
        bad.append("total")
    return bad


@pytest.mark.parametrize("name", ["input-lif-regular-async", "input-lif-regular-wave"])
def test_checkpoint_round_trip_growth_config(hip_ctx, tmp_path, name):
    print(f"\n[checkpoint {name}]", flush=True)
    cfg = H.load_config(name)
    ora = O.OracleNet(cfg)
    a = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    steps = H.steps_per_epoch(cfg)
    ms = H.max_spikes(cfg)
    try:
        a.advance(steps)
        a.update_connections()
        a.update_history()
        a.advance(1237)
        path = os.path.join(tmp_path, "net.ckpt")
        a.save(path)
        b = B.Net.load(hip_ctx, path)
        try:
            assert b.step == a.step and not _same(a, b, ms)
            for n in (a, b):
                n.advance(steps - 1237)
                n.update_connections()
                n.update_history()
                n.advance(777)
            bad = _same(a, b, ms)
            assert not bad, bad
            ra, qa = a.download_growth()
            rb, qb = b.download_growth()
            assert ra.tobytes() == rb.tobytes() and qa.tobytes() == qb.tobytes()
        finally:
            b.close()
    finally:
        a.close()
        ora.close()


def test_checkpoint_round_trip_stdp_with_poisson(hip_ctx, tmp_path):
    print("\n[checkpoint izh dynstdp poisson]", flush=True)
    a, meta = NW.izh_stdp(hip_ctx, 300, 9, k=30)
    a.upload_input(B.INPUT_POISSON, neuron_type=meta["ntype"], fr_mean=10.0, weight=25.0)
    try:
        a.advance(1500)
        path = os.path.join(tmp_path, "izh.ckpt")
        a.save(path)
        b = B.Net.load(hip_ctx, path)
        try:
            for n in (a, b):
                n.advance(2000)
                n.update_history()
                n.advance(333)
            bad = _same(a, b, meta["max_spikes"])
            assert not bad, bad
            assert a.download_neurons(meta["max_spikes"])["spikeCount"].sum() > 0
        finally:
            b.close()
    finally:
        a.close()


def test_damaged_checkpoints_are_refused(hip_ctx, tmp_path):
    cfg = H.load_config("conn-lif-ds")
    ora = O.OracleNet(cfg)
    a = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    try:
        path = os.path.join(tmp_path, "a.ckpt")
        a.save(path)
        data = open(path, "rb").read()
        for k, bad in enumerate([data[:-1], data + b"x", b"XXXXXXXX" + data[8:], data[:100]]):
            p = os.path.join(tmp_path, f"bad{k}.ckpt")
            open(p, "wb").write(bad)
            with pytest.raises(RuntimeError, match="net_load"):
                B.Net.load(hip_ctx, p)
    finally:
        a.close()
        ora.close()

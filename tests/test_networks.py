"""Sweep-workload network builders (braingrid_amd/networks.py) against the oracle's own creation path.

The static Izhikevich + DynamicSTDP builder writes the slot arrays itself; these tests pin every
field it writes to what the oracle's addSynapse/createSynapse chain produces for the same
(source, destination) list (Synapses/AllSynapses.cpp:357-377, Synapses/AllDSSynapses.cpp:180-218,
Synapses/AllSTDPSynapses.cpp:252-272), and the neuron constants to the oracle's
initNeuronConstsFromParamValues (Neurons/AllIFNeurons.cpp:479-503, Neurons/AllIZHNeurons.cpp:474-475).
"""
import numpy as np
import pytest

from braingrid_amd import networks as NW
from tests.helpers import izh_static_oracle

SYN_FIELDS = [("in_use", "in_use"), ("sourceNeuronIndex", "sourceNeuronIndex"), ("type", "type"), ("W", "W"),
              ("psr", "psr"), ("decay", "decay"), ("total_delay", "total_delay"), ("delayIdx", "delayIdx"),
              ("r", "r"), ("u_syn", "u"), ("U", "U"), ("D", "D"), ("F", "F"), ("lastSpike", "lastSpike"),
              ("Apos", "Apos"), ("Aneg", "Aneg"), ("STDPgap", "STDPgap"), ("tauspost", "tauspost"),
              ("tauspre", "tauspre"), ("taupos", "taupos"), ("tauneg", "tauneg"), ("Wex", "Wex"),
              ("mupos", "mupos"), ("muneg", "muneg"), ("useFroemkeDanSTDP", "useFroemkeDanSTDP")]


@pytest.fixture(scope="module")
def small():
    neu, syn, meta = NW.izh_stdp_arrays(150, 11, k=12)
    ora, cfg = izh_static_oracle(neu, syn, meta, 11)
    yield neu, syn, meta, ora
    ora.close()


def test_izh_builder_slots_match_oracle_creation(small):
    neu, syn, meta, ora = small
    use = meta["slot"]
    for of, bf in SYN_FIELDS:
        a, b = ora.array(of)[use], syn[bf][use]
        assert a.tobytes() == b.astype(a.dtype).tobytes(), of
    assert np.array_equal(ora.array("synapse_counts"), syn["synapse_counts"])
    assert int(ora.array("in_use").sum()) == meta["S_active"]


def test_izh_builder_neuron_constants_match_oracle(small):
    neu, syn, meta, ora = small
    for f in ("C1", "C2", "C3", "I0"):
        assert ora.setup_consts[f].tobytes() == neu[f].tobytes(), f


def test_izh_builder_structure():
    N, k = 400, 100
    neu, syn, meta = NW.izh_stdp_arrays(N, 3, k=k)
    src, dst = meta["src"], meta["dst"]
    assert len(src) == N * k
    assert not np.any(src == dst)
    assert np.all(np.bincount(src, minlength=N) == k)
    pairs = src.astype(np.int64) * N + dst
    assert len(np.unique(pairs)) == len(pairs)
    # per destination, ascending source = the order addSynapse is called in
    order = np.lexsort((src, dst))
    assert np.array_equal(order, np.arange(len(src)))
    assert meta["M"] > np.bincount(dst, minlength=N).max()
    # delays U{1..20} ms -> int(d/dt) + 1 = 11..201 steps; all fit the 256-bit ring
    assert meta["total_delay"].min() >= 11 and meta["total_delay"].max() <= 201
    assert set(np.unique(meta["total_delay"])) <= {int(np.float32(np.float32(d * 1e-3) / np.float32(1e-4))) + 1
                                                   for d in range(1, 21)}
    exc = meta["ntype"] == 2
    assert abs(exc.mean() - 0.8) < 0.01
    w = meta["W"] / NW.W_UNIT
    assert np.all((w[exc[src]] >= 0) & (w[exc[src]] <= 6 + 1e-6))
    assert np.all((w[~exc[src]] <= 0) & (w[~exc[src]] >= -5 - 1e-6))
    # Izhikevich 2003 parameter families
    a, b, c, d = (neu[f].astype(np.float64) for f in ("Aconst", "Bconst", "Cconst", "Dconst"))
    assert np.all(a[exc] == np.float32(0.02)) and np.all(b[exc] == np.float32(0.2))
    assert np.all((c[exc] >= -65) & (c[exc] <= -50)) and np.all((d[exc] >= 2) & (d[exc] <= 8))
    r_c, r_d = np.sqrt((c[exc] + 65) / 15), np.sqrt((8 - d[exc]) / 6)
    assert np.allclose(r_c, r_d, atol=1e-3)   # one r per neuron drives both c and d
    assert np.all((a[~exc] >= 0.02) & (a[~exc] <= 0.1)) and np.all(c[~exc] == -65) and np.all(d[~exc] == 2)
    assert np.allclose((a[~exc] - 0.02) / 0.08, (0.25 - b[~exc]) / 0.05, atol=1e-4)

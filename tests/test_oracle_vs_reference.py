"""The CPU oracle against the compiled reference, bit for bit.

Pinning of the oracle (oracle/bg_oracle.c), function by function:
  - setup draws (MTRand, Fixed/DynamicLayout, createAllNeurons for LIF and IZH), LIF and IZH
    advance with the Norm noise stream, spike recording, delay-ring insert/test, Spiking / DS /
    STDP / DynamicSTDP synapse advance, sequential summation, growth (rates, radii, frontier,
    overlap, weights, creation in the first free slot), createSynapseImap and the recorder's
    clamp/clear: every array after setup, after every epoch's steps and after every growth update
    equals the reference's dump (tests/golden/ref_*.npz, written by tests/golden/make_ref_fixtures.py
    from oracle/_ref/bgref_dump) on ten configs;
  - stimulus input: SInputPoisson (interval counters drawn from the global MTRand stream, masks,
    the per-neuron DS input synapse and its delay ring) and SInputRegular (10x10 value tiling,
    asynchronous shifts drawn before the model's setup draws, wave shifts, the cycle window) on
    four configs with the reference's -s stimulus files;
  - history compilation and output format: the formatted recorder output equals the reference's
    own 4-2016 goldens (old/validation/4-2016/growth/test-{tiny,small}-out.xml) byte for byte;
  - when oracle/_ref is present the same comparison also runs live against the reference.
Functions exercised only in deterministic mode (Philox noise, own math, wave64 summation) are
pinned in tests/test_detmath.py and tests/test_oracle_properties.py.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle_py as O
from tests import xmlout
from tests.helpers import INPUT_STIM

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
REF_BASE = os.path.join(GOLDEN, "ref")
CFG_BASE = os.path.join(GOLDEN, "configs")

CONFIGS = {"test-tiny": (os.path.join(REF_BASE, "configfiles/test-tiny.xml"), REF_BASE)}
for _n in ("conn-lif-ds", "conn-lif-ds-dense", "conn-lif-spiking", "conn-lif-stdp", "conn-lif-dynstdp",
           "conn-izh-ds", "conn-izh-stdp", "conn-lif-ds-dynamic", "static-izh-spiking", "static-lif-ds",
           "static-izh-all") + tuple(INPUT_STIM):
    CONFIGS[_n] = (os.path.join(CFG_BASE, _n + ".xml"), CFG_BASE)


def _cfg(name):
    path, base = CONFIGS[name] if name in CONFIGS else (os.path.join(REF_BASE, f"configfiles/{name}.xml"), REF_BASE)
    cfg = O.parse_config_xml(path, base)
    if "dynamic" in name:
        cfg.update(layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1)
    if name in INPUT_STIM:
        cfg["input"] = O.parse_input_xml(os.path.join(base, INPUT_STIM[name]), base)
    return cfg


def _records(npz):
    z = np.load(npz)
    n = len([k for k in z.files if k.endswith("__marker")])
    out = []
    for i in range(n):
        tag = bytes(z[f"{i:03d}__tag"]).decode().strip()
        arrs = {k[4:]: z[k] for k in z.files if k.startswith(f"{i:03d}_") and not k.startswith(f"{i:03d}__")}
        out.append((tag, int(z[f"{i:03d}__marker"][0]), arrs))
    return out


# Per-slot synapse arrays: compared on in-use slots only (free slots hold stale values).
SLOT_ARRAYS = {"sourceNeuronIndex", "destNeuronIndex", "W", "psr", "type", "decay", "total_delay", "delayQueue",
               "delayIdx", "r", "u_syn", "lastSpike", "delayQueuePost", "delayIdxPost"}


def _compare(net, arrs, where):
    bad = []
    use = arrs.get("in_use")
    m = use.astype(bool) if use is not None else None
    for k, v in arrs.items():
        try:
            mine = net.array(k)
        except KeyError:
            continue
        if m is not None and k in SLOT_ARRAYS:
            a, b = v[m], mine[m]
        else:
            a, b = v, mine[:v.size]
        if a.dtype.kind == "f":
            eq = np.array_equal(a.view(np.uint32), b.view(np.uint32))
        else:
            eq = np.array_equal(a, b)
        if not eq:
            bad.append(f"{where}:{k}")
    return bad


def _replay(name, recs):
    cfg = _cfg(name)
    net = O.OracleNet(cfg)
    steps = int(np.float32(np.float32(cfg["epoch_duration"]) / np.float32(cfg["dt"])))
    bad = []
    for tag, ep, arrs in recs:
        if tag == "advanced":
            net.advance(steps)
        elif tag == "grown":
            net.update_connections()
            net.update_history()
        bad += _compare(net, arrs, f"{tag}{ep}")
    return net, bad


@pytest.mark.parametrize("name", list(CONFIGS))
def test_oracle_matches_reference_fixtures(name):
    recs = _records(os.path.join(GOLDEN, f"ref_{name}.npz"))
    net, bad = _replay(name, recs)
    assert not bad, bad[:10]
    if name != "test-tiny" and not name.startswith("input-") or name.startswith("input-lif-regular"):
        assert net.total_synapses > 0


def test_oracle_matches_reference_test_small_fixture():
    recs = _records(os.path.join(GOLDEN, "ref_test-small.npz"))
    net, bad = _replay("test-small", recs)
    assert not bad, bad[:10]


@pytest.mark.parametrize("name", ["test-tiny", "test-small"])
def test_oracle_output_equals_2016_golden_xml(name):
    cfg = _cfg(name)
    net = O.OracleNet(cfg)
    for _ in range(cfg["num_epochs"]):
        net.run_epoch()
    mine = xmlout.oracle_growth_xml(net, cfg)
    with open(os.path.join(REF_BASE, f"{name}-out.xml")) as f:
        assert mine == f.read()


@pytest.mark.skipif(not O.ref_available() or not os.path.isdir("/root/reference"),
                    reason="compiled reference not present (built only where the reference sources are)")
@pytest.mark.parametrize("name", ["conn-lif-ds", "conn-izh-stdp", "conn-lif-ds-dynamic", "static-lif-ds"])
def test_oracle_matches_live_reference(name):
    path, base = CONFIGS[name]
    with tempfile.TemporaryDirectory() as td:
        dump = os.path.join(td, "d.bin")
        extra = ["--dynamic-layout", "0.9", "0.1"] if "dynamic" in name else []
        subprocess.run([os.path.join(O.REF_DIR, "bgref_dump"), path, dump] + extra, cwd=base, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        recs = [r for r in O.read_dump(dump) if r[0] != "header"]
        header = [r for r in O.read_dump(dump) if r[0] == "header"]
    net, bad = _replay(name, recs)
    assert not bad, bad[:10]
    if header:
        fresh = O.OracleNet(_cfg(name))
        assert not _compare(fresh, header[0][2], "header")


@pytest.mark.parametrize("name", ["conn-lif-ds", "input-lif-poisson"])
def test_fixture_comparison_is_not_vacuous(name):
    """Every fixture array is matched to an oracle array, and a one-ulp change in a neuron array,
    a synapse slot and an input counter is reported."""
    recs = _records(os.path.join(GOLDEN, f"ref_{name}.npz"))
    net, bad = _replay(name, recs[:4])   # header, setup, first epoch, first growth update
    assert not bad
    arrs = recs[3][2]
    for k in arrs:
        net.array(k)   # raises KeyError for a fixture array the comparison would skip
    assert len(arrs) >= 20
    plants = {"Vm": lambda a: np.nextafter(a, np.float32(1), dtype=np.float32)}
    if name.startswith("input"):
        plants["in_nISIs"] = lambda a: a + 1
    else:
        plants["sourceNeuronIndex"] = lambda a: a + 1
        assert arrs["in_use"].any()
    for k, f in plants.items():
        assert _compare(net, {**arrs, k: f(arrs[k])}, "plant") == [f"plant:{k}"], k


@pytest.mark.parametrize("name,k,thresh", [("static-izh-spiking", 10, 2.5), ("static-lif-ds", 14, 3.0)])
def test_static_fixtures_pin_the_tie_order(name, k, thresh):
    """The ConnStatic fixtures cut through groups of equal distance, so their synapse sets depend on
    the order std::sort leaves equal distances in: a selection that breaks ties by neuron index
    (a stable sort) connects a different set, while the distances each source reaches are the same."""
    recs = _records(os.path.join(GOLDEN, f"ref_{name}.npz"))
    setup = [a for tag, _, a in recs if tag == "setup"][0]
    use = setup["in_use"].astype(bool)
    cfg = _cfg(name)
    W, N = cfg["width"], cfg["width"] * cfg["height"]
    M = use.size // N
    ref = set(zip(setup["sourceNeuronIndex"][use].tolist(), np.repeat(np.arange(N), M)[use].tolist()))
    x = (np.arange(N) % W).astype(np.float32)
    y = (np.arange(N) // W).astype(np.float32)
    stable = set()
    for s in range(N):
        dx, dy = x[s] - x, y[s] - y
        d = np.sqrt(dx * dx + dy * dy, dtype=np.float32)
        cand = [j for j in np.argsort(d, kind="stable") if j != s and d[j] <= thresh][:k]
        stable |= {(s, int(j)) for j in cand}
        reached = sorted(d[j] for (a, j) in ref if a == s)
        assert reached == sorted(d[cand].tolist())
    assert len(ref) == len(stable) and len(ref ^ stable) > 50


def test_rate_fixture_is_the_oracles_reference_mode():
    """tests/golden/ref_rates_10k.npz (the compiled reference's per-neuron spike counts on the
    100 x 100 growth network, eight noise seeds) equals the oracle in reference mode, spike count
    for spike count, on the first seed's first epoch."""
    fx = np.load(os.path.join(GOLDEN, "ref_rates_10k.npz"))
    counts, seeds = fx["counts"], fx["seeds"]
    assert counts.shape == (8, 2, 10000)
    c = O.default_config()
    c.update(width=100, height=100, max_synapses_per_neuron=200, max_firing_rate=200, num_epochs=2,
             epoch_duration=10.0, seed=int(seeds[0]), layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1)
    ora = O.OracleNet(c)
    ora.run_epoch()
    got = np.rint(ora.array("ratesHistory")[10000:20000].astype(np.float64) * 10.0)
    # 1 s bins 0..8 are complete after the first epoch (bin 9 also receives step 100,000's spikes:
    # 100000 * dt, dt = float 1e-4 widened to double, is just below 10)
    burst = ora.array("burstinessHist")[:9].astype(np.int64)
    assert ora.total_synapses == 0
    ora.close()
    assert np.array_equal(got, counts[0, 0])
    assert np.array_equal(burst, fx["burst"][0, :9])

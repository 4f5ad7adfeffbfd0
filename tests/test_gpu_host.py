"""GPU tests of the host model (HipSpikingModel through include/bghost.h) and of growth at scale."""
import os
import tempfile

import numpy as np
import pytest

from braingrid_amd import host
from oracle import oracle_py as O
from tests import helpers as H
from tests import xmlout

pytestmark = pytest.mark.gpu

SETUP_ARRAYS = ("xloc", "yloc", "neuron_type_map", "starter_map", "Vthresh", "Vreset", "Vinit", "Iinject", "Inoise",
                "I0", "C1", "C2", "Trefract")
IZH_ARRAYS = ("Aconst", "Bconst", "Cconst", "Dconst", "C3")


def _host_cfg(name):
    c = host.read_config(os.path.join(H.CONFIG_DIR, name + ".xml"), H.CONFIG_DIR)
    if "dynamic" in name:
        c.update(layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1)
    return c


@pytest.mark.parametrize("name", ["conn-lif-ds", "conn-izh-stdp", "conn-lif-ds-dynamic"])
def test_host_setup_matches_oracle(hip_ctx, name):
    print(f"[host] setup {name}", flush=True)
    m = host.Model(hip_ctx, _host_cfg(name))
    ora = O.OracleNet(H.load_config(name))
    names = SETUP_ARRAYS + (IZH_ARRAYS if "izh" in name else ())
    bad = [k for k in names if not np.array_equal(m.array(k).view(np.uint8), ora.array(k).view(np.uint8))]
    m.close()
    assert not bad


def test_host_setup_matches_oracle_at_10k(hip_ctx):
    c = host.base_config()
    c.update(width=100, height=100, max_synapses_per_neuron=4, num_epochs=1, epoch_duration=0.01)
    m = host.Model(hip_ctx, c)
    oc = O.default_config()
    oc.update(width=100, height=100, max_synapses_per_neuron=4, num_epochs=1, epoch_duration=0.01,
              layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1)
    ora = O.OracleNet(oc)
    bad = [k for k in SETUP_ARRAYS if not np.array_equal(m.array(k).view(np.uint8), ora.array(k).view(np.uint8))]
    assert int((m.array("neuron_type_map") == 1).sum()) == 1000 and int(m.array("starter_map").sum()) == 1000
    m.close()
    assert not bad


@pytest.mark.parametrize("name", ["conn-lif-ds", "conn-izh-ds"])
def test_host_model_run_and_xml_output_match_oracle(hip_ctx, name):
    print(f"[host] run {name}", flush=True)
    cfg = _host_cfg(name)
    m = host.Model(hip_ctx, cfg)
    ocfg = H.load_config(name)
    ora = O.OracleNet(ocfg)
    for _ in range(cfg["num_epochs"]):
        m.run_epoch()
        ora.run_epoch()
    for k in ("radiiHistory", "ratesHistory", "burstinessHist", "spikesHistory"):
        assert np.array_equal(m.array(k).view(np.uint32), ora.array(k).view(np.uint32)), k
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "out.xml")
        m.save_xml(p)
        got = open(p).read()
    m.close()
    assert got == xmlout.oracle_growth_xml(ora, ocfg)


def test_pregrown_growth_update_matches_oracle(hip_ctx):
    """A 30x30 grid pre-grown to radius 3.1 (in-degree up to ~120), then two short epochs."""
    cfg = O.default_config()
    cfg.update(width=30, height=30, max_synapses_per_neuron=160, epoch_duration=0.05, num_epochs=2,
               layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1, max_firing_rate=1000,
               noise_mode=1, math_mode=1, sum_mode=1)
    ora = O.OracleNet(cfg)
    ora.set_radii(3.1)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    ora.update_connections()
    st = net.update_connections()
    assert st.total_synapses == ora.total_synapses > 900 * 80
    bad = H.compare_state(ora, net, cfg, "pregrown")
    for ep in range(2):
        ora.advance(H.steps_per_epoch(cfg))
        net.advance(H.steps_per_epoch(cfg))
        ora.update_connections()
        net.update_connections()
        ora.update_history()
        net.update_history()
        net.sync()
        bad += H.compare_state(ora, net, cfg, f"ep{ep}")
    net.close()
    assert not bad, bad[:8]


def test_pregrown_10k_synapse_count(hip_ctx):
    """r0 = 8.9 on the 100x100 grid connects every pair closer than 17.8 (area > 0): 8,440,432."""
    c = host.base_config()
    c.update(width=100, height=100, max_synapses_per_neuron=1000, num_epochs=1, epoch_duration=10.0,
             max_firing_rate=1000)
    m = host.Model(hip_ctx, c)
    m.set_radii(8.9)
    st = m.update_connections()
    g = np.arange(100)
    d2 = (g[:, None] - g[None, :]) ** 2
    # count pairs (i != j) with squared distance below 17.8^2 on the grid, by axis separation
    cnt_axis = np.bincount(np.abs(g[:, None] - g[None, :]).ravel(), minlength=100)
    want = 0
    for dx in range(100):
        for dy in range(100):
            if 0 < dx * dx + dy * dy < 17.8 ** 2:
                want += cnt_axis[dx] * cnt_axis[dy]
    m.close()
    assert st.total_synapses == want == 8440432
    assert st.created == want


def test_rates_agree_with_reference_noise_stream(hip_ctx):
    """Deterministic (Philox) GPU run vs the oracle in the reference's own Norm-noise mode: the
    noise streams differ, so only statistics can agree. 20x20 LIF growth network, 20 s."""
    c = host.base_config()
    c.update(width=20, height=20, max_synapses_per_neuron=40, num_epochs=2, epoch_duration=10.0, seed=3)
    m = host.Model(hip_ctx, c)
    for _ in range(2):
        m.run_epoch()
    gpu_spikes = float(m.array("burstinessHist").sum())
    m.close()
    oc = O.default_config()
    oc.update(width=20, height=20, max_synapses_per_neuron=40, num_epochs=2, epoch_duration=10.0, seed=3,
              layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1)
    ora = O.OracleNet(oc)
    for _ in range(2):
        ora.run_epoch()
    ref_spikes = float(ora.array("burstinessHist").sum())
    print(f"[stats] spikes gpu {gpu_spikes} reference-noise oracle {ref_spikes}", flush=True)
    assert ref_spikes > 1000
    assert abs(gpu_spikes - ref_spikes) / ref_spikes < 0.05


@pytest.mark.parametrize("name", ["conn-lif-ds-dynamic", "conn-izh-stdp"])
def test_host_model_serialize_deserialize_resumes_bit_for_bit(hip_ctx, tmp_path, name):
    """IModel::serialize after some epochs, IModel::deserialize into a freshly set-up model of the
    same configuration (the -w / -r flags of BrainGrid's driver): the resumed model's remaining
    epochs and XML output equal the uninterrupted model's byte for byte, and the oracle's."""
    print(f"[host] serialize {name}", flush=True)
    cfg = _host_cfg(name)
    a = host.Model(hip_ctx, cfg)
    b = None
    try:
        half = cfg["num_epochs"] // 2
        for _ in range(half):
            a.run_epoch()
        a.advance(123)    # a checkpoint need not fall on an epoch boundary
        path = os.path.join(tmp_path, "model.ckpt")
        a.serialize(path)
        b = host.Model(hip_ctx, cfg)
        b.deserialize(path)
        assert b.epoch == a.epoch == half and b.net.step == a.net.step
        for m in (a, b):
            m.advance(H.steps_per_epoch(H.load_config(name)) - 123)
            m.update_connections()
            m.update_history()
            for _ in range(half + 1, cfg["num_epochs"]):
                m.run_epoch()
        outs = []
        for m in (a, b):
            p = os.path.join(tmp_path, f"out{len(outs)}.xml")
            m.save_xml(p)
            outs.append(open(p).read())
        assert outs[0] == outs[1]
        ocfg = H.load_config(name)
        ora = O.OracleNet(ocfg)
        for _ in range(cfg["num_epochs"]):
            ora.run_epoch()
        assert outs[1] == xmlout.oracle_growth_xml(ora, ocfg)
        ora.close()
        other = dict(cfg, max_synapses_per_neuron=cfg["max_synapses_per_neuron"] + 1)
        c = host.Model(hip_ctx, other)
        with pytest.raises(RuntimeError, match="deserialize"):
            c.deserialize(path)
        c.close()
    finally:
        a.close()
        if b is not None:
            b.close()


@pytest.mark.parametrize("name", H.STATIC_CONFIGS)
def test_host_static_connections_match_oracle(hip_ctx, name):
    """The host model's ConnStatic (candidate lists sorted on worker threads, weights drawn in the
    reference's order, synapses created on the device by bghip_add_synapses) builds exactly the
    oracle's synapses, pinned to the reference's own tie order; the run and its XmlRecorder output
    then match the oracle."""
    print(f"[host] static {name}", flush=True)
    cfg = _host_cfg(name)
    ocfg = H.load_config(name)
    m = host.Model(hip_ctx, cfg)
    ora = O.OracleNet(ocfg)
    try:
        names = SETUP_ARRAYS + (IZH_ARRAYS if "izh" in name else ())
        bad = [k for k in names if not np.array_equal(m.array(k).view(np.uint8), ora.array(k).view(np.uint8))]
        bad += H.compare_state(ora, m.net, ocfg, "setup")
        assert not bad, bad
        assert m.net.total_synapses == ora.total_synapses > 0
        for ep in range(cfg["num_epochs"]):
            m.run_epoch()
            ora.run_epoch()
            bad = H.compare_state(ora, m.net, ocfg, f"epoch{ep}")
            assert not bad, bad
        assert ora.array("spikeCount").sum() + ora.array("spikeCountOffset").sum() > 0
        p = os.path.join(tempfile.mkdtemp(), "out.xml")
        m.save_xml(p)
        assert open(p).read() == xmlout.oracle_growth_xml(ora, ocfg)
    finally:
        m.close()
        ora.close()


@pytest.mark.parametrize("name", list(H.INPUT_STIM))
def test_host_model_with_stimulus_input_matches_oracle(hip_ctx, tmp_path, name):
    """The host model with a stimulus file (BrainGrid's -s): SInputRegular's asynchronous shifts
    drawn before the model's draws, the wave shifts, and SInputPoisson's masked DS input synapses,
    then every epoch's state, the input state and the XML output against the oracle."""
    print(f"[host] input {name}", flush=True)
    cfg = _host_cfg(name)
    cfg["input"] = host.read_stimulus(os.path.join(H.CONFIG_DIR, H.INPUT_STIM[name]), H.CONFIG_DIR)
    ocfg = H.load_config(name)
    m = host.Model(hip_ctx, cfg)
    ora = O.OracleNet(ocfg)
    try:
        names = SETUP_ARRAYS + (IZH_ARRAYS if "izh" in name else ())
        bad = [k for k in names if not np.array_equal(m.array(k).view(np.uint8), ora.array(k).view(np.uint8))]
        bad += H.compare_input(ora, m.net, ocfg, "setup")
        assert not bad, bad
        for ep in range(cfg["num_epochs"]):
            m.run_epoch()
            ora.run_epoch()
            bad = H.compare_state(ora, m.net, ocfg, f"epoch{ep}") + H.compare_input(ora, m.net, ocfg, f"epoch{ep}")
            assert not bad, bad
        p = os.path.join(tmp_path, "out.xml")
        m.save_xml(p)
        assert open(p).read() == xmlout.oracle_growth_xml(ora, ocfg)
    finally:
        m.close()
        ora.close()


def _burst_and_radii_agree(g_burst, r_burst, g_radii, r_radii, g_counts):
    """The recorder's burstinessHist (spikes per 1 s bin) and final radii of the eight seeds, against
    the compiled reference's. Tolerances: the burst totals equal the spike counts on each side;
    summed over the seeds, the whole histogram within 1% and every 1 s bin within 4% (a bin's
    eight-run sum varies by about 0.8% between independent noise streams, so 4% is a bound a
    mis-binned or missing second breaks and the noise does not); the mean final radius within 1e-5
    relative (the reference's seed-to-seed spread is about 1.3e-6)."""
    assert np.array_equal(g_burst.sum(axis=1), g_counts.sum(axis=(1, 2)))
    gb, rb = g_burst.sum(axis=0), r_burst.sum(axis=0)
    per_bin = (gb - rb) / rb
    print(f"[rates] burstinessHist: total {(gb.sum() - rb.sum()) / rb.sum() * 100:.3f}%, per 1 s bin "
          f"{np.round(per_bin * 100, 2).tolist()}%", flush=True)
    assert abs(gb.sum() - rb.sum()) / rb.sum() < 0.01
    assert np.all(np.abs(per_bin) < 0.04)
    rel = (g_radii.mean() - r_radii.mean()) / r_radii.mean()
    print(f"[rates] mean final radius {g_radii.mean():.7f} vs reference {r_radii.mean():.7f} ({rel:.2e} relative)",
          flush=True)
    assert abs(rel) < 1e-5


def test_firing_rates_within_one_percent_of_compiled_reference(hip_ctx):
    """Firing rates against the compiled reference's CPU path with its own noise stream. Workload:
    the first two 10 s epochs of the 10,000-neuron growth network (100 x 100, test-small parameter
    distributions, DynamicLayout 0.9 / 0.1), eight noise seeds. The reference side is the
    committed fixture tests/golden/ref_rates_10k.npz (oracle/_ref/bgref_dump, written by
    tests/golden/make_rate_fixture.py); the GPU side runs the host model with its counter-based
    noise. The streams differ, so the comparison is statistical. Tolerances, stated here:
      - population: the mean spike total over the eight seeds within 1% of the reference's
        (the per-run seed-to-seed spread is about 0.3%, so the difference of two eight-run means
        has a standard deviation of about 0.13%);
      - per neuron: the same neurons fire at all; for each firing neuron z = (g - r) / sqrt(g + r)
        over its summed counts has |mean| < 0.15 (no bias; 1,000 neurons give a standard error of
        0.03) and standard deviation < 1.2 (no excess per-neuron variance beyond Poisson)."""
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_rates_10k.npz"))
    ref, seeds = fx["counts"].astype(np.float64), [int(s) for s in fx["seeds"]]
    S, E, N = ref.shape
    gpu = np.zeros_like(ref)
    g_burst = np.zeros(fx["burst"].shape)
    g_radii = np.zeros(fx["radii"].shape)
    for k, s in enumerate(seeds):
        print(f"[rates] seed {s}", flush=True)
        c = host.base_config()
        c.update(width=100, height=100, max_synapses_per_neuron=200, max_firing_rate=200, num_epochs=E,
                 epoch_duration=10.0, seed=s, layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1)
        m = host.Model(hip_ctx, c)
        for _ in range(E):
            m.run_epoch()
        gpu[k] = np.rint(m.array("ratesHistory")[N:(E + 1) * N].astype(np.float64).reshape(E, N) * 10.0)
        g_burst[k] = m.array("burstinessHist")
        g_radii[k] = m.array("radiiHistory")[E * N:(E + 1) * N]
        m.close()
    _burst_and_radii_agree(g_burst, fx["burst"].astype(np.float64), g_radii, fx["radii"].astype(np.float64), gpu)
    g_tot, r_tot = gpu.sum(axis=(1, 2)), ref.sum(axis=(1, 2))
    rel = (g_tot.mean() - r_tot.mean()) / r_tot.mean()
    spread = np.std(r_tot / r_tot.mean(), ddof=1)
    print(f"[rates] spikes per seed: gpu {g_tot.tolist()} reference {r_tot.tolist()}; "
          f"mean difference {rel * 100:.3f}%, reference seed-to-seed spread {spread * 100:.3f}%", flush=True)
    assert abs(rel) < 0.01
    for e in range(E):
        re = (gpu[:, e].sum() - ref[:, e].sum()) / ref[:, e].sum()
        print(f"[rates] epoch {e + 1}: mean difference {re * 100:.3f}%", flush=True)
        assert abs(re) < 0.01
    cg, cr = gpu.sum(axis=(0, 1)), ref.sum(axis=(0, 1))
    assert np.array_equal(cg > 0, cr > 0)
    f = cr > 0
    z = (cg[f] - cr[f]) / np.sqrt(cg[f] + cr[f])
    print(f"[rates] {int(f.sum())} firing neurons: z mean {z.mean():.4f}, sd {z.std(ddof=1):.4f}", flush=True)
    assert abs(z.mean()) < 0.15
    assert z.std(ddof=1) < 1.2

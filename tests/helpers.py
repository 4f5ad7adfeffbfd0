"""Shared test helpers: oracle <-> GPU state transfer and bit-exact comparisons."""
from __future__ import annotations

import os

import numpy as np

from braingrid_amd import bghip as B
from braingrid_amd import model as MD
from oracle import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIG_DIR = os.path.join(ROOT, "tests", "golden", "configs")

# Validation configs with a stimulus input file (tests/golden/make_configs.py), relative to CONFIG_DIR.
INPUT_STIM = {"input-lif-poisson": "stim/poisson-50hz-w25-mask.xml", "input-izh-poisson": "stim/poisson-10hz-w25.xml",
              "input-lif-regular-async": "stim/regular-5ms-20ms-async.xml",
              "input-lif-regular-wave": "stim/regular-10ms-30ms-wave.xml"}

GROWTH_CONFIGS = ["conn-lif-ds", "conn-lif-ds-dense", "conn-lif-spiking", "conn-lif-stdp", "conn-lif-dynstdp",
                  "conn-izh-ds", "conn-izh-stdp", "conn-lif-ds-dynamic"]
# ConnStatic nearest-k configs (tests/golden/make_configs.py)
STATIC_CONFIGS = ["static-izh-spiking", "static-lif-ds", "static-izh-all"]

NEURON_UPLOAD = ("Vm", "Vthresh", "Vreset", "I0", "Inoise", "C1", "C2", "Trefract", "summation_map", "nStepsInRefr",
                 "spikeCount", "spikeCountOffset", "spike_history")
IZH_UPLOAD = ("Aconst", "Bconst", "Cconst", "Dconst", "u", "C3")
SYN_UPLOAD = ("in_use", "sourceNeuronIndex", "type", "W", "psr", "decay", "total_delay", "delayIdx", "delayQueue",
              "r", "U", "D", "F", "lastSpike", "delayQueuePost", "delayIdxPost", "Apos", "Aneg", "STDPgap",
              "tauspost", "tauspre", "taupos", "tauneg", "Wex", "mupos", "muneg", "useFroemkeDanSTDP")


def load_config(name: str, deterministic: bool = True) -> dict:
    cfg = O.parse_config_xml(os.path.join(CONFIG_DIR, name + ".xml"), CONFIG_DIR)
    if "dynamic" in name:
        cfg.update(layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1)
    if name in INPUT_STIM:
        cfg["input"] = O.parse_input_xml(os.path.join(CONFIG_DIR, INPUT_STIM[name]), CONFIG_DIR)
    if deterministic:
        cfg.update(noise_mode=1, math_mode=1, sum_mode=1)
    return cfg


def max_spikes(cfg) -> int:
    return int(np.float32(cfg["epoch_duration"]) * np.float32(cfg["max_firing_rate"]))


def steps_per_epoch(cfg) -> int:
    return int(np.float32(np.float32(cfg["epoch_duration"]) / np.float32(cfg["dt"])))


def gpu_net_from_oracle(ctx: B.Context, ora: O.OracleNet, cfg: dict, block_steps: int = 0) -> B.Net:
    """Create a device network holding exactly the oracle's current state."""
    N = cfg["width"] * cfg["height"]
    M = cfg["max_synapses_per_neuron"]
    nk = O.NEURON_CLASSES[cfg["neuron"]]
    sk = O.SYN_CLASSES[cfg["synapse"]]
    desc = MD.net_desc(N, M, nk, sk, max_spikes(cfg), cfg["dt"], cfg["seed"], cfg.get("ring_bits", 32),
                       block_steps, cfg.get("stdp"), cfg.get("delay_override"))
    net = B.Net(ctx, desc)
    neu = {f: ora.array(f) for f in NEURON_UPLOAD}
    if nk == O.IZH:
        neu.update({f: ora.array(f) for f in IZH_UPLOAD})
    net.upload_neurons(neu)
    syn = {}
    for f in SYN_UPLOAD:
        src = "u_syn" if f == "u" else f
        try:
            syn[f] = ora.array(src)
        except KeyError:
            pass
    syn["u"] = ora.array("u_syn")
    syn["synapse_counts"] = ora.array("synapse_counts")
    net.upload_synapses(syn, ora.step)
    net.upload_layout(ora.array("xloc"), ora.array("yloc"), ora.array("neuron_type_map"))
    g = MD.growth_params(cfg["epsilon"], cfg["beta"], cfg["rho"], cfg["target_rate"], cfg["min_radius"],
                         cfg["start_radius"], cfg["epoch_duration"])
    net.upload_growth(g, ora.array("radii"), ora.array("rates"))
    upload_input_from_oracle(net, ora, cfg)
    return net


INPUT_STATE = {"nISIs": "in_nISIs", "psr": "in_psr", "r": "in_r", "u": "in_u", "lastSpike": "in_lastSpike",
               "delayQueue": "in_delayQueue", "delayIdx": "in_delayIdx"}


def upload_input_from_oracle(net: B.Net, ora: O.OracleNet, cfg: dict):
    """Give the device network the oracle's stimulus input, state included."""
    inp = cfg.get("input")
    if not inp:
        return
    if inp["kind"] == "poisson":
        net.upload_input(B.INPUT_POISSON, neuron_type=ora.array("neuron_type_map"), fr_mean=inp["fr_mean"],
                         weight=inp["weight"], mask=ora.array("in_mask"),
                         state={k: ora.array(v) for k, v in INPUT_STATE.items()})
    else:
        net.upload_input(B.INPUT_REGULAR, duration=inp["duration"], interval=inp["interval"],
                         values=ora.array("in_values"), shift=ora.array("in_shift"),
                         state={"steps_in_cycle": int(ora.array("in_steps_in_cycle")[0])})


def compare_input(ora: O.OracleNet, net: B.Net, cfg: dict, where: str) -> list[str]:
    inp = cfg.get("input")
    if not inp:
        return []
    g = net.download_input()
    if inp["kind"] == "regular":
        ok = g["steps_in_cycle"] == int(ora.array("in_steps_in_cycle")[0])
        return [] if ok else [f"{where}:in_steps_in_cycle"]
    return [f"{where}:{v}" for k, v in INPUT_STATE.items() if not _eq(ora.array(v), g[k])]


def _eq(a, b):
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint32 if a.itemsize == 4 else np.uint64),
                              b.view(np.uint32 if b.itemsize == 4 else np.uint64))
    return np.array_equal(a, b)


def compare_state(ora: O.OracleNet, net: B.Net, cfg: dict, where: str, check_maps: bool = True) -> list[str]:
    """Bit-exact comparison of every state array; returns the names that differ."""
    bad = []
    ms = max_spikes(cfg)
    g = net.download_neurons(ms)
    names = ["Vm", "summation_map", "nStepsInRefr", "spikeCount", "spikeCountOffset", "spike_history"]
    if O.NEURON_CLASSES[cfg["neuron"]] == O.IZH:
        names.append("u")
    for f in names:
        if not _eq(ora.array(f), g[f]):
            bad.append(f"{where}:{f}")
    s = net.download_synapses()
    use = ora.array("in_use").astype(bool)
    if not np.array_equal(use, s["in_use"].astype(bool)):
        bad.append(f"{where}:in_use")
    if not np.array_equal(ora.array("synapse_counts"), s["synapse_counts"]):
        bad.append(f"{where}:synapse_counts")
    W32 = cfg.get("ring_bits", 32) // 32
    pairs = [("sourceNeuronIndex", "sourceNeuronIndex"), ("type", "type"), ("W", "W"), ("psr", "psr"),
             ("decay", "decay"), ("total_delay", "total_delay"), ("delayIdx", "delayIdx")]
    sk = O.SYN_CLASSES[cfg["synapse"]]
    if sk in (O.DS, O.DYNSTDP):
        pairs += [("r", "r"), ("u_syn", "u"), ("lastSpike", "lastSpike")]
    if sk in (O.STDP, O.DYNSTDP):
        pairs += [("delayQueuePost", "delayQueuePost"), ("delayIdxPost", "delayIdxPost")]
    for of, gf in pairs:
        if not _eq(ora.array(of)[use], s[gf][use]):
            bad.append(f"{where}:{of}")
    rq_o = ora.array("delayQueue").reshape(-1, W32)[use]
    rq_g = s["delayQueue"].reshape(-1, W32)[use]
    if not np.array_equal(rq_o, rq_g):
        bad.append(f"{where}:delayQueue")
    if cfg["conn"] == "ConnGrowth":
        r, q = net.download_growth()
        if not _eq(ora.array("radii"), r):
            bad.append(f"{where}:radii")
        if not _eq(ora.array("rates"), q):
            bad.append(f"{where}:rates")
    if check_maps and ora.total_synapses > 0:
        m = net.download_index_map()
        for f in m:
            if not np.array_equal(ora.array(f), m[f]):
                bad.append(f"{where}:{f}")
    return bad


def izh_static_oracle(neurons: dict, syn: dict, meta: dict, seed: int, epoch_s: float = 1.0,
                      max_rate: int = 200, **overrides) -> tuple[O.OracleNet, dict]:
    """An oracle network holding the static Izhikevich + DynamicSTDP network of
    braingrid_amd.networks.izh_stdp_arrays, built through the oracle's own addSynapse/createSynapse.
    Neuron constants drawn by the oracle's setup are replaced with the builder's; W, Wex and the
    per-synapse total_delay are written after creation, as ConnStatic writes W (ConnStatic.cpp:75-80)."""
    from braingrid_amd import networks as NW
    N = meta["N"]
    cfg = O.default_config()
    # connections come from the builder's list, so the oracle's own generators stay off
    cfg.update(width=N, height=1, neuron="AllIZHNeurons", synapse="AllDynamicSTDPSynapses", conn="none",
               layout="FixedLayout", inh_list=[int(i) for i in np.flatnonzero(meta["ntype"] == 1)], active_list=[],
               max_synapses_per_neuron=meta["M"], max_firing_rate=max_rate, num_epochs=1, epoch_duration=epoch_s,
               seed=seed, ring_bits=meta["ring_bits"], stdp=NW.STDP_CFG, noise_mode=1, math_mode=1, sum_mode=1,
               Iinject=NW.IZH_NEURON["Iinject"], Inoise=NW.IZH_NEURON["Inoise"], Vthresh=NW.IZH_NEURON["Vthresh"],
               Vresting=NW.IZH_NEURON["Vresting"], Vreset=NW.IZH_NEURON["Vreset"], Vinit=NW.IZH_NEURON["Vinit"])
    cfg.update(overrides)   # e.g. the reference's noise/math/summation modes, stimulus input
    ora = O.OracleNet(cfg)
    ora.setup_consts = {f: ora.array(f).copy() for f in ("C1", "C2", "C3", "I0")}
    for f in ("Vm", "Vthresh", "Vreset", "I0", "Inoise", "C1", "C2", "Trefract", "Aconst", "Bconst", "Cconst",
              "Dconst", "u", "C3"):
        ora.view(f)[:] = neurons[f]
    src, dst, slot = meta["src"], meta["dst"], meta["slot"]
    for q in range(len(src)):
        got = ora.add_synapse(int(src[q]), int(dst[q]), float(meta["W"][q]))
        assert got == slot[q], (q, got, slot[q])
    ora.view("total_delay")[slot] = meta["total_delay"]
    ora.view("Wex")[slot] = syn["Wex"][slot]
    ora.rebuild_index_map()
    return ora, cfg


def oracle_for_spec(spec: dict) -> tuple[O.OracleNet, dict]:
    """The oracle, in the deterministic mode the device computes in, holding the network of one
    sweep or served job specification (braingrid_amd.sweep): the growth family through the oracle's
    own setup from the same configuration as the host model's, the Izhikevich family through the
    oracle's addSynapse/createSynapse chain from the builder's draws, with its 10 Hz Poisson input.
    Recorder histories are sized for the job's whole duration."""
    from braingrid_amd import networks as NW
    from braingrid_amd import sweep as SW
    epochs = int(round(spec["sim_s"] / SW.EPOCH_S))
    if spec["kind"] == "lif_growth":
        c = O.default_config()
        c.update(width=spec["side"], height=spec["side"], layout="DynamicLayout", frac_exc=spec["fE"],
                 frac_starter=0.1, target_rate=spec["tR"], seed=spec["seed"], epoch_duration=SW.EPOCH_S,
                 num_epochs=epochs, max_synapses_per_neuron=SW.LIF_M, max_firing_rate=SW.LIF_MAX_RATE,
                 noise_mode=1, math_mode=1, sum_mode=1)
        return O.OracleNet(c), c
    N = spec["side"] ** 2
    neu, syn, meta = NW.izh_stdp_arrays(N, spec["seed"], k=min(100, N - 1), max_rate=SW.IZH_MAX_RATE,
                                        epoch_s=SW.EPOCH_S)
    return izh_static_oracle(neu, syn, meta, spec["seed"], epoch_s=SW.EPOCH_S, max_rate=SW.IZH_MAX_RATE,
                             num_epochs=epochs,
                             input=dict(kind="poisson", fr_mean=SW.POISSON_HZ, weight=SW.POISSON_WEIGHT))


def run_spec_in_oracle(spec: dict) -> dict:
    """A job of `spec` run start to end in the oracle, epoch by epoch as the served form and the
    sweep run it (advance one epoch, growth update for the growth family, history step); returns
    its recorder output (and final radii for the growth family), the served job's result fields."""
    from braingrid_amd import sweep as SW
    ora, _ = oracle_for_spec(spec)
    try:
        steps = int(round(SW.EPOCH_S / SW.DT))
        for _ in range(int(round(spec["sim_s"] / SW.EPOCH_S))):
            ora.advance(steps)
            if spec["kind"] == "lif_growth":
                ora.update_connections()
            ora.update_history()
        out = {k: ora.array(k).copy() for k in ("burstinessHist", "spikesHistory")}
        if spec["kind"] == "lif_growth":
            out["radii"] = ora.array("radii").copy()
        return out
    finally:
        ora.close()

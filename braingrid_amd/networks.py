"""Builders of the sweep workloads (batched and served forms).

Two network families, both uploaded through the C ABI:

* `lif_growth(...)` -- the plain method: a side x side grid of AllLIFNeurons with the test-small
  parameter distributions, DynamicLayout with excitatory fraction fE and 10% starters, DS synapses
  and ConnGrowth with target rate tR (HipSpikingModel does Model::setupSim on the host).
* `izh_stdp(...)` -- static Izhikevich + DynamicSTDP: 80% excitatory neurons with a = 0.02,
  b = 0.2, c = -65 + 15 r^2, d = 8 - 6 r^2 and 20% inhibitory with a = 0.02 + 0.08 r,
  b = 0.25 - 0.05 r, c = -65, d = 2 (r ~ U[0,1], Izhikevich 2003), the other neuron constants as
  configfiles/static_izh_1000.xml; k = 100 random distinct outgoing targets per neuron, no self
  connections; delays U{1..20} ms on a 256-bit delay ring; STDP with tau+ = tau- = 20 ms.
  Weights are given in Izhikevich's units and stored as W = w * 1e-8 (the reference's synapse
  strength unit, Synapses/AllSynapses.h SYNAPSE_STRENGTH_ADJUSTMENT): with C2 = Rm(1 - C1) ~ 3331
  and a 3-6 ms PSR time constant, one pre-spike of weight w moves an Izhikevich membrane by about
  w to 2w mV, which is the scale of Izhikevich's own weights. Excitatory w ~ U[0,6] and
  inhibitory w ~ U[-5,0]; A+ = 0.005 and A- = -0.00525 are in the same units, and Wex (the
  reference's signed weight bound, Synapses/AllSTDPSynapses.cpp:449-455) is +10 units for
  excitatory sources and -5 for inhibitory ones, so learning keeps each weight's sign.
  Froemke-Dan is off. Slots are filled the way the reference's ConnStatic does: createSynapse
  defaults (Synapses/AllDSSynapses.cpp:180-218, Synapses/AllSTDPSynapses.cpp:252-272) and
  addSynapse in ascending source order per destination (Connections/ConnStatic.cpp:42-80).
Both return (net, meta) where meta describes the network for drivers and tests.
"""
from __future__ import annotations

import numpy as np

from . import bghip as B
from . import host
from . import model as MD

F32 = np.float32
W_UNIT = 1e-8            # AllSynapses::SYNAPSE_STRENGTH_ADJUSTMENT (Synapses/AllSynapses.h:190)
STDP_CFG = dict(Apos=0.005 * W_UNIT, Aneg=-0.00525 * W_UNIT, taupos=20e-3, tauneg=20e-3, froemke_dan=False)
WEX_EXC, WEX_INH = 10 * W_UNIT, -5 * W_UNIT
IZH_NEURON = dict(Iinject=(13.5e-9, 13.5e-9), Inoise=(0.5e-6, 0.7329e-6), Vthresh=(30.0e-3, 30.0e-3),
                  Vresting=(0.0, 0.0), Vreset=(-0.065, -0.065), Vinit=(-0.065, -0.065))


def lif_growth(ctx: B.Context, side: int, target_rate: float, frac_exc: float, seed: int, epoch_s: float = 1.0,
               epochs: int = 5, max_synapses: int = 64, max_rate: int = 200):
    c = host.base_config()
    c.update(width=side, height=side, target_rate=target_rate, frac_exc=frac_exc, frac_starter=0.1, seed=seed,
             epoch_duration=epoch_s, num_epochs=epochs, max_synapses_per_neuron=max_synapses,
             max_firing_rate=max_rate, layout="DynamicLayout")
    m = host.Model(ctx, c)
    return m, dict(kind="lif_growth", N=side * side, side=side, steps_per_epoch=int(round(epoch_s / 1e-4)),
                   epochs=epochs, cfg=c)


def _random_targets(rng, N: int, k: int) -> np.ndarray:
    """k distinct targets per source, none equal to the source: row s holds k of the N - 1 other
    neurons. Drawn with replacement for all rows at once; rows that came out with a repeat are
    drawn again without replacement (a few percent of rows at k = 100)."""
    if k > N - 1:
        raise ValueError(f"{k} outgoing synapses need at least {k + 1} neurons")
    if 2 * k > N - 1:   # dense: a permutation per row
        t = np.argsort(rng.random((N, N - 1)), axis=1)[:, :k]
    else:
        t = rng.integers(0, N - 1, size=(N, k))
        srt = np.sort(t, axis=1)
        for r in np.flatnonzero((srt[:, 1:] == srt[:, :-1]).any(axis=1)):
            t[r] = rng.choice(N - 1, size=k, replace=False)
    return t + (t >= np.arange(N)[:, None])


def izh_stdp_draws(N: int, seed: int, k: int = 100, dt: float = 1e-4, ring_bits: int = 256, max_rate: int = 200,
                   epoch_s: float = 1.0):
    """The random draws of the static Izhikevich + STDP network: the neuron arrays, and in meta the
    synapse list (src, dst, W, total_delay, Wex) ordered per destination by ascending source, the
    order in which addSynapse fills the destination's slots."""
    rng = np.random.default_rng(seed)
    n_exc = int(round(0.8 * N))
    ntype = np.full(N, 2, np.int32)
    ntype[rng.permutation(N)[: N - n_exc]] = 1
    exc = ntype == 2
    r = rng.random(N).astype(np.float64)
    a = np.where(exc, 0.02, 0.02 + 0.08 * r)
    b = np.where(exc, 0.2, 0.25 - 0.05 * r)
    c = np.where(exc, -65 + 15 * r * r, -65.0)
    d = np.where(exc, 8 - 6 * r * r, 2.0)
    p = IZH_NEURON
    Iinject = np.full(N, F32(p["Iinject"][0]), F32)
    Inoise = (F32(p["Inoise"][0]) + rng.random(N).astype(F32) * F32(p["Inoise"][1] - p["Inoise"][0])).astype(F32)
    Vrest = np.zeros(N, F32)
    Rm = F32(1e6)
    Tau = F32(3e-8 * 1e6)
    C1 = np.full(N, MD.expf(F32(-F32(dt) / Tau)), F32)
    C2 = (Rm * (F32(1) - C1)).astype(F32)
    I0 = (Iinject + Vrest / Rm).astype(F32)
    Trefract = np.where(exc, F32(3e-3), F32(2e-3)).astype(F32)
    max_spikes = int(F32(epoch_s) * F32(max_rate))
    neurons = dict(Vm=np.full(N, F32(p["Vinit"][0]), F32), Vthresh=np.full(N, F32(p["Vthresh"][0]), F32),
                   Vreset=np.full(N, F32(p["Vreset"][0]), F32), I0=I0, Inoise=Inoise, C1=C1, C2=C2,
                   Trefract=Trefract, summation_map=np.zeros(N, F32), nStepsInRefr=np.zeros(N, np.int32),
                   Aconst=a.astype(F32), Bconst=b.astype(F32), Cconst=c.astype(F32), Dconst=d.astype(F32),
                   u=np.zeros(N, F32), C3=np.full(N, F32(F32(dt) * F32(1000)), F32),
                   spikeCount=np.zeros(N, np.int32), spikeCountOffset=np.zeros(N, np.int32),
                   spike_history=np.full(N * max_spikes, np.iinfo(np.uint64).max, np.uint64))
    targets = _random_targets(rng, N, k)
    w = np.where(exc[:, None], rng.uniform(0.0, 6.0, (N, k)), rng.uniform(-5.0, 0.0, (N, k))) * W_UNIT
    delay_ms = rng.integers(1, 21, (N, k))
    src = np.repeat(np.arange(N), k)
    dst = targets.ravel()
    order = np.lexsort((src, dst))            # per destination, ascending source = addSynapse order
    src, dst = src[order], dst[order]
    w = w.ravel()[order].astype(F32)
    delay_ms = delay_ms.ravel()[order]
    indeg = np.bincount(dst, minlength=N)
    M = int(indeg.max()) + 1                  # createSynapseImap asserts synapse_counts < M
    rank = np.arange(len(dst)) - np.repeat(np.cumsum(indeg) - indeg, indeg)
    slot = dst * M + rank
    S = N * M
    types = MD.synapse_types(dt, STDP_CFG)
    styp = np.where(ntype[src] == 1, np.where(ntype[dst] == 1, 0, 1), np.where(ntype[dst] == 1, 2, 3))
    tdel = np.array([MD.total_delay(v * 1e-3, dt) for v in range(21)])[delay_ms].astype(np.int32)
    wex = np.where(ntype[src] == 2, F32(WEX_EXC), F32(WEX_INH)).astype(F32)
    meta = dict(kind="izh_stdp", N=N, M=M, S=S, max_spikes=max_spikes, ring_bits=ring_bits, ntype=ntype, src=src,
                dst=dst, slot=slot, W=w, total_delay=tdel, Wex=wex, styp=styp, indeg=indeg, types=types,
                S_active=len(src))
    return neurons, meta


def izh_stdp_arrays(N: int, seed: int, k: int = 100, dt: float = 1e-4, ring_bits: int = 256, max_rate: int = 200,
                    epoch_s: float = 1.0):
    """Host arrays of the static Izhikevich + STDP network, every slot field as the reference's
    addSynapse/createSynapse leave it (used by tests to load the oracle)."""
    neurons, meta = izh_stdp_draws(N, seed, k, dt, ring_bits, max_rate, epoch_s)
    S, slot, src, dst, styp, types = meta["S"], meta["slot"], meta["src"], meta["dst"], meta["styp"], meta["types"]
    w, tdel, indeg = meta["W"], meta["total_delay"], meta["indeg"]
    syn = {f: np.zeros(S * (ring_bits // 32) if f == "delayQueue" else S, dt_) for f, dt_ in B.SYN_DTYPES.items()}
    syn["in_use"][slot] = 1
    syn["sourceNeuronIndex"][slot] = src
    syn["destNeuronIndex"][slot] = dst
    syn["type"][slot] = styp
    syn["W"][slot] = w
    syn["decay"][slot] = np.array(types.decay, F32)[styp]
    syn["total_delay"][slot] = tdel
    for f, v in (("Apos", types.Apos), ("Aneg", types.Aneg), ("STDPgap", types.STDPgap),
                 ("tauspost", types.tauspost), ("tauspre", types.tauspre), ("taupos", types.taupos),
                 ("tauneg", types.tauneg), ("Wex", types.Wex), ("mupos", types.mupos), ("muneg", types.muneg)):
        syn[f][slot] = F32(v)
    syn["useFroemkeDanSTDP"][slot] = types.useFroemkeDanSTDP
    syn["Wex"][slot] = meta["Wex"]
    syn["r"][slot] = F32(types.r0)
    syn["u"][slot] = F32(types.u0)
    syn["U"][slot] = np.array(types.U, F32)[styp]
    syn["D"][slot] = np.array(types.D, F32)[styp]
    syn["F"][slot] = np.array(types.F, F32)[styp]
    syn["lastSpike"][:] = np.iinfo(np.uint64).max
    syn["synapse_counts"] = indeg.astype(np.uint32)
    return neurons, syn, meta


def izh_stdp_desc(meta: dict, seed: int) -> B.NetDesc:
    return MD.net_desc(meta["N"], meta["M"], B.NEURON_IZH, B.SYN_DYNSTDP, meta["max_spikes"], 1e-4, seed,
                       meta["ring_bits"], 0, STDP_CFG)


def izh_stdp(ctx: B.Context, N: int, seed: int, k: int = 100, ring_bits: int = 256, max_rate: int = 200,
             epoch_s: float = 1.0, arrays=None):
    """Create the network on the device; `arrays` may hold izh_stdp_draws' result computed earlier.
    The synapses are created on the device from their list (bghip_add_synapses_ex: addSynapse in
    list order, createSynapse's per-type values, then the synapse's W, total delay and Wex), so the
    slots equal izh_stdp_arrays' without building them on the host. The neuron types reach the
    device as a layout (synapse types follow from them; the coordinates are not used: there is no
    growth)."""
    neurons, meta = arrays or izh_stdp_draws(N, seed, k, ring_bits=ring_bits, max_rate=max_rate, epoch_s=epoch_s)
    net = B.Net(ctx, izh_stdp_desc(meta, seed))
    net.upload_neurons(neurons)
    n = meta["N"]
    net.upload_layout(np.arange(n, dtype=F32), np.zeros(n, F32), meta["ntype"].astype(np.int32))
    net.add_synapses(meta["src"], meta["dst"], meta["W"], total_delay=meta["total_delay"], Wex=meta["Wex"])
    return net, meta

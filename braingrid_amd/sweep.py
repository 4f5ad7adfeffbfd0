"""Parameter sweeps of many independent networks: the batched form.

A sweep is a list of network specifications (grid side, family, growth target rate tR, excitatory
fraction fE, seed, simulated seconds). `build(ctx, spec)` creates one network on the device;
`run_alone` advances the networks one after another with their own launches, `run_batched`
advances all of them together with shared launches (bghip_batch_*) and does the per-network
epoch work (growth update, history) in between. Both produce bit-identical network states.

Families:
  lif_growth -- AllLIFNeurons + AllDSSynapses + ConnGrowth with the test-small parameter
                distributions (configfiles/test-small.xml), DynamicLayout with fraction fE and 10%
                starters, target rate tR, 1 s epochs, each followed by the growth update and the
                growth recorder's history step;
                With `r0` in the specification the network starts pre-grown: every radius is set to
                r0 and one growth update creates the synapses of overlapping neighbours before the
                first step (r0 = 2.8 gives about 100 synapses per interior neuron), so the network
                runs the fused Spiking/DS kernels from its first step; such networks fire near
                their refractory limit and keep 1000 spikes per neuron and epoch of history;
  izh_stdp   -- the static Izhikevich + DynamicSTDP network of braingrid_amd.networks with
                min(100, N - 1) outgoing synapses per neuron and a 10 Hz Poisson input of weight 25
                (configfiles/SInputPoisson/i_10hz_w25.xml) into every neuron; 1 s epochs, each
                followed by the history step.
"""
from __future__ import annotations

import hashlib
import time

import numpy as np

from . import bghip as B
from . import host
from . import model as MD
from . import networks as NW

TR_CHOICES = (0.1, 1.0, 1.9)
FE_CHOICES = (0.80, 0.90, 0.98)
EPOCH_S = 1.0
DT = 1e-4
LIF_M = 200
LIF_MAX_RATE = 200
LIF_PREGROWN_MAX_RATE = 1000
IZH_MAX_RATE = 400
POISSON_HZ, POISSON_WEIGHT = 10.0, 25.0
# batch launch-shape autotuner: timed steps per candidate shape (bghip_batch_autotune)
TUNE_STEPS = 8


def net_seed(global_seed: int, net_id: int) -> int:
    h = hashlib.blake2b(f"{global_seed}:{net_id}".encode(), digest_size=4).digest()
    return int.from_bytes(h, "little") or 1


def sweep_specs(n_nets: int = 256, seed: int = 2024, side_lo: int = 10, side_hi: int = 120, sim_s: float = 5.0,
                frac_lif: float = 0.5, lif_r0: float | None = None):
    """Sides log-uniform on [side_lo, side_hi]; families alternate by a seeded draw. With lif_r0
    the growth networks start pre-grown to that radius."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n_nets):
        side = int(round(float(np.exp(rng.uniform(np.log(side_lo), np.log(side_hi))))))
        kind = "lif_growth" if rng.random() < frac_lif else "izh_stdp"
        specs.append(dict(id=k, side=side, kind=kind, tR=float(rng.choice(TR_CHOICES)),
                          fE=float(rng.choice(FE_CHOICES)), seed=net_seed(seed, k), sim_s=float(sim_s)))
        if lif_r0 and kind == "lif_growth":
            specs[-1]["r0"] = float(lif_r0)
    return specs


def lif_max_rate(spec: dict) -> int:
    return LIF_PREGROWN_MAX_RATE if spec.get("r0") else LIF_MAX_RATE


MAX_SIDE = 512   # N = side^2 stays under the engine's 262,144-neuron limit


def lif_config(spec: dict) -> dict:
    c = host.base_config()
    c.update(width=spec["side"], height=spec["side"], layout="DynamicLayout", frac_exc=spec["fE"], frac_starter=0.1,
             target_rate=spec["tR"], seed=spec["seed"], epoch_duration=EPOCH_S,
             num_epochs=int(round(spec["sim_s"] / EPOCH_S)), max_synapses_per_neuron=LIF_M,
             max_firing_rate=lif_max_rate(spec))
    return c


def plan(spec: dict) -> dict:
    """Host-side preparation of a network: validates the specification, draws the static
    connectivity of an Izhikevich network and gives the device description (for the footprint).
    Raises ValueError for a specification the engine cannot run."""
    kind = spec.get("kind")
    side = spec.get("side")
    sim_s = spec.get("sim_s")
    if kind not in ("lif_growth", "izh_stdp"):
        raise ValueError(f"unknown network family {kind!r}")
    if not isinstance(side, (int, np.integer)) or not 2 <= side <= MAX_SIDE:
        raise ValueError(f"grid side {side!r} outside 2..{MAX_SIDE}")
    if not isinstance(sim_s, (int, float)) or sim_s <= 0 or abs(sim_s / EPOCH_S - round(sim_s / EPOCH_S)) > 1e-9:
        raise ValueError(f"simulated time {sim_s!r} is not a positive whole number of {EPOCH_S} s epochs")
    N = int(side) ** 2
    if kind == "lif_growth":
        if not (spec.get("tR", 0) > 0 and 0 < spec.get("fE", 0) <= 1):
            raise ValueError("lif_growth needs tR > 0 and 0 < fE <= 1")
        if spec.get("r0") is not None and not 0 < spec["r0"] <= 4.0:
            raise ValueError(f"pre-grown radius {spec['r0']!r} outside (0, 4] (the rows must fit {LIF_M} slots)")
        max_spikes = int(np.float32(EPOCH_S) * np.float32(lif_max_rate(spec)))
        desc = MD.net_desc(N, LIF_M, B.NEURON_LIF, B.SYN_DS, max_spikes, DT, spec["seed"])
        return dict(kind=kind, desc=desc, arrays=None)
    arrays = NW.izh_stdp_draws(N, spec["seed"], k=min(100, N - 1), max_rate=IZH_MAX_RATE, epoch_s=EPOCH_S)
    return dict(kind=kind, desc=NW.izh_stdp_desc(arrays[1], spec["seed"]), arrays=arrays)


def footprint(p: dict) -> int:
    """Device bytes of a planned network (its arena; growth is tiled, so no N^2 term)."""
    return int(B.footprint(p["desc"]))


class SweepNet:
    """One network of a sweep with its per-epoch work."""

    def __init__(self, ctx: B.Context, spec: dict, planned: dict | None = None):
        self.spec = spec
        self.kind = spec["kind"]
        self.N = spec["side"] ** 2
        self.epochs = int(round(spec["sim_s"] / EPOCH_S))
        self.steps_per_epoch = int(round(EPOCH_S / DT))
        self.model = None
        if self.kind == "lif_growth":
            self.model = host.Model(ctx, lif_config(spec))
            self.net = self.model.net
            self.max_spikes = int(np.float32(EPOCH_S) * np.float32(lif_max_rate(spec)))
            if spec.get("r0"):
                self.model.set_radii(spec["r0"])
                self.model.update_connections(stats=False)
        else:
            arrays = planned["arrays"] if planned else None
            self.net, meta = NW.izh_stdp(ctx, self.N, spec["seed"], k=min(100, self.N - 1), max_rate=IZH_MAX_RATE,
                                         epoch_s=EPOCH_S, arrays=arrays)
            self.net.upload_input(B.INPUT_POISSON, neuron_type=meta["ntype"], fr_mean=POISSON_HZ,
                                  weight=POISSON_WEIGHT)
            self.max_spikes = meta["max_spikes"]
            # the spike-count histograms of XmlRecorder, kept on the device
            self.net.recorder_enable(int(spec["sim_s"]), int(spec["sim_s"] * 100), self.epochs)
        self.synapses = self.net.total_synapses

    def save(self, path: str):
        """Checkpoint of the network (and, for growth networks, its recorder) at an epoch boundary."""
        if self.model is not None:
            self.model.serialize(path)
        else:
            self.net.save(path)

    @classmethod
    def restore(cls, ctx: B.Context, spec: dict, path: str) -> "SweepNet":
        """The network of `spec` continued from a checkpoint written by save()."""
        x = cls.__new__(cls)
        x.spec, x.kind, x.N = spec, spec["kind"], spec["side"] ** 2
        x.epochs = int(round(spec["sim_s"] / EPOCH_S))
        x.steps_per_epoch = int(round(EPOCH_S / DT))
        x.model = None
        if x.kind == "lif_growth":
            x.model = host.Model(ctx, lif_config(spec))
            try:
                x.model.deserialize(path)
            except Exception:
                x.model.close()
                raise
            x.net = x.model.net
        else:
            x.net = B.Net.load(ctx, path)
        x.max_spikes = x.net.desc.max_spikes
        x.synapses = x.net.total_synapses
        return x

    def summary(self) -> dict:
        """The job's output: XmlRecorder's spike-count histograms (1 s and 10 ms bins), and for
        growth networks the final radii; a few KB at most."""
        if self.model is not None:
            out = {"burstinessHist": self.model.array("burstinessHist"),
                   "spikesHistory": self.model.array("spikesHistory"), "radii": self.model.array("radii")}
        else:
            out = self.net.recorder_download()
        out["synapses"] = self.net.total_synapses
        return out

    def end_epoch(self):
        """Growth update and history step, queued on the network's stream (not waited for)."""
        if self.model is not None:
            self.model.update_connections(stats=False)
            self.model.update_history()
        else:
            self.net.update_history()

    def fingerprint(self) -> str:
        """Digest of the network's full neuron state and synapse weights/PSRs."""
        h = hashlib.blake2b(digest_size=16)
        neu = self.net.download_neurons(self.max_spikes)
        for f in ("Vm", "spikeCount", "spikeCountOffset", "spike_history", "nStepsInRefr"):
            h.update(np.ascontiguousarray(neu[f]).tobytes())
        syn = self.net.download_synapses()
        for f in ("in_use", "W", "psr"):
            h.update(np.ascontiguousarray(syn[f]).tobytes())
        if self.model is not None:
            r, q = self.net.download_growth()
            h.update(r.tobytes())
            h.update(q.tobytes())
        return h.hexdigest()

    def close(self):
        if self.model is not None:
            self.model.close()
        else:
            self.net.close()


def build(ctx: B.Context, specs, log=None):
    nets = []
    for s in specs:
        nets.append(SweepNet(ctx, s))
        if log and len(nets) % 32 == 0:
            log(f"built {len(nets)}/{len(specs)} networks")
    return nets


def run_alone(nets, log=None) -> float:
    """Each network on its own, one after another (plain per-network launches). Returns seconds."""
    for x in nets:
        x.net.sync()
    t0 = time.perf_counter()
    for k, x in enumerate(nets):
        for _ in range(x.epochs):
            x.net.advance(x.steps_per_epoch)
            x.end_epoch()
        x.net.sync()
        if log and (k + 1) % 32 == 0:
            log(f"alone: {k + 1}/{len(nets)} networks done, {time.perf_counter() - t0:.1f} s")
    return time.perf_counter() - t0


def run_batched(ctx: B.Context, nets, log=None, tune_steps: int = TUNE_STEPS, shapes: list | None = None) -> float:
    """All networks advanced by shared launches, epoch by epoch, with the batch's launch-shape
    autotuner at `tune_steps` timed steps per candidate (0: fixed default shapes). Returns seconds;
    the groups' launch shapes at the end are appended to `shapes` when given."""
    for x in nets:
        x.net.sync()
    batch = B.Batch(ctx)
    batch.autotune(tune_steps)
    t0 = time.perf_counter()
    live = [x for x in nets if x.epochs > 0]
    epoch = 0
    members = None
    while live:
        if members != len(live):
            batch.set([x.net for x in live])
            members = len(live)
        steps = live[0].steps_per_epoch
        batch.advance(steps)
        for x in live:
            x.end_epoch()
        epoch += 1
        live = [x for x in live if x.epochs > epoch]
        if log:
            log(f"batched: epoch {epoch} done, {time.perf_counter() - t0:.1f} s")
    batch.sync()
    for x in nets:
        x.net.sync()
    wall = time.perf_counter() - t0
    if shapes is not None:
        shapes.extend(batch.shapes())
    batch.close()
    return wall

"""Benchmark: synapse-updates/s of the 10k-neuron LIF + DS growth network on one MI355X.

Workload (the benchmark-scale plain method): a 100 x 100 grid of AllLIFNeurons with the
test-small parameter distributions (10% inhibitory, 10% starters, DynamicLayout fractions,
seed 1), AllDSSynapses, ConnGrowth (epsilon 0.6, beta 0.1, rho 1e-4, targetRate 1.9), FP32,
dt = 1e-4 s, epochs of 10 s (1e5 steps), maxSynapsesPerNeuron 1000, deterministic noise.
As configured the growth rule cannot form synapses within ten epochs, so the network is
pre-grown: every radius is set to 8.9 and one growth update runs on the GPU, which gives
8,440,432 DS synapses (mean in-degree 844). The timed region is whole epochs of the hot path:
`--steps` simulation steps, plus the growth update and history compilation whenever an epoch
boundary is crossed.

Output: one JSON line. `value` = synapse-updates/s = sum over timed steps of active synapses /
wall time. `roofline` prices the dominant kernel (the fused block step kernel, K = 9 steps per
launch) with the bytes a temporally blocked kernel must move per launch: each synapse's meta,
delay-ring word and psr read once and ring/psr written once per block (20 B), the DS state of
every synapse that receives a spike in the block (W, r, u, lastSpike read, r, u, lastSpike
written: 36 B; the count of such synapses is measured from the spike counts and out-degrees over
the timed launches) and the neuron state once per block (44 B); over the average launch duration
measured with HIP events on the network's stream. `per_step` does the same for the per-step
kernels that replace the reference's advanceNeurons / advanceSynapses one for one (block_steps =
1: 24 B per DS synapse-step, 44 B per LIF neuron-step). `cpu_baseline` times the CPU oracle (the
reference's single-threaded algorithm restated in C) on a bounded sample of the same network.
`batched` summarises the batched form (the 256-network sweep with shared launches against a
sample of its networks run alone: simulated seconds per wall hour both ways) and `served` the
served form (the first 200 requests of the open-loop job stream: jobs/s, p50/p99 completion
latency); `--form batched` / `--form served` run either form alone at full size.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
DS_SYN_STEP_BYTES = 24            # per-step kernel: psr r/w 8, decay 4, ring r/w 8, source index 4
LIF_NEURON_STEP_BYTES = 44        # Vm, Vthresh, sp, I0, Inoise, C1, C2, refr read; Vm, sp, refr written
DS_SYN_BLOCK_BYTES = 20           # blocked kernel, per synapse and block: meta 4, ring r/w 8, psr r/w 8
DS_DELIVERY_BYTES = 36            # per delivered spike: W, r, u (12) + lastSpike (8) read; r, u, lastSpike written
# implicit-ring blocked kernel: per synapse and block meta 4 and psr r/w 8 (no ring); per delivered
# spike W 4 (the r, u, lastSpike of a delivery are shared by the synapses of one source, type and
# cohort); per neuron and block its state (44) and fire word r/w (16), and per live cohort and target
# class its DS state r/w (32) and prepared first-delivery factor (8)
IMPL_SYN_BLOCK_BYTES = 12
IMPL_DELIVERY_BYTES = 4
IMPL_NEURON_BLOCK_BYTES = LIF_NEURON_STEP_BYTES + 16
IMPL_COHORT_CLASS_BYTES = 40
GRID = 100
R0 = 8.9
M = 1000
EPOCH = 10.0
# At r0 = 8.9 the growth weights (W = overlap area * 1e-8) drive every neuron to its refractory
# limit (about 310 Hz excitatory, 450 Hz inhibitory with its 2 ms refractory period), so the
# spike-history budget maxFiringRate must exceed that for the reference's
# assert(spikeCount < max_spikes) (Neurons/AllSpikingNeurons.cpp:112) to hold.
MAX_RATE = 1000


def workload_cfg():
    from braingrid_amd import host
    c = host.base_config()
    c.update(width=GRID, height=GRID, max_synapses_per_neuron=M, max_firing_rate=MAX_RATE, epoch_duration=EPOCH,
             num_epochs=10, layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1, seed=1)
    return c


REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "bgref_dump")
REF_CONFIG_TEMPLATE = os.path.join(ROOT, "tests", "golden", "configs", "conn-lif-ds-dynamic.xml")


def reference_config_xml() -> str:
    """The benchmark workload as a reference parameter file: the committed 8 x 8 conn-lif-ds-dynamic
    config (test-small parameter distributions, DS synapses, ConnGrowth) at 100 x 100, M = 1000,
    maxFiringRate 1000, 10 s epochs, seed 1 (DynamicLayout 0.9 / 0.1 comes from the harness)."""
    with open(REF_CONFIG_TEMPLATE) as f:
        x = f.read()
    reps = [
        ('<x name="x">8</x><y name="y">8</y>', f'<x name="x">{GRID}</x><y name="y">{GRID}</y>'),
        ('<Tsim name="Tsim">0.5</Tsim><numSims name="numSims">3</numSims>',
         f'<Tsim name="Tsim">{EPOCH}</Tsim><numSims name="numSims">10</numSims>'),
        ('<maxFiringRate name="maxFiringRate">200</maxFiringRate><maxSynapsesPerNeuron name="maxSynapsesPerNeuron">40'
         '</maxSynapsesPerNeuron>',
         f'<maxFiringRate name="maxFiringRate">{MAX_RATE}</maxFiringRate><maxSynapsesPerNeuron '
         f'name="maxSynapsesPerNeuron">{M}</maxSynapsesPerNeuron>'),
        ('<Seed name="Seed"><value name="value">7</value></Seed>', '<Seed name="Seed"><value name="value">1</value></Seed>'),
    ]
    for a, b in reps:
        if a not in x:
            raise RuntimeError(f"reference config template changed: {a!r} not found")
        x = x.replace(a, b)
    return x


class ReferenceBaseline:
    """Times the reference's own CPU step (oracle/_ref/bgref_dump, BrainGrid's single-threaded model
    compiled from its sources) on the same pre-grown network: the reference's growth update from
    radius 8.9, then `steps` timed calls of IModel::advance, on one core. Started before the GPU
    phases (its setup is the reference's O(N^2) growth update on dense matrices, about a minute)
    and collected after them; only its timed steps enter the figure."""

    def __init__(self, steps: int):
        import subprocess
        import tempfile
        self.td = tempfile.TemporaryDirectory()
        cfg = os.path.join(self.td.name, "bench10k.xml")
        with open(cfg, "w") as f:
            f.write(reference_config_xml())
        self.t0 = time.perf_counter()
        self.proc = subprocess.Popen([REF_DUMP, cfg, os.devnull, "--dynamic-layout", "0.9", "0.1", "--time-pregrown",
                                      str(R0), str(steps)], cwd=self.td.name, stdout=subprocess.PIPE,
                                     stderr=subprocess.DEVNULL, text=True)

    def result(self) -> dict:
        out, _ = self.proc.communicate(timeout=900)
        total = time.perf_counter() - self.t0
        self.td.cleanup()
        if self.proc.returncode != 0:
            raise RuntimeError(f"bgref_dump exited with {self.proc.returncode}")
        res = json.loads(out.strip().splitlines()[-1])
        S, dt = res["synapses"], res["seconds"]
        return {"value": S * res["steps"] / dt, "unit": "synapse-updates/s", "cores": 1, "kind": "reference",
                "sample": f"the reference's CPU model (BrainGrid's SingleThreadedSpikingModel compiled from its "
                          f"sources, Norm noise), {res['steps']} timed steps of the same pre-grown {GRID * GRID}-neuron "
                          f"network ({S} synapses) after its own growth update from radius {R0}, {dt:.1f} s of "
                          f"steps ({total:.0f} s with its setup, run beside the GPU phases)"}


def cpu_baseline(budget_s: float):
    """Times the oracle (reference algorithm, reference noise mode, one core) on the same network."""
    from oracle import oracle_py as O
    cfg = O.default_config()
    cfg.update(width=GRID, height=GRID, max_synapses_per_neuron=M, max_firing_rate=MAX_RATE, epoch_duration=EPOCH,
               num_epochs=10, layout="DynamicLayout", frac_exc=0.9, frac_starter=0.1, seed=1)
    net = O.OracleNet(cfg)
    net.set_radii(R0)
    net.update_connections()
    S = net.total_synapses
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        net.advance(5)
        steps += 5
    dt = time.perf_counter() - t0
    net.close()
    return {"value": S * steps / dt, "unit": "synapse-updates/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (reference algorithm, Norm noise, sequential summation, 1 thread), "
                      f"{steps} steps of the same pre-grown {GRID * GRID}-neuron network ({S} synapses), "
                      f"{dt:.1f} s"}


def oracle_sweep_net(spec: dict):
    """The oracle in the reference's mode (Norm noise, C library math, sequential summation) holding
    the network of one sweep or served job specification."""
    from oracle import oracle_py as O
    from braingrid_amd import networks as NW
    from braingrid_amd import sweep as SW
    if spec["kind"] == "lif_growth":
        c = O.default_config()
        c.update(width=spec["side"], height=spec["side"], layout="DynamicLayout", frac_exc=spec["fE"],
                 frac_starter=0.1, target_rate=spec["tR"], seed=spec["seed"], epoch_duration=SW.EPOCH_S,
                 num_epochs=int(round(spec["sim_s"] / SW.EPOCH_S)), max_synapses_per_neuron=SW.LIF_M,
                 max_firing_rate=SW.lif_max_rate(spec))
        net = O.OracleNet(c)
        if spec.get("r0"):
            net.set_radii(spec["r0"])
            net.update_connections()
        return net
    from tests import helpers as H
    N = spec["side"] ** 2
    neu, syn, meta = NW.izh_stdp_arrays(N, spec["seed"], k=min(100, N - 1), max_rate=SW.IZH_MAX_RATE,
                                        epoch_s=SW.EPOCH_S)
    ora, _ = H.izh_static_oracle(neu, syn, meta, spec["seed"], epoch_s=SW.EPOCH_S, max_rate=SW.IZH_MAX_RATE,
                                 noise_mode=0, math_mode=0, sum_mode=0,
                                 input=dict(kind="poisson", fr_mean=SW.POISSON_HZ, weight=SW.POISSON_WEIGHT))
    return ora


def forms_cpu_baseline(specs: list, sample: list, steps: int, threads: int) -> dict:
    """CPU baseline of a many-network workload: every sampled network in the oracle (reference
    mode, one thread each, `threads` at a time) runs `steps` steps and one epoch-end update; its
    CPU time per simulated second is steps scaled to a 1 s epoch plus that update. The workload's
    CPU time is the sample's scaled to all networks per family by neuron-seconds (neurons x
    simulated seconds; the Izhikevich networks have ~100 synapses per neuron), and its rate on
    `threads` cores assumes one network per core with perfect packing."""
    from concurrent.futures import ThreadPoolExecutor

    def cost(spec):
        ora = oracle_sweep_net(spec)
        t0 = time.thread_time()
        ora.advance(steps)
        t1 = time.thread_time()
        ora.update_connections()
        ora.update_history()
        t2 = time.thread_time()
        ora.close()
        return (t1 - t0) * (1.0 / (steps * 1e-4)) + (t2 - t1)   # core-seconds per simulated second

    t_start = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        costs = list(ex.map(cost, sample))
    wall = time.perf_counter() - t_start
    def nsec(s):
        return s["side"] ** 2 * s["sim_s"]

    core_s_all = 0.0
    for kind in {s["kind"] for s in specs}:
        pop = sum(nsec(s) for s in specs if s["kind"] == kind)
        got = [(c, s) for c, s in zip(costs, sample) if s["kind"] == kind]
        if not got:
            raise RuntimeError(f"CPU baseline sample has no {kind} network")
        core_s_all += sum(c * s["sim_s"] for c, s in got) * pop / sum(nsec(s) for _, s in got)
    sim_all = sum(s["sim_s"] for s in specs)
    return {"value": sim_all / (core_s_all / threads) * 3600.0, "unit": "sim-s/h", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (reference algorithm: Norm noise, sequential summation, Poisson input), "
                      f"{len(sample)} of {len(specs)} networks, {steps} steps and one epoch-end update each, one "
                      f"network per thread on {threads} threads ({wall:.1f} s); the workload's core-seconds scaled "
                      f"from the sample per family by neuron-seconds"}


KERNEL_SOURCES = ("bghip_kernels.h", "bghip_implicit.h", "bghip_device.h", "bghip_launch.h", "bghip_impl_lif.hip")


def kernel_source_sha() -> str:
    """Digest of the sources the headline kernel is compiled from (and the detmath header)."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "braingrid_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    with open(os.path.join(ROOT, "include", "bg_detmath.h"), "rb") as fh:
        h.update(fh.read())
    return h.hexdigest()[:16]


def load_traffic():
    """The counter-measured HBM bytes per launch of the headline kernel, if the profile was taken
    on the kernel sources of this tree (otherwise None: a stale figure is not reported)."""
    p = os.path.join(ROOT, "profiles", "fused_impl_10k_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    if t.get("kernel_source_sha") != kernel_source_sha():
        return None
    return t


def fused_phase(model, steps: int) -> dict:
    """Times `steps` steps of back-to-back fused launches with HIP events on the network's stream
    and counts the synapses that received a spike over them (spike counts before and after, times
    the out-degrees; the phase must lie inside one epoch)."""
    import numpy as np
    outdeg = model.net.download_index_map()["outgoingSynapseCount"].astype(np.int64)
    counts0 = model.net.download_spikes()["spikeCount"].astype(np.int64)
    model.net.timing(True)
    model.advance(steps)
    avg_ms, launches = model.net.timing(False)
    delivered = deliveries_between(model.net, counts0, outdeg)
    return {"avg_ms": avg_ms, "launches": launches, "delivered": delivered, "synapses": model.net.total_synapses,
            "implicit": model.net.implicit}


def blocked_bytes(S: int, delivered_per_launch: float, N: int, implicit: bool, cohorts: int = 1) -> float:
    """Minimum HBM bytes of one fused launch of the blocked kernel (implicit or per-synapse rings)."""
    if implicit:
        return (IMPL_SYN_BLOCK_BYTES * S + IMPL_DELIVERY_BYTES * delivered_per_launch
                + (IMPL_NEURON_BLOCK_BYTES + 2 * cohorts * IMPL_COHORT_CLASS_BYTES) * N)
    return DS_SYN_BLOCK_BYTES * S + DS_DELIVERY_BYTES * delivered_per_launch + LIF_NEURON_STEP_BYTES * N


def explicit_fused_roofline(steps: int) -> dict:
    """The same pre-grown network on the per-synapse ring kernels (a context created with
    BGHIP_IMPLICIT=0): its fused launch time, priced with the explicit blocked byte model."""
    from braingrid_amd import bghip as B
    from braingrid_amd import host
    old = os.environ.get("BGHIP_IMPLICIT")
    os.environ["BGHIP_IMPLICIT"] = "0"
    try:
        ctx = B.Context(0)
    finally:
        if old is None:
            del os.environ["BGHIP_IMPLICIT"]
        else:
            os.environ["BGHIP_IMPLICIT"] = old
    m = host.Model(ctx, workload_cfg())
    m.set_radii(R0)
    m.update_connections()
    m.advance(900)
    m.net.sync()
    ph = fused_phase(m, steps)
    m.close()
    ctx.close()
    N = GRID * GRID
    per = ph["delivered"] / max(1, ph["launches"])
    b = blocked_bytes(ph["synapses"], per, N, False)
    ach = b / (ph["avg_ms"] * 1e-3) if ph["avg_ms"] > 0 else 0.0
    return {"kernel": "k_fused_block<LIF,DS,1,9> (per-synapse rings)", "implicit": ph["implicit"],
            "avg_launch_ms": ph["avg_ms"], "launches": ph["launches"], "algorithmic_bytes_per_launch": b,
            "achieved": ach / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
            "steps_per_launch": steps / max(1, ph["launches"])}


def deliveries_between(net, counts0, outdeg):
    """Synapses that received a spike between two spike-count snapshots of the same epoch: every
    spike of a source reaches each of its outgoing synapses once (a source's refractory period is
    longer than a block, so no synapse receives two spikes in one block)."""
    import numpy as np
    c1 = net.download_spikes()["spikeCount"].astype(np.int64)
    return int(((c1 - counts0) * outdeg).sum())


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batched_main(args):
    print(json.dumps(batched_result(args)), flush=True)


def served_main(args):
    print(json.dumps(served_result(args)), flush=True)


def batched_result(args):
    """The batched form: a sweep of independent networks run one at a time with their own
    launches, then all together with shared launches; same initial states, same results."""
    from braingrid_amd import bghip as B
    from braingrid_amd import sweep as SW

    ctx = B.Context(0)
    specs = SW.sweep_specs(args.nets, args.seed, sim_s=args.sim_seconds, lif_r0=args.lif_r0)
    N_total = sum(s["side"] ** 2 for s in specs)
    sim_total = sum(s["sim_s"] for s in specs)
    log(f"sweep: {len(specs)} networks, {N_total} neurons, {sim_total} simulated s")
    # alone: every `alone_stride`-th network of the sweep, run one at a time; the sweep's alone
    # time is that sample's time scaled by the sampling ratio
    stride = max(1, args.alone_stride)
    sample = specs[::stride]
    nets = SW.build(ctx, sample, log)
    t_alone_sample = SW.run_alone(nets, log)
    fp_alone = {x.spec["id"]: x.fingerprint() for x in nets}
    for x in nets:
        x.close()
    t_alone = t_alone_sample * len(specs) / len(sample)
    nets = SW.build(ctx, specs, log)
    S_total = sum(x.synapses for x in nets)
    S_lif = sum(x.synapses for x in nets if x.kind == "lif_growth")
    shapes = []
    t_batched = SW.run_batched(ctx, nets, log, tune_steps=args.tune_steps, shapes=shapes)
    fp_batched = {x.spec["id"]: x.fingerprint() for x in nets if x.spec["id"] in fp_alone}
    errs = sum(x.net.error_bits() != 0 for x in nets)
    for x in nets:
        x.close()
    same = sum(fp_alone[k] == fp_batched[k] for k in fp_alone)
    value = sim_total / t_batched * 3600.0
    fused = fused_group_roofline(specs, S_lif, shapes) if args.lif_r0 else None
    out = {
        "metric": "sims/hr (simulated network-seconds per wall hour) of a batched sweep",
        "value": value,
        "unit": "sim-s/h",
        "n_gpus": 1,
        "steps": int(round(args.sim_seconds / 1e-4)),
        "warmup": 0,
        "ms_per_step": t_batched * 1e3 / (args.sim_seconds / 1e-4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sweep; deterministic Philox noise and Poisson input)",
        "config": {"workload": f"batched sweep: {len(specs)} networks, sides log-uniform 10..120, "
                               f"LIF+DS+growth{f' pre-grown to r0={args.lif_r0}' if args.lif_r0 else ''} / "
                               f"Izhikevich+DynamicSTDP+Poisson, {args.sim_seconds} s each",
                   "networks": len(specs), "neurons": N_total, "synapses": int(S_total),
                   "lif_growth": sum(s["kind"] == "lif_growth" for s in specs),
                   "izh_stdp": sum(s["kind"] == "izh_stdp" for s in specs), "seed": args.seed},
        "alone_wall_s": t_alone,
        "alone_sample": f"every {stride}th network ({len(sample)} of {len(specs)}) run alone in "
                        f"{t_alone_sample:.1f} s, scaled by {len(specs) / len(sample):.2f}",
        "batched_wall_s": t_batched,
        "alone_value": sim_total / t_alone * 3600.0,
        "speedup_vs_alone": t_alone / t_batched,
        "identical_results": f"{same}/{len(sample)} sampled networks bit-identical alone and batched",
        "networks_with_device_errors": int(errs),
        "autotune_steps": args.tune_steps,
        "launch_shapes": [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in s.items()} for s in shapes],
    }
    if fused is not None:
        out["fused_group"] = fused
    ctx.close()
    if not args.no_cpu:
        cs = specs[:: max(1, len(specs) // max(1, args.cpu_networks))]
        out["cpu_baseline"] = forms_cpu_baseline(specs, cs, args.cpu_steps, args.cpu_threads)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    return out


def fused_group_roofline(specs, S_lif: int, shapes: list) -> dict | None:
    """The batch's fused Spiking/DS group (the pre-grown growth networks, one shared implicit-ring
    launch per 9-step block): its time per step as the launch-shape autotuner measured it (HIP
    events on the batch stream over its trial steps), priced with the implicit byte model without
    the per-delivery term (a lower bound on the bytes)."""
    g = [x for x in shapes if x.get("mode") == 1]
    if not g or not g[0].get("step_us"):
        return None
    g = g[0]
    N = g["work_items"]
    K = 9
    launch_bytes = IMPL_SYN_BLOCK_BYTES * S_lif + (IMPL_NEURON_BLOCK_BYTES + 2 * IMPL_COHORT_CLASS_BYTES) * N
    step_s = g["step_us"] * 1e-6
    ach = launch_bytes / K / step_s
    return {"kernel": "k_fused_impl_batch<LIF,DS,9>", "networks": g["networks"], "neurons": N, "synapses": int(S_lif),
            "block": g["block"], "step_us": g["step_us"], "synapse_updates_per_s": S_lif / step_s,
            "byte_model": f"per 9-step launch: {IMPL_SYN_BLOCK_BYTES} B x synapses + "
                          f"{IMPL_NEURON_BLOCK_BYTES + 2 * IMPL_COHORT_CLASS_BYTES} B x neurons (deliveries not counted)",
            "achieved": ach / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
            "timing": "the autotuner's trial steps of the group's chosen shape (HIP events on the batch stream)"}


def served_result(args):
    """The served form: an open-loop Poisson stream of simulation jobs, continuously batched."""
    from braingrid_amd import bghip as B
    from braingrid_amd import serve as S

    ctx = B.Context(0)
    free, total = ctx.mem_info()
    budget = int(free * args.hbm_fraction)
    jobs = S.served_requests(args.requests, rate=args.rate, seed=args.seed)
    eng = S.DeviceEngine(ctx, tune_steps=args.tune_steps)
    srv = S.Server(engine=eng, budget=budget, log=log)
    log(f"served: {len(jobs)} requests at {args.rate}/s, budget {budget / 2**30:.1f} GiB of {total / 2**30:.1f} GiB")
    srv.run(jobs, progress_every=20.0)
    st = srv.stats(total_bytes=total)
    eng.close()
    alone = served_alone(ctx, jobs, S, max(1, args.served_alone_stride))
    ctx.close()
    out = {
        "metric": "completed simulation jobs/s of an open-loop Poisson job stream (continuous batching)",
        "value": st["throughput_jobs_per_s"],
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": int(sum(j.spec["sim_s"] for j in jobs if j.state == S.DONE) / 1e-4),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded request stream; deterministic Philox noise and Poisson input)",
        "config": {"workload": f"served: {args.requests} requests, Poisson arrivals at {args.rate}/s, sides "
                               "log-uniform 30..200, 70% LIF+DS+growth / 30% Izhikevich+DynamicSTDP+Poisson, "
                               "1..20 s each", "requests": args.requests, "rate": args.rate, "seed": args.seed,
                   "hbm_budget_bytes": budget},
        "latency_p50_s": st["latency_p50_s"],
        "latency_p99_s": st["latency_p99_s"],
        "latency_mean_s": st["latency_mean_s"],
        "sim_seconds_per_hour": st["sim_seconds_per_hour"],
        "hbm_occupancy_mean": st.get("hbm_occupancy_mean"),
        "hbm_occupancy_max": st.get("hbm_occupancy_max"),
        "wall_s": st["wall_s"],
        "counts": st["counts"],
        "phase_s": st["host_phase_s"],
    }
    out.update(alone)
    if alone.get("alone_sim_seconds_per_hour"):
        out["speedup_vs_alone"] = st["sim_seconds_per_hour"] / alone["alone_sim_seconds_per_hour"]
    if not args.no_cpu:
        specs = [j.spec for j in sorted(jobs, key=lambda j: j.id) if j.state == S.DONE]
        if specs:
            cs = specs[:: max(1, len(specs) // max(1, args.cpu_networks))]
            out["cpu_baseline"] = forms_cpu_baseline(specs, cs, args.cpu_steps, args.cpu_threads)
            out["speedup_vs_cpu"] = st["sim_seconds_per_hour"] / out["cpu_baseline"]["value"]
    return out


def served_alone(ctx, jobs, S, stride: int) -> dict:
    """Every `stride`-th completed job of the stream run again on its own (plain per-network
    launches, epoch work after every epoch, one job after another), timed, and its output compared
    with the one it got in the served mix."""
    from braingrid_amd import sweep as SW
    import numpy as np

    done = [j for j in sorted(jobs, key=lambda j: j.id) if j.state == S.DONE]
    sample = done[::stride]
    if not sample:
        return {}
    same = 0
    wall = 0.0
    for j in sample:
        x = SW.SweepNet(ctx, j.spec)
        x.net.sync()
        t0 = time.perf_counter()
        for _ in range(x.epochs):
            x.net.advance(x.steps_per_epoch)
            x.end_epoch()
        x.net.sync()
        wall += time.perf_counter() - t0
        r = x.summary()
        same += all(np.array_equal(np.asarray(r[k]), np.asarray(j.result[k])) for k in r if k in j.result)
        x.close()
    sim = sum(j.spec["sim_s"] for j in sample)
    return {"alone_sim_seconds_per_hour": sim / wall * 3600.0,
            "alone_sample": f"every {stride}th completed job ({len(sample)} of {len(done)}, {sim:.0f} simulated s) "
                            f"run alone in {wall:.1f} s",
            "identical_results": f"{same}/{len(sample)} sampled jobs with the same output alone and served"}


def per_step_roofline(ctx, steps: int) -> dict:
    """The same pre-grown network advanced with the per-step kernels (block_steps = 1: one neuron
    launch and one synapse launch per step, the reference's advanceNeurons / advanceSynapses
    shape), timed with HIP events over `steps` steps; priced at the per-step bytes."""
    from braingrid_amd import host
    cfg = workload_cfg()
    cfg["block_steps"] = 1
    m = host.Model(ctx, cfg)
    m.set_radii(R0)
    m.update_connections()
    m.advance(200)
    m.net.sync()
    S = m.net.total_synapses
    N = GRID * GRID
    m.net.timing(True)
    m.advance(steps)
    avg_ms, launches = m.net.timing(False)
    m.close()
    step_ms = avg_ms * launches / max(1, steps)
    step_bytes = DS_SYN_STEP_BYTES * S + LIF_NEURON_STEP_BYTES * N
    achieved = step_bytes / (step_ms * 1e-3) if step_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "kernels": "k_neuron_step<LIF> + k_synapse_step<DS,1> per step",
            "ms_per_step": step_ms, "steps": steps, "synapses": int(S),
            "synapse_updates_per_s": S / (step_ms * 1e-3) if step_ms > 0 else 0.0,
            "byte_model": f"per step: {DS_SYN_STEP_BYTES} B x synapses + {LIF_NEURON_STEP_BYTES} B x neurons",
            "algorithmic_bytes_per_step": step_bytes}


def forms_summary(args):
    """The batched and served forms at reduced sample sizes, summarised for the headline line: the
    full 256-network sweep batched against every `forms_alone_stride`-th network alone, and the
    first `forms_requests` requests of the served stream."""
    out = {}
    a = argparse.Namespace(**vars(args))
    a.alone_stride = args.forms_alone_stride
    try:
        r = batched_result(a)
        out["batched"] = {k: r[k] for k in ("value", "unit", "alone_value", "speedup_vs_alone", "batched_wall_s",
                                            "alone_wall_s", "alone_sample", "identical_results",
                                            "networks_with_device_errors")}
        out["batched"]["workload"] = r["config"]["workload"]
        for k in ("cpu_baseline", "speedup_vs_cpu"):
            if k in r:
                out["batched"][k] = r[k]
        out["batched"]["launch_shapes"] = [{k: s[k] for k in ("mode", "networks", "block", "dests_per_wave",
                                                              "block_stream", "syn_blocks_per_shard",
                                                              "sum_blocks_per_shard", "tuned", "step_us")}
                                           for s in r["launch_shapes"]]
    except Exception as e:  # the headline line is still printed
        out["batched"] = {"error": repr(e)}
    a = argparse.Namespace(**vars(args))
    a.requests = args.forms_requests
    try:
        r = served_result(a)
        out["served"] = {k: r[k] for k in ("value", "unit", "latency_p50_s", "latency_p99_s", "latency_mean_s",
                                           "sim_seconds_per_hour", "hbm_occupancy_mean", "hbm_occupancy_max",
                                           "wall_s", "counts", "alone_sim_seconds_per_hour", "alone_sample",
                                           "speedup_vs_alone", "identical_results", "cpu_baseline",
                                           "speedup_vs_cpu") if k in r}
        out["served"]["workload"] = r["config"]["workload"]
    except Exception as e:
        out["served"] = {"error": repr(e)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--form", choices=("single", "batched", "served"), default="single",
                    help="single: the 10k-neuron headline; batched: the multi-network sweep; "
                         "served: the open-loop job stream")
    ap.add_argument("--nets", type=int, default=256, help="batched form: networks in the sweep")
    ap.add_argument("--sim-seconds", type=float, default=5.0, help="batched form: simulated seconds per network")
    ap.add_argument("--seed", type=int, default=2024, help="batched form: sweep seed")
    ap.add_argument("--lif-r0", type=float, default=None,
                    help="batched form: start the growth networks pre-grown to this radius (2.8: about 100 "
                         "synapses per neuron), so they run the fused kernels")
    ap.add_argument("--requests", type=int, default=2000, help="served form: requests in the stream")
    ap.add_argument("--rate", type=float, default=20.0, help="served form: mean arrival rate (requests/s)")
    ap.add_argument("--hbm-fraction", type=float, default=0.9,
                    help="served form: share of the free HBM the resident networks may occupy")
    ap.add_argument("--alone-stride", type=int, default=4,
                    help="batched form: run every k-th network alone for the one-at-a-time comparison")
    ap.add_argument("--served-alone-stride", type=int, default=10,
                    help="served form: run every k-th completed job alone for the one-at-a-time comparison")
    ap.add_argument("--tune-steps", type=int, default=8,
                    help="batched/served forms: launch-shape autotuner, timed steps per candidate (0: off)")
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--roofline-steps", type=int, default=20000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-ref-steps", type=int, default=300,
                    help="single form: timed steps of the reference's own CPU model (oracle/_ref/bgref_dump)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-networks", type=int, default=12,
                    help="batched/served forms: networks of the workload timed in the CPU oracle")
    ap.add_argument("--cpu-steps", type=int, default=200,
                    help="batched/served forms: oracle steps per sampled network (plus one epoch-end update)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="batched/served forms: host threads of the CPU baseline (one network each)")
    ap.add_argument("--no-explicit", action="store_true",
                    help="single form: skip timing the per-synapse-ring fused kernel on the same network")
    ap.add_argument("--explicit-steps", type=int, default=9000,
                    help="single form: steps of the per-synapse-ring fused kernel's timing")
    ap.add_argument("--no-per-step", action="store_true",
                    help="single form: skip the per-step kernels' roofline on the same network")
    ap.add_argument("--per-step-steps", type=int, default=2000,
                    help="single form: steps of the per-step kernels' roofline sub-phase")
    ap.add_argument("--no-forms", action="store_true",
                    help="single form: skip the batched and served summaries added to the headline line")
    ap.add_argument("--forms-alone-stride", type=int, default=16,
                    help="single form's batched summary: run every k-th sweep network alone")
    ap.add_argument("--forms-requests", type=int, default=200,
                    help="single form's served summary: requests of the stream")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy 1 GiB device-to-device and exit (reference for HBM traffic counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this bench measures one GPU: --gpus above 1 is not supported (the batched "
                                   "and served forms would shard whole networks across GPUs; one network would "
                                   "split by destination neuron and exchange fire masks once per block)"}))
        sys.exit(2)
    if args.form == "batched":
        batched_main(args)
        return
    if args.form == "served":
        served_main(args)
        return

    from braingrid_amd import bghip as B
    from braingrid_amd import host

    # the reference's own CPU step on the same network, when its harness is built (it travels with
    # the tree): started now, collected after the GPU phases of this network
    ref_base = ReferenceBaseline(args.cpu_ref_steps) if not args.no_cpu and os.path.exists(REF_DUMP) else None
    ctx = B.Context(0)
    cfg = workload_cfg()
    model = host.Model(ctx, cfg)
    model.set_radii(R0)
    t_grow0 = time.perf_counter()
    st = model.update_connections()
    model.net.sync()
    t_pregrow = time.perf_counter() - t_grow0
    N = GRID * GRID
    S0 = st.total_synapses
    steps_per_epoch = int(round(EPOCH / 1e-4))

    # warm-up (also settles clocks); the epoch counter runs on
    model.advance(args.warmup)
    model.net.sync()
    done = args.warmup

    # roofline sub-phase: back-to-back fused launches only, timed with HIP events on the net stream;
    # the spike counts before and after give the synapses that received a spike in these blocks
    rl_steps = min(args.roofline_steps, steps_per_epoch - done % steps_per_epoch - 1)
    ph = fused_phase(model, rl_steps)
    avg_ms, launches, delivered, implicit = ph["avg_ms"], ph["launches"], ph["delivered"], ph["implicit"]
    done += rl_steps
    S_rl = ph["synapses"]
    K = max(1, round(rl_steps / max(1, launches)))

    # timed region: whole hot path including growth/history at epoch boundaries
    synapse_updates = 0
    growth_updates = 0
    model.net.sync()
    t0 = time.perf_counter()
    remaining = args.steps
    while remaining > 0:
        to_boundary = steps_per_epoch - (done % steps_per_epoch)
        n = min(remaining, to_boundary)
        S_now = model.net.total_synapses
        model.advance(n)
        synapse_updates += S_now * n
        done += n
        remaining -= n
        if done % steps_per_epoch == 0:
            model.update_connections()
            model.update_history()
            growth_updates += 1
    model.net.wait()   # the stream drained; its error word is read and checked just below, untimed
    wall = time.perf_counter() - t0
    model.net.sync()   # raises if the timed steps set a device error bit
    err = model.net.error_bits()
    neu = model.net.download_neurons(int(EPOCH * MAX_RATE))
    in_epoch = done % steps_per_epoch or steps_per_epoch
    rate_hz = float(neu["spikeCount"].mean()) / (in_epoch * 1e-4)

    value = synapse_updates / wall
    per_launch_delivered = delivered / max(1, launches)
    # the pre-growth update created every synapse: one cohort
    launch_bytes = blocked_bytes(S_rl, per_launch_delivered, N, implicit, cohorts=1)
    ref_bytes = blocked_bytes(S_rl, per_launch_delivered, N, False)
    achieved = launch_bytes / (avg_ms * 1e-3) if avg_ms > 0 else 0.0
    traffic = load_traffic()
    # the instantiation the full blocks run: exactly 9 steps for a 32-bit ring, else up to 16 / 32
    kb = 9 if K == 9 else (16 if K <= 16 else 32)
    kernel_name = ("k_fused_impl<LIF,DS,%d>" % kb) if implicit else ("k_fused_block<LIF,DS,1,%d>" % kb)
    out = {
        "metric": "synapse-updates/s at 10k-LIF growth; batched vs per-net sims/hr; p50/p99",
        "value": value,
        "unit": "synapse-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded parameter draws, deterministic Philox noise, pre-grown radii 8.9)",
        "config": {"workload": "10k LIF + DS + ConnGrowth, 100x100 grid pre-grown to r0=8.9, dt=1e-4 s, "
                               "epoch 1e5 steps", "neurons": N, "synapses": int(S0),
                   "max_synapses_per_neuron": M, "max_firing_rate": MAX_RATE, "block_steps": K,
                   "mean_rate_hz_last_epoch": round(rate_hz, 2), "growth_updates_timed": growth_updates,
                   "pregrow_s": round(t_pregrow, 4),
                   "deviations": "maxFiringRate 1000 (the pre-grown network fires at ~300 Hz); per-type "
                                 "delays of the reference (EE 1.5 ms = 16 steps, others 0.8 ms = 9 steps); "
                                 "INH Trefract 2 ms as in the reference"},
        "neuron_steps_per_s": N * args.steps / wall,
        "sim_seconds_per_hour": args.steps * 1e-4 / wall * 3600.0,
        "timed_region": f"{args.steps} steps in whole calls of the host model, growth update and history "
                        "compilation at each epoch boundary crossed, bracketed by device synchronisation",
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel": kernel_name, "avg_launch_ms": avg_ms, "launches": launches,
                     "steps_per_launch": rl_steps / max(1, launches),
                     "byte_model": (f"per launch: {IMPL_SYN_BLOCK_BYTES} B x synapses (meta, psr r/w once per block; "
                                    f"the delay rings are the sources' fire words) + {IMPL_DELIVERY_BYTES} B x synapses "
                                    f"receiving a spike in the block (W) + {IMPL_NEURON_BLOCK_BYTES} B x neurons "
                                    f"(state, fire word) + {2 * IMPL_COHORT_CLASS_BYTES} B x neurons per creation "
                                    f"cohort (DS state per source and target class, 1 cohort here)") if implicit else
                                   (f"per launch: {DS_SYN_BLOCK_BYTES} B x synapses (meta, ring r/w, psr r/w once "
                                    f"per block) + {DS_DELIVERY_BYTES} B x synapses receiving a spike in the block "
                                    f"(W, r, u, lastSpike read; r, u, lastSpike written) + "
                                    f"{LIF_NEURON_STEP_BYTES} B x neurons"),
                     "reference_blocked_bytes_per_launch": ref_bytes,
                     "reference_blocked_rate_frac": (ref_bytes / (avg_ms * 1e-3) / HBM_PEAK) if avg_ms > 0 else None,
                     "reference_blocked_note": "the bytes a per-synapse-ring blocked kernel moves for the same launch "
                                               "(20 B/synapse + 36 B/delivery + 44 B/neuron) over this launch's time: "
                                               "a rate of delivered work in those units, not a bandwidth",
                     "synapses": int(S_rl), "delivered_per_launch": per_launch_delivered,
                     "delivered_fraction": per_launch_delivered / max(1, S_rl),
                     "algorithmic_bytes_per_launch": launch_bytes,
                     "traffic_GBps": (traffic["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9) if traffic and avg_ms > 0 else None,
                     "traffic_frac": (traffic["bytes_per_launch"] / (avg_ms * 1e-3) / HBM_PEAK) if traffic and avg_ms > 0 else None,
                     "traffic_source": ("profiles/fused_impl_10k_traffic.json (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                                        "copy-calibrated, same kernel sources)") if traffic else
                                       "no counter profile of the current kernel sources"},
        "device_error_bits": int(err),
    }
    model.close()
    if implicit and not args.no_explicit:
        out["explicit_fused"] = explicit_fused_roofline(args.explicit_steps)
    if not args.no_per_step:
        out["per_step"] = per_step_roofline(ctx, args.per_step_steps)
    ctx.close()
    if not args.no_cpu:
        ref = None
        if ref_base is not None:
            try:
                ref = ref_base.result()
            except Exception as e:   # the port's figure stands; say why
                out["cpu_reference_error"] = str(e)
        port = cpu_baseline(args.cpu_seconds)
        if ref is not None:
            out["cpu_baseline"] = ref
            out["cpu_port"] = port
        else:
            out["cpu_baseline"] = port
    if not args.no_forms:
        out.update(forms_summary(args))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Served form: simulation jobs arriving over time, run with continuous batching on one GPU.

A job is one network specification (the families of braingrid_amd.sweep: grid side, lif_growth or
izh_stdp, tR, fE, seed, simulated seconds). The server keeps every admitted network resident in
HBM from admission to completion and advances all of them together, one 1 s epoch per tick, with
shared launches (bghip_batch_*); between ticks it runs each network's epoch work (growth update,
history), retires finished jobs, and admits queued ones. Admission is by device footprint: the sum
of the resident networks' arenas (bghip_net_footprint; growth is tiled, so there is no N^2 term)
never exceeds the HBM budget. A job whose footprint alone exceeds the budget is rejected; a job
whose specification is invalid, or that raises a device error, fails on its own; a job can be
cancelled while queued or running. A network's results do not depend on which other networks
shared its launches.

With eviction on, a queued job that has been passed over too often does not wait for residents to
finish: at the end of a tick (an epoch boundary) the most recently admitted residents are saved to
checkpoint files (bghip_net_save) and released until it fits; they rejoin the queue behind it and
continue from their checkpoints (bghip_net_load) bit for bit when readmitted. A job is evicted at
most `max_evictions` times.

The scheduling logic talks to an engine (DeviceEngine here; tests substitute a CPU one), so the
policy is testable without a GPU.
"""
from __future__ import annotations

import heapq
import os
import shutil
import tempfile
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import sweep as SW

QUEUED, ACTIVE, DONE, FAILED, REJECTED, CANCELLED = "queued", "active", "done", "failed", "rejected", "cancelled"


@dataclass
class Job:
    id: int
    spec: dict
    arrival: float                  # seconds after the stream started
    state: str = QUEUED
    plan: object = None
    handle: object = None
    footprint: int = 0
    admitted: float | None = None   # absolute clock times
    finished: float | None = None
    epochs_done: int = 0
    epochs: int = 0
    skipped: int = 0                # times a later job was admitted ahead of it
    error: str | None = None
    result: dict | None = None
    checkpoint: str | None = None   # file of an evicted job's state, until it is readmitted
    evictions: int = 0


def served_requests(n: int = 2000, rate: float = 20.0, seed: int = 7, side_lo: int = 30, side_hi: int = 200,
                    frac_lif: float = 0.7, dur_lo: int = 1, dur_hi: int = 20):
    """Open-loop request stream: exponential inter-arrival times of mean 1/rate, sides log-uniform
    on [side_lo, side_hi], 70% lif_growth / 30% izh_stdp, tR and fE as in the sweep, durations
    uniform on whole seconds dur_lo..dur_hi (the growth epoch is 1 s)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        side = int(round(float(np.exp(rng.uniform(np.log(side_lo), np.log(side_hi))))))
        kind = "lif_growth" if rng.random() < frac_lif else "izh_stdp"
        spec = dict(id=k, side=side, kind=kind, tR=float(rng.choice(SW.TR_CHOICES)),
                    fE=float(rng.choice(SW.FE_CHOICES)), seed=SW.net_seed(seed, k),
                    sim_s=float(rng.integers(dur_lo, dur_hi + 1)))
        out.append(Job(id=k, spec=spec, arrival=t))
    return out


class DeviceEngine:
    """Networks on the GPU through the C ABI. A job's result is its recorder output (spike-count
    histograms, final radii); with fingerprint=True it also carries a digest of the whole final
    network state (a full download, for comparisons in tests)."""

    def __init__(self, ctx, fingerprint: bool = False, tune_steps: int = SW.TUNE_STEPS):
        from . import bghip as B
        self.ctx = ctx
        self.batch = B.Batch(ctx)
        self.batch.autotune(tune_steps)
        self._members = None
        self.fingerprint = fingerprint

    def total_bytes(self) -> int:
        return self.ctx.mem_info()[1]

    def plan(self, spec):
        return SW.plan(spec)

    def footprint(self, plan) -> int:
        return SW.footprint(plan)

    def create(self, spec, plan):
        return SW.SweepNet(self.ctx, spec, plan)

    def save(self, h, path: str):
        h.save(path)

    def restore(self, spec, plan, path: str):
        return SW.SweepNet.restore(self.ctx, spec, path)

    def advance(self, handles, steps: int):
        # the members are compared by identity with references held here, so a released handle
        # cannot be mistaken for a new one that happens to reuse its address
        handles = list(handles)
        if (self._members is None or len(handles) != len(self._members)
                or any(a is not b for a, b in zip(handles, self._members))):
            self.batch.set([h.net for h in handles])
            self._members = handles
        self.batch.advance(steps)

    def end_epoch(self, h):
        h.end_epoch()

    def check(self, h):
        bits = h.net.error_bits()
        if bits:
            raise RuntimeError(f"device error bits 0x{bits:x}")

    def result(self, h) -> dict:
        out = h.summary()
        out["step"] = h.net.step
        if self.fingerprint:
            out["fingerprint"] = h.fingerprint()
        return out

    def release(self, h):
        if self._members is not None and any(m is h for m in self._members):
            # the batch drops a destroyed member itself; forget the cached membership too
            self._members = None
        h.close()

    def close(self):
        self.batch.close()


@dataclass
class Server:
    engine: object
    budget: int                          # device bytes the resident networks may occupy
    epoch_steps: int = int(round(SW.EPOCH_S / SW.DT))
    lookahead: int = 16                  # queued jobs considered per admission round
    max_skips: int = 32                  # after this many bypasses the oldest job blocks the queue
    evict: bool = False                  # make room for a blocked job by checkpointing residents
    max_evictions: int = 1
    spill_dir: str | None = None         # where checkpoints of evicted jobs go (a temporary folder)
    clock: object = time.perf_counter
    log: object = None
    plan_threads: int = 8                # queued jobs planned in parallel (numpy releases the lock)
    plan_ahead: int = 128                # queued jobs planned per tick
    jobs: dict = field(default_factory=dict)
    queue: list = field(default_factory=list)      # heap of (arrival, id)
    active: list = field(default_factory=list)     # admitted jobs in admission order
    used: int = 0
    t0: float = 0.0
    occupancy: list = field(default_factory=list)  # (clock, used bytes) per tick
    ticks: int = 0
    blocked: Job | None = None                     # the job eviction makes room for
    evicted: int = 0
    restored: int = 0
    spill_bytes: int = 0
    phase_s: dict = field(default_factory=lambda: {"admit": 0.0, "advance": 0.0, "epoch_issue": 0.0,
                                                   "epoch_wait": 0.0, "retire": 0.0, "evict": 0.0})

    # --- requests -----------------------------------------------------------------------------
    def submit(self, job: Job):
        self.jobs[job.id] = job
        job.epochs = int(round(job.spec.get("sim_s", 0) / SW.EPOCH_S)) if isinstance(job.spec.get("sim_s"), (int, float)) else 0
        heapq.heappush(self.queue, (job.arrival, job.id))

    def cancel(self, job_id: int) -> bool:
        job = self.jobs.get(job_id)
        if job is None or job.state not in (QUEUED, ACTIVE):
            return False
        if job.state == ACTIVE:
            self._release(job)
            self.active.remove(job)
        self._drop_checkpoint(job)
        job.state = CANCELLED
        job.finished = self.clock()
        return True

    # --- one scheduling round -----------------------------------------------------------------
    def _release(self, job: Job):
        if job.handle is not None:
            self.engine.release(job.handle)
            job.handle = None
        self.used -= job.footprint
        job.plan = None

    def _drop_checkpoint(self, job: Job):
        if job.checkpoint is not None:
            try:
                os.remove(job.checkpoint)
            except OSError:
                pass
            job.checkpoint = None

    def _fail(self, job: Job, state: str, err: str):
        self._drop_checkpoint(job)
        job.state = state
        job.error = err
        job.finished = self.clock()
        if self.log:
            self.log(f"job {job.id} {state}: {err}")

    def prepare(self, limit: int | None = None):
        """Plans (host-side setup of) up to `limit` (default plan_ahead) queued jobs in arrival
        order, on up to plan_threads threads."""
        limit = self.plan_ahead if limit is None else limit
        todo = []
        for arrival, jid in sorted(self.queue)[: 4 * limit]:
            job = self.jobs[jid]
            if job.state == QUEUED and job.plan is None and job.error is None:
                todo.append(job)
                if len(todo) >= limit:
                    break

        def plan_one(job):
            try:
                plan = self.engine.plan(job.spec)
                return plan, int(self.engine.footprint(plan)), None
            except Exception as e:
                return None, 0, f"{type(e).__name__}: {e}"

        if len(todo) > 1 and self.plan_threads > 1:
            with ThreadPoolExecutor(max_workers=min(self.plan_threads, len(todo))) as ex:
                results = list(ex.map(plan_one, todo))
        else:
            results = [plan_one(j) for j in todo]
        for job, (plan, fp, err) in zip(todo, results):
            job.plan, job.footprint, job.error = plan, fp, err

    def admit(self):
        """Admits queued jobs in arrival order while they fit; up to `lookahead` jobs that do not
        fit yet are passed over (and later ones admitted) before the round ends."""
        considered = 0
        kept = []
        head_blocked = False
        while self.queue and considered < self.lookahead:
            arrival, jid = heapq.heappop(self.queue)
            job = self.jobs[jid]
            if job.state != QUEUED:
                continue
            if head_blocked:
                considered += 1
                kept.append((arrival, jid))
                continue
            if job.error is not None:      # planning failed earlier
                self._fail(job, FAILED, job.error)
                continue
            try:
                if job.plan is None:
                    job.plan = self.engine.plan(job.spec)
                    job.footprint = int(self.engine.footprint(job.plan))
            except Exception as e:          # an invalid specification fails this job only
                self._fail(job, FAILED, f"{type(e).__name__}: {e}")
                continue
            if job.footprint > self.budget:
                job.plan = None
                self._fail(job, REJECTED, f"footprint {job.footprint} B exceeds the budget {self.budget} B")
                continue
            if self.used + job.footprint > self.budget:
                considered += 1
                job.skipped += 1
                kept.append((arrival, jid))
                # an old job that keeps being bypassed stops later ones from overtaking it
                if job.skipped > self.max_skips:
                    head_blocked = True
                    if self.blocked is None:
                        self.blocked = job
                continue
            try:
                if job.checkpoint is not None:
                    job.handle = self.engine.restore(job.spec, job.plan, job.checkpoint)
                    self._drop_checkpoint(job)
                    self.restored += 1
                else:
                    job.handle = self.engine.create(job.spec, job.plan)
            except Exception as e:
                job.plan = None
                self._fail(job, FAILED, f"{type(e).__name__}: {e}")
                continue
            if self.blocked is job:
                self.blocked = None
            self.used += job.footprint
            job.state = ACTIVE
            if job.admitted is None:
                job.admitted = self.clock()
            self.active.append(job)
        for item in kept:
            heapq.heappush(self.queue, item)

    def preempt(self):
        """Evicts the most recently admitted residents (each at most max_evictions times) until the
        blocked job fits; evicts nothing when that cannot free enough. Runs between ticks."""
        job = self.blocked
        if not self.evict or job is None:
            return
        if job.state != QUEUED:
            self.blocked = None
            return
        need = self.used + job.footprint - self.budget
        if need <= 0:
            return
        victims, freed = [], 0
        for r in reversed(self.active):
            if freed >= need:
                break
            if r.evictions < self.max_evictions:
                victims.append(r)
                freed += r.footprint
        if freed < need:
            return
        if self.spill_dir is None:
            self.spill_dir = tempfile.mkdtemp(prefix="bg_spill_")
        key = max(self.clock() - self.t0, job.arrival)   # behind the blocked job in the queue
        for r in victims:
            path = os.path.join(self.spill_dir, f"job{r.id}.ckpt")
            try:
                self.engine.save(r.handle, path)
            except Exception as e:
                self.active.remove(r)
                self._release(r)
                self._fail(r, FAILED, f"eviction: {type(e).__name__}: {e}")
                continue
            self.spill_bytes += os.path.getsize(path)
            self.engine.release(r.handle)
            r.handle = None
            self.used -= r.footprint
            self.active.remove(r)
            r.checkpoint = path
            r.evictions += 1
            r.state = QUEUED
            self.evicted += 1
            heapq.heappush(self.queue, (key, r.id))
            if self.log:
                self.log(f"job {r.id} evicted for job {job.id} after {r.epochs_done} epochs")

    def close(self):
        """Removes the spill folder of evicted jobs' checkpoints."""
        if self.spill_dir is not None:
            shutil.rmtree(self.spill_dir, ignore_errors=True)
            self.spill_dir = None

    def tick(self):
        """Advance every resident network by one epoch; while the device runs it, prepare and admit
        queued jobs (they join from the next tick on); then do the epoch work and retire."""
        c0 = self.clock()
        advanced = list(self.active)
        if advanced:
            # the advance call returns only when the device has taken most of the epoch's launches,
            # so the host-side work of queued jobs -- planning, then admitting those that fit and
            # creating their networks on their own streams -- runs beside it on a worker thread
            # (the C calls and numpy release the interpreter lock); admitted jobs join at the next
            # tick, and the networks advanced now are the ones listed above
            failure = []

            def host_work():
                try:
                    self.prepare()
                    t = self.clock()
                    self.admit()
                    self.phase_s["admit"] += self.clock() - t
                except BaseException as e:   # re-raised on the scheduling thread
                    failure.append(e)

            worker = threading.Thread(target=host_work)
            worker.start()
            try:
                self.engine.advance([j.handle for j in advanced], self.epoch_steps)
            finally:
                worker.join()
            if failure:
                raise failure[0]
            c1 = c2 = self.clock()
        else:
            c1 = self.clock()
            self.admit()
            c2 = self.clock()
            self.phase_s["admit"] += c2 - c1
        self.phase_s["advance"] += c1 - c0
        self.occupancy.append((c2, self.used))
        if not advanced:
            return
        # every network's epoch work is queued first, so the networks' streams overlap, then checked
        for job in advanced:
            if job.state == ACTIVE:
                self.engine.end_epoch(job.handle)
        c3 = self.clock()
        finished = []
        for job in advanced:
            if job.state != ACTIVE:
                continue
            try:
                self.engine.check(job.handle)
            except Exception as e:
                self.active.remove(job)
                self._release(job)
                self._fail(job, FAILED, f"{type(e).__name__}: {e}")
                continue
            job.epochs_done += 1
            if job.epochs_done >= job.epochs:
                finished.append(job)
        c4 = self.clock()
        for job in finished:
            job.result = self.engine.result(job.handle)
            self.active.remove(job)
            self._release(job)
            job.state = DONE
            job.finished = self.clock()
        c5 = self.clock()
        self.preempt()
        c6 = self.clock()
        self.phase_s["epoch_issue"] += c3 - c2
        self.phase_s["epoch_wait"] += c4 - c3
        self.phase_s["retire"] += c5 - c4
        self.phase_s["evict"] += c6 - c5
        self.ticks += 1

    # --- open-loop driver ---------------------------------------------------------------------
    def run(self, jobs, sleep=time.sleep, progress_every: float = 30.0):
        """Feed `jobs` at their arrival times (seconds after the start) and serve until every job
        has left the system."""
        pending = sorted(jobs, key=lambda j: j.arrival)
        self.t0 = self.clock()
        k = 0
        last_log = self.t0
        while k < len(pending) or self.queue or self.active:
            now = self.clock() - self.t0
            while k < len(pending) and pending[k].arrival <= now:
                self.submit(pending[k])
                k += 1
            if not self.active and not self.queue:
                if k < len(pending):
                    sleep(max(0.0, pending[k].arrival - (self.clock() - self.t0)))
                continue
            self.tick()
            if self.log and self.clock() - last_log >= progress_every:
                last_log = self.clock()
                c = self.counts()
                self.log(f"t={last_log - self.t0:.0f}s submitted={k} queued={len(self.queue)} "
                         f"active={len(self.active)} done={c.get(DONE, 0)} hbm={self.used / 2**30:.1f}GiB")

    # --- reporting ----------------------------------------------------------------------------
    def counts(self) -> dict:
        out = {}
        for j in self.jobs.values():
            out[j.state] = out.get(j.state, 0) + 1
        return out

    def stats(self, total_bytes: int | None = None) -> dict:
        done = [j for j in self.jobs.values() if j.state == DONE]
        lat = np.array([j.finished - (self.t0 + j.arrival) for j in done]) if done else np.zeros(0)
        span = (max(j.finished for j in done) - self.t0) if done else 0.0
        occ = np.array([u for _, u in self.occupancy], dtype=np.float64) if self.occupancy else np.zeros(1)
        sim_s = sum(j.spec["sim_s"] for j in done)
        out = {"completed": len(done), "counts": self.counts(), "wall_s": span,
               "throughput_jobs_per_s": len(done) / span if span > 0 else 0.0,
               "sim_seconds_per_hour": sim_s / span * 3600.0 if span > 0 else 0.0,
               "latency_p50_s": float(np.percentile(lat, 50)) if lat.size else None,
               "latency_p99_s": float(np.percentile(lat, 99)) if lat.size else None,
               "latency_mean_s": float(lat.mean()) if lat.size else None,
               "hbm_resident_mean_bytes": float(occ.mean()), "hbm_resident_max_bytes": float(occ.max()),
               "ticks": self.ticks, "budget_bytes": self.budget,
               "evicted": self.evicted, "restored": self.restored, "spill_bytes": self.spill_bytes,
               "host_phase_s": {k: round(v, 3) for k, v in self.phase_s.items()}}
        if total_bytes:
            out["hbm_occupancy_mean"] = float(occ.mean()) / total_bytes
            out["hbm_occupancy_max"] = float(occ.max()) / total_bytes
        return out

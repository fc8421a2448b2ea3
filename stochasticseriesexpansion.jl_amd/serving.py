"""Serving: continuous batching of simulation jobs on one GPU.

Jobs (a model, temperature, seed and a number of sweeps) arrive as an
open-loop stream.  Each job is one walker in a slot of a resident batch: it is
admitted when a slot is free and its estimated opstring+record footprint fits
the memory budget, initialised on the device while the other slots keep
sweeping, and retired after its last sweep; its binned results (device
accumulators) stream back as soon as it finishes.  The loop reports p50/p99
time-to-result beside the aggregate vertex-update throughput.

This replaces Carlo's task farming (one MC per task per rank) for the
serving use; each job's result is independent of what else runs, because a
walker's chain depends only on its own seed and parameters.
"""
from __future__ import annotations

import heapq
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .lattice import UnitCells
from .models import MagnetModel


@dataclass
class Job:
    job_id: int
    arrival: float          # seconds after start
    model_key: tuple        # (L, kind)
    T: float
    sweeps: int
    thermalization: int
    seed: int
    # filled by the loop
    slot: int = -1
    admitted: float = -1.0
    finished: float = -1.0
    result: tuple = None
    state: tuple = None     # final (scalars, ops, spins) when the loop keeps states


def magnet_square(L, kind):
    """Heisenberg (kind 'heis') or XXZ with Jz = 0.5 (kind 'xxz') on an LxL square lattice."""
    p = {"lattice": {"unitcell": UnitCells.square, "size": (L, L)}, "J": 1.0,
         "measure": ["magnetization", "staggered_magnetization"]}
    if kind == "xxz":
        p["d"] = -0.5
    return MagnetModel(p)


def poisson_jobs(n_jobs, rate=200.0, seed=7, L_range=(8, 48), beta_range=(1.0, 32.0),
                 sweeps_range=(1e3, 1e5), sweep_scale=1.0, therm_fraction=0.1):
    """Open-loop Poisson arrival stream: exponential gaps of mean 1/rate,
    L uniform, beta and sweeps log-uniform, Heisenberg/XXZ 50/50."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for j in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        kind = "heis" if rng.random() < 0.5 else "xxz"
        beta = float(np.exp(rng.uniform(np.log(beta_range[0]), np.log(beta_range[1]))))
        sw = float(np.exp(rng.uniform(np.log(sweeps_range[0]), np.log(sweeps_range[1]))))
        sw = max(2, int(round(sw * sweep_scale)))
        th = max(1, int(round(sw * therm_fraction)))
        jobs.append(Job(j, t, (L, kind), 1.0 / beta, sw, th, seed * 1000003 + j))
    return jobs


def footprint_bytes(n_sites, n_bonds, T):
    """Slab bytes a job needs at its steady-state cutoff: n ~ beta * (|E| +
    sum of bond offsets) bounded by beta * 0.375 * Nb * 2.5, L ~ 2.2 n."""
    beta = 1.0 / T
    n_est = beta * n_bonds * 0.375 * 2.5 + 64
    L = int(2.2 * n_est) + 128
    cap = L + L // 2 + 100
    return 24 * cap + 16 * (cap // 64 + 1)


class ServingLoop:
    def __init__(self, n_slots=512, max_L=48, memory_budget=8 << 30, sweeps_per_step=10,
                 lib=None, keep_state=False):
        self.keep_state = keep_state
        self.model_cache = {}
        self.flats = []
        proto = magnet_square(max_L, "heis").flat()
        # slot 0..n-1 start stopped; the prototype model sizes the per-walker arrays
        walkers = [capi.walker_desc(model=0, T=1.0, seed=0) for _ in range(n_slots)]
        self.batch = capi.Batch([proto], walkers, lib=lib)
        for s in range(n_slots):
            self.batch.slot_stop(s)
        self.model_cache[(max_L, "heis")] = 0
        self.n_slots = n_slots
        self.memory_budget = memory_budget
        self.sweeps_per_step = sweeps_per_step

    def model_index(self, key):
        if key not in self.model_cache:
            flat = magnet_square(*key).flat()
            self.model_cache[key] = self.batch.add_model(flat)
        return self.model_cache[key]

    def run(self, jobs, time_limit=None, log=None):
        """Serve `jobs` (sorted by arrival); returns the finished jobs and stats."""
        b = self.batch
        pending = list(jobs)
        queue = []           # arrived, not admitted (FIFO)
        free = list(range(self.n_slots))
        running = {}         # slot -> (job, footprint)
        used = 0
        done = []
        t0 = time.perf_counter()
        b.reset_accumulators()
        L0, n0, W0, s0 = b.work_counters()
        steps = 0
        while pending or queue or running:
            now = time.perf_counter() - t0
            if time_limit is not None and now > time_limit:
                break
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            # memory-based admission, FIFO
            while queue and free:
                job = queue[0]
                mi = self.model_index(job.model_key)
                flat = b.flats[mi]
                fp = footprint_bytes(flat.n_sites, flat.n_bonds, job.T)
                if used + fp > self.memory_budget and running:
                    break
                queue.pop(0)
                slot = free.pop()
                wd = capi.walker_desc(model=mi, T=job.T, seed=job.seed,
                                      thermalization=job.thermalization)
                b.slot_start(slot, wd, job.thermalization + job.sweeps)
                job.slot = slot
                job.admitted = time.perf_counter() - t0
                running[slot] = (job, fp)
                used += fp
            if not running:
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                continue
            b.run(self.sweeps_per_step, measure=True)
            steps += 1
            finished = b.slot_poll()
            if finished:
                acc = b.accumulators()
                tnow = time.perf_counter() - t0
                for slot in finished:
                    job, fp = running.pop(slot)
                    job.finished = tnow
                    sums, sq, cnt = acc[slot]
                    job.result = (sums.copy(), sq.copy(), cnt)
                    if self.keep_state:
                        job.state = b.get_walker(slot)
                    used -= fp
                    free.append(slot)
                    done.append(job)
                if log:
                    log(f"t={tnow:.2f}s done={len(done)} running={len(running)} "
                        f"queued={len(queue)}")
        wall = time.perf_counter() - t0
        L1, n1, W1, s1 = b.work_counters()
        lat = np.array([j.finished - j.arrival for j in done]) if done else np.zeros(1)
        stats = {"jobs_done": len(done), "wall_s": wall,
                 "vertex_updates_per_s": ((L1 - L0) + (W1 - W0)) / wall,
                 "walker_sweeps": s1 - s0,
                 "p50_time_to_result_s": float(np.percentile(lat, 50)),
                 "p99_time_to_result_s": float(np.percentile(lat, 99)),
                 "steps": steps}
        return done, stats

    def close(self):
        self.batch.close()


def n_estimate(L, T):
    """Operators a job on an LxL square lattice (2 L^2 bonds) holds at
    temperature T: the bound footprint_bytes uses."""
    return (1.0 / T) * 2 * L * L * 0.375 * 2.5 + 64


class _Class:
    """One size class: a resident batch of slots driven by its own host
    thread (and so its own HIP stream)."""

    def __init__(self, n_max, n_slots, sweeps_per_step, max_L, memory_budget, lib):
        import threading
        self.n_max = n_max
        self.max_L = max_L
        self.n_slots = n_slots
        self.sweeps_per_step = sweeps_per_step
        self.memory_budget = memory_budget
        proto = magnet_square(max_L, "heis").flat()
        walkers = [capi.walker_desc(model=0, T=1.0, seed=0) for _ in range(n_slots)]
        self.batch = capi.Batch([proto], walkers, lib=lib)
        # one stream per class: the classes themselves overlap on the device
        self.batch.set_option(capi.SSE_OPT_STREAMS, 1)
        self.batch.set_option(capi.SSE_OPT_TIMING, 0)
        for s in range(n_slots):
            self.batch.slot_stop(s)
        self.model_index = {(max_L, "heis"): 0}
        self.queue = []
        self.lock = threading.Lock()
        self.steps = 0

    def add_model(self, key, flat):
        if key not in self.model_index:
            self.model_index[key] = self.batch.add_model(flat)


class SizeClassServing:
    """Continuous batching with size classes.

    A job goes to the first class whose operator bound (n_estimate) and
    lattice size admit it.  Each class is a resident batch driven by its own
    host thread, so one class's step lasts only as long as its own slowest
    walker: small jobs never wait behind the worm tails of large ones, and
    the classes' kernels run side by side on the device.  Every job's chain
    still depends only on its seed and parameters (bit-identical to running
    it alone).  Models are built and uploaded before the clock starts
    (warm_models), as a service keeps its model cache warm."""

    def __init__(self, classes=((4000, 512, 10, 32), (float("inf"), 512, 2, 48)),
                 memory_budget=8 << 30, lib=None, keep_state=False):
        self.keep_state = keep_state
        total = sum(c[1] for c in classes)
        self.classes = [_Class(n_max, slots, sps, max_L, memory_budget * slots // total, lib)
                        for n_max, slots, sps, max_L in classes]
        self.flat_cache = {}

    def route(self, job):
        L = job.model_key[0]
        n = n_estimate(L, job.T)
        for k, c in enumerate(self.classes):
            if n <= c.n_max and L <= c.max_L:
                return k
        return len(self.classes) - 1

    def warm_models(self, jobs):
        for job in jobs:
            c = self.classes[self.route(job)]
            key = job.model_key
            if key not in c.model_index:
                if key not in self.flat_cache:
                    self.flat_cache[key] = magnet_square(*key).flat()
                c.add_model(key, self.flat_cache[key])

    def _serve(self, c, t0, done, done_lock, stop, time_limit):
        b = c.batch
        free = list(range(c.n_slots))
        running = {}
        used = 0
        while True:
            now = time.perf_counter() - t0
            if time_limit is not None and now > time_limit:
                break
            with c.lock:
                while c.queue and free:
                    job = c.queue[0]
                    mi = c.model_index[job.model_key]
                    flat = b.flats[mi]
                    fp = footprint_bytes(flat.n_sites, flat.n_bonds, job.T)
                    if used + fp > c.memory_budget and running:
                        break
                    c.queue.pop(0)
                    slot = free.pop()
                    wd = capi.walker_desc(model=mi, T=job.T, seed=job.seed,
                                          thermalization=job.thermalization)
                    b.slot_start(slot, wd, job.thermalization + job.sweeps)
                    job.slot = slot
                    job.admitted = time.perf_counter() - t0
                    running[slot] = (job, fp)
                    used += fp
                idle = not running
                empty = not c.queue
            if idle:
                if stop.is_set() and empty:
                    break
                time.sleep(0.0005)
                continue
            b.run(c.sweeps_per_step, measure=True)
            c.steps += 1
            finished = b.slot_poll()
            if finished:
                acc = b.accumulators()
                tnow = time.perf_counter() - t0
                out = []
                for slot in finished:
                    job, fp = running.pop(slot)
                    job.finished = tnow
                    sums, sq, cnt = acc[slot]
                    job.result = (sums.copy(), sq.copy(), cnt)
                    if self.keep_state:
                        job.state = b.get_walker(slot)
                    used -= fp
                    free.append(slot)
                    out.append(job)
                with done_lock:
                    done.extend(out)

    def run(self, jobs, time_limit=None, log=None):
        """Serve `jobs` (sorted by arrival); returns the finished jobs and
        stats (overall and per class).  log(msg), when given, gets a line
        every few seconds."""
        import threading
        self.warm_models(jobs)
        for c in self.classes:
            c.batch.reset_accumulators()
        w0 = [c.batch.work_counters() for c in self.classes]
        done, done_lock, stop = [], threading.Lock(), threading.Event()
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._serve, args=(c, t0, done, done_lock, stop,
                                                              time_limit), daemon=True)
                   for c in self.classes]
        for t in threads:
            t.start()
        pending = list(jobs)
        next_log = 5.0
        while pending or any(t.is_alive() for t in threads):
            now = time.perf_counter() - t0
            if time_limit is not None and now > time_limit:
                break
            while pending and pending[0].arrival <= now:
                job = pending.pop(0)
                c = self.classes[self.route(job)]
                with c.lock:
                    c.queue.append(job)
            if not pending:
                stop.set()
            if log and now >= next_log:
                with done_lock:
                    nd = len(done)
                log(f"t={now:.1f}s done={nd} submitted={len(jobs) - len(pending)}")
                next_log += 5.0
            time.sleep(max(0.0, min(0.002, pending[0].arrival - now)) if pending else 0.01)
        stop.set()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        w1 = [c.batch.work_counters() for c in self.classes]
        U = sum((b[0] - a[0]) + (b[2] - a[2]) for a, b in zip(w0, w1))
        sweeps = sum(b[3] - a[3] for a, b in zip(w0, w1))
        lat = np.array([j.finished - j.arrival for j in done]) if done else np.zeros(1)
        stats = {"jobs_done": len(done), "wall_s": wall, "vertex_updates_per_s": U / wall,
                 "walker_sweeps": sweeps,
                 "p50_time_to_result_s": float(np.percentile(lat, 50)),
                 "p99_time_to_result_s": float(np.percentile(lat, 99)),
                 "classes": []}
        for k, c in enumerate(self.classes):
            lk = np.array([j.finished - j.arrival for j in done if self.route(j) == k])
            stats["classes"].append({
                "n_max": c.n_max if c.n_max != float("inf") else None, "max_L": c.max_L,
                "slots": c.n_slots, "sweeps_per_step": c.sweeps_per_step, "steps": c.steps,
                "jobs_done": int(lk.size),
                "p50_time_to_result_s": float(np.percentile(lk, 50)) if lk.size else None,
                "p99_time_to_result_s": float(np.percentile(lk, 99)) if lk.size else None})
        return done, stats

    def close(self):
        for c in self.classes:
            c.batch.close()

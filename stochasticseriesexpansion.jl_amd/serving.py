"""Serving: continuous batching of simulation jobs on one GPU.

Jobs (a model, temperature, seed and a number of sweeps) arrive as an
open-loop stream.  Each job is one walker in a slot of a resident batch: it is
admitted when a slot is free and its estimated opstring+record footprint fits
the memory budget, initialised on the device while the other slots keep
sweeping, and retired after its last sweep.  With per-slot budgets (the
work-queue launch, sse_run_budgets) every slot advances by its own number of
sweeps per step, chosen so that a step lasts about one sweep of the class's
largest job; after thermalization that number is the job's bin length, and
every step streams one bin of each running job back (device accumulator
deltas, Carlo's internal bins).  The loop reports p50/p99 time-to-result over
all submitted jobs (jobs still running at the end count with their time so
far, and a percentile such a job can move is marked as a lower bound) beside
the aggregate vertex-update throughput.

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
    # budgeted stepping: sweeps per step (the bin length once thermalized),
    # sweeps run so far, the streamed bins (one mean row per bin, columns as
    # carlo.binned_names) and the accumulator totals at the last poll
    step_budget: int = 0
    done_sweeps: int = 0
    bins: list = field(default_factory=list)
    bin_counts: list = field(default_factory=list)   # measurements in each bin
    last_acc: tuple = None


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


def latency_stats(jobs, t_end):
    """p50/p99 time-to-result over every submitted job.  A job still running
    at t_end counts with t_end - arrival, a lower bound of its latency.  The
    percentile (linear interpolation between the sorted values at ranks
    floor(h) and ceil(h), h = q (N - 1)) is exact when no censored time sits
    at or below rank ceil(h): censored latencies can only move up, past the
    ranks that form it.  Otherwise it is marked as a lower bound."""
    fin = [j.finished - j.arrival for j in jobs if j.finished >= 0]
    cen = [max(0.0, t_end - j.arrival) for j in jobs if j.finished < 0 and j.arrival <= t_end]
    vals = np.array(fin + cen) if (fin or cen) else np.zeros(1)
    censored = np.array([False] * len(fin) + [True] * len(cen)) if (fin or cen) else np.zeros(1, bool)
    order = np.argsort(vals, kind="stable")
    out = {"jobs_submitted": len(fin) + len(cen), "jobs_done": len(fin),
           "jobs_unfinished": len(cen),
           "completion_fraction": len(fin) / max(1, len(fin) + len(cen))}
    for q in (50, 99):
        v = float(np.percentile(vals, q))
        hi = int(np.ceil(q / 100.0 * (len(vals) - 1)))
        exact = not censored[order[:hi + 1]].any()
        out[f"p{q}_time_to_result_s"] = v
        out[f"p{q}_is_lower_bound"] = bool(not exact)
    if fin:
        out["p50_done_only_s"] = float(np.percentile(fin, 50))
        out["p99_done_only_s"] = float(np.percentile(fin, 99))
    return out


def job_results(job, flat):
    """results.json entry of a served job from its streamed bins (internal
    bin length = the job's step budget; a shorter last bin is left out),
    merged as Carlo merges a task."""
    from . import carlo
    full = [r for r, k in zip(job.bins, job.bin_counts) if k == job.step_budget]
    if not full:
        return None
    bins = np.array(full)
    return {"task": f"job{job.job_id}",
            "parameters": {"L": job.model_key[0], "model": job.model_key[1], "T": job.T,
                           "sweeps": job.sweeps, "thermalization": job.thermalization,
                           "seed": job.seed},
            "results": carlo.evaluate_task(carlo.binned_names(flat), [bins], job.step_budget,
                                           flat.normalization_site_count)}


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
        stats = {"wall_s": wall, "vertex_updates_per_s": ((L1 - L0) + (W1 - W0)) / wall,
                 "walker_sweeps": s1 - s0, "steps": steps}
        stats.update(latency_stats(jobs, wall))
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

    def __init__(self, n_max, n_slots, sweeps_per_step, max_L, memory_budget, lib, budgets):
        import threading
        self.budgets = budgets
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
        # operator count a step is sized for: the class bound, or the largest
        # job the lattice bound admits (beta = 32) for the open-ended class
        self.n_ref = min(n_max, n_estimate(max_L, 1.0 / 32.0))

    def add_model(self, key, flat):
        if key not in self.model_index:
            self.model_index[key] = self.batch.add_model(flat)


class SizeClassServing:
    """Continuous batching with size classes.

    A job goes to the first class whose operator bound (n_estimate) and
    lattice size admit it.  Each class is a resident batch driven by its own
    host thread, so one class's step lasts only as long as its own slowest
    walker: small jobs never wait behind the worm tails of large ones, and
    the classes' kernels run side by side on the device.  With budgets (the
    default) each class steps through the work-queue launch with a budget per
    slot: a job of expected operator count n gets about sweeps_per_step *
    n_ref / n sweeps per step (at most max_budget), so the class's jobs finish
    their step together, and each step streams one bin per running job
    (on_bin(job, measurements, row) is called from the class's thread).
    Every job's chain still depends only on its seed and parameters
    (bit-identical to running it alone).  Models are built and uploaded
    before the clock starts (warm_models), as a service keeps its model cache
    warm."""

    def __init__(self, classes=((4000, 512, 10, 32), (float("inf"), 512, 2, 48)),
                 memory_budget=8 << 30, lib=None, keep_state=False, budgets=True, max_budget=1000,
                 on_bin=None):
        self.keep_state = keep_state
        self.max_budget = max_budget
        self.on_bin = on_bin
        total = sum(c[1] for c in classes)
        self.classes = [_Class(n_max, slots, sps, max_L, memory_budget * slots // total, lib,
                               budgets)
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

    def step_budget(self, c, job):
        n = n_estimate(job.model_key[0], job.T)
        return int(max(1, min(self.max_budget, round(c.sweeps_per_step * c.n_ref / n))))

    def _stream_bins(self, c, b, running, finished):
        """One bin per job that measured in the last step: the accumulator
        and WormLengthFraction deltas since the previous poll."""
        acc = b.accumulators()
        wlf = b.worm_fractions()
        for slot, (job, _) in list(running.items()) + [(s, f) for s, f in finished]:
            sums, sq, cnt = acc[slot]
            ws, _, wc = wlf[slot]
            last = job.last_acc or (np.zeros_like(sums), 0, 0.0, 0)
            dc = cnt - last[1]
            if dc > 0:
                row = np.empty(len(sums) + 1)
                row[:-1] = (sums - last[0]) / dc
                dwc = wc - last[3]
                row[-1] = (ws - last[2]) / dwc if dwc else float("nan")
                job.bins.append(row)
                job.bin_counts.append(int(dc))
                if self.on_bin:
                    self.on_bin(job, dc, row)
            job.last_acc = (sums.copy(), cnt, ws, wc)
        return acc

    def _serve(self, c, t0, done, done_lock, stop, time_limit):
        b = c.batch
        free = list(range(c.n_slots))
        running = {}
        used = 0
        bud = np.zeros(c.n_slots, dtype=np.int64)
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
                    job.step_budget = self.step_budget(c, job) if c.budgets else c.sweeps_per_step
                    job.done_sweeps = 0
                    job.last_acc = None
                    running[slot] = (job, fp)
                    used += fp
                idle = not running
                empty = not c.queue
            if idle:
                if stop.is_set() and empty:
                    break
                time.sleep(0.0005)
                continue
            if c.budgets:
                # thermalization ends on a step boundary, so every later step
                # measures exactly step_budget sweeps (one bin)
                bud[:] = 0
                for slot, (job, _) in running.items():
                    total = job.thermalization + job.sweeps
                    lim = job.thermalization if job.done_sweeps < job.thermalization else total
                    k = min(job.step_budget, lim - job.done_sweeps)
                    bud[slot] = k
                    job.done_sweeps += k
                b.run_budgets(bud, measure=True)
            else:
                b.run(c.sweeps_per_step, measure=True)
            c.steps += 1
            fin_slots = b.slot_poll()
            finished = [(s, running.pop(s)) for s in fin_slots]
            acc = self._stream_bins(c, b, running, [(s, f) for s, f in finished])
            if finished:
                tnow = time.perf_counter() - t0
                out = []
                for slot, (job, fp) in finished:
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
        submitted = [j for j in jobs if j.arrival <= wall]
        stats = {"wall_s": wall, "vertex_updates_per_s": U / wall, "walker_sweeps": sweeps,
                 "bins_streamed": int(sum(len(j.bins) for j in submitted)), "classes": []}
        stats.update(latency_stats(submitted, wall))
        for k, c in enumerate(self.classes):
            st = latency_stats([j for j in submitted if self.route(j) == k], wall)
            st.update({"n_max": c.n_max if c.n_max != float("inf") else None, "max_L": c.max_L,
                       "slots": c.n_slots, "sweeps_per_step": c.sweeps_per_step,
                       "steps": c.steps, "budgets": c.budgets})
            stats["classes"].append(st)
        return done, stats

    def close(self):
        for c in self.classes:
            c.batch.close()

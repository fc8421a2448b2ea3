#!/usr/bin/env python3
"""Batched SSE sweep throughput on one MI355X.

Workload (the batched Heisenberg workload): 4096 independent walkers of the
S=1/2 Heisenberg antiferromagnet (J=1) on a 16x16 periodic square lattice at
beta=8, Philox seeds 0..4095, random initial spins, cutoff grown by the
reference rule, 2000 thermalization sweeps (untimed setup).  One step = one
sweep (diagonal update + vertex list + worm update) of every walker followed
by the measurement (energy, uniform and staggered magnetization family),
accumulated on the device.

Reported beside it, in the same run:
  * the plain GPU form: the same walkers (thermalized states) run one walker
    per launch with the lane kernel, and with the wave kernels;
  * the CPU baseline: the C restatement oracle (one walker per thread) on
    thermalized states copied from the GPU, for a bounded time;
  * the roofline of the dominant kernel from HIP events on the batch stream.

Vertex updates per walker-sweep U = L + W (W = scatter calls); algorithmic
HBM bytes per walker-sweep B = 16L + 16n + 12W + 3N.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(sse, L, beta, walkers, thermalization):
    model = sse.MagnetModel({"lattice": {"unitcell": sse.UnitCells.square, "size": (L, L)},
                             "J": 1.0, "measure": ["magnetization", "staggered_magnetization"]})
    flat = model.flat()
    wds = [sse.walker_desc(T=1.0 / beta, seed=s, thermalization=thermalization)
           for s in range(walkers)]
    return flat, wds


def kernel_bytes(Ls, ns, Ws, N, walker_sweeps):
    """Algorithmic bytes per kernel: diagonal+vertex list 8L + 4L + 16n + 2N
    (spins in/out), worm 12W + N, measure 4L."""
    return {"diag_vlist": 12 * Ls + 16 * ns + 2 * N * walker_sweeps,
            "worm": 12 * Ws + N * walker_sweeps,
            "measure": 4 * Ls}


def plain_form(sse, flat, wds, states, sweeps, mode):
    """One walker per launch: each walker its own batch, sweeps timed with the
    device between launches drained (sse_sweep synchronizes)."""
    tot_U, tot_t = 0, 0.0
    for k, (sc, ops, spins) in enumerate(states):
        b = sse.Batch([flat], [wds[k]], mode=mode)
        b.init()
        b.set_walker(0, sc, ops, spins)
        b.sweep(1)
        b.reset_accumulators()
        t0 = time.perf_counter()
        for _ in range(sweeps):
            b.sweep(1)
        dt = time.perf_counter() - t0
        Ls, ns, Ws, nsw = b.work_counters()
        tot_U += Ls + Ws
        tot_t += dt
        b.close()
    return tot_U / tot_t


def cpu_baseline(sse, flat, wds, states, seconds, threads):
    from oracle import pyoracle
    walkers = []
    for k in range(threads):
        sc, ops, spins = states[k % len(states)]
        o = pyoracle.OracleWalker(sse.capi, flat, wds[k % len(wds)])
        o.init()
        o.set_state(sc, ops, spins)
        o.sweep(1)
        o.work_counters()
        walkers.append(o)
    res = [None] * threads

    def run(i):
        o = walkers[i]
        t0 = time.perf_counter()
        sweeps = 0
        while time.perf_counter() - t0 < seconds:
            o.run_timed(5, measure=True)
            sweeps += 5
        dt = time.perf_counter() - t0
        Ls, ns, Ws, nsw = o.work_counters()
        res[i] = (Ls + Ws, dt, nsw, ns)

    th = [threading.Thread(target=run, args=(i,)) for i in range(threads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    U = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    sweeps = sum(r[2] for r in res)
    nops = sum(r[3] for r in res)
    return {"value": U / wall, "unit": "vertex-updates/s", "cores": threads, "kind": "port",
            "sample": f"C restatement oracle, {threads} walkers (one per thread) on thermalized "
                      f"16x16 beta=8 states from the GPU run, {sweeps} walker-sweeps with measure "
                      f"in {wall:.1f} s",
            "ns_per_operator_sweep": wall * threads / nops * 1e9 if nops else None}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--walkers", type=int, default=4096)
    ap.add_argument("--lattice", type=int, default=16)
    ap.add_argument("--beta", type=float, default=8.0)
    ap.add_argument("--thermalization", type=int, default=2000)
    ap.add_argument("--plain-walkers", type=int, default=2)
    ap.add_argument("--plain-sweeps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-baselines", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); walkers are independent "
              f"chains, so N GPUs run N independent batches.", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__
    sse = __graft_entry__.load_package()
    flat, wds = build_workload(sse, args.lattice, args.beta, args.walkers, args.thermalization)
    N = flat.n_sites
    log(f"workload: {args.walkers} walkers, {args.lattice}x{args.lattice}, beta={args.beta}")
    batch = sse.Batch([flat], wds, mode=sse.SSE_MODE_WAVE)
    t0 = time.perf_counter()
    batch.init()
    done = 0
    while done < args.thermalization:
        k = min(250, args.thermalization - done)
        batch.sweep(k)
        done += k
        sc, _, _ = batch.get_walker(0, with_ops=False, with_spins=False)
        log(f"thermalization {done}/{args.thermalization} sweeps ({time.perf_counter() - t0:.1f} s)"
            f" walker0 L={sc['cutoff']} n={sc['num_operators']} worms={sc['num_worms']:.2f}")
    for _ in range(args.warmup):
        batch.run(1, measure=True)
    batch.reset_accumulators()

    step_ms = []
    ktimes = np.zeros(3)
    for s in range(args.steps):
        t1 = time.perf_counter()
        batch.run(1, measure=True)
        step_ms.append((time.perf_counter() - t1) * 1e3)
        ktimes += batch.phase_times()
    total_s = sum(step_ms) / 1e3
    Ls, ns, Ws, wsweeps = batch.work_counters()
    U = Ls + Ws
    B = 16 * Ls + 16 * ns + 12 * Ws + 3 * N * wsweeps
    value = U / total_s
    log(f"timed {args.steps} steps: {total_s:.3f} s, {wsweeps} walker-sweeps, "
        f"<L>={Ls / wsweeps:.0f} <n>={ns / wsweeps:.0f} <W>={Ws / wsweeps:.0f}")

    kb = kernel_bytes(Ls, ns, Ws, N, wsweeps)
    names = ["diag_vlist", "worm", "measure"]
    kms = {n: ktimes[i] / args.steps for i, n in enumerate(names)}
    dom = max(names, key=lambda n: kms[n])
    per_launch_bytes = kb[dom] / args.steps
    achieved = per_launch_bytes / (kms[dom] * 1e-3) / 1e9
    log(f"kernel ms/launch: " + ", ".join(f"{n}={kms[n]:.3f}" for n in names))

    states = [batch.get_walker(k) for k in range(max(args.plain_walkers, 8))]
    baselines = {}
    if not args.no_baselines:
        log("plain form: one walker per launch (lane kernel)")
        plain_lane = plain_form(sse, flat, wds, states[:args.plain_walkers], args.plain_sweeps,
                                sse.SSE_MODE_LANE)
        log("plain form: one walker per launch (wave kernels)")
        plain_wave = plain_form(sse, flat, wds, states[:args.plain_walkers], args.plain_sweeps,
                                sse.SSE_MODE_WAVE)
        baselines = {"plain_lane_vertex_updates_per_s": plain_lane,
                     "plain_wave_vertex_updates_per_s": plain_wave,
                     "batched_over_plain_lane": value / plain_lane,
                     "batched_over_plain_wave": value / plain_wave}
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads x {args.cpu_seconds} s")
        cpu = cpu_baseline(sse, flat, wds, states, args.cpu_seconds, threads)
    else:
        cpu = None
    batch.close()

    out = {
        "metric": "vertex-updates/s batched (4096 walkers) vs plain GPU form; HBM GB/s; p99 latency",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(step_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic (Philox-seeded random initial spins, thermalized chains)",
        "config": {"workload": f"batched Heisenberg {args.lattice}x{args.lattice} beta={args.beta:g}",
                   "walkers": args.walkers, "thermalization_sweeps": args.thermalization,
                   "parallelism": "one wavefront per walker"},
        "sweeps_per_s": wsweeps / total_s,
        "hbm_gbs_algorithmic": B / total_s / 1e9,
        "hbm_roofline_frac": B / total_s / 1e9 / HBM_PEAK_GBS,
        "p50_step_ms": float(np.percentile(step_ms, 50)),
        "p99_step_ms": float(np.percentile(step_ms, 99)),
        "per_walker_sweep": {"L": Ls / wsweeps, "n": ns / wsweeps, "W": Ws / wsweeps,
                             "U": U / wsweeps, "B": B / wsweeps},
        "kernel_ms_per_launch": kms,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
    }
    out.update(baselines)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

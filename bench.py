#!/usr/bin/env python3
"""Batched SSE sweep throughput on one MI355X.

Workload (the batched Heisenberg workload): 4096 independent walkers of the
S=1/2 Heisenberg antiferromagnet (J=1) on a 16x16 periodic square lattice at
beta=8, Philox seeds 0..4095, random initial spins, cutoff grown by the
reference rule, 2000 thermalization sweeps (untimed setup).  One step = one
sweep (diagonal update + vertex list + worm update) of every walker followed
by the measurement (energy, uniform and staggered magnetization family),
accumulated on the device.  The timed steps run as one call (or in calls of
--chunk sweeps); the walkers are split into --streams groups, each swept on
its own HIP stream so one group's slow worm-update tail overlaps the others'
sweeps.  The process asks for --hw-queues (8) hardware queues
(GPU_MAX_HW_QUEUES) and by default the library makes one group per 512
walkers up to one per queue (8 for 4096 walkers); more streams than queues
would share queues and serialize.

Reported beside it, in the same run:
  * the serving stream (open-loop Poisson arrivals of mixed jobs, 20 s of
    arrivals by default): p50/p99 time-to-result of the jobs done;
  * the plain GPU form: the same walkers (thermalized states) run one walker
    per launch with the lane kernel, and with the wave kernels;
  * the CPU baseline: the C restatement oracle (one walker per thread) on
    thermalized states copied from the GPU, for a bounded time;
  * the roofline of the dominant kernel (the worm update): three extra
    sweeps on one stream (untimed diagnostic) with HIP events around each
    kernel give its launch duration for the whole batch.

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


def measured_traffic(kernel, walkers, lattice, beta):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated) when it covers this
    workload, else None."""
    import re
    path = os.path.join(ROOT, "profiles",
                        "worm_kernel_hbm_traffic_heisenberg16x16_beta8_4096walkers.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    want = f"traffic_workload: walkers={walkers} lattice={lattice} beta={beta:g} kernel={kernel}"
    m = re.search(r"traffic_bytes_per_launch: (\d+)", text)
    return int(m.group(1)) if (want in text and m) else None


def copy_calibration(mb):
    """Device-to-device hipMemcpy of `mb` MiB, three times: a copy kernel of
    known traffic (mb MiB read + mb MiB written per launch) that scales the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 x {mb} MiB device-to-device")


# ---------------------------------------------------------------- extras
def _xxz_types(args):
    """VertexData of the two square-lattice bond types for one walker's Jz
    (J=1, d=Jz-1, h split over the 4 bonds of a site)."""
    jz, h = args
    import __graft_entry__
    sse = __graft_entry__.load_package()
    from sse_amd.models import MagnetBondParams, generate_bond_hamiltonian
    bp = MagnetBondParams(1.0, jz - 1.0, (0.0, 0.0), (0.0, 0.0), (h / 4, h / 4))
    dims, H, f = generate_bond_hamiltonian(None, bp, (2, 2))
    vd = sse.VertexData(dims, H, energy_offset_factor=f)
    return [vd, vd]


def _ladder_flat(jl):
    import __graft_entry__
    sse = __graft_entry__.load_package()
    m = sse.ClusterModel({
        "inner_model": sse.MagnetModel, "cluster_bases": (sse.ClusterBases.dimer,),
        "lattice": {"unitcell": sse.UnitCells.ladder, "size": (64, 1)},
        "measure_quantum_numbers": [{"name": "", "quantum_number": 2}],
        "parameter_map": {"J": ["Jr", "Jl", "Jl"]}, "Jr": 1.0, "Jl": jl})
    return m.flat()


def run_batch(sse, flats, wds, thermalization, steps, what):
    batch = sse.Batch(flats, wds)
    t0 = time.perf_counter()
    batch.init()
    done = 0
    while done < thermalization:
        k = min(100, thermalization - done)
        batch.sweep(k)
        done += k
    log(f"{what}: thermalized {thermalization} sweeps in {time.perf_counter() - t0:.1f} s")
    batch.run(1, measure=True)
    batch.reset_accumulators()
    t1 = time.perf_counter()
    batch.set_option(sse.capi.SSE_OPT_TIMING, 0)
    batch.run(steps, measure=True)
    dt = time.perf_counter() - t1
    Ls, ns, Ws, wsw = batch.work_counters()
    batch.close()
    return {"vertex_updates_per_s": (Ls + Ws) / dt, "sweeps_per_s": wsw / dt,
            "walkers": len(wds), "timed_sweeps": steps, "thermalization_sweeps": thermalization,
            "mean_L": Ls / wsw, "mean_n": ns / wsw, "mean_W": Ws / wsw}


def extra_heterogeneous(sse, walkers=2048, thermalization=500, steps=20, seed=3):
    """2048 walkers, square L in {8,12,16,24,32}, beta log-uniform [0.5,16],
    XXZ with Jz ~ U[0.5,1.5], Jxy=1, h=0.1; one model per walker."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([8, 12, 16, 24, 32], size=walkers)
    betas = np.exp(rng.uniform(np.log(0.5), np.log(16.0), size=walkers))
    jzs = rng.uniform(0.5, 1.5, size=walkers)
    base = {L: sse.MagnetModel({"lattice": {"unitcell": sse.UnitCells.square, "size": (L, L)},
                                "J": 1.0, "d": 0.0, "hz": 0.1,
                                "measure": ["magnetization", "staggered_magnetization"]})
            for L in (8, 12, 16, 24, 32)}
    base_flat = {L: m.flat() for L, m in base.items()}
    t0 = time.perf_counter()
    types = [_xxz_types((float(j), 0.1)) for j in jzs]  # native transition LPs
    log(f"heterogeneous: {walkers} per-walker tables built in {time.perf_counter() - t0:.1f} s")
    from sse_amd.sse_data import FlatModel, SSEData
    bonds = {L: [sse.SSEBond(b.type, (b.i, b.j)) for b in m.lattice.bonds] for L, m in base.items()}
    flats, wds = [], []
    for k in range(walkers):
        L = int(sizes[k])
        m = base[L]
        flats.append(FlatModel(SSEData(types[k], bonds[L]), m.opstring_estimators,
                               m.normalization_site_count(),
                               (base_flat[L].est_coef, base_flat[L].est_mval)))
        wds.append(sse.walker_desc(model=k, T=1.0 / float(betas[k]), seed=100000 + k,
                                   thermalization=thermalization))
    r = run_batch(sse, flats, wds, thermalization, steps, "heterogeneous")
    r["workload"] = ("heterogeneous XXZ batch: 2048 walkers, L in {8,12,16,24,32}, "
                     "beta log-uniform [0.5,16], Jz ~ U[0.5,1.5], h=0.1")
    return r


def extra_ladder(sse, walkers=1024, thermalization=500, steps=20, seed=4):
    """1024 walkers of a 64-rung ladder in the dimer basis (local dim 4),
    J_rung=1, J_leg ~ U[0.2,1.0], beta=10."""
    rng = np.random.default_rng(seed)
    jls = rng.uniform(0.2, 1.0, size=walkers)
    t0 = time.perf_counter()
    flats = [_ladder_flat(float(j)) for j in jls]  # native transition LPs
    log(f"ladder: {walkers} per-walker models built in {time.perf_counter() - t0:.1f} s")
    wds = [sse.walker_desc(model=k, T=0.1, seed=200000 + k, thermalization=thermalization)
           for k in range(walkers)]
    r = run_batch(sse, flats, wds, thermalization, steps, "ladder")
    r["workload"] = "two-leg ladder, 64 rungs, dimer basis, 1024 walkers, J_leg ~ U[0.2,1], beta=10"
    return r


def extra_serving(sse, seconds=30.0, sweep_scale=1e-3, slots=1024):
    """Open-loop Poisson stream (200 jobs/s, seed 7) of Heisenberg/XXZ jobs on
    L in 8..48, beta log-uniform [1,32]; requested sweeps log-uniform
    [1e3,1e5] scaled by sweep_scale to bound the run."""
    from importlib import import_module
    sv = import_module("sse_amd.serving")
    jobs = sv.poisson_jobs(int(200 * seconds), rate=200.0, seed=7, sweep_scale=sweep_scale)
    # size classes by estimated operator count, each its own batch, host
    # thread and stream (small jobs do not wait behind large walkers)
    loop = sv.SizeClassServing(classes=((2000, slots * 3 // 16, 10, 48),
                                        (8000, slots * 5 // 16, 4, 48),
                                        (30000, slots * 5 // 16, 2, 48),
                                        (60000, slots * 2 // 16, 1, 48),
                                        (float("inf"), slots * 1 // 16, 1, 48)))
    done, stats = loop.run(jobs, time_limit=seconds * 3,
                           log=lambda m: print(f"serving: {m}", file=sys.stderr, flush=True))
    loop.close()
    stats["workload"] = (f"serving: Poisson 200 jobs/s seed 7 for {seconds:g} s (results "
                         f"counted for {3 * seconds:g} s), L 8..48, beta 1..32, sweeps "
                         f"log-uniform [1e3,1e5] x {sweep_scale:g}, {slots} slots in 5 size "
                         f"classes (n <= 2000 / 8000 / 30000 / 60000 / more), models prebuilt")
    stats["jobs_submitted"] = len(jobs)
    return stats


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--chunk", type=int, default=0,
                    help="sweeps per timed call (0: all timed steps in one call)")
    ap.add_argument("--streams", type=int, default=0,
                    help="walker groups, each swept on its own HIP stream (0: the "
                         "library's choice, one per 512 walkers up to one per hardware queue)")
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="GPU_MAX_HW_QUEUES for this process (HIP's default is 4): set "
                         "before the HIP runtime starts, so each walker group's stream "
                         "gets a hardware queue of its own")
    ap.add_argument("--walkers", type=int, default=4096)
    ap.add_argument("--lattice", type=int, default=16)
    ap.add_argument("--beta", type=float, default=8.0)
    ap.add_argument("--thermalization", type=int, default=2000)
    ap.add_argument("--plain-walkers", type=int, default=2)
    ap.add_argument("--plain-sweeps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="end with three device-to-device copies of this many MiB "
                         "(reference kernel for HBM traffic counters)")
    ap.add_argument("--serving-seconds", type=float, default=20.0,
                    help="arrival window of the serving stream reported in the default line "
                         "(p50/p99 time-to-result; 0 skips it, --extra runs 30 s)")
    ap.add_argument("--extra", action="store_true",
                    help="also run the heterogeneous, ladder and serving workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); walkers are independent "
              f"chains, so N GPUs run N independent batches.", file=sys.stderr)
        sys.exit(2)

    if not 1 <= args.hw_queues <= 32:
        print("--hw-queues takes 1..32", file=sys.stderr)
        sys.exit(2)
    # read by the HIP runtime when it starts (the library's first call below)
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
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
    # warm-up and timed region with the walkers split over `streams` stream
    # groups; the timed region is `steps` sweeps (+ fused measurement) of every
    # walker in calls of `chunk` sweeps
    batch.set_option(sse.capi.SSE_OPT_STREAMS, args.streams)
    queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4)
    groups = args.streams or min(queues, max(2, args.walkers // 512))  # the library's rule
    if args.warmup:
        batch.run(args.warmup, measure=True)
    batch.reset_accumulators()
    chunk = max(1, min(args.chunk if args.chunk > 0 else args.steps, args.steps))
    call_ms, dev_ms = [], 0.0
    done = 0
    while done < args.steps:
        k = min(chunk, args.steps - done)
        t1 = time.perf_counter()
        batch.run(k, measure=True)
        call_ms.append((time.perf_counter() - t1) * 1e3 / k)
        ms, sw = batch.call_time()
        dev_ms += ms
        done += k
    Ls, ns, Ws, wsweeps = batch.work_counters()
    total_s = sum(call_ms[i] * min(chunk, args.steps - i * chunk) for i in range(len(call_ms))) / 1e3
    U = Ls + Ws
    B = 16 * Ls + 16 * ns + 12 * Ws + 3 * N * wsweeps
    value = U / total_s
    log(f"timed {args.steps} steps: {total_s:.3f} s wall, {dev_ms / 1e3:.3f} s device, "
        f"{wsweeps} walker-sweeps, <L>={Ls / wsweeps:.0f} <n>={ns / wsweeps:.0f} "
        f"<W>={Ws / wsweeps:.0f}")

    # roofline of the dominant kernel: three more sweeps on one stream with
    # HIP events around each kernel (untimed diagnostic); bytes per launch
    # from the per-walker-sweep averages of the timed region
    batch.set_option(sse.capi.SSE_OPT_STREAMS, 1)
    ktimes = np.zeros(3)
    for _ in range(3):
        batch.run(1, measure=True)
        ktimes += batch.phase_times()
    batch.set_option(sse.capi.SSE_OPT_STREAMS, args.streams)
    lds = batch.lds_bytes()
    names = ["diag_vlist", "worm", "measure"]
    kms = {n: ktimes[i] / 3 for i, n in enumerate(names)}
    nw = args.walkers
    per_launch = {"diag_vlist": nw * (12 * Ls + 16 * ns) / wsweeps + 2 * N * nw,
                  "worm": nw * 12 * Ws / wsweeps + N * nw,
                  "measure": nw * 4 * Ls / wsweeps}
    dom = max(names, key=lambda n: kms[n])
    kname = {"diag_vlist": "wave_diag_vlist_kernel", "worm": "wave_worm_kernel",
             "measure": "wave_measure_kernel"}[dom]
    achieved = per_launch[dom] / (kms[dom] * 1e-3) / 1e9
    log("single-stream kernels ms/launch: " + ", ".join(f"{n}={kms[n]:.3f}" for n in names))

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
        "ms_per_step": total_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic (Philox-seeded random initial spins, thermalized chains)",
        "config": {"workload": f"batched Heisenberg {args.lattice}x{args.lattice} beta={args.beta:g}",
                   "walkers": args.walkers, "thermalization_sweeps": args.thermalization,
                   "parallelism": f"one wavefront per walker, {groups} stream groups on "
                                  f"{os.environ.get('GPU_MAX_HW_QUEUES')} hardware queues"},
        "sweeps_per_s": wsweeps / total_s,
        "hbm_gbs_algorithmic": B / total_s / 1e9,
        "hbm_roofline_frac": B / total_s / 1e9 / HBM_PEAK_GBS,
        "step_ms_per_call": {"chunk": chunk, "p50": float(np.percentile(call_ms, 50)),
                             "p99": float(np.percentile(call_ms, 99))},
        "device_ms_per_step": dev_ms / args.steps,
        "per_walker_sweep": {"L": Ls / wsweeps, "n": ns / wsweeps, "W": Ws / wsweeps,
                             "U": U / wsweeps, "B": B / wsweeps},
        "kernel_ms_per_launch_single_stream": kms,
        "lds_bytes_per_walker": {"diag_vlist": lds[0], "worm": lds[1], "worm_walkers_per_cu": lds[2]},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kname, args.walkers, args.lattice, args.beta),
                     "algorithmic_bytes_per_launch": per_launch[dom]},
        "cpu_baseline": cpu,
    }
    out.update(baselines)
    if args.serving_seconds > 0 and not args.extra:
        log(f"serving stream: {args.serving_seconds:g} s of arrivals")
        try:
            st = extra_serving(sse, seconds=args.serving_seconds)
            out["serving"] = {k: st[k] for k in ("p50_time_to_result_s", "p99_time_to_result_s",
                                                 "jobs_done", "jobs_submitted", "wall_s",
                                                 "vertex_updates_per_s", "workload")}
        except Exception as e:  # the headline line is reported either way
            out["serving"] = {"error": f"{type(e).__name__}: {e}"}
    if args.extra:
        extras = {}
        for name, fn in (("heterogeneous", extra_heterogeneous), ("ladder", extra_ladder),
                         ("serving", extra_serving)):
            log(f"extra workload: {name}")
            extras[name] = fn(sse)
        out["extra_workloads"] = extras
    if args.calibrate_copy_mb:
        copy_calibration(args.calibrate_copy_mb)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Batched SSE sweep throughput on one MI355X.

Workload (the batched Heisenberg workload): 4096 independent walkers of the
S=1/2 Heisenberg antiferromagnet (J=1) on a 16x16 periodic square lattice at
beta=8, Philox seeds 0..4095, random initial spins, cutoff grown by the
reference rule, 2000 thermalization sweeps (untimed setup).  One step = one
sweep (diagonal update + vertex list + worm update) of every walker followed
by the measurement (energy, uniform and staggered magnetization family),
accumulated on the device.  After thermalization the library's autotuner
measures its launch forms (stream groups at several counts, the persistent
work-queue launch at two LDS sizes, the record-lookahead chase) on real
sweeps (untimed setup) and keeps the fastest; then the warm-up steps and
the timed steps, in one call.  The process asks for --hw-queues (8) hardware
queues (GPU_MAX_HW_QUEUES) so up to 8 stream groups get a queue each.

Reported beside it, in the same run:
  * the roofline of the dominant kernel in the timed configuration: the same
    number of steps again with HIP events around every kernel launch on the
    stream it runs on (direct launches), giving each kernel's average launch
    duration and the algorithmic bytes it moved;
  * the serving stream (open-loop Poisson arrivals at 200 jobs/s of jobs
    with 10^3-10^5 sweeps, 20 s of arrivals, results counted for 60 s):
    p50/p99 time-to-result over all submitted jobs (unfinished ones at their
    censoring time), completion fraction, bins streamed;
  * the plain GPU form: the same walkers (thermalized states) run one walker
    per launch with the lane kernel, and with the wave kernels;
  * the plain single walker (8x8, beta=2, cutoff from 64) on the GPU and on
    one CPU thread;
  * the CPU baseline: the C restatement oracle, one walker per thread on the
    process's CPU share, on thermalized states copied from the GPU, for a
    bounded time.

With --extra [NAME,...] it also runs the heterogeneous XXZ batch (per-walker
budgets), the dim-4 ladder and the three-site-bond trimer ring, each with its
own per-kernel times and roofline from HIP events around every launch.

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


def measured_traffic(kernel, walkers_per_launch):
    """HBM bytes per launch of `kernel` over `walkers_per_launch` walkers of
    the batched 16x16 beta=8 workload, from the committed rocprofv3 counter
    summary of the timed configuration (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated): (bytes, note), or (None, reason).  The summary was
    taken at one walker count per launch; at another count (the tuner may
    pick another group count) the bytes are scaled by the walker count, since
    every byte a launch moves belongs to one of its walkers (records,
    operator string, spins)."""
    import re
    path = os.path.join(ROOT, "profiles",
                        "hbm_traffic_timed_configuration_heisenberg16x16_beta8_4096walkers.md")
    try:
        text = open(path).read()
    except OSError:
        return None, "no counter summary"
    m = re.search(rf"traffic: kernel={kernel} walkers_per_launch=(\d+) bytes_per_launch=(\d+)",
                  text)
    if not m:
        return None, f"counter summary does not cover {kernel}"
    w0, b0 = int(m.group(1)), int(m.group(2))
    src = os.path.relpath(path, ROOT)
    if w0 == walkers_per_launch:
        return b0, f"{src}: measured at {w0} walkers per launch"
    return (int(round(b0 * walkers_per_launch / w0)),
            f"{src}: measured at {w0} walkers per launch, scaled to {walkers_per_launch}")


def issue_bound(kernel):
    """Instruction-issue figures of `kernel` from the committed rocprofv3
    counter summary of the same workload (SQ_INSTS_SALU / SQ_INSTS_VALU per
    launch, the launch's duration): the rate at which the launch issued
    scalar and vector instructions against MI355X's issue peak (256 CUs, one
    scalar and one wave-wide vector instruction per CU per cycle at 2.4 GHz),
    or None when the summary does not cover the kernel."""
    import re
    path = os.path.join(ROOT, "profiles", "issue_counters_heisenberg16x16_beta8_4096walkers.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(rf"issue: kernel={kernel} salu_per_launch=(\d+) valu_per_launch=(\d+) "
                  rf"launch_ms=([0-9.]+)", text)
    if not m:
        return None
    salu, valu, ms = int(m.group(1)), int(m.group(2)), float(m.group(3))
    peak = 256 * 2.4e9  # instructions per second, per type
    return {"salu_per_launch": salu, "valu_per_launch": valu, "launch_ms": ms,
            "salu_frac": salu / (ms * 1e-3) / peak, "valu_frac": valu / (ms * 1e-3) / peak,
            "peak_per_type": peak, "unit": "wave-instructions/s",
            "source": os.path.relpath(path, ROOT)}


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


def tune(batch, what, max_calls=12):
    """Run measured sweeps until the library's autotuner has chosen (setup,
    untimed); returns its choice."""
    for _ in range(max_calls):
        batch.run(8, measure=True)
        choice = batch.tuning(0)
        if choice is not None:
            log(f"{what}: tuned {choice}")
            return choice
    return None


def parallelism(choice, walkers, generic=False):
    if not choice:
        return "one wavefront per walker (untuned defaults)"
    chase = "record-lookahead chase" if choice["chase"] else "scalar chase"
    lds = ("worm LDS sized from the largest walker" if not choice["k2_lds"]
           else f"worm LDS {choice['k2_lds']} B")
    if generic:
        chase = "chase codes in LDS"
        lds = "worm LDS sized from the largest walker"
    if choice["persistent"]:
        return (f"autotuned: persistent work-queue launch (one workgroup per walker task, "
                f"{walkers} tasks per call, HIP-graph replay), {lds}, {chase}")
    return (f"autotuned: {choice['groups']} stream groups of per-sweep kernels (one stream and "
            f"hardware queue per group, direct launches), one wavefront per walker, {lds}, "
            f"{chase}")


def kernel_region(sse, batch, flats, wds, steps, budgets=None):
    """The timed calls once more with HIP events around every kernel launch on
    its stream (SSE_OPT_TIMING 2, direct launches): per kernel the launches,
    mean launch time and algorithmic bytes per launch, and the HBM roofline of
    the kernel with the most time.  Bytes per walker-sweep, with P = 2 x sites
    per bond link words per operator and N sites: diagonal update + vertex list
    12L + 4Pn + 2N, worm update 12W + N, measurement 4L; the persistent launch
    runs all three."""
    batch.reset_accumulators()
    batch.set_option(sse.capi.SSE_OPT_TIMING, 2)
    ms = np.zeros(4)
    cnt = np.zeros(4, dtype=np.int64)
    calls = 1 if budgets is None else steps
    for _ in range(calls):
        if budgets is None:
            batch.run(steps, measure=True)
        else:
            batch.run_budgets(budgets, measure=True)
        m, c = batch.kernel_times()
        ms += m
        cnt += c
    batch.set_option(sse.capi.SSE_OPT_TIMING, 1)
    L, n, W, wsw = batch.work_counters()
    sites = np.array([flats[w.model].n_sites for w in wds], dtype=np.float64)
    sweeps = np.full(len(wds), steps) if budgets is None else np.asarray(budgets) * steps
    Nsw = float(np.sum(sites * sweeps))
    P = 2 * max(f.bond_n_sites for f in flats)
    generic = P != 4
    names = ("diag_vlist", "worm", "measure", "persistent")
    kernels = dict(zip(names, ("waven_diag_vlist_kernel", "waven_worm_kernel",
                               "waven_measure_kernel") if generic else
                       ("wave_diag_vlist_kernel", "wave_worm_kernel", "wave_measure_kernel")))
    kernels["persistent"] = "wave_persist_kernel"
    region = {"diag_vlist": 12 * L + 4 * P * n + 2 * Nsw, "worm": 12 * W + Nsw, "measure": 4 * L}
    region["persistent"] = sum(region.values())
    per_kernel = {}
    for i, nm in enumerate(names):
        if cnt[i]:
            avg = ms[i] / cnt[i]
            bpl = region[nm] / cnt[i]
            per_kernel[nm] = {"kernel": kernels[nm], "launches": int(cnt[i]), "avg_launch_ms": avg,
                              "algorithmic_bytes_per_launch": bpl,
                              "achieved_gbs_per_launch": bpl / (avg * 1e-3) / 1e9}
    if not per_kernel:
        return {}, None
    dom = max(per_kernel, key=lambda nm: ms[names.index(nm)])
    d = per_kernel[dom]
    roof = {"bound": "hbm", "kernel": d["kernel"], "achieved": d["achieved_gbs_per_launch"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["achieved_gbs_per_launch"] / HBM_PEAK_GBS,
            "traffic": None, "walker_sweeps": int(wsw),
            "avg_launch_ms": d["avg_launch_ms"], "launches": d["launches"]}
    log("kernels: " + ", ".join(f"{v['kernel']} {v['launches']} x {v['avg_launch_ms']:.3f} ms"
                                for v in per_kernel.values()))
    return per_kernel, roof


def run_batch(sse, flats, wds, thermalization, steps, what, budgets=None, refine=None):
    """Thermalize, tune, then time `steps` sweeps (equal budgets) or one
    budgeted call per step (budgets: per-walker sweeps per call; refine, when
    given, replaces them after tuning: refine(batch, choice) -> budgets)."""
    batch = sse.Batch(flats, wds)
    t0 = time.perf_counter()
    batch.init()
    done = 0
    while done < thermalization:
        k = min(100, thermalization - done)
        batch.sweep(k)
        done += k
    log(f"{what}: thermalized {thermalization} sweeps in {time.perf_counter() - t0:.1f} s")
    if budgets is None:
        choice = tune(batch, what)
        batch.run(2, measure=True)
    else:
        choice = None
        for _ in range(12):  # budgeted calls tune over their first calls
            batch.run_budgets(budgets, measure=True)
            choice = batch.tuning(1)
            if choice is not None:
                break
        log(f"{what}: tuned {choice}")
        if refine is not None:
            budgets = refine(batch, choice)
            batch.run_budgets(budgets, measure=True)
    batch.reset_accumulators()
    t1 = time.perf_counter()
    if budgets is None:
        batch.run(steps, measure=True)
    else:
        for _ in range(steps):
            batch.run_budgets(budgets, measure=True)
    dt = time.perf_counter() - t1
    Ls, ns, Ws, wsw = batch.work_counters()
    per_kernel, roof = kernel_region(sse, batch, flats, wds, steps, budgets)
    batch.close()
    generic = any(f.bond_n_sites != 2 for f in flats)
    return {"vertex_updates_per_s": (Ls + Ws) / dt, "sweeps_per_s": wsw / dt,
            "kernels_timed_configuration": per_kernel, "roofline": roof,
            "walkers": len(wds), "thermalization_sweeps": thermalization,
            "timed": (f"{steps} sweeps of every walker" if budgets is None else
                      f"{steps} budgeted calls, {int(np.sum(budgets))} walker-sweeps each"),
            "mean_L": Ls / wsw, "mean_n": ns / wsw, "mean_W": Ws / wsw,
            "parallelism": parallelism(choice, len(wds), generic) if budgets is None else
            (f"persistent work-queue launch with per-walker budgets (costliest walker "
             f"{int(np.min(budgets))} sweep(s) per call, cheapest {int(np.max(budgets))}), tuned "
             f"{choice}")}


def heterogeneous_models(sse, walkers=2048, seed=3):
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
    flats = []
    for k in range(walkers):
        L = int(sizes[k])
        m = base[L]
        flats.append(FlatModel(SSEData(types[k], bonds[L]), m.opstring_estimators,
                               m.normalization_site_count(),
                               (base_flat[L].est_coef, base_flat[L].est_mval)))
    return flats, betas


def sweep_budgets(cost, factor=4, cap=400):
    """Per-walker sweeps per call of about `factor` sweeps of the costliest
    walker's work, so that every walker's budget takes about the same time."""
    cost = np.asarray(cost, dtype=np.float64)
    return np.minimum(cap, np.maximum(1, np.round(factor * cost.max() / cost))).astype(np.int64)


def operator_count_costs(batch, n_walkers, k2_lds, slow=3.0):
    """Work per sweep of each walker from its chain: the operator count n,
    times `slow` when the worm update cannot hold the walker's records in
    LDS (8-byte records need n <= 4096 and 8n <= the worm LDS per walker;
    beyond, links are read from L2 at about three times the time per step)."""
    ns = np.array([batch.get_walker(k, with_ops=False, with_spins=False)[0]["num_operators"]
                   for k in range(n_walkers)], dtype=np.float64)
    in_lds = ns <= min(4096, k2_lds // 8)
    return np.maximum(ns, 1.0) * np.where(in_lds, 1.0, slow), ns, in_lds


def extra_heterogeneous(sse, walkers=2048, thermalization=300, steps=6, seed=3):
    """The heterogeneous XXZ batch, each walker advancing by its own budget per
    call: about 4 sweeps of the costliest walker's work per call.  The
    autotuner measures on budgets from the expected work (bonds x beta); the
    timed calls use budgets from each walker's operator count after
    thermalization, weighted by the worm update's record layout."""
    flats, betas = heterogeneous_models(sse, walkers, seed)
    wds = [sse.walker_desc(model=k, T=1.0 / float(betas[k]), seed=100000 + k,
                           thermalization=thermalization) for k in range(walkers)]
    budgets = sweep_budgets([flats[k].n_bonds * float(betas[k]) for k in range(walkers)])
    info = {}

    def refine(batch, choice):
        k2 = int(choice["k2_lds"]) if choice and choice["k2_lds"] else int(batch.lds_bytes()[1])
        cost, ns, in_lds = operator_count_costs(batch, walkers, k2)
        info.update({"worm_lds_bytes": k2, "walkers_records_in_lds": int(in_lds.sum()),
                     "n_quantiles_50_90_99_100": np.percentile(ns, [50, 90, 99, 100]).tolist()})
        log(f"heterogeneous: budgets from operator counts, {int(in_lds.sum())} of {walkers} "
            f"walkers with records in LDS ({k2} B)")
        return sweep_budgets(cost)

    r = run_batch(sse, flats, wds, thermalization, steps, "heterogeneous", budgets=budgets,
                  refine=refine)
    r["budgets"] = dict(info, rule="about 4 sweeps of the costliest walker's work per call, "
                                   "cost = operator count x (3 if the records exceed the worm "
                                   "LDS), at most 400 sweeps")
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


def _trimer_ring_flat(sse, n, J=-1.0, h=0.15):
    """Spin-1/2 ring whose bonds are three consecutive sites (SSEBond{3}):
    H_b = J (S_a.S_b + S_b.S_c + S_a.S_c) - (h/3) sum S^z, ferromagnetic, with
    a magnetization estimator."""
    from sse_amd.sse_data import FlatModel, SSEBond, SSEData
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    eye = np.eye(2)

    def lift(op, k):
        mats = [eye, eye, eye]
        mats[k] = op
        return np.kron(np.kron(mats[0], mats[1]), mats[2])

    Hb = np.zeros((8, 8))
    for a, b in ((0, 1), (1, 2), (0, 2)):
        Hb += J * (lift(sz, a) @ lift(sz, b)
                   + 0.5 * (lift(sp, a) @ lift(sp.T, b) + lift(sp.T, a) @ lift(sp, b)))
    for a in range(3):
        Hb -= (h / 3.0) * lift(sz, a)
    vd = sse.VertexData((2, 2, 2), Hb)
    bonds = [SSEBond(1, (i + 1, (i + 1) % n + 1, (i + 2) % n + 1)) for i in range(n)]

    class _Est:
        prefix = ""
    mval = np.tile(np.array([0.5, -0.5]), (1, n, 1))
    return FlatModel(SSEData([vd], bonds), [_Est()], n, (np.ones((1, n)), mval))


def extra_trimer(sse, walkers=4096, thermalization=1000, steps=20):
    """Bonds of three sites through the generic wave kernels: 4096 walkers of
    a 256-site trimer ring at beta = 8."""
    flat = _trimer_ring_flat(sse, 256)
    wds = [sse.walker_desc(T=0.125, seed=300000 + k, thermalization=thermalization)
           for k in range(walkers)]
    r = run_batch(sse, [flat], wds, thermalization, steps, "trimer")
    r["workload"] = ("three-site bonds (SSEBond{3}): 256-site spin-1/2 trimer ring, J=-1, h=0.15, "
                     "4096 walkers, beta=8, generic wave kernels")
    return r


EXTRA_WORKLOADS = {"heterogeneous": extra_heterogeneous, "ladder": extra_ladder,
                   "trimer": extra_trimer}


def serving_stream(sse, seconds=20.0, count_for=60.0, sweep_scale=1.0, slots=1024):
    """Open-loop Poisson stream (200 jobs/s, seed 7) of Heisenberg/XXZ jobs on
    L in 8..48, beta log-uniform [1,32], sweeps log-uniform [1e3,1e5] (times
    sweep_scale); arrivals for `seconds`, results counted until `count_for`."""
    from importlib import import_module
    sv = import_module("sse_amd.serving")
    jobs = sv.poisson_jobs(int(200 * seconds), rate=200.0, seed=7, sweep_scale=sweep_scale)
    # size classes by estimated operator count, each its own batch, host
    # thread and stream; within a class per-slot budgets balance the jobs
    loop = sv.SizeClassServing(classes=((2000, slots * 3 // 16, 10, 48),
                                        (8000, slots * 5 // 16, 4, 48),
                                        (30000, slots * 5 // 16, 2, 48),
                                        (60000, slots * 2 // 16, 1, 48),
                                        (float("inf"), slots * 1 // 16, 1, 48)))
    done, stats = loop.run(jobs, time_limit=count_for,
                           log=lambda m: print(f"serving: {m}", file=sys.stderr, flush=True))
    loop.close()
    scale = "" if sweep_scale == 1.0 else f" x {sweep_scale:g}"
    stats["workload"] = (f"serving: Poisson 200 jobs/s seed 7 for {seconds:g} s of arrivals "
                         f"(results counted for {count_for:g} s), Heisenberg/XXZ on L 8..48, "
                         f"beta log-uniform 1..32, sweeps log-uniform [1e3,1e5]{scale}, "
                         f"{slots} slots in 5 size classes (n <= 2000 / 8000 / 30000 / 60000 / "
                         f"more) with per-slot budgets, models prebuilt")
    return stats


def carlo_boundary(sse, flat, wds, batch, tasks=16, steps=30, warm=15):
    """The batched workload driven through Carlo's AbstractMC methods
    (carlo_mc): `tasks` MC objects, each carrying len(wds)/tasks runs, in
    one shared batch, stepped by step_all (sweep! of every MC, then
    measure! of every MC: one launch per Carlo sweep, measurements handed
    to each MC's context).  The walkers start from `batch`'s thermalized
    states.  Also the same walkers as one MC per run (the slot-per-MC
    pattern), whose host side is one Python object per walker."""
    cm = sse.carlo_mc
    n = len(wds)
    # the direct path at one sweep per call: the per-call floor any
    # one-sweep-per-Carlo-sweep driver pays
    batch.run(2, measure=True)
    Lw = batch.work_counters()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run(1, measure=True)
    dt = time.perf_counter() - t0
    Lw1 = batch.work_counters()
    out = {"direct_one_sweep_per_call": {
        "vertex_updates_per_s": (Lw1[0] - Lw[0] + Lw1[2] - Lw[2]) / dt,
        "ms_per_call": dt * 1e3 / steps, "calls": steps}}
    states = [batch.get_walker(k) for k in range(n)]
    for name, n_mc in (("mc_with_runs", tasks), ("mc_per_run", n)):
        R = n // n_mc
        sh = cm.SharedBatch(flat, n)
        mcs = []
        for t in range(n_mc):
            mc = cm.MC({"model": flat, "T": wds[t * R].T, "seed": 10 ** 6 + t * R,
                        "thermalization": 0}, sh, runs=R)
            for r, slot in enumerate(mc.slots):
                sc, ops, spins = states[t * R + r]
                sh.batch.set_walker(int(slot), sc, ops, spins)
            mcs.append(mc)
        ctxs = [cm.MCContext(0) for _ in mcs]
        k_warm, k_steps = (warm, steps) if n_mc == tasks else (3, 5)
        for _ in range(k_warm):
            cm.step_all(mcs, ctxs)
        sh.work[:] = 0
        l0 = sh.launches
        t0 = time.perf_counter()
        for _ in range(k_steps):
            cm.step_all(mcs, ctxs)
        dt = time.perf_counter() - t0
        Ls, ns, Ws, wsw = (int(x) for x in sh.work)
        out[name] = {"vertex_updates_per_s": (Ls + Ws) / dt, "ms_per_carlo_sweep": dt * 1e3 / k_steps,
                     "mc_objects": n_mc, "runs_per_mc": R, "launches_per_carlo_sweep":
                     (sh.launches - l0) / k_steps, "walker_sweeps": wsw, "timed_steps": k_steps,
                     "samples_per_mc": len(ctxs[0].series("Sign"))}
        sh.close()
    out["workload"] = (f"{n} thermalized walkers of the batched workload as Carlo MC objects in "
                       f"one shared batch; step_all per Carlo sweep (every sweep! posted, one "
                       f"launch, then every measure!)")
    return out


def single_walker(sse, thermalization=10000, sweeps=5000):
    """The plain single walker: 8x8 Heisenberg, beta=2, seed 1, initial cutoff
    64, on the GPU (wave and lane kernels, one walker per batch) and on one
    CPU thread (the oracle)."""
    from oracle import pyoracle
    model = sse.MagnetModel({"lattice": {"unitcell": sse.UnitCells.square, "size": (8, 8)},
                             "J": 1.0, "measure": ["magnetization", "staggered_magnetization"]})
    flat = model.flat()
    wd = sse.walker_desc(T=0.5, seed=1, init_opstring_cutoff=64, thermalization=thermalization)
    out = {"workload": f"8x8 Heisenberg beta=2, seed 1, cutoff from 64, {thermalization} "
                       f"thermalization sweeps, {sweeps} timed sweeps with measurement"}
    for name, mode in (("gpu_wave", sse.SSE_MODE_WAVE), ("gpu_lane", sse.SSE_MODE_LANE)):
        b = sse.Batch([flat], [wd], mode=mode)
        b.set_option(sse.capi.SSE_OPT_TIMING, 0)
        b.init()
        b.run(thermalization, measure=False)
        b.reset_accumulators()
        t0 = time.perf_counter()
        b.run(sweeps, measure=True)
        dt = time.perf_counter() - t0
        Ls, ns, Ws, nsw = b.work_counters()
        out[name] = {"vertex_updates_per_s": (Ls + Ws) / dt, "sweeps_per_s": nsw / dt,
                     "mean_n": ns / nsw}
        b.close()
    o = pyoracle.OracleWalker(sse.capi, flat, wd)
    o.init()
    o.run_timed(thermalization, measure=False)
    o.work_counters()
    t, _ = o.run_timed(sweeps, measure=True)
    Ls, ns, Ws, nsw = o.work_counters()
    out["cpu_one_thread"] = {"vertex_updates_per_s": (Ls + Ws) / t, "sweeps_per_s": nsw / t,
                             "kind": "port (C restatement oracle)"}
    return out


def dispatch_concurrency(spans):
    """Overlap of kernel dispatches given as (start, end) timestamps in ns: the
    window from the first start to the last end, the time at least one runs
    (busy), the time-weighted mean number running over the window and over the
    busy time, and the largest number running at once."""
    events = sorted([(t0, 1) for t0, _ in spans] + [(t1, -1) for _, t1 in spans],
                    key=lambda e: (e[0], e[1]))
    running = peak = 0
    area = busy = 0
    prev = events[0][0]
    for t, d in events:
        if running:
            area += running * (t - prev)
            busy += t - prev
        prev = t
        running += d
        peak = max(peak, running)
    window = events[-1][0] - events[0][0]
    return {"dispatches": len(spans), "window_ms": window * 1e-6, "busy_ms": busy * 1e-6,
            "mean_running": area / window if window else 0.0,
            "mean_running_busy": area / busy if busy else 0.0, "max_running": peak}


def summarize_rocprof(directory, last=0, regex=None):
    """Markdown tables from rocprofv3 CSV output under `directory`: per kernel
    and launch grid, the kernel-trace durations (calls, mean/min/max ms), the
    overlap of the last `last` matched dispatches of each trace, and, for
    counter collections, each counter's mean per dispatch over the last `last`
    dispatches of that kernel and grid (all when 0)."""
    import csv
    import glob
    import re
    pat = re.compile(regex) if regex else None
    lines = []

    def col(row, *names):
        for n in names:
            if n in row:
                return row[n]
        return None

    for path in sorted(glob.glob(os.path.join(directory, "**", "*kernel_trace.csv"), recursive=True)):
        groups = {}
        spans = []
        for r in csv.DictReader(open(path)):
            name = (col(r, "Kernel_Name") or "").split("(")[0]
            if pat and not pat.search(name):
                continue
            grid = col(r, "Grid_Size_X", "Grid_Size") or "?"
            t0, t1 = int(col(r, "Start_Timestamp")), int(col(r, "End_Timestamp"))
            groups.setdefault((name, grid), []).append((t1 - t0) * 1e-6)
            spans.append((t0, t1))
        lines.append(f"kernel trace {os.path.relpath(path, directory)}:")
        lines.append("| kernel | grid | calls | avg ms | min ms | max ms | avg ms, last |")
        lines.append("|---|---|---|---|---|---|---|")
        for (name, grid), ts in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
            tail = ts[-last:] if last else ts
            lines.append(f"| {name} | {grid} | {len(ts)} | {np.mean(ts):.4f} | {min(ts):.4f} | "
                         f"{max(ts):.4f} | {np.mean(tail):.4f} ({len(tail)}) |")
        if spans:
            c = dispatch_concurrency(sorted(spans)[-last:] if last else spans)
            lines.append(f"concurrency over the last {c['dispatches']} dispatches: window "
                         f"{c['window_ms']:.3f} ms, busy {c['busy_ms']:.3f} ms, "
                         f"mean dispatches running {c['mean_running']:.2f} "
                         f"({c['mean_running_busy']:.2f} while any runs), max {c['max_running']}")
    for path in sorted(glob.glob(os.path.join(directory, "**", "*counter_collection.csv"),
                                 recursive=True)):
        groups = {}
        for r in csv.DictReader(open(path)):
            name = (col(r, "Kernel_Name") or "").split("(")[0]
            if pat and not pat.search(name):
                continue
            grid = col(r, "Grid_Size", "Grid_Size_X") or "?"
            disp = int(col(r, "Dispatch_Id", "Correlation_Id") or 0)
            cname = col(r, "Counter_Name")
            groups.setdefault((name, grid, cname), {}).setdefault(disp, 0.0)
            groups[(name, grid, cname)][disp] += float(col(r, "Counter_Value"))
        lines.append(f"counters {os.path.relpath(path, directory)}:")
        lines.append("| kernel | grid | counter | dispatches | mean per dispatch (last) |")
        lines.append("|---|---|---|---|---|")
        for (name, grid, cname), d in sorted(groups.items()):
            vals = [d[k] for k in sorted(d)]
            tail = vals[-last:] if last else vals
            lines.append(f"| {name} | {grid} | {cname} | {len(vals)} | {np.mean(tail):.6g} "
                         f"({len(tail)}) |")
    return "\n".join(lines)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-rocprof":
        sp = argparse.ArgumentParser(description="summarize rocprofv3 CSV output")
        sp.add_argument("--summarize-rocprof", dest="directory", required=True)
        sp.add_argument("--last", type=int, default=0)
        sp.add_argument("--kernel-regex", default=None)
        a = sp.parse_args()
        print(summarize_rocprof(a.directory, a.last, a.kernel_regex))
        return
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--streams", type=int, default=0,
                    help="fix the walker groups (each swept on its own HIP stream) instead of "
                         "letting the autotuner choose the launch form (0: autotuned)")
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="GPU_MAX_HW_QUEUES for this process (HIP's default is 4): set "
                         "before the HIP runtime starts, so each walker group's stream "
                         "gets a hardware queue of its own")
    ap.add_argument("--walkers", type=int, default=4096)
    ap.add_argument("--lattice", type=int, default=16)
    ap.add_argument("--beta", type=float, default=8.0)
    ap.add_argument("--thermalization", type=int, default=2000)
    ap.add_argument("--plain-walkers", type=int, default=8)
    ap.add_argument("--plain-sweeps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads of the CPU baseline (0: the process's CPU share, "
                         "min(nproc, 16))")
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--no-carlo", action="store_true",
                    help="skip the Carlo-boundary measurement (MC objects over a shared batch)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="end with three device-to-device copies of this many MiB "
                         "(reference kernel for HBM traffic counters)")
    ap.add_argument("--serving-seconds", type=float, default=20.0,
                    help="arrival window of the serving stream (0 skips it); results are "
                         "counted for three times as long")
    ap.add_argument("--extra", nargs="?", const=",".join(EXTRA_WORKLOADS), default="",
                    help="also run the heterogeneous, ladder and three-site-bond (trimer) "
                         "workloads, or the comma-separated ones named")
    args = ap.parse_args()
    unknown = [x for x in args.extra.split(",") if x and x not in EXTRA_WORKLOADS]
    if unknown:
        ap.error(f"--extra: unknown workload(s) {unknown}; choose from {list(EXTRA_WORKLOADS)}")
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
    if args.streams:
        batch.set_option(sse.capi.SSE_OPT_STREAMS, args.streams)
        choice = None
    else:
        choice = tune(batch, "batched workload")
    if args.warmup:
        batch.run(args.warmup, measure=True)
    batch.reset_accumulators()
    t1 = time.perf_counter()
    batch.run(args.steps, measure=True)
    total_s = time.perf_counter() - t1
    dev_ms, _ = batch.call_time()
    Ls, ns, Ws, wsweeps = batch.work_counters()
    U = Ls + Ws
    B = 16 * Ls + 16 * ns + 12 * Ws + 3 * N * wsweeps
    value = U / total_s
    log(f"timed {args.steps} steps: {total_s:.3f} s wall, {dev_ms / 1e3:.3f} s device, "
        f"{wsweeps} walker-sweeps, <L>={Ls / wsweeps:.0f} <n>={ns / wsweeps:.0f} "
        f"<W>={Ws / wsweeps:.0f}")

    # roofline of the dominant kernel in the timed configuration: the same
    # number of steps again with events around every launch on its stream
    batch.reset_accumulators()
    batch.set_option(sse.capi.SSE_OPT_TIMING, 2)
    batch.run(args.steps, measure=True)
    batch.set_option(sse.capi.SSE_OPT_TIMING, 1)
    diag_ms, _ = batch.call_time()
    kms, klaunch = batch.kernel_times()
    dL, dn, dW, dsw = batch.work_counters()
    names = ["diag_vlist", "worm", "measure", "persistent"]
    kernels = {"diag_vlist": "wave_diag_vlist_kernel", "worm": "wave_worm_kernel",
               "measure": "wave_measure_kernel", "persistent": "wave_persist_kernel"}
    # algorithmic bytes of each kernel over the region (per walker-sweep:
    # K1 12L + 16n + 2N, K2 12W + N, K3 4L; the persistent launch all of B)
    region_bytes = {"diag_vlist": 12 * dL + 16 * dn + 2 * N * dsw, "worm": 12 * dW + N * dsw,
                    "measure": 4 * dL,
                    "persistent": 16 * dL + 16 * dn + 12 * dW + 3 * N * dsw}
    per_kernel = {}
    for i, nm in enumerate(names):
        if klaunch[i]:
            avg = kms[i] / klaunch[i]
            bpl = region_bytes[nm] / klaunch[i]
            per_kernel[nm] = {"launches": int(klaunch[i]), "avg_launch_ms": avg,
                              "algorithmic_bytes_per_launch": bpl,
                              "achieved_gbs_per_launch": bpl / (avg * 1e-3) / 1e9}
    dom = max(per_kernel, key=lambda nm: kms[names.index(nm)])
    d = per_kernel[dom]
    if dom == "persistent":
        walkers_per_launch = args.walkers
    else:
        g = args.streams or (choice["groups"] if choice else 1)
        walkers_per_launch = args.walkers // max(1, g)
    log("kernels in the timed configuration: " + ", ".join(
        f"{nm} {v['launches']} x {v['avg_launch_ms']:.3f} ms" for nm, v in per_kernel.items()))
    traffic = (measured_traffic(kernels[dom], walkers_per_launch)
               if (args.walkers, args.lattice, args.beta) == (4096, 16, 8.0)
               else (None, "counter summary covers the 4096-walker 16x16 beta=8 workload only"))
    lds = batch.lds_bytes()
    states = [batch.get_walker(k) for k in range(max(args.plain_walkers, 16))]
    baselines = {}
    if not args.no_baselines:
        log("plain form: one walker per launch (lane kernel)")
        plain_lane = plain_form(sse, flat, wds, states[:args.plain_walkers],
                                max(1, args.plain_sweeps // 2), sse.SSE_MODE_LANE)
        log("plain form: one walker per launch (wave kernels)")
        plain_wave = plain_form(sse, flat, wds, states[:args.plain_walkers], args.plain_sweeps,
                                sse.SSE_MODE_WAVE)
        baselines = {"plain_lane_vertex_updates_per_s": plain_lane,
                     "plain_wave_vertex_updates_per_s": plain_wave,
                     "plain_sample": f"{args.plain_walkers} thermalized walkers, each its own "
                                     f"batch: {args.plain_sweeps} sweeps (wave kernels) / "
                                     f"{max(1, args.plain_sweeps // 2)} (lane kernel) each",
                     "batched_over_plain_lane": value / plain_lane,
                     "batched_over_plain_wave": value / plain_wave}
        nproc = os.cpu_count() or 1
        threads = args.cpu_threads or min(16, nproc)
        log(f"cpu baseline: {threads} threads x {args.cpu_seconds} s")
        cpu = cpu_baseline(sse, flat, wds, states, args.cpu_seconds, threads)
        cpu["host_nproc"] = nproc
        cpu["sample"] += (f"; {threads} threads = the process's CPU share on the GPU host "
                          f"(min(nproc, 16), nproc = {nproc})")
        log("plain single walker (8x8, beta=2)")
        baselines["single_walker"] = single_walker(sse)
    else:
        cpu = None
    if not args.no_carlo:
        log("Carlo boundary: MC objects over one shared batch")
        try:
            cb = carlo_boundary(sse, flat, wds, batch)
            for v in cb.values():
                if isinstance(v, dict):
                    v["fraction_of_direct"] = v["vertex_updates_per_s"] / value
            baselines["carlo_boundary"] = cb
        except Exception as e:  # the headline line is reported either way
            baselines["carlo_boundary"] = {"error": f"{type(e).__name__}: {e}"}
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
                   "parallelism": parallelism(choice, args.walkers) if not args.streams else
                   f"{args.streams} stream groups (fixed), one wavefront per walker"},
        "sweeps_per_s": wsweeps / total_s,
        "hbm_gbs_algorithmic": B / total_s / 1e9,
        "hbm_roofline_frac": B / total_s / 1e9 / HBM_PEAK_GBS,
        "device_ms_per_step": dev_ms / args.steps,
        "per_walker_sweep": {"L": Ls / wsweeps, "n": ns / wsweeps, "W": Ws / wsweeps,
                             "U": U / wsweeps, "B": B / wsweeps},
        "tuning": choice,
        "kernels_timed_configuration": per_kernel,
        "kernel_region_ms_per_step": diag_ms / args.steps,
        "lds_bytes_per_walker": {"diag_vlist": lds[0], "worm": lds[1], "worm_walkers_per_cu": lds[2]},
        "roofline": {"bound": "hbm", "kernel": kernels[dom],
                     "achieved": d["achieved_gbs_per_launch"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": d["achieved_gbs_per_launch"] / HBM_PEAK_GBS,
                     "traffic": traffic[0], "traffic_source": traffic[1],
                     "algorithmic_bytes_per_launch": d["algorithmic_bytes_per_launch"],
                     "avg_launch_ms": d["avg_launch_ms"], "launches": d["launches"],
                     "aggregate_achieved_gbs": region_bytes[dom] / (diag_ms * 1e-3) / 1e9,
                     "issue": issue_bound(kernels[dom])},
        "cpu_baseline": cpu,
    }
    out.update(baselines)
    if args.serving_seconds > 0:
        log(f"serving stream: {args.serving_seconds:g} s of arrivals")
        try:
            st = serving_stream(sse, seconds=args.serving_seconds,
                                count_for=3 * args.serving_seconds)
            keys = ("p50_time_to_result_s", "p99_time_to_result_s", "p50_is_lower_bound",
                    "p99_is_lower_bound", "jobs_submitted", "jobs_done", "jobs_unfinished",
                    "completion_fraction", "p50_done_only_s", "p99_done_only_s", "wall_s",
                    "vertex_updates_per_s", "bins_streamed", "workload")
            out["serving"] = {k: st.get(k) for k in keys}
        except Exception as e:  # the headline line is reported either way
            out["serving"] = {"error": f"{type(e).__name__}: {e}"}
    if args.extra:
        extras = {}
        for name in args.extra.split(","):
            if name:
                log(f"extra workload: {name}")
                extras[name] = EXTRA_WORKLOADS[name](sse)
        out["extra_workloads"] = extras
    if args.calibrate_copy_mb:
        copy_calibration(args.calibrate_copy_mb)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

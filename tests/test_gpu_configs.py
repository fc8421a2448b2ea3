"""The single-walker and serving workloads at their stated sizes, bit for bit
against the oracle through the C ABI.

* Plain single walker: S=1/2 Heisenberg J=1 on the 8x8 periodic square
  lattice, beta=2, Philox key 1, initial cutoff 64 grown by the reference rule
  (sse.jl:138-145), one walker per batch, 220 sweeps (controller adapting
  for the first 100), compared after every sweep; worldline consistency
  (test/test_sse.jl:5-28) every 20 sweeps; the fused measurement sums against
  the oracle's per-sweep series.
* Serving at the stated job sizes: Heisenberg/XXZ jobs on square lattices up
  to L=48 at beta up to 32 (operator counts up to ~1e5, worm records in global
  memory) through the size-class serving loop, each job's final state and
  measurement sums against the oracle, and the same job run alone equal to
  its run in the mix, bit for bit (sse.jl:47-68 per job)."""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from fixtures import isconsistent, square_heisenberg

pytestmark = pytest.mark.gpu

KEYS = ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter",
        "last_scatter_calls", "sweeps")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", [0, 1])
def test_single_walker_8x8_beta2(sse_pkg, oracle_mod, mode):
    flat = square_heisenberg(sse_pkg, 8).flat()
    wd = sse_pkg.walker_desc(T=0.5, seed=1, init_opstring_cutoff=64, thermalization=100)
    _say(f"single walker 8x8 beta=2 cutoff 64: mode={mode}")
    b = sse_pkg.Batch([flat], [wd], mode=mode)
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
    assert b.get_walker(0, with_ops=False, with_spins=False)[0]["cutoff"] == 64
    b.init()
    o.init()
    # the five warm-up diagonal updates of init! already grow the cutoff
    cutoffs = {b.get_walker(0, with_ops=False, with_spins=False)[0]["cutoff"]}
    assert cutoffs == {o.scalars()["cutoff"]}
    series = []
    for s in range(220):
        b.run(1, measure=True)
        series.append(o.run_series(1))
        gs, gops, gspins = b.get_walker(0)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops), f"sweep {s}: ops"
        assert np.array_equal(gspins, ospins), f"sweep {s}: spins"
        for key in KEYS:
            assert gs[key] == os_[key], f"sweep {s}: {key} {gs[key]} != {os_[key]}"
        cutoffs.add(gs["cutoff"])
        if s % 20 == 19:
            assert isconsistent(flat, gops, gspins), f"sweep {s}: worldlines"
    # the cutoff left 64 along the reference growth sequence 64 -> 196 -> 394 ...
    grown = {64}
    x = 64
    for _ in range(8):
        x = (3 * x) // 2 + 100
        grown.add(x)
    assert cutoffs <= grown and 64 not in cutoffs
    rows = np.concatenate([r for r in series if len(r)])
    s1, s2, cnt = b.accumulators()[0]
    assert cnt == len(rows) == 220 - 100
    assert np.allclose(s1, rows.sum(axis=0), rtol=1e-12, atol=1e-12)
    assert np.allclose(s2, (rows ** 2).sum(axis=0), rtol=1e-12, atol=1e-12)
    b.close()


def _c5_jobs(sv):
    """Jobs at the serving workload's extremes: L up to 48, beta up to 32.
    The two L=48 jobs run few sweeps: after 15-20 thermalization sweeps the
    worm-number controller (attenuation 0.01, sse.jl:204-217) still makes
    ~50 n scatters per sweep (W ~ 5e6 at n ~ 1e5), so one sweep of the beta=32
    job is seconds of one walker's pointer chase."""
    spec = [((48, "heis"), 32.0, 4, 20), ((48, "xxz"), 16.0, 5, 15),
            ((33, "xxz"), 12.0, 10, 15), ((24, "heis"), 8.0, 40, 20),
            ((12, "xxz"), 4.0, 300, 40), ((8, "heis"), 1.0, 1000, 100)]
    return [sv.Job(j, 0.0, key, 1.0 / beta, sw, th, 7 * 1000003 + j)
            for j, (key, beta, sw, th) in enumerate(spec)]


@pytest.mark.timeout(600)
def test_serving_stated_sizes(sse_pkg, oracle_mod):
    from importlib import import_module
    sv = import_module("sse_amd.serving")
    jobs = _c5_jobs(sv)

    def oracle_job(job):
        flat = sv.magnet_square(*job.model_key).flat()
        wd = sse_pkg.walker_desc(T=job.T, seed=job.seed, thermalization=job.thermalization)
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
        o.init()
        rows = o.run_series(job.thermalization + job.sweeps)
        return o.state(), rows

    # the oracle chains run on host threads while the GPU serves
    pool = ThreadPoolExecutor(max_workers=len(jobs))
    futs = {job.job_id: pool.submit(oracle_job, job) for job in jobs}
    _say("serving at stated sizes: mix of 6 jobs in 3 size classes")
    loop = sv.SizeClassServing(classes=((8000, 3, 4, 48), (60000, 2, 2, 48),
                                        (float("inf"), 1, 1, 48)), keep_state=True)
    routes = {loop.route(j) for j in jobs}
    assert routes == {0, 1, 2}
    done, stats = loop.run(jobs, time_limit=400, log=_say)
    loop.close()
    assert stats["jobs_done"] == len(jobs)
    by_id = {j.job_id: j for j in done}
    assert by_id[0].state[0]["num_operators"] > 80000, "the beta=32 L=48 job reaches n ~ 1e5"
    for job in jobs:
        (os_, oops, ospins), rows = futs[job.job_id].result()
        gs, gops, gspins = by_id[job.job_id].state
        assert np.array_equal(gops, oops), f"job {job.job_id}: ops"
        assert np.array_equal(gspins, ospins), f"job {job.job_id}: spins"
        for key in KEYS:
            assert gs[key] == os_[key], f"job {job.job_id}: {key}"
        s, s2, cnt = by_id[job.job_id].result
        assert cnt == len(rows) == job.sweeps
        assert np.allclose(s, rows.sum(axis=0), rtol=1e-11, atol=1e-9), f"job {job.job_id}"
        _say(f"job {job.job_id} {job.model_key} beta={1 / job.T:g}: n={gs['num_operators']} "
             f"cutoff={gs['cutoff']} equal to the oracle")
    pool.shutdown()
    # alone == mix for an L=48 job and a small one
    for jid in (1, 4):
        job = jobs[jid]
        alone = sv.ServingLoop(n_slots=1, max_L=48, sweeps_per_step=2, keep_state=True)
        d, _ = alone.run([sv.Job(job.job_id, 0.0, job.model_key, job.T, job.sweeps,
                                 job.thermalization, job.seed)])
        alone.close()
        s, s2, cnt = d[0].result
        ms, ms2, mcnt = by_id[jid].result
        assert mcnt == cnt and np.array_equal(ms, s) and np.array_equal(ms2, s2), f"job {jid}"
        assert np.array_equal(d[0].state[1], by_id[jid].state[1]), f"job {jid}: ops"
        _say(f"job {jid}: alone == mix")

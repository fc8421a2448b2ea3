"""Serving: continuous batching of jobs in resident slots.

A job's chain depends only on its seed and parameters, so its result must be
bit-identical whether it runs alone or mixed with other jobs that arrive,
start and finish around it (slot reuse, pool compaction, different LDS
sizing); and it must equal the oracle's sums of per-sweep measurements."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(sv, n, sweeps=30, therm=5):
    jobs = []
    rng = np.random.default_rng(3)
    for j in range(n):
        L = int(rng.integers(4, 9))
        kind = "heis" if j % 2 == 0 else "xxz"
        beta = float(np.exp(rng.uniform(np.log(1.0), np.log(8.0))))
        jobs.append(sv.Job(j, 0.0, (L, kind), 1.0 / beta, sweeps + 3 * j, therm, 500 + j))
    return jobs


def test_alone_equals_mix_and_oracle(sse_pkg, oracle_mod):
    from importlib import import_module
    sv = import_module("sse_amd.serving")
    print("serving: alone vs mix", flush=True)
    jobs = _jobs(sv, 7)
    alone = {}
    for job in jobs[:3]:
        loop = sv.ServingLoop(n_slots=1, max_L=8, sweeps_per_step=4)
        done, _ = loop.run([sv.Job(job.job_id, 0.0, job.model_key, job.T, job.sweeps,
                                   job.thermalization, job.seed)])
        loop.close()
        assert len(done) == 1
        alone[job.job_id] = done[0].result
    # mix: 3 slots, 7 jobs arriving at once -> queueing and slot reuse
    loop = sv.ServingLoop(n_slots=3, max_L=8, sweeps_per_step=3)
    done, stats = loop.run(jobs)
    loop.close()
    assert stats["jobs_done"] == 7
    by_id = {j.job_id: j for j in done}
    for jid, (s, s2, cnt) in alone.items():
        ms, ms2, mcnt = by_id[jid].result
        assert mcnt == cnt == jobs[jid].sweeps
        assert np.array_equal(ms, s) and np.array_equal(ms2, s2), f"job {jid} differs in the mix"
    # every job against the oracle (sum of per-sweep measurements)
    for job in done:
        flat = sv.magnet_square(*job.model_key).flat()
        wd = sse_pkg.walker_desc(T=job.T, seed=job.seed, thermalization=job.thermalization)
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
        o.init()
        rows = o.run_series(job.thermalization + job.sweeps)
        s, s2, cnt = job.result
        assert cnt == len(rows)
        assert np.allclose(s, rows.sum(axis=0), rtol=1e-12, atol=1e-10)


def test_serving_stream_small(sse_pkg):
    """A short Poisson stream (scaled sweeps) completes with latency stats."""
    from importlib import import_module
    sv = import_module("sse_amd.serving")
    print("serving: poisson stream", flush=True)
    jobs = sv.poisson_jobs(40, rate=400.0, L_range=(4, 12), beta_range=(1.0, 8.0),
                           sweeps_range=(20, 200), seed=7)
    loop = sv.ServingLoop(n_slots=16, max_L=12, sweeps_per_step=5)
    done, stats = loop.run(jobs, time_limit=120)
    loop.close()
    assert stats["jobs_done"] == 40
    assert stats["p99_time_to_result_s"] >= stats["p50_time_to_result_s"] > 0
    for j in done:
        assert j.result[2] == j.sweeps


def test_size_classes_equal_alone(sse_pkg):
    """Size-class serving (one batch, host thread and stream per class):
    every job's result is bit-identical to running it alone."""
    from importlib import import_module
    sv = import_module("sse_amd.serving")
    print("serving: size classes", flush=True)
    jobs = _jobs(sv, 8, sweeps=20)
    routes = set()
    loop = sv.SizeClassServing(classes=((200, 3, 3, 8), (float("inf"), 3, 2, 8)))
    for j in jobs:
        routes.add(loop.route(j))
    assert routes == {0, 1}, "the jobs should exercise both classes"
    done, stats = loop.run(jobs, time_limit=120)
    loop.close()
    assert stats["jobs_done"] == 8
    assert sum(c["jobs_done"] for c in stats["classes"]) == 8
    by_id = {j.job_id: j for j in done}
    for job in jobs:
        alone = sv.ServingLoop(n_slots=1, max_L=8, sweeps_per_step=4)
        d, _ = alone.run([sv.Job(job.job_id, 0.0, job.model_key, job.T, job.sweeps,
                                 job.thermalization, job.seed)])
        alone.close()
        s, s2, cnt = d[0].result
        ms, ms2, mcnt = by_id[job.job_id].result
        assert mcnt == cnt == job.sweeps
        assert np.array_equal(ms, s) and np.array_equal(ms2, s2), f"job {job.job_id} differs"


def test_streamed_bins_add_up(sse_pkg):
    """With per-slot budgets every step streams one bin per measuring job:
    bins hold exactly the job's step budget of measurements (a shorter last
    one at most), their weighted sum is the job's final accumulator, and the
    bins merge into a Carlo results entry."""
    from importlib import import_module
    sv = import_module("sse_amd.serving")
    print("serving: streamed bins", flush=True)
    jobs = _jobs(sv, 6, sweeps=40)
    seen = []
    loop = sv.SizeClassServing(classes=((200, 3, 3, 8), (float("inf"), 3, 2, 8)),
                               on_bin=lambda job, k, row: seen.append((job.job_id, k)))
    done, stats = loop.run(jobs, time_limit=120)
    loop.close()
    assert stats["jobs_done"] == 6 and stats["jobs_unfinished"] == 0
    assert stats["bins_streamed"] == len(seen) > 6
    for job in done:
        s, s2, cnt = job.result
        assert sum(job.bin_counts) == cnt == job.sweeps
        assert all(k == job.step_budget for k in job.bin_counts[:-1])
        assert 1 <= job.bin_counts[-1] <= job.step_budget
        total = sum(k * row[:-1] for k, row in zip(job.bin_counts, job.bins))
        assert np.allclose(total, s, rtol=1e-10, atol=1e-10), f"job {job.job_id}"
        flat = sv.magnet_square(*job.model_key).flat()
        entry = sv.job_results(job, flat)
        if entry is not None:
            assert "Energy" in entry["results"] and "SignMag" in entry["results"]

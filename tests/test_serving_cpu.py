"""Host logic of the serving loop: time-to-result percentiles over all
submitted jobs with unfinished jobs counted at their censoring time, and
the per-slot step budgets."""
import numpy as np

from importlib import import_module


def _job(sv, jid, arrival, finished):
    j = sv.Job(jid, arrival, (8, "heis"), 1.0, 10, 1, jid)
    j.finished = finished
    return j


def test_percentiles_all_done(sse_pkg):
    sv = import_module("sse_amd.serving")
    jobs = [_job(sv, k, 0.0, 1.0 + k) for k in range(100)]
    st = sv.latency_stats(jobs, t_end=500.0)
    assert st["jobs_unfinished"] == 0 and st["completion_fraction"] == 1.0
    assert st["p50_time_to_result_s"] == np.percentile(np.arange(100) + 1.0, 50)
    assert not st["p50_is_lower_bound"] and not st["p99_is_lower_bound"]


def test_percentiles_censored(sse_pkg):
    sv = import_module("sse_amd.serving")
    # 95 jobs done within 1..95 s, 5 still running at t_end = 200 after
    # arriving at 0: their latencies are at least 200 s
    jobs = [_job(sv, k, 0.0, 1.0 + k) for k in range(95)]
    jobs += [_job(sv, 95 + k, 0.0, -1.0) for k in range(5)]
    st = sv.latency_stats(jobs, t_end=200.0)
    assert st["jobs_submitted"] == 100 and st["jobs_unfinished"] == 5
    assert abs(st["completion_fraction"] - 0.95) < 1e-12
    # the median lies among finished jobs below every censored time: exact
    assert not st["p50_is_lower_bound"]
    # p99 falls among the censored jobs: only a lower bound, at least 200 s
    assert st["p99_is_lower_bound"] and st["p99_time_to_result_s"] >= 200.0
    # the done-only p99 (the biased figure) is reported separately
    assert st["p99_done_only_s"] < 100.0


def test_percentiles_late_censored_job(sse_pkg):
    sv = import_module("sse_amd.serving")
    # a job that arrived just before the end is censored at 0.5 s, below the
    # median of the finished ones: the median is then only a lower bound
    jobs = [_job(sv, k, 0.0, 10.0) for k in range(10)] + [_job(sv, 10, 99.5, -1.0)]
    st = sv.latency_stats(jobs, t_end=100.0)
    assert st["p50_is_lower_bound"]


def test_step_budget_balances_cost(sse_pkg):
    sv = import_module("sse_amd.serving")
    loop = sv.SizeClassServing.__new__(sv.SizeClassServing)
    loop.max_budget = 1000

    class C:
        sweeps_per_step = 2
        n_ref = sv.n_estimate(32, 1.0 / 16)

    small = sv.Job(0, 0.0, (8, "heis"), 1.0, 10, 1, 0)
    big = sv.Job(1, 0.0, (32, "heis"), 1.0 / 16, 10, 1, 1)
    assert loop.step_budget(C, big) == 2
    bs = loop.step_budget(C, small)
    # expected operators times sweeps per step roughly equal
    ratio = bs * sv.n_estimate(8, 1.0) / (2 * sv.n_estimate(32, 1.0 / 16))
    assert 0.5 < ratio < 1.5 or bs == 1000

"""Parallel tempering in T on the GPU (the reference's exchange hooks,
src/sse.jl:390-405, with the exchange round on the device) against the oracle:
bit-exact walkers, temperatures, slot orders and swap counters after runs with
exchange rounds between sweeps, accumulator rows that follow the
temperatures, and every temperature slot of a PT run against exact
diagonalisation (test/test_ed_compare.jl:57 gate)."""
import sys

import numpy as np
import pytest

import ed
import stats
from fixtures import TS, job_magnet_square, square_heisenberg

pytestmark = pytest.mark.gpu

KEYS = ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter",
        "last_scatter_calls", "sweeps")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


LADDER_T = [0.3, 0.4, 0.55, 0.75, 1.0]


def _pt_parity(sse, oracle_mod, mode, streams, graph):
    flat = square_heisenberg(sse, 4).flat()
    # two ladders of 5 walkers (walkers 0-4 and 6-10, in T order) and two
    # walkers outside any ladder (5 and 11)
    Ts = LADDER_T + [0.5] + LADDER_T + [0.5]
    walkers = [sse.walker_desc(T=T, seed=40 + i, thermalization=6) for i, T in enumerate(Ts)]
    ladders = [[0, 1, 2, 3, 4], [6, 7, 8, 9, 10]]
    seed, interval = 1234, 2
    _say(f"pt parity: mode={mode} streams={streams} graph={graph} walkers={len(walkers)}")
    batch = sse.Batch([flat], walkers, mode=mode)
    batch.set_option(sse.capi.SSE_OPT_STREAMS, streams)
    batch.set_option(sse.capi.SSE_OPT_GRAPH, graph)
    oracles = [oracle_mod.OracleWalker(sse.capi, flat, w) for w in walkers]
    batch.init()
    for o in oracles:
        o.init()
    batch.pt_set_ladders(ladders, seed, interval)
    sw = [np.array(l, dtype=np.int32) for l in ladders]
    acc = [np.zeros(5, dtype=np.int64) for _ in ladders]
    prop = [np.zeros(5, dtype=np.int64) for _ in ladders]
    width = flat.measure_width
    rows_sum = np.zeros((len(walkers), width))
    rows_sq = np.zeros((len(walkers), width))
    rows_n = np.zeros(len(walkers), dtype=np.int64)
    row_of = np.arange(len(walkers))  # accumulator row of each walker's current T
    state = {"since": 0, "round": 0}

    def oracle_exchange():
        for l in range(len(ladders)):
            acc_before = acc[l].copy()
            oracle_mod.pt_exchange(sse.capi, oracles, sw[l], seed, l, state["round"], acc[l],
                                   prop[l])
            # rows follow the temperatures: the two walkers of an accepted
            # pair swap rows
            for k in np.nonzero(acc[l] != acc_before)[0]:
                a, b = int(sw[l][k]), int(sw[l][k + 1])
                row_of[a], row_of[b] = row_of[b], row_of[a]
        state["round"] += 1

    def oracle_run(n):
        for _ in range(n):
            for i, o in enumerate(oracles):
                r = o.run_series(1)
                if len(r):
                    rows_sum[row_of[i]] += r[0]
                    rows_sq[row_of[i]] += r[0] * r[0]
                    rows_n[row_of[i]] += 1
            state["since"] += 1
            if state["since"] == interval:
                state["since"] = 0
                oracle_exchange()

    for n in (25, 7):
        batch.run(n, measure=True)
        oracle_run(n)
    batch.pt_exchange()
    oracle_exchange()

    gsw, gacc, gprop, grounds = batch.pt_ladders()
    assert grounds == state["round"] == 17
    for l in range(len(ladders)):
        assert np.array_equal(gsw[l], sw[l]), f"ladder {l} slots {gsw[l]} != {sw[l]}"
        assert np.array_equal(gacc[l], acc[l]), f"ladder {l} accepted {gacc[l]} != {acc[l]}"
        assert np.array_equal(gprop[l], prop[l]), f"ladder {l} proposed"
    assert sum(int(a.sum()) for a in acc) > 0, "no swap accepted: the test would pin nothing"
    for k, o in enumerate(oracles):
        assert batch.pt_get_parameter(k) == o.T, f"walker {k}: T"
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops), f"walker {k}: ops"
        assert np.array_equal(gspins, ospins), f"walker {k}: spins"
        for key in KEYS:
            assert gs[key] == os_[key], f"walker {k}: {key}"
        # the reference hook on the device state
        assert batch.pt_log_weight_ratio(k, 0.6) == pytest.approx(o.pt_log_weight_ratio(0.6),
                                                                  rel=1e-14, abs=1e-300)
    accs = batch.accumulators()
    for r in range(len(walkers)):
        s, s2, c = accs[r]
        assert c == rows_n[r], f"row {r}: count {c} != {rows_n[r]}"
        assert np.allclose(s, rows_sum[r], rtol=1e-12, atol=1e-12), f"row {r}: sums"
        assert np.allclose(s2, rows_sq[r], rtol=1e-12, atol=1e-12), f"row {r}: sums of squares"
    # the walkers outside the ladders kept their temperature and rows
    assert batch.pt_get_parameter(5) == 0.5 and batch.pt_get_parameter(11) == 0.5
    batch.close()


@pytest.mark.parametrize("mode,streams,graph", [(0, 2, 1), (0, 1, 1), (0, 1, 0), (1, 1, 1)])
def test_pt_parity(sse_pkg, oracle_mod, mode, streams, graph):
    _pt_parity(sse_pkg, oracle_mod, mode, streams, graph)


def test_pt_hooks_and_errors(sse_pkg):
    flat = square_heisenberg(sse_pkg, 4).flat()
    walkers = [sse_pkg.walker_desc(T=0.5 + 0.1 * i, seed=i) for i in range(4)]
    batch = sse_pkg.Batch([flat], walkers)
    batch.init()
    batch.sweep(5)
    n = batch.get_walker(2, with_ops=False, with_spins=False)[0]["num_operators"]
    assert batch.pt_log_weight_ratio(2, 1.4) == pytest.approx(-n * np.log(1.4 / 0.7), rel=1e-14)
    batch.pt_change_parameter(2, 0.9)
    assert batch.pt_get_parameter(2) == 0.9
    with pytest.raises(sse_pkg.capi.SSEError):
        batch.pt_change_parameter(2, -1.0)
    with pytest.raises(sse_pkg.capi.SSEError):
        batch.pt_exchange()  # no ladders
    with pytest.raises(sse_pkg.capi.SSEError):
        batch.pt_set_ladders([[0, 1], [1, 2]], 1)  # walker twice
    with pytest.raises(sse_pkg.capi.SSEError):
        batch.pt_set_ladders([[0, 7]], 1)  # out of range
    batch.pt_set_ladders([[0, 1], [2, 3]], 1, interval=0)
    batch.pt_exchange()
    sw, acc, prop, rounds = batch.pt_ladders()
    assert rounds == 1 and prop[:, 0].tolist() == [1, 1] and prop[:, 1].tolist() == [0, 0]
    assert sorted(sw.ravel().tolist()) == [0, 1, 2, 3]
    batch.pt_set_ladders([], 1)
    batch.close()


PT_LADDERS = 32
PT_THERM = 1000
PT_SWEEPS = 3000


def test_pt_ed_compare_gpu(sse_pkg):
    """32 independent ladders over the 7 temperatures of the reference's ED
    jobs (testjob_magnet_square), one exchange round after every sweep; each
    temperature row (the accumulators follow the temperatures) against exact
    diagonalisation, jackknife over ladders, |z| <= 4."""
    model = job_magnet_square(sse_pkg, TS[0])
    flat = model.flat()
    exact = ed.magnet_ed(model, TS, sse_pkg.spin_operators)
    K = len(TS)
    walkers, ladders = [], []
    for l in range(PT_LADDERS):
        ladders.append([l * K + k for k in range(K)])
        for k, T in enumerate(TS):
            walkers.append(sse_pkg.walker_desc(T=float(T), seed=7000 + l * K + k,
                                               thermalization=PT_THERM))
    _say(f"pt ed gpu: {len(walkers)} walkers in {PT_LADDERS} ladders")
    batch = sse_pkg.Batch([flat], walkers)
    batch.init()
    batch.pt_set_ladders(ladders, 31337, interval=1)
    batch.run(PT_THERM + PT_SWEEPS, measure=True)
    _, acc, prop, rounds = batch.pt_ladders()
    assert rounds == PT_THERM + PT_SWEEPS
    assert np.all(acc[:, 1:-1].sum(axis=0) > 0)
    accs = batch.accumulators()
    names = flat.observable_names()
    within2 = total = 0
    for k, T in enumerate(TS):
        means = []
        for l in range(PT_LADDERS):
            s, _, c = accs[l * K + k]
            assert c == PT_SWEEPS
            means.append(s / c)
        res = stats.jackknife(names, np.array(means), flat.normalization_site_count)
        for name, ev in exact.items():
            mean, err = res[name]
            err = err if err != 0 else 1e-8
            z = abs(mean - ev[k]) / err
            total += 1
            within2 += z <= 2
            assert z <= 4, f"T={T:.2f} {name}: PT GPU {mean}+-{err} ED {ev[k]} z={z:.2f}"
    _say(f"pt ed gpu: {within2}/{total} observables within 2 sigma")
    assert within2 >= 0.8 * total
    batch.close()

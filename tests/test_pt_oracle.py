"""Parallel tempering in T on the oracle side: the two reference hooks
(src/sse.jl:390-405), the series logarithm both sides use in the exchange
test, the exchange rule, and the physics of a PT run against exact
diagonalisation (test/test_jobs.jl:4-29 job, test/test_ed_compare.jl gate)."""
import math

import numpy as np
import pytest

import ed
import stats
from fixtures import TS, job_magnet_square


def test_series_log_matches_libm(sse_pkg, oracle_mod):
    rng = np.random.default_rng(5)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 4000)), rng.uniform(0, 1, 2000) + 1e-300,
                         1 + rng.uniform(-1e-3, 1e-3, 500), [0.5, 1.0, 2.0, math.e, 10.0,
                                                             np.nextafter(1.0, 2.0),
                                                             np.nextafter(1.0, 0.0),
                                                             2.0 ** -53, 2.0 ** 1000]])
    got = np.array([oracle_mod.log(sse_pkg.capi, float(x)) for x in xs])
    want = np.log(xs)
    err = np.abs(got - want)
    # within 4 ulp of the correctly rounded value (absolute floor at log ~ 0)
    bound = 4 * np.spacing(np.maximum(np.abs(want), 1e-300)) + 1e-300
    assert np.all(err <= np.maximum(bound, 4 * 2.0 ** -53 * np.abs(xs - 1))), \
        xs[np.argmax(err / np.maximum(bound, 1e-300))]
    assert oracle_mod.log(sse_pkg.capi, 1.0) == 0.0


def _walker_with_n(sse_pkg, oracle_mod, flat, T, n, seed=1):
    """An oracle walker holding n diagonal operators on bond 0 (the hooks read
    only n and T)."""
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, sse_pkg.walker_desc(T=T, seed=seed))
    o.init()
    L = max(2 * n, 8)
    ops = np.zeros(L, dtype=np.uint32)
    ops[:n] = 1 | (1 << 1)  # bond 0, diagonal vertex code 1
    sc = o.scalars()
    sc["num_operators"] = n
    assert o.set_state(sc, ops, o.spins()) == 0
    return o


def _flat(sse_pkg):
    return job_magnet_square(sse_pkg, 1.0).flat()


def test_log_weight_ratio_closed_form(sse_pkg, oracle_mod):
    """-n log(T'/T) (sse.jl:395): sign and orientation of the ratio."""
    flat = _flat(sse_pkg)
    for n, T, Tn in [(10, 1.0, 2.0), (3, 0.5, 0.25), (0, 1.0, 3.0), (1234, 0.1, 0.11)]:
        o = _walker_with_n(sse_pkg, oracle_mod, flat, T, n)
        want = -n * math.log(Tn / T)
        assert o.pt_log_weight_ratio(Tn) == pytest.approx(want, rel=1e-14, abs=1e-300)
    # raising T lowers the weight of a configuration with operators
    o = _walker_with_n(sse_pkg, oracle_mod, flat, 1.0, 5)
    assert o.pt_log_weight_ratio(2.0) < 0 < o.pt_log_weight_ratio(0.5)


def test_change_parameter_is_a_walker_at_the_new_T(sse_pkg, oracle_mod):
    """change_parameter! (sse.jl:398-405) then sweeping equals a walker
    created at the new T with the same state and RNG position."""
    flat = _flat(sse_pkg)
    a = oracle_mod.OracleWalker(sse_pkg.capi, flat, sse_pkg.walker_desc(T=0.7, seed=11))
    a.init()
    a.sweep(30)
    sc, ops, spins = a.state()
    a.pt_change_parameter(0.3)
    assert a.T == 0.3
    b = oracle_mod.OracleWalker(sse_pkg.capi, flat, sse_pkg.walker_desc(T=0.3, seed=11))
    b.init()
    assert b.set_state(sc, ops, spins) == 0
    a.sweep(20)
    b.sweep(20)
    sa, oa, pa = a.state()
    sb, ob, pb = b.state()
    assert np.array_equal(oa, ob) and np.array_equal(pa, pb)
    assert sa == sb


def test_exchange_acceptance_probability(sse_pkg, oracle_mod):
    """A single proposal between (n=3, T=1) and (n=1, T=2) has
    lwr sum = -3 ln 2 + ln 2 = -2 ln 2: accepted with probability 1/4.  A
    flipped ratio or a flipped Metropolis test would accept always."""
    flat = _flat(sse_pkg)
    trials = 4000
    acc_total = 0
    for t in range(trials):
        a = _walker_with_n(sse_pkg, oracle_mod, flat, 1.0, 3)
        b = _walker_with_n(sse_pkg, oracle_mod, flat, 2.0, 1)
        sw = np.array([0, 1], dtype=np.int32)
        acc = np.zeros(2, dtype=np.int64)
        prop = np.zeros(2, dtype=np.int64)
        oracle_mod.pt_exchange(sse_pkg.capi, [a, b], sw, 1000 + t, 0, 0, acc, prop)
        assert prop[0] == 1 and prop[1] == 0
        acc_total += int(acc[0])
        if acc[0]:
            assert (a.T, b.T) == (2.0, 1.0) and list(sw) == [1, 0]
        else:
            assert (a.T, b.T) == (1.0, 2.0) and list(sw) == [0, 1]
    rate = acc_total / trials
    sd = math.sqrt(0.25 * 0.75 / trials)
    assert abs(rate - 0.25) < 5 * sd, rate


def test_exchange_rounds_alternate_parity(sse_pkg, oracle_mod):
    """Round r proposes the pairs (k, k+1) with k % 2 == r % 2; the
    temperatures stay a permutation of the ladder's."""
    flat = _flat(sse_pkg)
    Ts = [0.3, 0.4, 0.5, 0.6, 0.7]
    ws = [_walker_with_n(sse_pkg, oracle_mod, flat, T, n) for T, n in zip(Ts, [9, 7, 6, 4, 2])]
    sw = np.arange(5, dtype=np.int32)
    acc = np.zeros(5, dtype=np.int64)
    prop = np.zeros(5, dtype=np.int64)
    for r in range(40):
        oracle_mod.pt_exchange(sse_pkg.capi, ws, sw, 77, 3, r, acc, prop)
        assert sorted(w.T for w in ws) == Ts
        assert [ws[sw[k]].T for k in range(5)] == Ts
    assert list(prop) == [20, 20, 20, 20, 0]
    # equal-n neighbours always swap
    ws2 = [_walker_with_n(sse_pkg, oracle_mod, flat, T, 4) for T in Ts[:2]]
    sw2 = np.arange(2, dtype=np.int32)
    a2 = np.zeros(2, dtype=np.int64)
    p2 = np.zeros(2, dtype=np.int64)
    oracle_mod.pt_exchange(sse_pkg.capi, ws2, sw2, 5, 0, 0, a2, p2)
    assert a2[0] == 1


PT_SWEEPS = 30000
PT_THERM = 3000
PT_BIN = 1000


def test_pt_ladder_matches_ed(sse_pkg, oracle_mod):
    """One ladder over the 7 temperatures of the reference's ED jobs with an
    exchange round after every sweep: every temperature slot's observables
    match exact diagonalisation (|z| <= 4 as test/test_ed_compare.jl:57).  A
    wrong acceptance orientation lets high-n configurations drift to high T
    and biases the slots."""
    model = job_magnet_square(sse_pkg, TS[0])
    flat = model.flat()
    exact = ed.magnet_ed(model, TS, sse_pkg.spin_operators)
    ws = []
    for i, T in enumerate(TS):
        w = oracle_mod.OracleWalker(sse_pkg.capi, flat,
                                    sse_pkg.walker_desc(T=float(T), seed=500 + i,
                                                        thermalization=PT_THERM))
        w.init()
        ws.append(w)
    K = len(TS)
    sw = np.arange(K, dtype=np.int32)
    acc = np.zeros(K, dtype=np.int64)
    prop = np.zeros(K, dtype=np.int64)
    width = flat.measure_width
    rows = [[] for _ in range(K)]
    for s in range(PT_THERM + PT_SWEEPS):
        slot_of = np.empty(K, dtype=np.int64)
        slot_of[sw] = np.arange(K)
        for wi, w in enumerate(ws):
            r = w.run_series(1)
            if len(r):
                rows[slot_of[wi]].append(r[0])
        oracle_mod.pt_exchange(sse_pkg.capi, ws, sw, 99, 0, s, acc, prop)
    assert all(len(r) == PT_SWEEPS for r in rows)
    # the widely spaced lowest pair (T 0.04 / 0.70) never swaps; the others do
    assert np.all(acc[1:-1] > 0)
    names = flat.observable_names()
    for k in range(K):
        bins = stats.bin_rows(np.array(rows[k]).reshape(-1, width), PT_BIN)
        res = stats.jackknife(names, bins, flat.normalization_site_count)
        for name, ev in exact.items():
            mean, err = res[name]
            err = err if err != 0 else 1e-8
            z = abs(mean - ev[k]) / err
            assert z <= 4, f"slot {k} T={TS[k]:.2f} {name}: PT {mean}+-{err} ED {ev[k]} z={z:.2f}"

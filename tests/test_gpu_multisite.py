"""Bonds of one, three and four sites on the GPU (SSEData{NSites},
sse_data.jl:15-42), in both kernel families: the lane kernels and the
generic wave kernels (one wavefront per walker; a batch with any model of
other than two sites runs all its walkers through them, two-site ones
included).  Bit-exact with the oracle after init and after every sweep
(first the diagonal update and vertex list alone), measurements included,
in single-model and mixed batches; energies agree with exact
diagonalisation."""
import sys

import numpy as np
import pytest

import stats
from fixtures import (bond_model_ed, field_spins, isconsistent, plaquette_ring,
                      square_heisenberg, trimer_ring)
from test_gpu_parity import KEYS, _compare

pytestmark = pytest.mark.gpu

LANE = 1
MODES = [0, 1]  # SSE_MODE_WAVE, SSE_MODE_LANE


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _flat(sse, kind):
    if kind == "trimer":
        return trimer_ring(sse, 7)[0]
    if kind == "plaquette":
        return plaquette_ring(sse, 9)[0]
    return field_spins(sse, 5)


@pytest.mark.parametrize("kind", ["trimer", "plaquette", "field"])
def test_wave_diagonal_and_vertex_list(sse_pkg, oracle_mod, kind):
    """The generic diagonal-update + vertex-list kernel alone (phases 1|2):
    operator strings, spins, RNG counters and the vertex list equal the
    oracle's after each of 12 such phases (the worm update never runs)."""
    flat = _flat(sse_pkg, kind)
    walkers = [sse_pkg.walker_desc(T=T, seed=70 + k) for k, T in enumerate((0.3, 0.7, 1.5))]
    _say(f"multisite wave diag+vertex list: {kind} bond_n_sites={flat.bond_n_sites}")
    batch = sse_pkg.Batch([flat], walkers, mode=0)
    oracles = [oracle_mod.OracleWalker(sse_pkg.capi, flat, w) for w in walkers]
    batch.init()
    for o in oracles:
        o.init()
    _compare(batch, oracles, f"{kind} init")
    for s in range(12):
        batch.sweep_phases(1 | 2)
        for o in oracles:
            assert o.sweep_phases(1 | 2) == 0
        for k, o in enumerate(oracles):
            gs, gops, gspins = batch.get_walker(k)
            os_, oops, ospins = o.state()
            assert np.array_equal(gops, oops), f"{kind} phase {s} walker {k}: ops"
            assert np.array_equal(gspins, ospins), f"{kind} phase {s} walker {k}: spins"
            assert gs["rng_counter"] == os_["rng_counter"]
            links, first, last = batch.vertex_list(k)
            olinks, ofirst, olast = o.vertex_list()
            assert np.array_equal(links, olinks), f"{kind} phase {s} walker {k}: links"
            assert np.array_equal(first, ofirst) and np.array_equal(last, olast)
    batch.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kind", ["trimer", "plaquette", "field"])
def test_parity_every_sweep(sse_pkg, oracle_mod, kind, mode):
    flat = _flat(sse_pkg, kind)
    walkers = [sse_pkg.walker_desc(T=T, seed=40 + k, thermalization=5)
               for k, T in enumerate((0.25, 0.5, 1.0, 2.0))]
    _say(f"multisite parity: {kind} bond_n_sites={flat.bond_n_sites} mode={mode}")
    batch = sse_pkg.Batch([flat], walkers, mode=mode)
    oracles = [oracle_mod.OracleWalker(sse_pkg.capi, flat, w) for w in walkers]
    batch.init()
    for o in oracles:
        o.init()
    _compare(batch, oracles, f"{kind} init")
    for s in range(25):
        batch.sweep(1)
        for o in oracles:
            o.sweep(1)
        _compare(batch, oracles, f"{kind} sweep {s}")
    gm = batch.measure()
    for k, o in enumerate(oracles):
        assert np.allclose(gm[k], o.measure(), rtol=1e-12, atol=1e-13)
    for k in range(len(walkers)):
        sc, ops, spins = batch.get_walker(k)
        assert isconsistent(flat, ops, spins)
    batch.close()


def test_wave_lookups_from_global_memory(sse_pkg, oracle_mod):
    """A model too large for the generic diagonal-update kernel to stage its
    lookups in LDS (a 2048-site trimer ring: the kernel's own layout alone
    exceeds the 32 KiB limit) reads them from global memory; bit-exact with
    the oracle after every sweep, as the staged path is in the other tests."""
    flat = trimer_ring(sse_pkg, 2048)[0]
    walkers = [sse_pkg.walker_desc(T=1.0, seed=80 + k, thermalization=3) for k in range(2)]
    _say("multisite wave, lookups from global memory: 2048-site trimer ring")
    batch = sse_pkg.Batch([flat], walkers, mode=0)
    oracles = [oracle_mod.OracleWalker(sse_pkg.capi, flat, w) for w in walkers]
    batch.init()
    for o in oracles:
        o.init()
    _compare(batch, oracles, "init")
    for s in range(8):
        batch.sweep(1)
        for o in oracles:
            o.sweep(1)
        _compare(batch, oracles, f"sweep {s}")
    batch.close()


@pytest.mark.parametrize("mode", MODES)
def test_vertex_list_legs(sse_pkg, oracle_mod, mode):
    """sse_get_vertex_list returns 2 NSites legs per slot, equal to the
    oracle's list of the same operator string."""
    flat = _flat(sse_pkg, "trimer")
    w = sse_pkg.walker_desc(T=0.5, seed=3)
    _say(f"multisite vertex list mode={mode}")
    batch = sse_pkg.Batch([flat], [w], mode=mode)
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, w)
    batch.init()
    o.init()
    batch.sweep(6)
    o.sweep(6)
    links, first, last = batch.vertex_list(0)
    olinks, ofirst, olast = o.vertex_list()
    assert len(links) == 6 * o.scalars()["cutoff"]
    assert np.array_equal(links, olinks)
    assert np.array_equal(first, ofirst) and np.array_equal(last, olast)
    batch.close()


@pytest.mark.parametrize("mode", MODES)
def test_mixed_site_counts_run(sse_pkg, oracle_mod, mode):
    """One batch holding one-, two-, three- and four-site models, run by
    sse_run with fused measurement (in wave mode the two-site walker runs
    through the generic kernels too): accumulators equal the oracle's
    series summed, per walker."""
    flats = [_flat(sse_pkg, "field"), square_heisenberg(sse_pkg, 4).flat(),
             _flat(sse_pkg, "trimer"), _flat(sse_pkg, "plaquette")]
    walkers = [sse_pkg.walker_desc(model=k, T=0.4 + 0.2 * k, seed=500 + k, thermalization=8)
               for k in range(4)]
    _say(f"multisite mixed batch mode={mode}")
    batch = sse_pkg.Batch(flats, walkers, mode=mode)
    batch.init()
    batch.run(30, measure=True)
    acc = batch.accumulators()
    for k, w in enumerate(walkers):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flats[k], sse_pkg.walker_desc(
            T=w.T, seed=w.seed, thermalization=8))
        o.init()
        rows = o.run_series(30)
        s, s2, cnt = acc[k]
        assert cnt == len(rows) == 30 - 8
        assert np.allclose(s, rows.sum(axis=0), rtol=1e-12, atol=1e-12)
        assert np.allclose(s2, (rows ** 2).sum(axis=0), rtol=1e-12, atol=1e-12)
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins)
        for key in KEYS:
            assert gs[key] == os_[key]
    batch.close()


@pytest.mark.parametrize("width", [1, 7, 64])
def test_lane_width_equal_results(sse_pkg, width):
    """SSE_OPT_LANE_WIDTH (walkers per wavefront in lane mode) changes the
    launch shape only: 40 walkers of two models, 30 measured sweeps, walkers
    and accumulators equal those of the default shape bit for bit."""
    flats = [_flat(sse_pkg, "trimer"), square_heisenberg(sse_pkg, 4).flat()]
    walkers = [sse_pkg.walker_desc(model=k % 2, T=0.3 + 0.1 * (k % 5), seed=800 + k,
                                   thermalization=6) for k in range(40)]
    _say(f"lane width {width}")
    a = sse_pkg.Batch(flats, walkers, mode=LANE)
    b = sse_pkg.Batch(flats, walkers, mode=LANE)
    a.set_option(sse_pkg.capi.SSE_OPT_LANE_WIDTH, width)
    for x in (a, b):
        x.init()
        x.run(30, measure=True)
        x.run_budgets(np.arange(40) % 4, measure=True)
    for k in range(40):
        sa, oa, pa = a.get_walker(k)
        sb, ob, pb = b.get_walker(k)
        assert np.array_equal(oa, ob) and np.array_equal(pa, pb), k
        for key in KEYS:
            assert sa[key] == sb[key], (k, key)
    for (s1, q1, c1), (s2, q2, c2) in zip(a.accumulators(), b.accumulators()):
        assert c1 == c2 and np.array_equal(s1, s2) and np.array_equal(q1, q2)
    with pytest.raises(sse_pkg.capi.SSEError):
        a.set_option(sse_pkg.capi.SSE_OPT_LANE_WIDTH, 65)
    a.close()
    b.close()


def test_wave_budgets_and_work_queue_limits(sse_pkg, oracle_mod):
    """sse_run_budgets on a generic wave batch (per-sweep launches, walkers
    past their budget idle) against the oracle; the work-queue launch and
    the allocation-free entry point refuse such a batch."""
    flats = [_flat(sse_pkg, "trimer"), square_heisenberg(sse_pkg, 4).flat()]
    walkers = [sse_pkg.walker_desc(model=k % 2, T=0.4 + 0.1 * k, seed=610 + k, thermalization=4)
               for k in range(6)]
    _say("multisite wave budgets")
    batch = sse_pkg.Batch(flats, walkers, mode=0)
    batch.init()
    b1 = np.array([3, 0, 7, 5, 1, 9])
    b2 = np.array([4, 6, 0, 2, 8, 1])
    batch.run_budgets(b1, measure=True)
    batch.run_budgets(b2, measure=True)
    acc = batch.accumulators()
    for k, w in enumerate(walkers):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flats[k % 2], sse_pkg.walker_desc(
            T=w.T, seed=w.seed, thermalization=4))
        o.init()
        rows = o.run_series(int(b1[k] + b2[k]))
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins), k
        for key in KEYS:
            assert gs[key] == os_[key], (k, key)
        s1, _, cnt = acc[k]
        assert cnt == len(rows)
        if cnt:
            assert np.allclose(s1, rows.sum(axis=0), rtol=1e-12, atol=1e-12)
    with pytest.raises(sse_pkg.capi.SSEError):
        batch.set_option(sse_pkg.capi.SSE_OPT_PERSISTENT, 1)
    with pytest.raises(sse_pkg.capi.SSEError):
        batch.run_budgets_noalloc(b1, measure=True)
    batch.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kind,n,T", [("trimer", 6, 0.5), ("plaquette", 8, 0.7)])
def test_energy_against_ed(sse_pkg, kind, n, T, mode):
    """64 independent walkers, 3000 measured sweeps each; jackknife over
    walkers; |E - E_ED| <= 4 sigma."""
    make = trimer_ring if kind == "trimer" else plaquette_ring
    flat, Hb, bs = make(sse_pkg, n)
    exact = bond_model_ed(n, 2, bs, Hb, [T])[0]
    walkers = [sse_pkg.walker_desc(T=T, seed=900 + k, thermalization=500) for k in range(64)]
    _say(f"multisite ED: {kind} n={n} T={T} mode={mode}")
    batch = sse_pkg.Batch([flat], walkers, mode=mode)
    batch.init()
    batch.run(3500, measure=True)
    acc = batch.accumulators()
    means = np.array([a[0] / a[2] for a in acc])
    res = stats.jackknife(flat.observable_names(), means, flat.normalization_site_count)
    mean, err = res["Energy"]
    _say(f"  E = {mean:.5f} +- {err:.5f}  ED {exact:.5f}")
    assert abs(mean - exact) <= 4 * err
    batch.close()

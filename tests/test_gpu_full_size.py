"""The heterogeneous and cluster-basis workloads at their full sizes, checked
through size-independent properties (the 16x16 beta=8 workload has its own
test in test_gpu_parity.py).

C3: 2048 walkers, one XXZ model per walker (square L in {8,12,16,24,32},
Jz ~ U[0.5,1.5], h = 0.1, beta log-uniform on [0.5,16]).  C4: 1024 walkers of
a 64-rung two-leg ladder in the rung-dimer basis (local dimension 4, 36
vertices), J_leg ~ U[0.2,1] per walker, T = 0.1.  Both run with the launch
schedule and stream groups of the library defaults; sampled walkers must be
bit-exact against the oracle (operators, spins, controller, stream position,
accumulated measurements), every walker's operator count must match its
string, a sample must be worldline-consistent, and each walker must have
one measurement per thermalized sweep."""
import sys

import numpy as np
import pytest

from fixtures import isconsistent, ladder_dimer, square_heisenberg

pytestmark = pytest.mark.gpu

KEYS = ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter",
        "last_scatter_calls", "sweeps")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(sse_pkg, oracle_mod, flats, walkers, batch, sweeps, therm, sample):
    acc = batch.accumulators()
    assert all(a[2] == sweeps - therm for a in acc)
    for k in sample:
        flat = flats[walkers[k].model]
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, walkers[k])
        o.init()
        rows = o.run_series(sweeps)
        gs, gops, gspins = batch.get_walker(k)
        os_, oops, ospins = o.state()
        assert np.array_equal(gops, oops) and np.array_equal(gspins, ospins), f"walker {k}"
        for key in KEYS:
            assert gs[key] == os_[key], f"walker {k}: {key}"
        assert np.allclose(acc[k][0], rows.sum(axis=0), rtol=1e-11, atol=1e-9), f"walker {k}"
    for k in range(len(walkers)):
        sc, ops, spins = batch.get_walker(k, with_spins=(k % 32 == 0))
        assert sc["num_operators"] == int((ops != 0).sum()), f"walker {k}"
        if k % 32 == 0:
            assert isconsistent(flats[walkers[k].model], ops, spins), f"walker {k}"


def test_full_size_c3_heterogeneous(sse_pkg, oracle_mod):
    nw, sweeps, therm = 2048, 60, 20
    rng = np.random.default_rng(3)
    sizes = rng.choice([8, 12, 16, 24, 32], size=nw)
    betas = np.exp(rng.uniform(np.log(0.5), np.log(16.0), size=nw))
    jzs = rng.uniform(0.5, 1.5, size=nw)
    _say(f"C3 full size: building {nw} models")
    flats = []
    for k in range(nw):
        flats.append(square_heisenberg(sse_pkg, int(sizes[k]), d=float(jzs[k] - 1.0),
                                       hz=0.1).flat())
        if k % 512 == 511:
            _say(f"C3 full size: {k + 1} models")
    walkers = [sse_pkg.walker_desc(model=k, T=1.0 / float(betas[k]), seed=100000 + k,
                                   thermalization=therm) for k in range(nw)]
    _say(f"C3 full size: {nw} walkers, {sweeps} sweeps")
    batch = sse_pkg.Batch(flats, walkers)
    batch.init()
    for _ in range(sweeps // 20):
        batch.run(20, measure=True)
    big = int(np.argmax(sizes * betas))
    sample = sorted({0, nw - 1, big} | set(rng.integers(0, nw, 3).tolist()))
    _say(f"C3 full size: oracle on walkers {sample}")
    _check(sse_pkg, oracle_mod, flats, walkers, batch, sweeps, therm, sample)
    batch.close()


def test_full_size_c4_ladder(sse_pkg, oracle_mod):
    nw, sweeps, therm = 1024, 60, 20
    rng = np.random.default_rng(4)
    jls = rng.uniform(0.2, 1.0, size=nw)
    _say(f"C4 full size: building {nw} ladder models")
    flats = [ladder_dimer(sse_pkg, 64, float(j)).flat() for j in jls]
    walkers = [sse_pkg.walker_desc(model=k, T=0.1, seed=200000 + k, thermalization=therm)
               for k in range(nw)]
    _say(f"C4 full size: {nw} walkers, {sweeps} sweeps")
    batch = sse_pkg.Batch(flats, walkers)
    batch.init()
    for _ in range(sweeps // 20):
        batch.run(20, measure=True)
    sample = sorted({0, nw - 1} | set(rng.integers(0, nw, 4).tolist()))
    _say(f"C4 full size: oracle on walkers {sample}")
    _check(sse_pkg, oracle_mod, flats, walkers, batch, sweeps, therm, sample)
    batch.close()

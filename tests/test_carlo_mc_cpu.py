"""The host protocol of carlo_mc (Carlo's AbstractMC methods, sse.jl:26-107,
over one shared batch), run on oracle walkers: what every MC hands its
context equals the reference's per-sweep measure! series of its own run,
whatever the call order; a phase-grouped round is one launch; MCs carrying
R runs report the run mean; finished MCs free their slots for new ones;
checkpoints written mid-run resume bit for bit."""
import functools
import os

import numpy as np
import pytest

from fixtures import square_heisenberg
from oracle_batch import OracleBatch


@pytest.fixture
def cm(sse_pkg):
    return sse_pkg.carlo_mc


def _shared(sse_pkg, oracle_mod, cm, proto, n_slots):
    cls = functools.partial(OracleBatch, oracle_mod, sse_pkg.capi)
    return cm.SharedBatch(proto, n_slots, batch_cls=cls)


def _expected(sse_pkg, oracle_mod, flat, T, seed, therm, n_steps, runs=1, k=1):
    """Per-step rows a lone reference run reports: the mean over R runs and
    k sweeps of the per-sweep measurement rows, and the WormLengthFraction
    series."""
    rows, wsum, wcnt = [], np.zeros(n_steps), np.zeros(n_steps)
    for r in range(runs):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, sse_pkg.walker_desc(
            T=T, seed=seed + r, thermalization=k * therm))
        o.init()
        rr = []
        for i in range(n_steps * k):
            o.sweep(1)
            sc = o.scalars()
            if sc["sweeps"] > k * therm:
                rr.append(o.measure())
                f = sc["last_worm_length_fraction"]
                if not np.isnan(f):   # measured only for n != 0
                    wsum[i // k] += f
                    wcnt[i // k] += 1
        rows.append(np.array(rr))
    m = np.mean([r.reshape(-1, k, r.shape[1]).mean(axis=1) for r in rows], axis=0)
    return m, wsum[wcnt > 0] / wcnt[wcnt > 0]


def _tasks(sse_pkg, n):
    models = [square_heisenberg(sse_pkg, 4), square_heisenberg(sse_pkg, 6, hz=0.2)]
    return [{"model": models[t % 2], "T": 0.3 + 0.15 * t, "seed": 100 + 10 * t,
             "thermalization": 3 + t % 3} for t in range(n)]


def _check(ctx, flat, exp_rows, exp_wlf, rtol=0.0):
    for j, name in enumerate(flat.observable_names()):
        got = ctx.series(name)
        assert len(got) == len(exp_rows), name
        if rtol:
            assert np.allclose(got, exp_rows[:, j], rtol=rtol, atol=rtol), name
        else:
            assert np.array_equal(got, exp_rows[:, j]), name
    got = ctx.series("WormLengthFraction")
    assert np.allclose(got, exp_wlf, rtol=1e-12, atol=0)


@pytest.mark.parametrize("grouped", [True, False])
def test_series_equal_lone_runs(sse_pkg, oracle_mod, cm, grouped):
    tasks = _tasks(sse_pkg, 5)
    sh = _shared(sse_pkg, oracle_mod, cm, tasks[1]["model"].flat(), 8)
    mcs = [cm.MC(t, sh) for t in tasks]
    ctxs = [cm.MCContext(t["thermalization"]) for t in tasks]
    for mc, ctx in zip(mcs, ctxs):
        mc.init(ctx)
    steps = 12
    for _ in range(steps):
        if grouped:
            cm.step_all(mcs, ctxs)
        else:
            for mc, ctx in zip(mcs, ctxs):
                cm.step_one(mc, ctx)
    assert sh.work[3] == steps * len(mcs)
    if grouped:
        assert sh.launches == steps
    else:
        # a measure! runs the owed sweeps at once; unmeasured (thermalization)
        # sweeps wait and share the launch of the next run that measures
        assert steps < sh.launches <= steps * len(mcs)
    for t, mc, ctx in zip(tasks, mcs, ctxs):
        rows, wlf = _expected(sse_pkg, oracle_mod, mc.flat, t["T"], t["seed"],
                              t["thermalization"], steps)
        _check(ctx, mc.flat, rows, wlf)
    # two model objects with equal content share one device model
    assert len(sh.models) == 2


def test_runs_per_mc_and_sweeps_per_step(sse_pkg, oracle_mod, cm):
    tasks = _tasks(sse_pkg, 2)
    sh = _shared(sse_pkg, oracle_mod, cm, tasks[1]["model"].flat(), 8)
    a = cm.MC(tasks[0], sh, runs=3)
    b = cm.MC(tasks[1], sh, runs=2, sweeps_per_step=4)
    ca, cb = cm.MCContext(tasks[0]["thermalization"]), cm.MCContext(tasks[1]["thermalization"])
    for _ in range(9):
        cm.step_all([a, b], [ca, cb])
    assert sh.launches == 9
    ra, wa = _expected(sse_pkg, oracle_mod, a.flat, tasks[0]["T"], tasks[0]["seed"],
                       tasks[0]["thermalization"], 9, runs=3)
    rb, wb = _expected(sse_pkg, oracle_mod, b.flat, tasks[1]["T"], tasks[1]["seed"],
                       tasks[1]["thermalization"], 9, runs=2, k=4)
    _check(ca, a.flat, ra, wa, rtol=1e-13)
    _check(cb, b.flat, rb, wb, rtol=1e-13)


def test_finished_runs_free_slots(sse_pkg, oracle_mod, cm):
    """Six tasks of different lengths through four slots: a task that is
    done releases its slot and the next task starts there mid-run."""
    tasks = _tasks(sse_pkg, 6)
    lengths = [5, 9, 7, 12, 4, 6]
    sh = _shared(sse_pkg, oracle_mod, cm, tasks[1]["model"].flat(), 4)
    queue = list(range(6))
    live = {}
    ctxs = {}
    while queue or live:
        while queue and sh.free:
            t = queue.pop(0)
            live[t] = cm.MC(tasks[t], sh)
            ctxs[t] = cm.MCContext(tasks[t]["thermalization"])
        ids = sorted(live)
        cm.step_all([live[t] for t in ids], [ctxs[t] for t in ids])
        for t in ids:
            if ctxs[t].sweeps == tasks[t]["thermalization"] + lengths[t]:
                live.pop(t).close()
    for t in range(6):
        rows, wlf = _expected(sse_pkg, oracle_mod, tasks[t]["model"].flat(), tasks[t]["T"],
                              tasks[t]["seed"], tasks[t]["thermalization"],
                              tasks[t]["thermalization"] + lengths[t])
        _check(ctxs[t], tasks[t]["model"].flat(), rows, wlf)
    assert sorted(sh.free) == [0, 1, 2, 3]


def test_measure_before_thermalization_raises(sse_pkg, oracle_mod, cm):
    t = _tasks(sse_pkg, 1)[0]
    sh = _shared(sse_pkg, oracle_mod, cm, t["model"].flat(), 2)
    mc = cm.MC(t, sh)
    ctx = cm.MCContext(t["thermalization"])
    mc.sweep(ctx)
    with pytest.raises(RuntimeError, match="not thermalized"):
        mc.measure(ctx)


def test_checkpoint_resume_in_another_batch(sse_pkg, oracle_mod, cm, tmp_path):
    tasks = _tasks(sse_pkg, 3)
    sh = _shared(sse_pkg, oracle_mod, cm, tasks[1]["model"].flat(), 4)
    mcs = [cm.MC(t, sh, runs=2 if k == 2 else 1) for k, t in enumerate(tasks)]
    ctxs = [cm.MCContext(t["thermalization"]) for t in tasks]
    for _ in range(8):
        cm.step_all(mcs, ctxs)
    path = os.path.join(str(tmp_path), "ck.h5")
    with sse_pkg.h5lite.File(path, "w") as f:
        for k, mc in enumerate(mcs):
            with f.create_group(f"task{k}") as g:
                mc.write_checkpoint(g)
    sh2 = _shared(sse_pkg, oracle_mod, cm, tasks[1]["model"].flat(), 6)
    sh2.attach(tasks[0]["model"].flat(), sse_pkg.walker_desc(T=1.0, seed=5))  # shift slots
    mcs2 = [cm.MC(t, sh2, runs=2 if k == 2 else 1) for k, t in enumerate(tasks)]
    ctxs2 = [cm.MCContext(t["thermalization"]) for t in tasks]
    with sse_pkg.h5lite.File(path, "r") as f:
        for k, mc in enumerate(mcs2):
            with f.group(f"task{k}") as g:
                mc.read_checkpoint(g)
            ctxs2[k].sweeps = ctxs[k].sweeps
    for _ in range(6):
        cm.step_all(mcs, ctxs)
        cm.step_all(mcs2, ctxs2)
    for k in range(3):
        for name in mcs[k].names:
            assert np.array_equal(ctxs[k].series(name)[-6:], ctxs2[k].series(name)), name


def test_parallel_tempering_hooks(sse_pkg, oracle_mod, cm):
    """The hooks flush owed sweeps first, then act on the run's walker
    (-n log(T'/T), sse.jl:390-405)."""
    t = _tasks(sse_pkg, 1)[0]
    sh = _shared(sse_pkg, oracle_mod, cm, t["model"].flat(), 2)
    mc = cm.MC(t, sh)
    ctx = cm.MCContext(t["thermalization"])
    for _ in range(5):
        cm.step_one(mc, ctx)
    mc.sweep(ctx)
    ctx.sweeps += 1
    r = mc.parallel_tempering_log_weight_ratio("T", 0.5)
    n = sh.batch.get_walker(mc.slots[0])[0]["num_operators"]
    assert r == pytest.approx(-n * np.log(0.5 / t["T"]), rel=1e-14)
    mc.parallel_tempering_change_parameter("T", 0.5)
    assert sh.batch.pt_get_parameter(mc.slots[0]) == 0.5
    with pytest.raises(ValueError):
        mc.parallel_tempering_change_parameter("J", 1.0)

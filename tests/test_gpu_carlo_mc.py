"""Carlo's AbstractMC methods over one shared device batch (carlo_mc), on
the GPU through the C ABI: every MC's reported series equals the oracle
run's per-sweep measure! series (1e-12), one launch per phase-grouped
round, R runs per MC, slot reuse, checkpoints across batches, multi-site
models in lane mode."""
import os
import sys

import numpy as np
import pytest

from fixtures import trimer_ring
from test_carlo_mc_cpu import _check, _expected, _tasks

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("grouped", [True, False])
def test_series_equal_oracle_runs(sse_pkg, oracle_mod, grouped):
    cm = sse_pkg.carlo_mc
    tasks = _tasks(sse_pkg, 6)
    _say(f"carlo shared batch: 6 tasks, grouped={grouped}")
    sh = cm.SharedBatch(tasks[1]["model"].flat(), 8)
    mcs = [cm.MC(t, sh) for t in tasks]
    ctxs = [cm.MCContext(t["thermalization"]) for t in tasks]
    steps = 15
    for _ in range(steps):
        if grouped:
            cm.step_all(mcs, ctxs)
        else:
            for mc, ctx in zip(mcs, ctxs):
                cm.step_one(mc, ctx)
    assert sh.work[3] == steps * len(mcs)
    if grouped:
        assert sh.launches == steps
    for t, mc, ctx in zip(tasks, mcs, ctxs):
        rows, wlf = _expected(sse_pkg, oracle_mod, mc.flat, t["T"], t["seed"],
                              t["thermalization"], steps)
        _check(ctx, mc.flat, rows, wlf, rtol=1e-12)
    sh.close()


def test_runs_and_sweeps_per_step(sse_pkg, oracle_mod):
    cm = sse_pkg.carlo_mc
    tasks = _tasks(sse_pkg, 2)
    _say("carlo shared batch: runs=3, sweeps_per_step=4")
    sh = cm.SharedBatch(tasks[1]["model"].flat(), 8)
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
    _check(ca, a.flat, ra, wa, rtol=1e-12)
    _check(cb, b.flat, rb, wb, rtol=1e-12)
    sh.close()


def test_slot_reuse_and_checkpoint(sse_pkg, oracle_mod, tmp_path):
    """Tasks of different lengths through three slots (a finished MC frees
    its slot for the next task), then one task checkpointed mid-run into
    HDF5 and resumed in another shared batch at another slot."""
    cm = sse_pkg.carlo_mc
    tasks = _tasks(sse_pkg, 5)
    lengths = [6, 3, 8, 5, 4]
    _say("carlo shared batch: slot reuse")
    sh = cm.SharedBatch(tasks[1]["model"].flat(), 3)
    queue, live, ctxs = list(range(5)), {}, {}
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
    for t in range(5):
        flat = tasks[t]["model"].flat()
        rows, wlf = _expected(sse_pkg, oracle_mod, flat, tasks[t]["T"], tasks[t]["seed"],
                              tasks[t]["thermalization"],
                              tasks[t]["thermalization"] + lengths[t])
        _check(ctxs[t], flat, rows, wlf, rtol=1e-12)
    _say("carlo shared batch: checkpoint")
    mc = cm.MC(tasks[2], sh, runs=2)
    ctx = cm.MCContext(tasks[2]["thermalization"])
    for _ in range(7):
        cm.step_one(mc, ctx)
    path = os.path.join(str(tmp_path), "ck.h5")
    with sse_pkg.h5lite.File(path, "w") as f:
        with f.create_group("task") as g:
            mc.write_checkpoint(g)
    sh2 = cm.SharedBatch(tasks[1]["model"].flat(), 4)
    sh2.attach(tasks[0]["model"].flat(), sse_pkg.walker_desc(T=1.0, seed=5))
    mc2 = cm.MC(tasks[2], sh2, runs=2)
    ctx2 = cm.MCContext(tasks[2]["thermalization"])
    with sse_pkg.h5lite.File(path, "r") as f:
        with f.group("task") as g:
            mc2.read_checkpoint(g)
    ctx2.sweeps = ctx.sweeps
    for _ in range(6):
        cm.step_all([mc, mc2], [ctx, ctx2])
    for name in mc.names:
        assert np.array_equal(ctx.series(name)[-6:], ctx2.series(name)), name
    sh.close()
    sh2.close()


def test_lane_mode_three_site_model(sse_pkg, oracle_mod):
    cm = sse_pkg.carlo_mc
    flat = trimer_ring(sse_pkg, 6)[0]
    tasks = [{"model": flat, "T": 0.5 + 0.25 * k, "seed": 60 + k, "thermalization": 2}
             for k in range(3)]
    _say("carlo shared batch: lane mode, three-site bonds")
    sh = cm.SharedBatch(flat, 4, mode=1)
    mcs = [cm.MC(t, sh) for t in tasks]
    ctxs = [cm.MCContext(2) for _ in tasks]
    for _ in range(10):
        cm.step_all(mcs, ctxs)
    for t, ctx in zip(tasks, ctxs):
        rows, wlf = _expected(sse_pkg, oracle_mod, flat, t["T"], t["seed"], 2, 10)
        _check(ctx, flat, rows, wlf, rtol=1e-12)
    sh.close()

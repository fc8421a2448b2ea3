"""Serving slots through the C ABI: restarts, budgets, measurement order and
model uploads, each checked bit for bit against the oracle.

* A slot restarted with more diagonal warm-up sweeps than any walker the
  batch was created with gets all of them (Carlo.init!, sse.jl:47-60).
* sse_measure right after sse_slot_start measures the initialised job, not
  the previous job's operator string (init! precedes measure!).
* sse_sweep(k) with k beyond a slot's budget stops the slot after its last
  budgeted sweep.
* A model upload that fails part-way leaves the batch's model list as it
  was: the next model gets the next index and runs bit-exact."""
import sys
from fractions import Fraction

import numpy as np
import pytest

from fixtures import square_heisenberg

pytestmark = pytest.mark.gpu

MODES = [0, 1]
KEYS = ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter",
        "last_scatter_calls", "sweeps")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same_as_oracle(batch, slot, o, tag):
    gs, gops, gspins = batch.get_walker(slot)
    os_, oops, ospins = o.state()
    assert np.array_equal(gops, oops), f"{tag}: ops"
    assert np.array_equal(gspins, ospins), f"{tag}: spins"
    for key in KEYS:
        assert gs[key] == os_[key], f"{tag}: {key} {gs[key]} != {os_[key]}"


@pytest.mark.parametrize("mode", MODES)
def test_restart_with_more_warmup_sweeps(sse_pkg, oracle_mod, mode):
    flat = square_heisenberg(sse_pkg, 6).flat()
    created = [sse_pkg.walker_desc(T=0.4, seed=10 + s) for s in range(3)]  # 5 warm-up sweeps
    _say(f"slots: restart with 13 warm-up sweeps, mode={mode}")
    b = sse_pkg.Batch([flat], created, mode=mode)
    b.init()
    b.sweep(4)
    wd = sse_pkg.walker_desc(T=0.3, seed=77, diagonal_warmup_sweeps=13, thermalization=3)
    b.slot_start(1, wd, 9)
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
    o.init()
    b.sweep(0)  # initialises the pending slot only
    _same_as_oracle(b, 1, o, "after init")
    b.run(9, measure=True)
    rows = o.run_series(9)
    _same_as_oracle(b, 1, o, "after 9 sweeps")
    s, s2, cnt = b.slot_result(1)
    assert cnt == len(rows) == 9 - 3
    assert np.allclose(s, rows.sum(axis=0), rtol=1e-12, atol=1e-12)
    b.close()


@pytest.mark.parametrize("mode", MODES)
def test_measure_right_after_slot_start(sse_pkg, oracle_mod, mode):
    """The slot's previous job ran an 8x8 lattice (128 bonds); the restarted
    job runs a 4x4 XXZ model (32 bonds).  Measuring before any sweep must
    read the new job's initialised state, never the old operator string."""
    big = square_heisenberg(sse_pkg, 8).flat()
    small = square_heisenberg(sse_pkg, 4, d=0.3, hz=0.1).flat()
    _say(f"slots: measure after slot_start, mode={mode}")
    b = sse_pkg.Batch([big], [sse_pkg.walker_desc(T=0.5, seed=3 + s) for s in range(2)],
                      mode=mode)
    b.init()
    b.sweep(6)
    mi = b.add_model(small)
    wd = sse_pkg.walker_desc(model=mi, T=0.3, seed=901)
    b.slot_start(0, wd, 5)
    got = b.measure()
    o = oracle_mod.OracleWalker(sse_pkg.capi, small, sse_pkg.walker_desc(T=0.3, seed=901))
    o.init()
    assert np.allclose(got[0], o.measure(), rtol=1e-12, atol=1e-13)
    _same_as_oracle(b, 0, o, "initialised by measure")
    b.close()


@pytest.mark.parametrize("mode", MODES)
def test_sweep_stops_at_budget(sse_pkg, oracle_mod, mode):
    flat = square_heisenberg(sse_pkg, 4).flat()
    _say(f"slots: sweep beyond the budget, mode={mode}")
    b = sse_pkg.Batch([flat], [sse_pkg.walker_desc(T=0.5, seed=20 + s) for s in range(3)],
                      mode=mode)
    b.init()
    wd = sse_pkg.walker_desc(T=0.5, seed=321, thermalization=2)
    b.slot_start(2, wd, 5)
    b.sweep(12)
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
    o.init()
    o.sweep(5)
    sc, _, _ = b.get_walker(2, with_ops=False, with_spins=False)
    assert sc["sweeps"] == 5
    _same_as_oracle(b, 2, o, "budget of 5")
    assert b.slot_poll() == [2]
    # the other slots kept sweeping
    assert b.get_walker(0, with_ops=False, with_spins=False)[0]["sweeps"] == 12
    b.close()


def _broken_model(sse_pkg):
    """A model that passes validation but whose upload fails: one transition
    with 4096 exits (the device tables hold at most 4095)."""
    flat = square_heisenberg(sse_pkg, 4).flat()
    t = dict(flat.types[0])
    ne = 4096
    t["n_exits"] = ne
    t["exit_cumprob"] = np.linspace(1.0 / ne, 1.0, ne)
    t["exit_target"] = np.full(ne, int(t["exit_target"][0]), dtype=np.int32)
    t["exit_leg"] = np.zeros(ne, dtype=np.uint8)
    t["exit_worm"] = np.ones(ne, dtype=np.uint8)
    off = np.array(t["transition_offset"]).copy()
    cnt = np.array(t["transition_count"]).copy()
    off[off >= 0] = 0
    cnt[cnt > 0] = 1
    off.reshape(-1)[0] = 0
    cnt.reshape(-1)[0] = ne
    t["transition_offset"], t["transition_count"] = off, cnt
    flat.types = [t, flat.types[1]]
    return flat


def test_failed_model_upload_keeps_model_list(sse_pkg, oracle_mod):
    base = square_heisenberg(sse_pkg, 4).flat()
    # mixed spin-1/2 and spin-1 sites, the same estimator width as the slots
    nxt = sse_pkg.MagnetModel({"lattice": {"unitcell": sse_pkg.UnitCells.honeycomb, "size": (2, 2)},
                               "parameter_map": {"S": ["Sa", "Sb"]}, "Sa": Fraction(1, 2),
                               "Sb": 1, "J": 1.0, "d": 0.1,
                               "measure": ["magnetization"]}).flat()
    _say("slots: failed sse_add_model")
    b = sse_pkg.Batch([base], [sse_pkg.walker_desc(T=0.5, seed=s) for s in range(2)])
    b.init()
    with pytest.raises(sse_pkg.capi.SSEError, match="4095"):
        b.add_model(_broken_model(sse_pkg))
    mi = b.add_model(nxt)
    assert mi == 1
    wd = sse_pkg.walker_desc(model=mi, T=0.4, seed=55)
    b.slot_start(1, wd, 8)
    b.sweep(8)
    o = oracle_mod.OracleWalker(sse_pkg.capi, nxt, sse_pkg.walker_desc(T=0.4, seed=55))
    o.init()
    o.sweep(8)
    _same_as_oracle(b, 1, o, "model after a failed upload")
    # a restored state is checked against the right model's tables
    sc, ops, spins = b.get_walker(1)
    b.set_walker(1, sc, ops, spins)
    bad = spins.copy()
    bad[int(np.argmin(nxt.site_dims))] = 3  # a spin-1/2 site cannot hold state 3
    with pytest.raises(sse_pkg.capi.SSEError, match="local state"):
        b.set_walker(1, sc, ops, bad)
    b.close()

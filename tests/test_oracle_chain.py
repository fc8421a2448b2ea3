"""The CPU restatement oracle against the reference's fixtures and invariants:
golden vertex list (test/test_vertex_list.jl), worldline consistency after
every sweep (test/test_sse.jl:63-94), growth, checkpoint/resume."""
import numpy as np
import pytest

from fixtures import isconsistent, job_honeycomb, square_heisenberg, vertex_list_fixture


def test_vertex_list_golden(sse_pkg, oracle_mod):
    n_sites, bond_sites, ops, links, first, last = vertex_list_fixture()
    got_links, got_first, got_last = oracle_mod.make_vertex_list(sse_pkg.capi, n_sites,
                                                                 bond_sites, ops)
    assert np.array_equal(got_links, links)
    assert np.array_equal(got_first, first)
    assert np.array_equal(got_last, last)


def _walker(sse_pkg, oracle_mod, model, **kw):
    flat = model.flat()
    wd = sse_pkg.walker_desc(**kw)
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat, wd)
    o.init()
    return flat, o


def test_chain_consistency_mixed_spins(sse_pkg, oracle_mod):
    """Mixed spin-1/2 / spin-1 honeycomb 4x4 at T=0.1 (test/test_sse.jl:63-94)."""
    from fractions import Fraction
    model = sse_pkg.MagnetModel({"lattice": {"unitcell": sse_pkg.UnitCells.honeycomb, "size": (4, 4)},
                                 "parameter_map": {"S": ["S1", "S2"]}, "J": 1.4,
                                 "S1": Fraction(1, 2), "S2": 1, "measure": []})
    flat, o = _walker(sse_pkg, oracle_mod, model, T=0.1, seed=11, thermalization=100)
    for s in range(300):
        sc, ops, spins = o.state()
        assert isconsistent(flat, ops, spins), f"sweep {s}"
        o.sweep(1)
    sc, ops, spins = o.state()
    assert isconsistent(flat, ops, spins)
    assert np.all((spins >= 1) & (spins <= flat.site_dims))
    assert sc["num_operators"] == int((ops != 0).sum())


def test_growth_from_empty_cutoff(sse_pkg, oracle_mod):
    model = square_heisenberg(sse_pkg, 4)
    flat, o = _walker(sse_pkg, oracle_mod, model, T=0.25, seed=3, init_opstring_cutoff=0)
    Ls = [o.scalars()["cutoff"]]
    for _ in range(50):
        o.sweep(1)
        Ls.append(o.scalars()["cutoff"])
    # every growth step is L -> floor(1.5 L + 100) (sse.jl:138-145)
    for a, b in zip(Ls, Ls[1:]):
        assert b == a or b == (3 * a) // 2 + 100
    assert Ls[-1] > 100


def test_checkpoint_resume_equals_uninterrupted(sse_pkg, oracle_mod):
    model = square_heisenberg(sse_pkg, 4)
    flat, a = _walker(sse_pkg, oracle_mod, model, T=0.3, seed=21, thermalization=10)
    a.sweep(7)
    sc, ops, spins = a.state()
    flat2, b = _walker(sse_pkg, oracle_mod, model, T=0.3, seed=21, thermalization=10)
    assert b.set_state(sc, ops, spins) == 0
    a.sweep(20)
    b.sweep(20)
    sa, oa, pa = a.state()
    sb, ob, pb = b.state()
    assert np.array_equal(oa, ob) and np.array_equal(pa, pb)
    for k in ("cutoff", "num_operators", "num_worms", "avg_worm_length", "rng_counter"):
        assert sa[k] == sb[k]
    assert np.array_equal(a.measure(), b.measure())


def test_worm_count_controller(sse_pkg, oracle_mod):
    """Worm-length fraction settles near the controller target 2.0 (sse.jl:34,204-217)."""
    model = square_heisenberg(sse_pkg, 6)
    flat, o = _walker(sse_pkg, oracle_mod, model, T=0.2, seed=5, thermalization=2000)
    o.sweep(2000)
    o.sweep(1)  # thermalized now: WormLengthFraction measured
    w = []
    for _ in range(400):
        o.sweep(1)
        w.append(o.scalars()["last_worm_length_fraction"])
    assert 1.5 < np.mean(w) < 2.6

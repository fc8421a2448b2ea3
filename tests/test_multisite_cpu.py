"""Bonds of one, three and four sites (SSEData{NSites}, sse_data.jl:15-42):
the oracle's worldlines stay consistent after every sweep, its vertex list
has 2 NSites legs per vertex (vertex_list.jl:24-62), and its energy and
specific heat agree with exact diagonalisation within 4 sigma, as the
reference's ED comparison does (test/test_ed_compare.jl).  Host-side
flattening rejects mixed site counts."""
import numpy as np
import pytest

import stats
from fixtures import bond_model_ed, field_spins, isconsistent, plaquette_ring, trimer_ring


def _oracle(sse_pkg, oracle_mod, flat, T, seed, therm=500):
    o = oracle_mod.OracleWalker(sse_pkg.capi, flat,
                                sse_pkg.walker_desc(T=T, seed=seed, thermalization=therm))
    o.init()
    return o


@pytest.mark.parametrize("kind", ["trimer", "plaquette", "field"])
def test_worldlines_consistent(sse_pkg, oracle_mod, kind):
    if kind == "trimer":
        flat = trimer_ring(sse_pkg, 7)[0]
    elif kind == "plaquette":
        flat = plaquette_ring(sse_pkg, 8)[0]
    else:
        flat = field_spins(sse_pkg, 5)
    o = _oracle(sse_pkg, oracle_mod, flat, 0.4, 5)
    for s in range(150):
        o.sweep(1)
        sc, ops, spins = o.state()
        assert isconsistent(flat, ops, spins), f"{kind} sweep {s}"
        assert sc["num_operators"] == int((ops != 0).sum())
    links, first, last = o.vertex_list()
    legs = 2 * flat.bond_n_sites
    assert len(links) == legs * sc["cutoff"]
    # every linked leg points back to itself (vertex_list.jl:47-58)
    for k, t in enumerate(links):
        if t >= 0:
            assert links[t] == k


def test_vertex_list_three_site_bond(sse_pkg, oracle_mod):
    """Two three-site operators, on bonds (0,1,2) at slot 0 and (1,2,3) at
    slot 2, of 4 sites; slot 1 is the identity.  Leg 6p+s is the lower leg
    of the bond's site s, 6p+3+s its upper leg.  Expected links worked out
    by hand from the linking rule of vertex_list.jl:24-62."""
    bond_sites = np.array([[0, 1, 2], [1, 2, 3]], dtype=np.int32)
    ops = np.array([1 | (0 << 11), 0, 1 | (1 << 11)], dtype=np.uint32)
    links, first, last = oracle_mod.make_vertex_list(sse_pkg.capi, 4, bond_sites, ops)
    assert links.tolist() == [3, 15, 16, 0, 12, 13, -1, -1, -1, -1, -1, -1,
                              4, 5, 17, 1, 2, 14]
    assert first.tolist() == [0, 1, 2, 14]
    assert last.tolist() == [3, 15, 16, 17]


def test_vertex_list_one_site_bond(sse_pkg, oracle_mod):
    """One-site bonds: two legs per vertex; a site's operators chain in
    imaginary time and its last upper leg wraps to its first lower leg."""
    bond_sites = np.array([[0], [1]], dtype=np.int32)
    ops = np.array([1 | (1 << 11), 1 | (0 << 11), 1 | (1 << 11)], dtype=np.uint32)
    links, first, last = oracle_mod.make_vertex_list(sse_pkg.capi, 3, bond_sites, ops)
    assert links.tolist() == [5, 4, 3, 2, 1, 0]
    assert first.tolist() == [2, 0, -1]
    assert last.tolist() == [3, 5, -1]


def _energy(sse_pkg, oracle_mod, flat, T, seed, sweeps, therm=2000, binsize=500):
    o = _oracle(sse_pkg, oracle_mod, flat, T, seed, therm)
    rows = o.run_series(sweeps + therm + 1)
    bins = stats.bin_rows(rows, binsize)
    return stats.jackknife(flat.observable_names(), bins, flat.normalization_site_count)


@pytest.mark.parametrize("kind,n,T", [("trimer", 6, 0.5), ("trimer", 6, 1.5),
                                      ("plaquette", 8, 0.7)])
def test_energy_against_ed(sse_pkg, oracle_mod, kind, n, T):
    make = trimer_ring if kind == "trimer" else plaquette_ring
    flat, Hb, bs = make(sse_pkg, n)
    exact = bond_model_ed(n, 2, bs, Hb, [T])[0]
    res = _energy(sse_pkg, oracle_mod, flat, T, 77 + n, 30000)
    mean, err = res["Energy"]
    assert abs(mean - exact) <= 4 * err, f"{kind} T={T}: {mean}+-{err} ED {exact}"
    assert err < 0.01


def test_diagonal_field_closed_form(sse_pkg, oracle_mod):
    """One-site bonds, H_i = -hz S^z_i: energy per site
    -(hz/2) tanh(hz / 2T); the spins move only through empty sites
    (sse.jl:239-245 redraws them)."""
    hz, T = 0.9, 0.5
    flat = field_spins(sse_pkg, 6, hx=0.0, hz=hz)
    res = _energy(sse_pkg, oracle_mod, flat, T, 3, 30000)
    mean, err = res["Energy"]
    exact = -(hz / 2) * np.tanh(hz / (2 * T))
    assert abs(mean - exact) <= 4 * err, f"{mean}+-{err} exact {exact}"


def test_one_site_worm_keeps_vertices_diagonal(sse_pkg, oracle_mod):
    """Spin-1/2 one-site bonds in a transverse field: a worm through a
    two-leg vertex bounces or flips both legs, so no off-diagonal vertex is
    ever created (an algorithmic property of the abstract-loop update,
    vertex_data.jl transitions + abstract_loop_update.jl), and the energy is
    that of the diagonal part alone."""
    hx, hz, T = 0.8, 0.3, 0.5
    flat = field_spins(sse_pkg, 4, hx=hx, hz=hz)
    o = _oracle(sse_pkg, oracle_mod, flat, T, 9)
    for _ in range(300):
        o.sweep(1)
        ops = o.ops()
        assert np.all((ops == 0) | ((ops >> 1) & 1 == 1))

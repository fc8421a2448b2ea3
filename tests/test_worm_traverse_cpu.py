"""The reference's own worm_traverse! fixture (test/test_sse.jl:30-60),
restated on the oracle: a 4x4 spin-1/2 Heisenberg model at T = 0.1, two tiny
operator strings -- one operator on bond 1 with diagonal vertex number 3,
and two copies of vertex (diagonal, index 1) on bond 1 -- and one worm
started at (leg 1, worm 1, operator 1).  After the traverse the operator
string must be consistent with the state read off the first vertex of
every site (sites without operators in state 1).  The reference draws the
scatters from one Xoshiro stream; here many Philox seeds cover the possible
exit choices (at this coupling most transitions are deterministic, so most
seeds take the same path).  The fixture is weak on its own: following the
link of the wrong leg of the exit site keeps these two strings consistent
(checked by planting that mistake in the oracle); the worldline test over
300 sweeps of a honeycomb model (test_oracle_chain.py) catches it."""
import numpy as np
import pytest

from fixtures import isconsistent, square_heisenberg


def _opword(bond, vcode):
    return 1 | (vcode << 1) | (bond << 11)


def _state0(flat, ops, first):
    legs = 2 * flat.bond_n_sites
    vd = flat.types[0]
    out = np.ones(flat.n_sites, dtype=np.uint8)
    for site, f in enumerate(first):
        if f < 0:
            continue
        p, leg = divmod(int(f), legs)
        vidx = (int(ops[p]) >> 1 & 1023) >> 1
        out[site] = vd["leg_states"][vidx][leg]
    return out


@pytest.mark.parametrize("case", ["one_operator", "two_operators"])
def test_worm_traverse_reference_fixture(sse_pkg, oracle_mod, case):
    model = square_heisenberg(sse_pkg, 4, measure=(), T=0.1)
    flat = model.flat()
    diag = flat.types[0]["diagonal_vertices"]
    if case == "one_operator":
        ops = np.array([_opword(0, int(diag[2]))], dtype=np.uint32)
    else:
        v1 = 1 | (0 << 1)  # VertexCode(true, 1)
        ops = np.array([_opword(0, v1), _opword(0, v1)], dtype=np.uint32)
    _, first0, _ = oracle_mod.make_vertex_list(sse_pkg.capi, flat.n_sites, flat.bond_sites, ops)
    spins = _state0(flat, ops, first0)
    assert isconsistent(flat, ops, spins)
    for seed in range(200):
        o = oracle_mod.OracleWalker(sse_pkg.capi, flat, sse_pkg.walker_desc(T=0.1, seed=seed))
        o.init()
        sc = {"cutoff": len(ops), "num_operators": len(ops), "avg_worm_length": 1.0,
              "num_worms": 1.0, "sweeps": 0, "rng_counter": 0}
        assert o.set_state(sc, ops, spins) == 0
        assert o.worm_traverse_from(0, 0, 1) >= 1
        new_ops = o.ops()
        _, first, _ = o.vertex_list()
        assert isconsistent(flat, new_ops, _state0(flat, new_ops, first)), f"seed {seed}"
        assert np.all((new_ops >> 11) == 0)
        assert np.all(new_ops != 0)

"""The native transition builder (sse_vertex_transitions, C++ dense simplex)
against scipy's HiGHS on the same linear programs (src/vertex_data.jl:213-416).

Both must reach the same optimum: the total bounce weight (the LP objective,
summed over all tables) agrees to 1e-9.  The native tables satisfy detailed
balance and normalisation; where the builder reports every optimum unique,
its tables equal HiGHS's entry for entry.  Models: the reference's test jobs
(spin-1/2 square with field, mixed spin-1/2 / spin-1 honeycomb with Dz and
Dx, dimer-basis bilayer), the dimer ladder, and random Hamiltonians (also of
a three-site bond: the tables are generic in the number of bond sites, the
device kernels are not)."""
import itertools

import numpy as np
import pytest

from fixtures import job_bilayer, job_honeycomb, job_magnet_square, ladder_dimer, square_heisenberg
from test_vertex_data import detailed_balance

import highs_transitions


def _bond_hamiltonians(sse_pkg):
    VD = sse_pkg.vertex_data.VertexData
    captured = []
    orig = VD.__init__

    def cap(self, dims, H, **kw):
        captured.append((tuple(dims), np.array(H, dtype=np.float64), dict(kw)))
        orig(self, dims, H, **kw)

    VD.__init__ = cap
    try:
        for m in (square_heisenberg(sse_pkg, 4), job_magnet_square(sse_pkg, 0.5),
                  job_honeycomb(sse_pkg, 0.5), ladder_dimer(sse_pkg, 4, 0.6),
                  job_bilayer(sse_pkg, 0.5)):
            m.generate_sse_data()
    finally:
        VD.__init__ = orig
    sp, sz = sse_pkg.spin_operators(2)
    heis = np.kron(sz, sz) + 0.5 * (np.kron(sp, sp.T) + np.kron(sp.T, sp))
    captured.append(((2, 2), heis, {"energy_offset_factor": 0.0}))
    rng = np.random.default_rng(17)
    for dims in ((2, 2), (2, 3), (3, 3), (2, 4), (2, 2, 2), (4, 4)):
        D = int(np.prod(dims))
        H = rng.random((D, D))
        captured.append((dims, H + H.T, {}))
    return captured


def _bounce(sse_pkg, vd):
    u = sse_pkg.util
    tot = 0.0
    for (leg, worm, v), (off, ln) in vd.transitions.items():
        dim = vd.dims[u.site_of_leg(leg, len(vd.dims)) - 1]
        cum = vd.transition_cumprobs[off - 1:off + ln]
        probs = np.diff(np.concatenate([[0.0], cum]))
        for p, (lo, wo) in zip(probs, vd.transition_step_outs[off - 1:off + ln]):
            if lo == leg and wo == u.worm_inverse(worm, dim):
                tot += p * vd.weights[v - 1]
    return tot


@pytest.fixture(scope="module")
def hamiltonians(sse_pkg):
    return _bond_hamiltonians(sse_pkg)


def test_native_matches_highs_optimum(sse_pkg, hamiltonians):
    VD = sse_pkg.vertex_data.VertexData
    unique = 0
    for dims, H, kw in hamiltonians:
        a = VD(dims, H, lp_solver="native", **kw)
        b = VD(dims, H, lp_solver=highs_transitions.construct_transitions, **kw)
        assert np.array_equal(a.weights, b.weights) and a.diagonal_vertices == b.diagonal_vertices
        ba, bb = _bounce(sse_pkg, a), _bounce(sse_pkg, b)
        assert abs(ba - bb) <= 1e-9 * max(1.0, bb), (dims, ba, bb)
        detailed_balance(sse_pkg, a)
        assert set(a.transitions) == set(b.transitions)
        for (offset, length) in a.transitions.values():
            assert length < 0 or abs(a.transition_cumprobs[offset + length - 1] - 1) < 1e-7
        if not a.lp_flags & 4:
            unique += 1
            for key, (oa, la) in a.transitions.items():
                ob, lb = b.transitions[key]
                assert la == lb, key
                assert np.allclose(a.transition_cumprobs[oa - 1:oa + la],
                                   b.transition_cumprobs[ob - 1:ob + lb], atol=1e-9)
                assert a.transition_targets[oa - 1:oa + la] == b.transition_targets[ob - 1:ob + lb]
                assert (a.transition_step_outs[oa - 1:oa + la]
                        == b.transition_step_outs[ob - 1:ob + lb])
    assert unique >= 1  # the epsilon = 0 Heisenberg bond has a unique optimum


def test_native_builder_rejects_bad_input(sse_pkg):
    lib = sse_pkg.capi.load_library()
    with pytest.raises(sse_pkg.capi.SSEError):
        sse_pkg.capi.vertex_transitions(np.array([0, 2]), np.ones((4, 1), np.uint8),
                                        np.ones(1), 1e-7, 1e-10)
    assert lib.sse_vertex_transitions(0, None, 0, None, None, 0.0, 0.0, None, None, 0, None,
                                      None, None, None, None, None) == 1


def test_empty_vertex_set(sse_pkg):
    vd = sse_pkg.VertexData((3, 2), np.eye(6))
    assert vd.vertex_count() == 0 and vd.transitions == {} and vd.transition_cumprobs == []


def test_every_step_has_a_transition(sse_pkg, hamiltonians):
    """Each (vertex, step) whose inverse step reaches a vertex has an entry."""
    u = sse_pkg.util
    VD = sse_pkg.vertex_data.VertexData
    for dims, H, kw in hamiltonians[-6:]:
        vd = VD(dims, H, **kw)
        ns = len(dims)
        for v in range(1, vd.vertex_count() + 1):
            for leg, worm in itertools.product(range(1, 2 * ns + 1), range(1, max(dims))):
                if worm <= u.worm_count(dims[u.site_of_leg(leg, ns) - 1]):
                    assert (leg, worm, v) in vd.transitions

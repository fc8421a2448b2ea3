"""Index helpers, worm algebra, operator encoding and lattices.

Known answers and invariants of the reference's test/test_util.jl,
test/test_worms.jl, test/test_opercode.jl and test/test_lattice.jl.
"""
import itertools
import math
import random

import numpy as np
import pytest


def test_split_join(sse_pkg):
    u = sse_pkg.util
    assert u.split_idx((5, 2), 2) == (2, 1)
    assert u.split_idx((5, 2), 6) == (1, 2)
    dims = (6, 3, 1, 5)
    for i in range(1, int(np.prod(dims)) + 1):
        assert u.join_idx(dims, u.split_idx(dims, i)) == i
    rng = random.Random(3)
    for _ in range(10):
        dims = tuple(rng.randint(1, 400) for _ in range(3))
        for _ in range(10):
            i = rng.randint(1, int(np.prod(dims)))
            assert u.join_idx(dims, u.split_idx(dims, i)) == i


@pytest.mark.parametrize("dim", [2, 4, 8])
def test_worm_algebra(sse_pkg, dim):
    u = sse_pkg.util
    for s_in in range(1, dim + 1):
        outs = []
        for worm in range(1, u.worm_count(dim) + 1):
            s_out = u.worm_action(worm, s_in, dim)
            assert u.worm_action(u.worm_inverse(worm, dim), s_out, dim) == s_in
            outs.append(s_out)
        outs.append(s_in)
        assert sorted(outs) == list(range(1, dim + 1))


def _ref_opcode(bond1, diag, vidx1):
    """Reference UInt64 OperCode (opercode.jl:18-21,43-47): 1-based bond, vidx."""
    v = int(diag) | (vidx1 << 1)
    return 1 | (v << 1) | (bond1 << 26)


def test_opcode_roundtrip_reference_layout(sse_pkg):
    lib = sse_pkg.capi.load_library()
    rng = random.Random(5)
    for bond in rng.sample(range(1, 100 * 100), 20):
        for vidx in rng.sample(range(1, 78), 20):
            for diag in (True, False):
                ref = _ref_opcode(bond, diag, vidx)
                dev = lib.sse_opcode_from_reference(ref)
                assert dev & 1  # non-identity
                assert (dev >> 11) == bond - 1
                assert ((dev >> 1) & 1) == int(diag)
                assert ((dev >> 2) & 511) == vidx - 1
                assert lib.sse_opcode_to_reference(dev) == ref
                # get_vertex masks 25 bits then >>1 (opercode.jl:61-62)
                assert ((ref & ((1 << 25) - 1)) >> 1) == (int(diag) | (vidx << 1))
                assert ref >> 26 == bond
    assert lib.sse_opcode_to_reference(0) == 0 and lib.sse_opcode_from_reference(0) == 0


def test_unit_cells(sse_pkg):
    UC = sse_pkg.UnitCells
    assert UC.square.sites[0].sublattice_sign == 1 and UC.square.sites[0].coordination == 4
    for uc in (UC.columnar_dimer,):
        assert [s.sublattice_sign for s in uc.sites] == [1, -1]
        assert [s.coordination for s in uc.sites] == [4, 4]
    assert [s.sublattice_sign for s in UC.honeycomb.sites] == [1, -1]
    assert [s.coordination for s in UC.honeycomb.sites] == [3, 3]


def _adjacency(lat):
    n = lat.site_count()
    A = np.zeros((n, n), dtype=np.int64)
    for b in lat.bonds:
        A[b.i - 1, b.j - 1] = 1
        A[b.j - 1, b.i - 1] = 1
    return A


def _sq(n):
    return math.comb(n, n // 2) ** 2 if n % 2 == 0 else 0


def _hc(n):
    return sum(math.comb(n // 2, k) ** 2 * math.comb(2 * k, k) for k in range(n // 2 + 1)) \
        if n % 2 == 0 else 0


def _tri(n):
    return sum((-2) ** (n - k) * math.comb(n, k) * sum(math.comb(k, j) ** 3 for j in range(k + 1))
               for k in range(n + 1))


@pytest.mark.parametrize("Ls", [(3, 4), (8, 4), (5, 5), (8, 8)])
def test_lattice_walks(sse_pkg, Ls):
    UC = sse_pkg.UnitCells
    cells = [UC.square, UC.columnar_dimer, UC.honeycomb, UC.triangle]
    walks = [_sq, _sq, _hc, _tri]
    for uc, nw in zip(cells, walks):
        lat = sse_pkg.Lattice(uc, Ls)
        A = _adjacency(lat)
        assert np.trace(A @ A) // 2 == len(lat.bonds)
        for n in range(2, 7):
            if any(L <= n for L in Ls):
                continue
            An = np.linalg.matrix_power(A, n)
            assert np.trace(An) // lat.site_count() == nw(n)
            assert np.trace(An) % lat.site_count() == 0


def test_neel_vectors(sse_pkg):
    UC = sse_pkg.UnitCells
    assert sse_pkg.neel_vector(UC.square) == ((1, 1), False)
    assert sse_pkg.neel_vector(UC.columnar_dimer) == ((1, 0), True)
    assert sse_pkg.neel_vector(UC.honeycomb) == ((0, 0), True)
    assert sse_pkg.neel_vector(UC.triangle) is None


def test_bond_order_square(sse_pkg):
    """Bond order fixes the meaning of the random bond draw (lattice.jl:227-236)."""
    lat = sse_pkg.Lattice(sse_pkg.UnitCells.square, (3, 2))
    # r = (x, y) with x fastest; bonds per cell: (0,1) then (1,0)
    got = [(b.type, b.i, b.j) for b in lat.bonds[:4]]
    assert got == [(1, 1, 4), (2, 1, 2), (1, 2, 5), (2, 2, 3)]
    assert len(lat.bonds) == 12

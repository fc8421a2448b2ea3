"""Unit cells and lattices (host-side model input).

Restates src/models/common/lattice.jl of StochasticSeriesExpansion.jl.  Site
and bond indices are 1-based here, exactly as in the reference, because the
bond order fixes the meaning of the diagonal update's random bond draw
(lattice.jl:227-240).  Conversion to the 0-based device layout happens in
sse_data.flatten_model.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

from .util import join_idx, split_idx


@dataclass(frozen=True)
class UCBond:
    """Unit-cell bond (lattice.jl:14-18): site `iuc` to site `juc` of the cell
    displaced by the non-negative offsets `jd`."""
    iuc: int
    jd: tuple
    juc: int


@dataclass
class UCSite:
    """Site inside the unit cell (lattice.jl:34-40)."""
    pos: tuple
    sublattice_sign: int = 0
    coordination: int = 0


def calculate_uc_signs(bonds, num_sites):
    """Bipartite sublattice signs from intra-cell bonds (lattice.jl:74-102)."""
    signs = [0] * num_sites
    signs[0] = 1
    tries = 0
    guard = 0
    while any(s == 0 for s in signs):
        guard += 1
        if guard > 4 * (len(bonds) ** 2 + 1):
            break
        for b in bonds:
            if any(d != 0 for d in b.jd):
                continue
            i, j = b.iuc - 1, b.juc - 1
            if signs[i] != 0 and signs[j] == 0:
                signs[j] = -signs[i]
            elif signs[j] != 0 and signs[i] == 0:
                signs[i] = -signs[j]
            elif signs[i] == signs[j]:
                signs = [1] * num_sites  # lattice not bipartite
                break
            tries += 1
    if any(s == 0 for s in signs):
        signs = [1] * num_sites
    return signs


def calculate_uc_coordinations(bonds, num_sites):
    """Coordination number per unit-cell site (lattice.jl:104-113)."""
    c = [0] * num_sites
    for b in bonds:
        c[b.iuc - 1] += 1
        c[b.juc - 1] += 1
    return c


@dataclass
class UnitCell:
    """Unit cell (lattice.jl:67-132); lattice vectors are the columns."""
    lattice_vectors: np.ndarray
    sites: list
    bonds: list
    name: str = ""

    def __post_init__(self):
        signs = calculate_uc_signs(self.bonds, len(self.sites))
        coords = calculate_uc_coordinations(self.bonds, len(self.sites))
        self.sites = [UCSite(tuple(s.pos), sg, c) for s, sg, c in zip(self.sites, signs, coords)]
        self.lattice_vectors = np.asarray(self.lattice_vectors, dtype=float)

    @property
    def dimension(self):
        return len(self.sites[0].pos)


def neel_vector(uc: UnitCell):
    """Ordering vector q (in units of pi) and stagger_uc flag for which every
    bond connects opposite signs, or None (lattice.jl:134-150)."""
    D = uc.dimension
    for stagger_uc in (False, True):
        for q in itertools.product((False, True), repeat=D):
            # Iterators.product varies the FIRST index fastest
            q = tuple(reversed(q))
            ok = True
            for b in uc.bonds:
                si = uc.sites[b.iuc - 1].sublattice_sign ** int(stagger_uc)
                sj = uc.sites[b.juc - 1].sublattice_sign ** int(stagger_uc) * \
                    (-1) ** sum(jd * int(qq) for jd, qq in zip(b.jd, q))
                if not (si != sj):
                    ok = False
                    break
            if ok:
                return tuple(int(x) for x in q), stagger_uc
    return None


@dataclass(frozen=True)
class LatticeBond:
    """Lattice bond (lattice.jl:165-169): translation class `type`, sites i, j (1-based)."""
    type: int
    i: int
    j: int


@dataclass(frozen=True)
class LatticeSite:
    iuc: int
    ix: tuple


class Lattice:
    """D-dimensional periodic lattice built from a unit cell (lattice.jl:210-243).

    Bond order: for r in product(1:L1, 1:L2, ...) with the first coordinate
    fastest, for each unit-cell bond in order (lattice.jl:227-236)."""

    def __init__(self, uc: UnitCell, Ls):
        Ls = tuple(int(x) for x in Ls)
        assert len(Ls) == uc.dimension
        self.uc = uc
        self.Ls = Ls
        dims = (len(uc.sites),) + Ls
        bonds = []
        sites = []
        # Iterators.product: first index fastest
        for r_rev in itertools.product(*[range(1, L + 1) for L in reversed(Ls)]):
            r = tuple(reversed(r_rev))
            for bond_type, b in enumerate(uc.bonds, start=1):
                assert 1 <= b.iuc <= len(uc.sites)
                assert 1 <= b.juc <= len(uc.sites)
                i = join_idx(dims, (b.iuc,) + r)
                jr = tuple((rr + d - 1) % L + 1 for rr, d, L in zip(r, b.jd, Ls))
                j = join_idx(dims, (b.juc,) + jr)
                bonds.append(LatticeBond(bond_type, i, j))
            for iuc in range(1, len(uc.sites) + 1):
                sites.append(LatticeSite(iuc, r))
        self.bonds = bonds
        self.sites = sites

    @classmethod
    def from_params(cls, p):
        if isinstance(p, dict):
            return cls(p["unitcell"], p["size"])
        return cls(p.unitcell, p.size)

    @property
    def dimension(self):
        return len(self.Ls)

    def site_count(self):
        return len(self.uc.sites) * int(np.prod(self.Ls))

    def split_idx(self, site_idx):
        r = split_idx((len(self.uc.sites),) + self.Ls, site_idx)
        return r[0], r[1:]

    def staggered_sign(self, ordering_vector, stagger_uc, site_idx):
        """Sign of site `site_idx` (1-based) under ordering vector q
        (lattice.jl:257-267)."""
        s = self.sites[site_idx - 1]
        sign = self.uc.sites[s.iuc - 1].sublattice_sign if stagger_uc else 1
        tot = sum(int(q) * x for q, x in zip(ordering_vector, s.ix))
        sign *= 1 - 2 * (tot % 2)
        return sign


def _uc(vectors, sites, bonds, name):
    return UnitCell(np.array(vectors, dtype=float), [UCSite(tuple(p)) for p in sites],
                    [UCBond(i, tuple(jd), j) for (i, jd, j) in bonds], name)


class UnitCells:
    """Predefined unit cells (lattice.jl:280-331)."""
    square = _uc([[1.0, 0.0], [0.0, 1.0]], [(0.0, 0.0)],
                 [(1, (0, 1), 1), (1, (1, 0), 1)], "square")
    columnar_dimer = _uc([[1.0, 0.0], [0.0, 2.0]], [(0.0, 0.0), (0.0, 0.5)],
                         [(1, (0, 0), 2), (1, (1, 0), 1), (2, (1, 0), 2), (2, (0, 1), 1)],
                         "columnar_dimer")
    honeycomb = _uc([[math.sqrt(3) / 2, math.sqrt(3 / 2)], [-0.5, 0.5]],
                    [(0.0, 0.0), (1 / 3, 1 / 3)],
                    [(1, (0, 0), 2), (2, (0, 1), 1), (2, (1, 0), 1)], "honeycomb")
    triangle = _uc([[1.0, -0.5], [0.0, math.sqrt(3 / 2)]], [(0.0, 0.0)],
                   [(1, (0, 1), 1), (1, (1, 0), 1), (1, (1, 1), 1)], "triangle")
    fully_frust_square_bilayer = _uc(
        [[1.0, 0.0], [0.0, 1.0]], [(0.0, 0.0), (0.0, 0.0)],
        [(1, (0, 0), 2), (1, (0, 1), 1), (2, (0, 1), 2), (1, (1, 0), 1), (2, (1, 0), 2),
         (1, (0, 1), 2), (2, (0, 1), 1), (1, (1, 0), 2), (2, (1, 0), 1)],
        "fully_frust_square_bilayer")
    # two-site rung cell of a ladder: rung bond, then the two leg bonds
    # (the configuration the ladder workload uses; lattice size (Lx, 1))
    ladder = _uc([[1.0, 0.0], [0.0, 1.0]], [(0.0, 0.0), (0.0, 0.5)],
                 [(1, (0, 0), 2), (1, (1, 0), 1), (2, (1, 0), 2)], "ladder")

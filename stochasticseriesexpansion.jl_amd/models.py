"""Models: MagnetModel, ClusterModel and the MagnetizationEstimator family.

Host-side restatement of src/models/common/operators.jl,
src/models/magnet/magnet.jl, src/models/cluster/cluster.jl and the
observable naming / per-site tables of
src/models/common/magnetization_estimator.jl.  A model turns task
parameters into an SSEData (bond graph + VertexData tables) which the device
path consumes unchanged, so any model that supplies a bond Hamiltonian
matrix runs on it.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

from .lattice import Lattice, neel_vector
from .sse_data import FlatModel, SSEBond, SSEData
from .vertex_data import VertexData


# ---------------------------------------------------------------- operators
def spin_operators(dimension):
    """(S+, Sz) for S = (dimension-1)/2 (operators.jl:4-22)."""
    S = (dimension - 1) / 2
    sz = np.zeros((dimension, dimension))
    splus = np.zeros((dimension, dimension))
    for i in range(1, dimension + 1):
        m = S - i + 1
        sz[i - 1, i - 1] = m
        if i < dimension:
            splus[i - 1, i] = np.sqrt((S - m + 1) * (S + m))
    return splus, sz


def bosonic_a_operator(dimension):
    """Bosonic annihilation operator with cutoff n < dimension (operators.jl:25-29)."""
    return np.diag(np.sqrt(np.arange(1, dimension, dtype=float)), 1)


# ------------------------------------------------------------ estimators
def magnetization_estimator_standard_prefix(q, stagger_uc):
    """Conventional observable prefix (magnetization_estimator.jl:172-180)."""
    if not any(q) and not stagger_uc:
        return ""
    names = "".join(chr((ord("X") - ord("A") + i) % 26 + ord("A"))
                    for i, qq in enumerate(q) if qq)
    return "Stag" + names + ("uc" if stagger_uc else "")


@dataclass(frozen=True)
class MagnetizationEstimator:
    """MagnetizationEstimator{q, stagger_uc, Model, prefix, tag} type parameters
    (magnetization_estimator.jl:32-48)."""
    q: tuple
    stagger_uc: bool
    prefix: str
    tag: object = None


def all_magnetization_estimators(dimension, tag=None):
    """Every (q, stagger_uc) combination in the reference's array order
    (magnetization_estimator.jl:55-70): stagger_uc fastest, then q with its
    first component fastest."""
    out = []
    for q_rev in itertools.product((False, True), repeat=dimension):
        q = tuple(reversed(q_rev))
        for stagger_uc in (False, True):
            out.append(MagnetizationEstimator(q, stagger_uc,
                                              magnetization_estimator_standard_prefix(q, stagger_uc),
                                              tag))
    return out


class ParameterMap:
    """Parameter maps (magnet.jl:63-78)."""

    def __init__(self, m):
        self.map = m

    def get(self, name, index):
        if self.map is None:
            return name
        if name not in self.map:
            return name
        seq = self.map[name]
        if not (1 <= index <= len(seq)):
            return name
        return seq[index - 1]


def _num(x):
    return float(x) if not isinstance(x, Fraction) else float(x)


@dataclass
class MagnetBondParams:
    J: float
    d: float
    Dx: tuple
    Dz: tuple
    hz: tuple


class MagnetModel:
    """Arbitrary-spin quantum magnet (magnet.jl:26-120):
    H = sum_<ij> J (S_i.S_j + d S_i^z S_j^z) + sum_i h S_i^z + Dz (S_i^z)^2 + Dx (S_i^x)^2.

    Task parameters: lattice, S, J, d, hz, Dz, Dx, parameter_map, measure."""

    leg_count = 4

    def __init__(self, params):
        self.params = params
        lattice = params["lattice"]
        if not isinstance(lattice, Lattice):
            lattice = Lattice.from_params(lattice)
        self.lattice = lattice
        assert len(lattice.bonds) > 0
        pm = ParameterMap(params.get("parameter_map")).get

        def split_site(param, bond, default):
            iuc, _ = lattice.split_idx(bond.i)
            juc, _ = lattice.split_idx(bond.j)
            first = _num(params.get(pm(param, iuc), default)) / lattice.uc.sites[iuc - 1].coordination
            second = _num(params.get(pm(param, juc), default)) / lattice.uc.sites[juc - 1].coordination
            return (first, second)

        self.bond_params = [
            MagnetBondParams(_num(params[pm("J", b.type)]), _num(params.get(pm("d", b.type), 0.0)),
                             split_site("Dx", b, 0.0), split_site("Dz", b, 0.0),
                             split_site("hz", b, 0.0))
            for b in lattice.bonds]
        per_cell = [int(Fraction(params.get(pm("S", i), Fraction(1, 2))) * 2 + 1)
                    for i in range(1, len(lattice.uc.sites) + 1)]
        self.spin_states = per_cell * int(np.prod(lattice.Ls))
        self.opstring_estimators = self._gen_opstring_estimators(params.get("measure", []))

    def _gen_opstring_estimators(self, measure):
        """gen_opstring_estimators (magnet.jl:186-209)."""
        ests = []
        D = self.lattice.dimension
        for est in measure:
            if est == "magnetization":
                ests.append(MagnetizationEstimator(tuple([False] * D), False, ""))
            elif est == "staggered_magnetization":
                neel = neel_vector(self.lattice.uc)
                if neel is None:
                    raise ValueError("selected :staggered_magnetization measurement, but "
                                     "lattice does not have a Neel vector.")
                q, stagger_uc = neel
                ests.append(MagnetizationEstimator(tuple(bool(x) for x in q), stagger_uc, "Stag"))
            elif isinstance(est, MagnetizationEstimator):
                ests.append(est)
            else:
                raise ValueError(f"Unrecognized measure option '{est}'")
        return ests

    # AbstractModel interface (abstract_model.jl:17-54)
    def normalization_site_count(self):
        return self.lattice.site_count()

    def get_opstring_estimators(self):
        return self.opstring_estimators

    def magnetization_state(self, tag, site_idx, state_idx):
        """m of 1-based state at 1-based site (magnet.jl:122-130)."""
        return (self.spin_states[site_idx - 1] - 1) * 0.5 - state_idx + 1

    def staggered_sign(self, q, stagger_uc, site_idx):
        return self.lattice.staggered_sign(q, stagger_uc, site_idx)

    def generate_sse_data(self):
        """One VertexData per unit-cell bond, zipped with the first lattice
        bonds' parameters (magnet.jl:169-184)."""
        vds = []
        for uc_bond, bond in zip(self.lattice.uc.bonds, self.bond_params):
            dims, H, factor = generate_bond_hamiltonian(
                uc_bond, bond, (self.spin_states[uc_bond.iuc - 1], self.spin_states[uc_bond.juc - 1]))
            vds.append(VertexData(dims, H, energy_offset_factor=factor))
        bonds = [SSEBond(b.type, (b.i, b.j)) for b in self.lattice.bonds]
        return SSEData(vds, bonds)

    def estimator_tables(self):
        return _estimator_tables(self, self.lattice.site_count(), self.spin_states)

    def flat(self, sse_data=None):
        sse_data = sse_data or self.generate_sse_data()
        return FlatModel(sse_data, self.opstring_estimators, self.normalization_site_count(),
                         self.estimator_tables())


def _estimator_tables(model, n_sites, site_dims):
    ests = model.get_opstring_estimators()
    maxdim = max(site_dims) if len(site_dims) else 1
    coef = np.zeros((len(ests), n_sites))
    mval = np.zeros((len(ests), n_sites, maxdim))
    for e, est in enumerate(ests):
        for site in range(1, n_sites + 1):
            coef[e, site - 1] = model.staggered_sign(est.q, est.stagger_uc, site)
            for s in range(1, site_dims[site - 1] + 1):
                mval[e, site - 1, s - 1] = model.magnetization_state(est.tag, site, s)
    return coef, mval


def generate_bond_hamiltonian(uc_bond, bond, spin_states):
    """Bond Hamiltonian in the kron convention (magnet.jl:133-165).

    The energy offset factor is 0.25 in every case: the reference's special
    case for S=1/2 compares the field tuples to a scalar, which is never true
    (magnet.jl:158-162)."""
    dimi, dimj = spin_states
    splusi, szi = spin_operators(dimi)
    splusj, szj = spin_operators(dimj)
    idi = np.eye(dimi)
    idj = np.eye(dimj)
    H = (bond.J * (0.5 * (np.kron(splusi.T, splusj) + np.kron(splusi, splusj.T))
                   + (1 + bond.d) * np.kron(szi, szj))
         + bond.hz[0] * np.kron(idi, szj)
         + bond.hz[1] * np.kron(szi, idj)
         + bond.Dx[0] / 4 * np.kron(np.linalg.matrix_power(splusi + splusi.T, 2), idj)
         + bond.Dx[1] / 4 * np.kron(idi, np.linalg.matrix_power(splusj + splusj.T, 2))
         + bond.Dz[0] * np.kron(szi @ szi, idj)
         + bond.Dz[1] * np.kron(idi, szj @ szj))
    return (dimi, dimj), H, 0.25


# ---------------------------------------------------------------- cluster
@dataclass
class ClusterBasis:
    """Cluster basis: quantum numbers per state and the unitary whose columns
    are the basis states (cluster.jl:1-11)."""
    quantum_numbers: list
    transformation: np.ndarray


class ClusterBases:
    s2 = 1 / np.sqrt(2)
    dimer = ClusterBasis([(0.0, 0.0), (1.0, 1.0), (1.0, 0.0), (1.0, -1.0)],
                         np.array([[0, 1, 0, 0], [s2, 0, s2, 0], [-s2, 0, s2, 0], [0, 0, 0, 1]]))


def lift_twobody_operator(op, site_dims, sites):
    """Lift a two-site operator to the product space of site_dims (kron order,
    leftmost slowest) acting on the 1-based `sites` (cluster.jl:93-126)."""
    site_dims = tuple(site_dims)
    n = len(site_dims)
    D = int(np.prod(site_dims))
    d1, d2 = site_dims[sites[0] - 1], site_dims[sites[1] - 1]
    op4 = np.asarray(op).reshape(d1, d2, d1, d2)  # row (a1,a2) col (b1,b2), a1 slowest
    res = np.zeros((D, D), dtype=np.result_type(op, np.float64))
    for I in itertools.product(*[range(d) for d in site_dims]):
        a1, a2 = I[sites[0] - 1], I[sites[1] - 1]
        row = np.ravel_multi_index(I, site_dims)
        for b1 in range(d1):
            for b2 in range(d2):
                K = list(I)
                K[sites[0] - 1] = b1
                K[sites[1] - 1] = b2
                col = np.ravel_multi_index(tuple(K), site_dims)
                res[row, col] = op4[a1, a2, b1, b2]
    return res


class ClusterModel:
    """Wraps a MagnetModel and simulates it in a cluster basis (cluster.jl:26-62)."""

    leg_count = 4

    def __init__(self, params):
        inner_cls = params["inner_model"]
        self.inner_model = inner_cls(params)
        self.basis = tuple(params["cluster_bases"])
        pm = ParameterMap(params.get("parameter_map")).get
        lat = self.inner_model.lattice
        cids = [params.get(pm("cluster_id", i), 1) for i in range(1, len(lat.uc.sites) + 1)]
        if len(set(cids)) != len(self.basis):
            raise ValueError(f"Number of cluster bases ({len(self.basis)}) does not match number "
                             f"of distinct cluster ids ({len(set(cids))}).")
        self.uc_cluster_ids = cids
        self.cluster_ids = cids * int(np.prod(lat.Ls))
        self.opstring_estimators = [
            MagnetizationEstimator(tuple([False] * lat.dimension), False, m["name"],
                                   m["quantum_number"])
            for m in params.get("measure_quantum_numbers", [])]
        self._sse_sites = None

    def normalization_site_count(self):
        return self.inner_model.normalization_site_count()

    def get_opstring_estimators(self):
        return self.opstring_estimators

    def magnetization_state(self, tag, site_idx, state_idx):
        """Quantum number `tag` of the cluster state (cluster.jl:66-73)."""
        return self.basis[self.cluster_ids[site_idx - 1] - 1].quantum_numbers[state_idx - 1][tag - 1]

    def staggered_sign(self, q, stagger_uc, site_idx):
        return 1

    def build_cluster_hamiltonians(self):
        """Intra- and inter-cluster bond Hamiltonians (cluster.jl:128-196)."""
        mag = self.inner_model
        cids = self.uc_cluster_ids
        uniq = list(dict.fromkeys(cids))
        clusters = [[i for i in range(1, len(cids) + 1) if cids[i - 1] == cid] for cid in uniq]
        cluster_dims = [tuple(mag.spin_states[s - 1] for s in cl) for cl in clusters]
        ordering = [0] * len(cids)
        for cl in clusters:
            for k, s in enumerate(cl, start=1):
                ordering[s - 1] = k
        intra = [np.zeros((int(np.prod(d)), int(np.prod(d)))) for d in cluster_dims]
        inter = {}
        for uc_bond, bond in zip(mag.lattice.uc.bonds, mag.bond_params):
            (di, dj), H, _ = generate_bond_hamiltonian(
                uc_bond, bond, (mag.spin_states[uc_bond.iuc - 1], mag.spin_states[uc_bond.juc - 1]))
            ci, cj = cids[uc_bond.iuc - 1], cids[uc_bond.juc - 1]
            if ci == cj and all(x == 0 for x in uc_bond.jd):
                if uc_bond.iuc == uc_bond.juc:
                    raise ValueError("found bond in Magnet connecting site to itself... not "
                                     "supported by ClusterModel")
                intra[ci - 1] += lift_twobody_operator(
                    H, cluster_dims[ci - 1], (ordering[uc_bond.iuc - 1], ordering[uc_bond.juc - 1]))
            else:
                key = (ci, cj, tuple(uc_bond.jd))
                bd = cluster_dims[ci - 1] + cluster_dims[cj - 1]
                dim = int(np.prod(bd))
                if key not in inter:
                    inter[key] = np.zeros((dim, dim))
                inter[key] += lift_twobody_operator(
                    H, bd, (ordering[uc_bond.iuc - 1],
                            len(cluster_dims[ci - 1]) + ordering[uc_bond.juc - 1]))
        return intra, inter

    def generate_sse_data(self):
        """Cluster-basis SSEData (cluster.jl:245-306).  Bond types follow the
        insertion order of the inter-cluster dictionary."""
        mag = self.inner_model
        intra, inter = self.build_cluster_hamiltonians()
        keys = list(inter.keys())
        nclusters = len(intra)
        nuc = len(mag.lattice.uc.sites)
        bonds = []
        seen = set()
        for b in mag.lattice.bonds:
            ucb = mag.lattice.uc.bonds[b.type - 1]
            iuc = self.uc_cluster_ids[ucb.iuc - 1]
            juc = self.uc_cluster_ids[ucb.juc - 1]
            i = nclusters * ((b.i - 1) // nuc) + iuc
            j = nclusters * ((b.j - 1) // nuc) + juc
            if iuc == juc and all(x == 0 for x in ucb.jd):
                continue
            bt = next(k for k, key in enumerate(keys, start=1)
                      if key[0] == iuc and key[1] == juc and key[2] == tuple(ucb.jd))
            sb = SSEBond(bt, (i, j))
            if sb not in seen:
                seen.add(sb)
                bonds.append(sb)
        coord = [sum((k[0] == c) + (k[1] == c) for k in keys) for c in range(1, nclusters + 1)]
        vds = []
        for key in keys:
            ci, cj, _ = key
            H = inter[key] + np.kron(intra[ci - 1], np.eye(intra[cj - 1].shape[0]) / coord[ci - 1]) \
                + np.kron(np.eye(intra[ci - 1].shape[0]) / coord[cj - 1], intra[cj - 1])
            U = np.kron(self.basis[ci - 1].transformation, self.basis[cj - 1].transformation)
            vds.append(VertexData((intra[ci - 1].shape[0], intra[cj - 1].shape[0]),
                                  U.T @ H @ U, energy_offset_factor=0.25))
        return SSEData(vds, bonds)

    def estimator_tables(self):
        n = len(self.cluster_ids)
        dims = [len(self.basis[c - 1].quantum_numbers) for c in self.cluster_ids]
        return _estimator_tables(self, n, dims)

    def flat(self, sse_data=None):
        sse_data = sse_data or self.generate_sse_data()
        return FlatModel(sse_data, self.opstring_estimators, self.normalization_site_count(),
                         self.estimator_tables())

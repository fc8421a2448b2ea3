"""Shared test inputs: the reference's test jobs and small workloads."""
import json
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def vertex_list_fixture():
    """Reference golden vertex list converted to 0-based leg ids 4p+leg."""
    d = json.load(open(os.path.join(HERE, "golden", "vertex_list_fixture.json")))
    ops = []
    for o in d["operators"]:
        if o is None:
            ops.append(0)
        else:
            bond, diag, vidx = o
            vcode = int(diag) | ((vidx - 1) << 1)
            ops.append(1 | (vcode << 1) | ((bond - 1) << 11))
    L = len(ops)

    def leg_id(pair):
        leg, p = pair
        return -1 if p == -1 else 4 * (p - 1) + (leg - 1)

    links = np.full(4 * L, -1, dtype=np.int64)
    for p, legs in d["expected_vertices"].items():
        for k, pair in enumerate(legs):
            links[4 * (int(p) - 1) + k] = leg_id(pair)
    first = np.array([leg_id(x) for x in d["expected_v_first"]], dtype=np.int64)
    last = np.array([leg_id(x) for x in d["expected_v_last"]], dtype=np.int64)
    bond_sites = np.array(d["bonds"], dtype=np.int32) - 1
    return d["site_count"], bond_sites, np.array(ops, dtype=np.uint32), links, first, last


def isconsistent(flat, ops, state0):
    """Worldline consistency (test/test_sse.jl:5-28): propagating the state
    through the operators, every operator's lower legs match the state."""
    state = np.array(state0, dtype=np.int64).copy()
    if not np.all((state >= 1) & (state <= flat.site_dims)):
        return False
    ns = flat.bond_n_sites
    for op in ops:
        op = int(op)
        if op == 0:
            continue
        b = op >> 11
        v = (op >> 2) & 511
        t = flat.types[flat.bond_type[b]]
        ls = t["leg_states"][v]
        sites = flat.bond_sites[b]
        if any(state[s] != ls[i] for i, s in enumerate(sites)):
            return False
        for i, s in enumerate(sites):
            state[s] = ls[ns + i]
    return bool(np.array_equal(state, state0))


TS = np.linspace(0.04, 4.00, 7)


def job_magnet_square(sse, T):
    """testjob_magnet_square (test/test_jobs.jl:4-29)."""
    return sse.MagnetModel({"lattice": {"unitcell": sse.UnitCells.square, "size": (2, 4)},
                            "measure": sse.all_magnetization_estimators(2), "J": 1.23, "hz": -0.2,
                            "T": T})


def job_honeycomb(sse, T):
    """testjob_honeycomb (test/test_jobs.jl:31-67)."""
    return sse.MagnetModel({"lattice": {"unitcell": sse.UnitCells.honeycomb, "size": (2, 2)},
                            "parameter_map": {"S": ["Sa", "Sb"], "J": ["J1", "J2", "J3"]},
                            "measure": sse.all_magnetization_estimators(2), "J1": 1.0, "J2": 0.5,
                            "J3": 1.0, "d": 0.2, "Dz": 0.2, "Dx": 0.5, "Sa": Fraction(1, 2),
                            "Sb": 1, "T": T})


def job_bilayer(sse, T):
    """testjob_fully_frustrated_bilayer (test/test_jobs.jl:69-110)."""
    return sse.ClusterModel({
        "inner_model": sse.MagnetModel, "cluster_bases": (sse.ClusterBases.dimer,),
        "lattice": {"unitcell": sse.UnitCells.fully_frust_square_bilayer, "size": (2, 2)},
        "measure_quantum_numbers": [{"name": "", "quantum_number": 2}],
        "parameter_map": {"S": ["Sa", "Sb"],
                          "J": ["JD", "JxP", "JxP", "JyP", "JyP", "JxX", "JxX", "JyX", "JyX"]},
        "JD": 1, "JxP": 0.55, "JyP": 0.6, "JxX": 0.7, "JyX": 0.3, "d": -0.2, "Dz": 0.2, "Dx": 0.5,
        "Sa": Fraction(1, 2), "Sb": Fraction(1, 2), "T": T})


def square_heisenberg(sse, L, measure=("magnetization", "staggered_magnetization"), **kw):
    p = {"lattice": {"unitcell": sse.UnitCells.square, "size": (L, L)}, "J": 1.0,
         "measure": list(measure)}
    p.update(kw)
    return sse.MagnetModel(p)


def ladder_dimer(sse, Lx, J_leg, J_rung=1.0):
    """Two-leg ladder in the rung-dimer basis (cluster basis, local dim 4)."""
    return sse.ClusterModel({
        "inner_model": sse.MagnetModel, "cluster_bases": (sse.ClusterBases.dimer,),
        "lattice": {"unitcell": sse.UnitCells.ladder, "size": (Lx, 1)},
        "measure_quantum_numbers": [{"name": "", "quantum_number": 2}],
        "parameter_map": {"J": ["Jr", "Jl", "Jl"]}, "Jr": J_rung, "Jl": J_leg})


class _Est:
    """Name prefix of an estimator row of a hand-built FlatModel."""

    def __init__(self, prefix):
        self.prefix = prefix


def _spin_half():
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    return sp, sz


def trimer_ring(sse, n, J=-1.0, h=0.15):
    """Spin-1/2 ring whose bonds are three consecutive sites (SSEBond{3}):
    H_b = J (S_a.S_b + S_b.S_c + S_a.S_c) - (h/3)(S_a^z + S_b^z + S_c^z), the
    field split over the three bonds that touch a site.  Ferromagnetic J < 0:
    every vertex weight is positive.  Returns (flat, bond matrices, bond
    site tuples 0-based)."""
    sp, sz = _spin_half()
    I2 = np.eye(2)

    def lift(op, k):
        mats = [I2, I2, I2]
        mats[k] = op
        return np.kron(np.kron(mats[0], mats[1]), mats[2])

    Hb = np.zeros((8, 8))
    for a, b in ((0, 1), (1, 2), (0, 2)):
        Hb += J * (lift(sz, a) @ lift(sz, b) + 0.5 * (lift(sp, a) @ lift(sp.T, b)
                                                        + lift(sp.T, a) @ lift(sp, b)))
    for a in range(3):
        Hb -= (h / 3.0) * lift(sz, a)
    # VertexData takes the bond matrix in kron order, the first site slowest
    # (vertex_data.jl:137-169 splits the index over reversed dims)
    vd = sse.VertexData((2, 2, 2), Hb)
    bonds = [sse.SSEBond(1, (i + 1, (i + 1) % n + 1, (i + 2) % n + 1)) for i in range(n)]
    coef = np.ones((1, n))
    mval = np.tile(np.array([0.5, -0.5]), (1, n, 1))
    flat = sse.FlatModel(sse.SSEData([vd], bonds), [_Est("")], n, (coef, mval))
    return flat, Hb, [tuple(x - 1 for x in b.sites) for b in bonds]


def field_spins(sse, n, hx=0.8, hz=0.3):
    """Spin-1/2 sites in a tilted field, one-site bonds (SSEBond{1}, two legs
    per vertex): H_i = -hx S^x_i - hz S^z_i.  A worm through a two-leg vertex
    either bounces or passes and flips both legs, so it never turns a
    diagonal vertex into an off-diagonal one: with one-site bonds alone the
    abstract-loop chain (reference included) is not ergodic and does not
    reach -(|h|/2) tanh(beta |h| / 2).  The model exercises the leg-count-2
    code paths for bit-exact parity only."""
    sp, sz = _spin_half()
    sx = 0.5 * (sp + sp.T)
    Hb = -hx * sx - hz * sz
    vd = sse.VertexData((2,), Hb)
    bonds = [sse.SSEBond(1, (i + 1,)) for i in range(n)]
    coef = np.ones((1, n))
    mval = np.tile(np.array([0.5, -0.5]), (1, n, 1))
    return sse.FlatModel(sse.SSEData([vd], bonds), [_Est("")], n, (coef, mval))


def bond_model_ed(n, site_dim, bond_sites, Hb, Ts):
    """Exact energy per site of H = sum over bonds of Hb acting on the bond's
    sites (Hb in kron order, first bond site slowest); returns E(T) array."""
    D = site_dim ** n
    H = np.zeros((D, D))
    for sites in bond_sites:
        H += _embed(Hb, sites, n, site_dim)
    w = np.linalg.eigvalsh(H)
    out = []
    for T in Ts:
        x = np.exp(-(w - w.min()) / T)
        out.append(float((w * x).sum() / x.sum()) / n)
    return np.array(out)


def _embed(Hb, sites, n, d):
    """Hb (kron order over `sites`) as an operator on n sites of dim d (site 0
    slowest)."""
    k = len(sites)
    D = d ** n
    H = np.zeros((D, D))
    for i in range(D):
        di = [(i // d ** (n - 1 - s)) % d for s in range(n)]
        bi = 0
        for s in sites:
            bi = bi * d + di[s]
        for bj in range(d ** k):
            v = Hb[bi, bj]
            if v == 0.0:
                continue
            dj = list(di)
            rem = bj
            for s in reversed(sites):
                dj[s] = rem % d
                rem //= d
            j = 0
            for s in range(n):
                j = j * d + dj[s]
            H[i, j] += v
    return H


def plaquette_ring(sse, n, J=-0.7, h=0.1):
    """Spin-1/2 ring whose bonds are four consecutive sites (SSEBond{4}):
    H_b = J sum over the six site pairs of S.S - (h/4) sum S^z, ferromagnetic
    (all vertex weights positive).  Returns (flat, Hb, bond sites 0-based)."""
    sp, sz = _spin_half()
    I2 = np.eye(2)

    def lift(op, k):
        mats = [I2, I2, I2, I2]
        mats[k] = op
        out = mats[0]
        for mm in mats[1:]:
            out = np.kron(out, mm)
        return out

    Hb = np.zeros((16, 16))
    for a in range(4):
        for b in range(a + 1, 4):
            Hb += J * (lift(sz, a) @ lift(sz, b) + 0.5 * (lift(sp, a) @ lift(sp.T, b)
                                                            + lift(sp.T, a) @ lift(sp, b)))
        Hb -= (h / 4.0) * lift(sz, a)
    vd = sse.VertexData((2, 2, 2, 2), Hb)
    bonds = [sse.SSEBond(1, tuple((i + k) % n + 1 for k in range(4))) for i in range(n)]
    coef = np.ones((1, n))
    mval = np.tile(np.array([0.5, -0.5]), (1, n, 1))
    flat = sse.FlatModel(sse.SSEData([vd], bonds), [_Est("")], n, (coef, mval))
    return flat, Hb, [tuple(x - 1 for x in b.sites) for b in bonds]

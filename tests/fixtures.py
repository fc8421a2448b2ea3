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
    for op in ops:
        op = int(op)
        if op == 0:
            continue
        b = op >> 11
        v = (op >> 2) & 511
        t = flat.types[flat.bond_type[b]]
        ls = t["leg_states"][v]
        s0, s1 = flat.bond_sites[b]
        if state[s0] != ls[0] or state[s1] != ls[1]:
            return False
        state[s0], state[s1] = ls[2], ls[3]
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

"""SSEData (bond graph + vertex tables) and its flattening into the C ABI.

SSEBond/SSEData restate src/sse_data.jl; flatten_model packs an SSEData plus
the magnetization-estimator tables into the plain-array layout declared in
include/ssegpu.h (0-based sites, bonds, vertex indices and legs).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class SSEBond:
    """Bond of the SSE graph (sse_data.jl:10-13): VertexData `type` (1-based)
    acting on `sites` (1-based)."""
    type: int
    sites: tuple


@dataclass(frozen=True)
class SSESite:
    dim: int


def generate_sites_from_bonds(vertex_data, bonds):
    """Site dimensions implied by the bonds; inconsistent dims raise
    (sse_data.jl:44-63)."""
    max_site = max(max(b.sites) for b in bonds)
    dims = [0] * max_site
    for b in bonds:
        for is_, s in enumerate(b.sites):
            new_dim = vertex_data[b.type - 1].dims[is_]
            if dims[s - 1] == 0:
                dims[s - 1] = new_dim
            elif dims[s - 1] != new_dim:
                raise ValueError(
                    "SSEData: site dimensions set by VertexData are inconsistent!\n"
                    f"Site {s}: {dims[s - 1]}, Bond {b.type}: {new_dim}")
    return [SSESite(d) for d in dims]


class SSEData:
    """Everything the abstract-loop SSE needs about a model (sse_data.jl:15-42)."""

    def __init__(self, vertex_data, bonds):
        self.vertex_data = list(vertex_data)
        self.bonds = list(bonds)
        self.energy_offset = 0.0
        for b in self.bonds:  # sum(b -> offset, bonds), left to right
            self.energy_offset += self.vertex_data[b.type - 1].energy_offset
        self.sites = generate_sites_from_bonds(self.vertex_data, self.bonds)

    def get_vertex_data(self, bond_idx):
        """VertexData of 1-based bond `bond_idx` (sse_data.jl:70-71)."""
        return self.vertex_data[self.bonds[bond_idx - 1].type - 1]


def _vd_arrays(vd):
    """Flatten one VertexData into the device table layout (legs
    2 * NSites per vertex)."""
    nv = vd.vertex_count()
    mw = vd.max_worm_count
    legs = 2 * len(vd.dims)
    diag = np.full(int(np.prod(vd.dims)), -1, dtype=np.int32)
    for k, v in enumerate(vd.diagonal_vertices):
        if v is not None:
            diag[k] = (v & 1) | (((v >> 1) - 1) << 1)
    toff = np.full((nv, mw, legs), -1, dtype=np.int32)
    tcnt = np.zeros((nv, mw, legs), dtype=np.int32)
    for (leg, worm, v), (offset, length) in vd.transitions.items():
        toff[v - 1, worm - 1, leg - 1] = offset - 1
        tcnt[v - 1, worm - 1, leg - 1] = length + 1
    ne = len(vd.transition_cumprobs)
    cum = np.array(vd.transition_cumprobs, dtype=np.float64).reshape(ne)
    tgt = np.array([(t & 1) | (((t >> 1) - 1) << 1) for t in vd.transition_targets],
                   dtype=np.int32).reshape(ne)
    eleg = np.array([so[0] - 1 for so in vd.transition_step_outs], dtype=np.uint8).reshape(ne)
    ewrm = np.array([so[1] for so in vd.transition_step_outs], dtype=np.uint8).reshape(ne)
    ls = np.ascontiguousarray(vd.leg_states.T.astype(np.uint8))  # [nv][2 NSites]
    return dict(dims=tuple(vd.dims), energy_offset=float(vd.energy_offset), n_vertices=nv,
                weights=np.ascontiguousarray(vd.weights, dtype=np.float64),
                signs=np.ascontiguousarray(vd.signs, dtype=np.int8), leg_states=ls,
                diagonal_vertices=diag, max_worm_count=mw, transition_offset=toff,
                transition_count=tcnt, n_exits=ne, exit_cumprob=cum, exit_target=tgt,
                exit_leg=eleg, exit_worm=ewrm)


class FlatModel:
    """Plain-array form of an SSEData plus estimator tables (include/ssegpu.h
    sse_model_desc).  Arrays are numpy; capi.model_desc builds the ctypes
    struct and keeps them alive."""

    def __init__(self, sse_data, estimators=(), normalization_site_count=None,
                 estimator_tables=None):
        # SSEData{NSites}: every bond and bond type has the same site count
        ns = {len(b.sites) for b in sse_data.bonds} | {len(vd.dims) for vd in sse_data.vertex_data}
        if len(ns) != 1:
            raise ValueError("all bonds and bond types of a model need the same number of sites")
        self.bond_n_sites = ns.pop()
        if not 1 <= self.bond_n_sites <= 4:
            raise ValueError("bonds of 1 to 4 sites are supported")
        self.n_sites = len(sse_data.sites)
        self.site_dims = np.array([s.dim for s in sse_data.sites], dtype=np.uint8)
        self.n_bonds = len(sse_data.bonds)
        self.bond_type = np.array([b.type - 1 for b in sse_data.bonds], dtype=np.int32)
        self.bond_sites = np.array([[s - 1 for s in b.sites] for b in sse_data.bonds],
                                   dtype=np.int32).reshape(self.n_bonds, self.bond_n_sites)
        self.types = [_vd_arrays(vd) for vd in sse_data.vertex_data]
        self.energy_offset = float(sse_data.energy_offset)
        self.normalization_site_count = float(
            normalization_site_count if normalization_site_count is not None else self.n_sites)
        if estimator_tables is None:
            estimator_tables = (np.zeros((0, self.n_sites)), np.zeros((0, self.n_sites, 1)))
        coef, mval = estimator_tables
        self.est_coef = np.ascontiguousarray(coef, dtype=np.float64)
        self.est_mval = np.ascontiguousarray(mval, dtype=np.float64)
        self.n_estimators = int(self.est_coef.shape[0])
        self.est_max_dim = int(self.est_mval.shape[2]) if self.n_estimators else 1
        self.estimators = list(estimators)

    @property
    def measure_width(self):
        return 5 + 5 * self.n_estimators

    def observable_names(self):
        names = ["Sign", "OperatorCount", "SignOperatorCount", "SignOperatorCount2", "SignEnergy"]
        for e in self.estimators:
            p = e.prefix
            names += [f"Sign{p}Mag", f"Sign{p}AbsMag", f"Sign{p}Mag2", f"Sign{p}Mag4",
                      f"Sign{p}MagChi"]
        return names

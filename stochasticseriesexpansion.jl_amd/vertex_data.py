"""VertexData: vertices and abstract-loop transition tables of one bond type.

Host-side restatement of src/vertex_data.jl (StochasticSeriesExpansion.jl).
Conventions follow the reference: vertex indices and states are 1-based,
a vertex code is ``diag | idx << 1`` (opercode.jl:18-21) and ``None`` is the
invalid vertex (opercode.jl:16).  The transition LP (vertex_data.jl:213-416)
is solved by the library's C++ builder (sse_vertex_transitions: a dense
two-phase simplex); `lp_solver` may also be a callable with the signature
of that builder (the tests pass a HiGHS-based one as a cross-check).  When
an LP optimum is degenerate the builder may return a different optimal
vertex than the reference's HiGHS.jl build; every returned table is checked
for detailed balance and normalisation by the tests instead.
"""
from __future__ import annotations

import warnings

import numpy as np

from .util import join_idx, site_of_leg, split_idx, worm_action, worm_count, worm_inverse


def vertex_code(diagonal, idx):
    """VertexCode(diagonal, idx) with 1-based idx (opercode.jl:18-21)."""
    return int(bool(diagonal)) | (int(idx) << 1)


def vcode_is_diagonal(v):
    return bool(v & 1)


def vcode_idx(v):
    return v >> 1


def calc_energy_offset(H, energy_offset_factor):
    """-(hmax + (hmax-hmin)*factor) over the diagonal (vertex_data.jl:129-135)."""
    d = np.diag(H)
    hmin = d.min()
    hmax = d.max()
    epsilon = (hmax - hmin) * energy_offset_factor
    return -(hmax + epsilon)


def construct_vertices(dims, H, energy_offset, tolerance):
    """Vertices = matrix elements with |w| > tolerance, w = -H (minus the offset
    on the diagonal), in column-major order (vertex_data.jl:137-169)."""
    n = H.shape[0]
    rdims = tuple(reversed(dims))
    diagonal_vertices = [None] * n
    weights, signs, leg_states = [], [], []
    for j in range(1, n + 1):          # CartesianIndices: first index fastest
        for i in range(1, n + 1):
            w = -H[i - 1, j - 1]
            if i == j:
                w -= energy_offset
            if abs(w) > tolerance:
                si = tuple(reversed(split_idx(rdims, i)))
                sj = tuple(reversed(split_idx(rdims, j)))
                if i == j:
                    reversed_i = join_idx(dims, si)
                    diagonal_vertices[reversed_i - 1] = vertex_code(True, len(weights) + 1)
                leg_states.append(si + sj)
                weights.append(abs(w))
                signs.append(1 if w >= 0 else -1)
    ls = np.array(leg_states, dtype=np.uint8).reshape(len(weights), 2 * len(dims)).T.copy()
    return diagonal_vertices, np.array(weights, dtype=np.float64), ls, np.array(signs, dtype=np.int8)


def wrap_vertex_idx(leg_states, vertex_idx):
    """VertexCode with the diagonal flag computed from the leg states (vertex_data.jl:171-187)."""
    if vertex_idx is None:
        return None
    leg_count = leg_states.shape[0]
    ls = leg_states[:, vertex_idx - 1]
    diagonal = all(ls[s] == ls[leg_count // 2 + s] for s in range(leg_count // 2))
    return vertex_code(diagonal, vertex_idx)


class _LegStateIndex:
    """findfirst over leg-state columns (vertex_data.jl:210) as a dict lookup."""

    def __init__(self, leg_states):
        self.map = {}
        for v in range(leg_states.shape[1]):
            key = tuple(int(x) for x in leg_states[:, v])
            self.map.setdefault(key, v + 1)

    def find(self, state):
        return self.map.get(tuple(int(x) for x in state))


def vertex_apply_change(leg_states, dims, vertex, step_in, step_out, index=None):
    """Vertex reached by applying step_in then step_out to `vertex`, or None
    (vertex_data.jl:192-211)."""
    ns = len(dims)
    new = [int(x) for x in leg_states[:, vertex - 1]]
    (leg_in, worm_in) = step_in
    (leg_out, worm_out) = step_out
    dim_in = dims[site_of_leg(leg_in, ns) - 1]
    dim_out = dims[site_of_leg(leg_out, ns) - 1]
    new[leg_in - 1] = worm_action(worm_in, new[leg_in - 1], dim_in)
    new[leg_out - 1] = worm_action(worm_out, new[leg_out - 1], dim_out)
    if index is None:
        index = _LegStateIndex(leg_states)
    return index.find(new)


def construct_transitions_native(weights, leg_states, max_worm_count, dims, tolerance,
                                 lp_tolerance):
    """construct_transitions by the library's C++ builder (sse_vertex_transitions:
    the same LPs, orders and rules, solved by a dense simplex on the host).
    Also returns the builder's flags (bit 2: some LP optimum may not be
    unique)."""
    from .capi import vertex_transitions
    off, ln, cum, tg, ol, ow, flags = vertex_transitions(dims, leg_states, weights, tolerance,
                                                         lp_tolerance)
    transitions = {}
    nv, mw, legs = off.shape
    for v, worm, leg in zip(*np.nonzero(off)):
        transitions[(int(leg) + 1, int(worm) + 1, int(v) + 1)] = (int(off[v, worm, leg]),
                                                                  int(ln[v, worm, leg]))
    if flags & 2:
        warnings.warn("normalization error in a transition exit list")
    if flags & 1:
        warnings.warn("had to truncate some probabilities in a possibly inaccurate way!")
    return (transitions, [float(x) for x in cum], [int(t) for t in tg],
            [(int(a), int(b)) for a, b in zip(ol, ow)], flags)


class VertexData:
    """Probability tables for one bond type (vertex_data.jl:13-88).

    Fields mirror the reference: energy_offset, dims, diagonal_vertices
    (list indexed by compound state - 1, entries vertex codes or None),
    signs, weights, transitions {(leg, worm, v): (offset, length)},
    transition_cumprobs, transition_targets, transition_step_outs and
    leg_states (2*NSites x n_vertices, 1-based states)."""

    def __init__(self, dims, bond_hamiltonian, energy_offset_factor=0.25, tolerance=1e-7,
                 lp_tolerance=1e-10, lp_solver="native"):
        H = np.asarray(bond_hamiltonian, dtype=np.float64)
        dims = tuple(int(d) for d in dims)
        total_dim = int(np.prod(dims))
        assert len(dims) >= 1
        assert total_dim == H.shape[0] == H.shape[1]
        self.dims = dims
        self.energy_offset = calc_energy_offset(H, energy_offset_factor)
        self.max_worm_count = max(worm_count(d) for d in dims)
        (self.diagonal_vertices, self.weights, self.leg_states, self.signs) = \
            construct_vertices(dims, H, self.energy_offset, tolerance)
        assert self.leg_states.shape == (2 * len(dims), len(self.weights))
        args = (self.weights, self.leg_states, self.max_worm_count, dims, tolerance, lp_tolerance)
        self.lp_solver = lp_solver
        self.lp_flags = 0
        if lp_solver == "native":
            (self.transitions, self.transition_cumprobs, self.transition_targets,
             self.transition_step_outs, self.lp_flags) = construct_transitions_native(*args)
        elif callable(lp_solver):
            (self.transitions, self.transition_cumprobs, self.transition_targets,
             self.transition_step_outs) = lp_solver(*args)
        else:
            raise ValueError(f"lp_solver {lp_solver!r}: 'native' or a callable "
                             "(weights, leg_states, max_worm_count, dims, tolerance, lp_tolerance)")

    # accessors (vertex_data.jl:91-104)
    def vertex_count(self):
        return len(self.weights)

    def get_diagonal_vertex(self, compound_state_idx):
        return self.diagonal_vertices[compound_state_idx - 1]

    def get_vertex_weight(self, v):
        return 0.0 if v is None else float(self.weights[vcode_idx(v) - 1])

    def get_sign(self, v):
        return int(self.signs[vcode_idx(v) - 1])

    def get_leg_state(self, v):
        return self.leg_states[:, vcode_idx(v) - 1]

    def get_transition(self, leg_in, worm_in, vidx):
        return self.transitions.get((leg_in, worm_in, vidx))

    def scatter(self, v, leg_in, worm_in, random):
        """(leg_out, worm_out, new vertex) or (-1, -1, None) (vertex_data.jl:106-125)."""
        t = self.transitions[(leg_in, worm_in, vcode_idx(v))]
        offset, length = t
        for out in range(offset, offset + length + 1):
            if random < self.transition_cumprobs[out - 1]:
                leg_out, worm_out = self.transition_step_outs[out - 1]
                return leg_out, worm_out, self.transition_targets[out - 1]
        return -1, -1, None

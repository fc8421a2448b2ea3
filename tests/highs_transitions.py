"""construct_transitions (vertex_data.jl:213-416) with every LP solved by
scipy's HiGHS: a cross-check of the library's native builder
(sse_vertex_transitions), test infrastructure only.

It follows the reference step by step (the same step, variable and class
orders, normalisation, clamping and pruning), because the tables can only
agree with the reference's where those orders agree.  Pass it as
VertexData(..., lp_solver=construct_transitions).
"""
import warnings

import numpy as np


def _pkg():
    import __graft_entry__
    return __graft_entry__.load_package()


def _u():
    return _pkg().util


def _vd():
    return _pkg().vertex_data


def _solve_lp(cost, A, b, lp_tolerance):
    from scipy.optimize import linprog  # only the lp_solver="highs" path needs scipy
    res = linprog(cost, A_eq=A, b_eq=b, bounds=[(0, None)] * len(cost), method="highs-ds",
                  options={"primal_feasibility_tolerance": lp_tolerance,
                           "dual_feasibility_tolerance": lp_tolerance})
    if res.status != 0:
        raise RuntimeError(f"transition probability optimization failed: {res.message}")
    return res.x


def construct_transitions(weights, leg_states, max_worm_count, dims, tolerance, lp_tolerance):
    """Transition tables from one LP per class of vertices connected by a
    single worm step (vertex_data.jl:213-416), each LP solved by scipy's HiGHS.

    Returns (transitions, cumprobs, targets, step_outs) where transitions maps
    (leg_in, worm_in, vertex) -> (offset, length) with a 1-based offset and
    length = number of exits - 1, as the reference's Transition."""
    ns = len(dims)
    leg_count = 2 * ns
    transitions = {}
    cumprobs, targets, step_outs = [], [], []

    steps = []
    inv_steps = {}
    step_idx = {}
    for worm in range(1, max_worm_count + 1):
        for leg in range(1, leg_count + 1):
            dim = dims[_u().site_of_leg(leg, ns) - 1]
            if worm <= _u().worm_count(dim):
                steps.append((leg, worm))
                step_idx[(leg, worm)] = len(steps)
                inv_steps[(leg, worm)] = (leg, _u().worm_inverse(worm, dim))

    def inverse(vc):
        return (inv_steps[vc[1]], inv_steps[vc[0]])

    variables = []
    var_set = set()
    for s_in in steps:
        for s_out in steps:
            vc = (s_in, s_out)
            if inverse(vc) not in var_set:
                variables.append(vc)
                var_set.add(vc)
    # variable of (in, out): findfirst over variables equal to it or whose
    # inverse is (vertex_data.jl:359-365)
    var_lookup = {}
    for s_in in steps:
        for s_out in steps:
            key = (s_in, s_out)
            cands = [k for k, vc in enumerate(variables) if vc == key or inverse(vc) == key]
            var_lookup[key] = cands[0]

    cost = np.array([1.0 if vc == inverse(vc) else 0.0 for vc in variables])
    A = np.zeros((len(steps), len(variables)))
    for r, step in enumerate(steps):
        for k, vc in enumerate(variables):
            if vc[0] == step or inverse(vc)[0] == step:
                A[r, k] = 1.0

    index = _vd()._LegStateIndex(leg_states)
    nv = len(weights)
    used_inaccurate = False
    while True:
        v = None
        step_in = None
        for empty_v in range(1, nv + 1):
            found = None
            for s in steps:
                if (s[0], s[1], empty_v) not in transitions:
                    found = s
                    break
            if found is not None:
                v = empty_v
                step_in = inv_steps[found]
                break
        if v is None:
            break

        tmap = {}
        constraints = np.zeros(len(steps))
        for s_out in steps:
            target = _vd().vertex_apply_change(leg_states, dims, v, step_in, s_out, index)
            tmap[s_out] = target
            constraints[step_idx[s_out] - 1] = weights[target - 1] if target is not None else 0.0

        solution = _solve_lp(cost, A, constraints, lp_tolerance)

        for s_in in steps:
            if tmap[s_in] is None:
                continue
            c = constraints[step_idx[s_in] - 1]
            norm = 1.0 if c == 0 else c
            offset = len(cumprobs) + 1
            length = -1
            for s_out in steps:
                var = var_lookup[(s_in, s_out)]
                x = solution[var]
                prob = x / norm
                if not x >= -lp_tolerance:
                    raise AssertionError(f"LP solution below bound: {x}")
                if prob < 0.0:
                    used_inaccurate = prob < -tolerance
                    prob = 0.0
                if not x <= norm + lp_tolerance:
                    raise AssertionError(f"LP solution above bound: {x} > {norm}")
                if prob > 1.0:
                    used_inaccurate = prob > 1 + tolerance
                    prob = 1.0
                if prob > tolerance / len(steps):
                    cumprobs.append(prob)
                    vc = variables[var]
                    in_inv = inverse(vc)[0] if vc[0] == s_in else vc[0]
                    t = _vd().wrap_vertex_idx(leg_states, tmap[in_inv])
                    if t is None:
                        raise AssertionError("transition to invalid vertex")
                    targets.append(t)
                    step_outs.append(s_out)
                    length += 1
            transitions[(s_in[0], s_in[1], tmap[s_in])] = (offset, length)
            acc = 0.0
            for k in range(offset - 1, offset + length):
                acc = acc + cumprobs[k]
                cumprobs[k] = acc
            if length >= 0 and abs(cumprobs[offset + length - 1] - 1) > tolerance:
                warnings.warn(f"normalization error: {cumprobs[offset + length - 1]} != 1")

    if used_inaccurate:
        warnings.warn("had to truncate some probabilities in a possibly inaccurate way!")
    return transitions, cumprobs, targets, step_outs

// Abstract-loop transition tables of one bond type, built on the host in C++:
// the linear programs of src/vertex_data.jl:213-416 solved by a dense
// two-phase simplex.  Host code only (no device calls); included by capi.hip
// for the sse_vertex_transitions entry point.
//
// The LPs are small (one row per worm step: 4 for spin-1/2 bonds, 12 for two
// dim-4 sites; one column per (step in, step out) pair up to inversion), so a
// dense tableau with Bland's pivoting rule (no cycling) is exact enough and
// takes microseconds.  Every class of vertices connected by one worm step has
// its own right-hand side (the weights of the vertices it reaches); the
// objective (the number of bounces, steps that return along their own
// inverse) and the constraint matrix are shared.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace ssevd {

enum { kLpOptimal = 0, kLpInfeasible = 1, kLpUnbounded = 2, kLpIterations = 3 };

// min c.x subject to A x = b, x >= 0 (A: m x n row-major, b >= 0).
// On kLpOptimal x holds an optimal basic solution; *alt_optima is set when a
// non-basic column has zero reduced cost at the optimum (another optimal
// vertex may exist, so a different solver may return a different table).
inline int lp_solve(int m, int n, const double *A, const double *b, const double *c, double tol,
                    double *x, bool *alt_optima)
{
    const int W = n + m + 1;  // columns: x (n), artificials (m), rhs
    std::vector<double> T((size_t)(m + 1) * W, 0.0);
    std::vector<int> basis(m);
    std::vector<char> dead(m, 0);
    auto at = [&](int r, int col) -> double & { return T[(size_t)r * W + col]; };
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < n; ++j)
            at(i, j) = A[(size_t)i * n + j];
        at(i, n + i) = 1.0;
        at(i, W - 1) = b[i];
        basis[i] = n + i;
    }
    auto pivot = [&](int pr, int pc) {
        const double p = at(pr, pc);
        for (int col = 0; col < W; ++col)
            at(pr, col) /= p;
        for (int r = 0; r <= m; ++r) {
            if (r == pr)
                continue;
            const double f = at(r, pc);
            if (f != 0.0)
                for (int col = 0; col < W; ++col)
                    at(r, col) -= f * at(pr, col);
        }
        basis[pr] = pc;
    };
    // Bland: entering = lowest column with negative reduced cost; leaving =
    // minimum ratio, ties to the lowest basic column.
    auto iterate = [&](int ncols) -> int {
        for (int it = 0; it < 50000; ++it) {
            int pc = -1;
            for (int j = 0; j < ncols; ++j)
                if (at(m, j) < -tol) {
                    pc = j;
                    break;
                }
            if (pc < 0)
                return kLpOptimal;
            int pr = -1;
            double best = 0.0;
            for (int i = 0; i < m; ++i) {
                if (dead[i] || at(i, pc) <= tol)
                    continue;
                const double ratio = at(i, W - 1) / at(i, pc);
                if (pr < 0 || ratio < best - tol ||
                    (ratio <= best + tol && basis[i] < basis[pr])) {
                    pr = i;
                    best = ratio;
                }
            }
            if (pr < 0)
                return kLpUnbounded;
            pivot(pr, pc);
        }
        return kLpIterations;
    };
    // phase 1: minimise the sum of the artificials
    for (int col = 0; col < W; ++col) {
        double s = 0.0;
        for (int i = 0; i < m; ++i)
            s += at(i, col);
        at(m, col) = (col >= n && col < n + m) ? 0.0 : -s;
    }
    int rc = iterate(n + m);
    if (rc != kLpOptimal)
        return rc;
    double bsum = 0.0;
    for (int i = 0; i < m; ++i)
        bsum += std::fabs(b[i]);
    if (-at(m, W - 1) > tol * (1.0 + bsum))
        return kLpInfeasible;
    // drive artificials out of the basis; a row with no structural entry is
    // redundant and drops out
    for (int i = 0; i < m; ++i) {
        if (basis[i] < n)
            continue;
        int pc = -1;
        for (int j = 0; j < n; ++j)
            if (std::fabs(at(i, j)) > tol) {
                pc = j;
                break;
            }
        if (pc >= 0)
            pivot(i, pc);
        else
            dead[i] = 1;
    }
    // phase 2: reduced costs of the structural columns
    for (int col = 0; col < W; ++col)
        at(m, col) = col < n ? c[col] : 0.0;
    for (int i = 0; i < m; ++i) {
        if (dead[i])
            continue;
        const double cb = c[basis[i]];
        if (cb != 0.0)
            for (int col = 0; col < W; ++col)
                at(m, col) -= cb * at(i, col);
    }
    rc = iterate(n);
    if (rc != kLpOptimal)
        return rc;
    for (int j = 0; j < n; ++j)
        x[j] = 0.0;
    std::vector<char> is_basic(n, 0);
    for (int i = 0; i < m; ++i)
        if (!dead[i] && basis[i] < n) {
            x[basis[i]] = at(i, W - 1);
            is_basic[basis[i]] = 1;
        }
    bool alt = false;
    for (int j = 0; j < n && !alt; ++j)
        alt = !is_basic[j] && std::fabs(at(m, j)) <= tol;
    *alt_optima = alt;
    return kLpOptimal;
}

struct Step {
    int leg, worm;  // 1-based
    bool operator==(const Step &o) const { return leg == o.leg && worm == o.worm; }
    bool operator<(const Step &o) const { return leg != o.leg ? leg < o.leg : worm < o.worm; }
};

// Output of build_transitions: the reference's Transition table
// (vertex_data.jl:13-28) with 1-based offsets into the exit lists.
struct Transitions {
    // [vertex-1][worm-1][leg-1]: 1-based offset (0 = no entry), exit count - 1
    std::vector<int32_t> offset, length;
    std::vector<double> cumprobs;
    std::vector<int32_t> targets;  // vertex codes diag | idx << 1 (1-based idx)
    std::vector<int32_t> out_leg, out_worm;
    bool inaccurate = false;     // a probability had to be clamped beyond tolerance
    bool norm_warning = false;   // an exit list does not sum to 1 within tolerance
    int64_t alt_optima = 0;      // LPs whose optimum may not be unique
};

enum { kVdOk = 0, kVdBadInput = 1, kVdLpFailed = 2, kVdBelowBound = 3, kVdAboveBound = 4,
       kVdInvalidTarget = 5 };

// Restates construct_transitions (src/vertex_data.jl:213-416) step for step:
// the same step, variable and class orders, the same normalisation,
// clamping, pruning (prob > tolerance / #steps) and cumulative sums.
inline int build_transitions(int ns, const int32_t *dims, int nv, const uint8_t *leg_states,
                             const double *weights, double tolerance, double lp_tolerance,
                             Transitions &out)
{
    if (ns < 1 || ns > 8 || nv < 0)
        return kVdBadInput;
    const int legs = 2 * ns;
    int max_worm = 0;
    for (int s = 0; s < ns; ++s) {
        if (dims[s] < 1 || dims[s] > 255)  // leg states are bytes
            return kVdBadInput;
        max_worm = dims[s] - 1 > max_worm ? dims[s] - 1 : max_worm;
    }
    auto dim_of_leg = [&](int leg) { return dims[(leg > ns ? leg - ns : leg) - 1]; };
    // steps in the reference order: worm outer, leg inner
    std::vector<Step> steps;
    std::map<Step, int> step_idx;  // 0-based
    for (int worm = 1; worm <= max_worm; ++worm)
        for (int leg = 1; leg <= legs; ++leg)
            if (worm <= dim_of_leg(leg) - 1) {
                step_idx[{leg, worm}] = (int)steps.size();
                steps.push_back({leg, worm});
            }
    const int S = (int)steps.size();
    out = Transitions();
    if (S == 0 || nv == 0)
        return kVdOk;  // no worm steps (all sites of dimension 1) or no vertices: empty tables
    auto inv_step = [&](const Step &s) { return Step{s.leg, dim_of_leg(s.leg) - s.worm}; };
    typedef std::pair<int, int> VC;  // (step in, step out) as step indices
    auto inverse = [&](const VC &v) {
        return VC(step_idx[inv_step(steps[v.second])], step_idx[inv_step(steps[v.first])]);
    };
    std::vector<VC> variables;
    std::map<VC, int> var_of;
    for (int a = 0; a < S; ++a)
        for (int b = 0; b < S; ++b) {
            VC v(a, b);
            if (!var_of.count(inverse(v))) {
                var_of[v] = (int)variables.size();
                variables.push_back(v);
            }
        }
    const int n = (int)variables.size();
    // variable of (in, out): the first variable equal to it or to its inverse
    std::vector<int> var_lookup((size_t)S * S, -1);
    for (int a = 0; a < S; ++a)
        for (int b = 0; b < S; ++b)
            for (int k = 0; k < n; ++k)
                if (variables[k] == VC(a, b) || inverse(variables[k]) == VC(a, b)) {
                    var_lookup[(size_t)a * S + b] = k;
                    break;
                }
    std::vector<double> cost(n), A((size_t)S * n, 0.0);
    for (int k = 0; k < n; ++k) {
        cost[k] = variables[k] == inverse(variables[k]) ? 1.0 : 0.0;
        for (int r = 0; r < S; ++r)
            if (variables[k].first == r || inverse(variables[k]).first == r)
                A[(size_t)r * n + k] = 1.0;
    }
    // leg-state lookup: first vertex with the given leg states
    std::map<std::vector<uint8_t>, int> vindex;
    for (int v = 0; v < nv; ++v) {
        std::vector<uint8_t> key(legs);
        for (int l = 0; l < legs; ++l)
            key[l] = leg_states[(size_t)l * nv + v];
        vindex.emplace(key, v + 1);
    }
    auto apply_change = [&](int v, const Step &sin, const Step &sout) -> int {
        std::vector<uint8_t> st(legs);
        for (int l = 0; l < legs; ++l)
            st[l] = leg_states[(size_t)l * nv + v - 1];
        const int di = dim_of_leg(sin.leg), dout = dim_of_leg(sout.leg);
        st[sin.leg - 1] = (uint8_t)((st[sin.leg - 1] + sin.worm - 1) % di + 1);
        st[sout.leg - 1] = (uint8_t)((st[sout.leg - 1] + sout.worm - 1) % dout + 1);
        auto it = vindex.find(st);
        return it == vindex.end() ? 0 : it->second;
    };
    auto wrap = [&](int v) -> int32_t {
        bool diag = true;
        for (int s = 0; s < ns; ++s)
            diag = diag && leg_states[(size_t)s * nv + v - 1] == leg_states[(size_t)(ns + s) * nv + v - 1];
        return (int32_t)((diag ? 1 : 0) | (v << 1));
    };
    out.offset.assign((size_t)nv * max_worm * legs, 0);
    out.length.assign((size_t)nv * max_worm * legs, 0);
    std::vector<char> present((size_t)nv * S, 0);  // [vertex-1][step]
    auto tix = [&](int v, const Step &s) { return ((size_t)(v - 1) * max_worm + s.worm - 1) * legs + s.leg - 1; };
    std::vector<int> tmap(S);
    std::vector<double> rhs(S), sol(n);
    for (;;) {
        int v = 0, found = -1;
        for (int cand = 1; cand <= nv && !v; ++cand)
            for (int s = 0; s < S; ++s)
                if (!present[(size_t)(cand - 1) * S + s]) {
                    v = cand;
                    found = s;
                    break;
                }
        if (!v)
            break;
        const Step step_in = inv_step(steps[found]);
        for (int so = 0; so < S; ++so) {
            tmap[so] = apply_change(v, step_in, steps[so]);
            rhs[so] = tmap[so] ? weights[tmap[so] - 1] : 0.0;
        }
        bool alt = false;
        if (lp_solve(S, n, A.data(), rhs.data(), cost.data(), lp_tolerance, sol.data(), &alt) !=
            kLpOptimal)
            return kVdLpFailed;
        out.alt_optima += alt;
        for (int si = 0; si < S; ++si) {
            if (!tmap[si])
                continue;
            const double cval = rhs[si];
            const double norm = cval == 0.0 ? 1.0 : cval;
            const int32_t offset = (int32_t)out.cumprobs.size() + 1;
            int32_t length = -1;
            for (int so = 0; so < S; ++so) {
                const int var = var_lookup[(size_t)si * S + so];
                const double xv = sol[var];
                double prob = xv / norm;
                if (!(xv >= -lp_tolerance))
                    return kVdBelowBound;
                if (prob < 0.0) {
                    out.inaccurate = prob < -tolerance;
                    prob = 0.0;
                }
                if (!(xv <= norm + lp_tolerance))
                    return kVdAboveBound;
                if (prob > 1.0) {
                    out.inaccurate = prob > 1 + tolerance;
                    prob = 1.0;
                }
                if (prob > tolerance / S) {
                    out.cumprobs.push_back(prob);
                    const VC &vc = variables[var];
                    const int in_inv = vc.first == si ? inverse(vc).first : vc.first;
                    if (!tmap[in_inv])
                        return kVdInvalidTarget;
                    out.targets.push_back(wrap(tmap[in_inv]));
                    out.out_leg.push_back(steps[so].leg);
                    out.out_worm.push_back(steps[so].worm);
                    ++length;
                }
            }
            const size_t t = tix(tmap[si], steps[si]);
            out.offset[t] = offset;
            out.length[t] = length;
            present[(size_t)(tmap[si] - 1) * S + si] = 1;
            double acc = 0.0;
            for (int32_t k = offset - 1; k < offset + length; ++k) {
                acc = acc + out.cumprobs[k];
                out.cumprobs[k] = acc;
            }
            if (length >= 0 && std::fabs(out.cumprobs[offset + length - 1] - 1) > tolerance)
                out.norm_warning = true;
        }
    }
    return kVdOk;
}

}  // namespace ssevd

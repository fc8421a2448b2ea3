// sse_lane.h -- one lane per walker: the sweep in the reference's order, slot
// by slot (diagonal update sse.jl:137-191, make_vertex_list! vertex_list.jl:15-54,
// worm_update sse.jl:193-303, measure sse.jl:70-87,305-376).  This is the
// plain GPU form: launched with one walker it is the one-walker-per-launch
// baseline; launched over a batch every lane runs its own chain.
#pragma once
#include "sse_common.h"
#include "../../include/ssegpu.h"

// Bonds act on nsb = m.nsb sites (SSEData{NSites}); an operator has
// legs = 2 * nsb legs, leg ids are legs * p + leg, legs 0..nsb-1 lower and
// nsb..2nsb-1 upper (site_of_leg(l) = l mod nsb, vertex_data.jl:190).

// Grow the cutoff if n >= L/2, reallocating the slab from the pool when the
// capacity is exceeded.  Returns false (walker untouched) on pool exhaustion.
__device__ bool lane_grow(DevWalker &w, DevPool *pool, int legs)
{
    if ((double)w.n >= (double)w.L * 0.5) {
        int64_t nl = grown_cutoff(w.L);
        if (nl > w.cap) {
            int64_t cap = slab_capacity(nl);
            char *s = pool_alloc(pool, slab_bytes(cap, legs));
            if (!s) {
                w.error = SSE_ERR_OOM;
                return false;
            }
            uint32_t *nops = (uint32_t *)s;
            for (int64_t p = 0; p < w.L; ++p)
                nops[p] = w.ops[p];
            w.ops = nops;
            w.links = (int32_t *)(s + (size_t)cap * 4);
            w.cap = cap;
        }
        for (int64_t p = w.L; p < nl; ++p)
            w.ops[p] = 0u;
        w.L = nl;
    }
    return true;
}

// compound state index of the bond's sites, the first site fastest (join_idx)
__device__ __forceinline__ int lane_join(const DevModel &m, const DevVD &vd, uint32_t b,
                                         const uint8_t *spins)
{
    const int32_t *bs = m.bond_sites + (size_t)m.nsb * b;
    int idx = 0, stride = 1;
    for (int i = 0; i < m.nsb; ++i) {
        idx += (spins[bs[i]] - 1) * stride;
        stride *= vd.dims[i];
    }
    return idx;
}

__device__ void lane_diagonal(DevWalker &w, const DevModel &m, LaneRng &rng)
{
    const double pm_raw = (double)m.n_bonds / w.T;
    const double pr_raw = w.T / (double)m.n_bonds;
    const int nsb = m.nsb;
    uint32_t *ops = w.ops;
    uint8_t *spins = w.spins;
    int64_t n = w.n;
    const int64_t L = w.L;
    for (int64_t p = 0; p < L; ++p) {
        uint32_t op = ops[p];
        if (op == 0u) {
            uint32_t b = (uint32_t)rng.range((uint64_t)m.n_bonds);
            const DevVD &vd = m.types[m.bond_type[b]];
            int32_t dv = vd.diag_vertex[lane_join(m, vd, b, spins)];
            double weight = dv < 0 ? 0.0 : vd.weights[dv >> 1];
            double p_make = pm_raw / (double)(L - n);
            if (rng.next_float() < p_make * weight) {
                ops[p] = op_make(b, (uint32_t)dv);
                n += 1;
            }
        } else {
            uint32_t b = op_bond(op);
            const DevVD &vd = m.types[m.bond_type[b]];
            uint32_t vc = op_vcode(op);
            if (vc & 1u) {
                double weight = vd.weights[vc >> 1];
                double p_remove = (double)(L - n + 1) * pr_raw;
                if (rng.next_float() * weight < p_remove) {
                    ops[p] = 0u;
                    n -= 1;
                }
            } else {
                const uint8_t *ls = vd.leg_states + 2 * nsb * (vc >> 1);
                const int32_t *bs = m.bond_sites + (size_t)nsb * b;
                for (int i = 0; i < nsb; ++i)
                    spins[bs[i]] = ls[nsb + i];
            }
        }
    }
    w.n = n;
}

__device__ void lane_vertex_list(DevWalker &w, const DevModel &m)
{
    const int nsb = m.nsb, legs = 2 * nsb;
    int32_t *links = w.links;
    const uint32_t *ops = w.ops;
    for (int64_t k = 0; k < legs * w.L; ++k)
        links[k] = -1;
    for (int i = 0; i < m.n_sites; ++i)
        w.first[i] = w.last[i] = -1;
    for (int64_t p = 0; p < w.L; ++p) {
        uint32_t op = ops[p];
        if (op == 0u)
            continue;
        const int32_t *bs = m.bond_sites + (size_t)nsb * op_bond(op);
        for (int s = 0; s < nsb; ++s) {
            int site = bs[s];
            int32_t lst = w.last[site];
            int32_t me = (int32_t)(legs * p + s);
            if (lst != -1) {
                links[lst] = me;
                links[me] = lst;
            } else {
                w.first[site] = me;
            }
            w.last[site] = (int32_t)(legs * p + nsb + s);
        }
    }
    for (int i = 0; i < m.n_sites; ++i) {
        int32_t f = w.first[i];
        if (f != -1) {
            links[f] = w.last[i];
            links[w.last[i]] = f;
        }
    }
}

// scatter (vertex_data.jl:106-125); returns false on fall-through
__device__ __forceinline__ bool lane_scatter(const DevVD &vd, int legs, uint32_t vcode, int leg_in,
                                             int worm_in, double r, int &leg_out, int &worm_out,
                                             uint32_t &nv)
{
    int k = ((int)(vcode >> 1) * vd.max_worm + (worm_in - 1)) * legs + leg_in;
    int32_t off = vd.toff[k];
    int32_t cnt = vd.tcnt[k];
    if (off < 0)
        return false;
    for (int e = off; e < off + cnt; ++e) {
        if (r < vd.cum[e]) {
            leg_out = vd.eleg[e];
            worm_out = vd.ewrm[e];
            nv = (uint32_t)vd.tgt[e];
            return true;
        }
    }
    return false;
}

__device__ int64_t lane_worm_traverse(DevWalker &w, const DevModel &m, LaneRng &rng, int64_t &scat)
{
    if (w.n == 0)
        return 0;
    const int nsb = m.nsb, legs = 2 * nsb;
    int64_t p0;
    int l0;
    for (;;) {
        p0 = (int64_t)rng.range((uint64_t)w.L);
        l0 = (int)rng.range((uint64_t)legs);
        if (w.links[legs * p0 + l0] >= 0)
            break;
    }
    int site0 = m.bond_sites[(size_t)nsb * op_bond(w.ops[p0]) + l0 % nsb];
    int wf0 = 1 + (int)rng.range((uint64_t)(m.site_dims[site0] - 1));
    int leg_in = l0, wf = wf0;
    int64_t p = p0, len = 1;
    for (;;) {
        uint32_t op = w.ops[p];
        uint32_t b = op_bond(op);
        const DevVD &vd = m.types[m.bond_type[b]];
        int leg_out, wf_out;
        uint32_t nv;
        double r = rng.next_float();
        scat += 1;
        if (!lane_scatter(vd, legs, op_vcode(op), leg_in, wf, r, leg_out, wf_out, nv)) {
            w.error = SSE_ERR_SCATTER;
            return len;
        }
        w.ops[p] = op_make(b, nv);
        int site_out = m.bond_sites[(size_t)nsb * b + leg_out % nsb];
        int dim_out = m.site_dims[site_out];
        if (p == p0 && leg_out == l0 && wf_out == dim_out - wf0)
            break;
        len += 1;
        wf = wf_out;
        int32_t nx = w.links[legs * p + leg_out];
        leg_in = nx % legs;
        p = nx / legs;
        if (p == p0 && leg_in == l0 && wf == wf0)
            break;
    }
    return len;
}

__device__ void lane_worm_update(DevWalker &w, const DevModel &m, LaneRng &rng, bool therm)
{
    const int legs = 2 * m.nsb;
    double total = 1.0;
    int64_t nworms = (int64_t)ceil(w.num_worms);
    int64_t scat = 0;
    for (int64_t k = 0; k < nworms && !w.error; ++k)
        total += (double)lane_worm_traverse(w, m, rng, scat);
    w.last_scatter = scat;
    worm_controller(w, total, therm);
    for (int i = 0; i < m.n_sites; ++i) {
        int32_t f = w.first[i];
        if (f < 0) {
            w.spins[i] = (uint8_t)(1 + rng.range(m.site_dims[i]));
        } else {
            uint32_t op = w.ops[f / legs];
            const DevVD &vd = m.types[m.bond_type[op_bond(op)]];
            w.spins[i] = vd.leg_states[legs * (op_vcode(op) >> 1) + (f % legs)];
        }
    }
}

// phases: 1 diagonal, 2 vertex list, 4 worm; phases==7 is a full sweep and
// advances the walker's sweep counter.  Walkers with an error flag are skipped.
// With budget non-null this launch is sweep `step` of a budgeted call: walker
// wi takes part only while step < budget[wi].
__global__ void lane_sweep_kernel(DevWalker *W, const DevModel *models, DevPool *pool, int w_begin,
                                  int w_end, int phases, const int32_t *budget, int step)
{
    int wi = w_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= w_end)
        return;
    if (budget && step >= budget[wi])
        return;
    DevWalker w = W[wi];
    // a slot whose sweep budget ran out (retire set) waits for its
    // retirement; it takes no further sweep
    if (w.error || !w.active || w.retire)
        return;
    const DevModel &m = models[w.model];
    bool therm = w.sweeps > w.thermalization;
    LaneRng rng;
    rng.init(w.key, w.ctr);
    if (phases & 1) {
        if (!lane_grow(w, pool, 2 * m.nsb)) {
            W[wi].error = w.error;
            flag_error(pool, w.error);
            return;
        }
        lane_diagonal(w, m, rng);
    }
    if (phases & 2)
        lane_vertex_list(w, m);
    if (phases & 4)
        lane_worm_update(w, m, rng, therm);
    w.ctr = rng.ctr;
    if (w.error)
        flag_error(pool, w.error);
    if ((phases & 7) == 7) {
        w.sweeps += 1;
        w.acc_L += w.L;
        w.acc_n += w.n;
        w.acc_W += w.last_scatter;
        w.acc_sweeps += 1;
        if (w.remaining > 0 && --w.remaining == 0)
            w.retire = 1;
    }
    W[wi] = w;
}

// Carlo.init! (sse.jl:47-60): spins, identity opstring, diagonal warm-up.
__global__ void lane_init_kernel(DevWalker *W, const DevModel *models, DevPool *pool, int w_begin,
                                 int w_end, int pending_only)
{
    int wi = w_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= w_end)
        return;
    DevWalker w = W[wi];
    if (pending_only && !w.pending)
        return;
    const DevModel &m = models[w.model];
    LaneRng rng;
    rng.init(w.key, 0);
    for (int i = 0; i < m.n_sites; ++i)
        w.spins[i] = (uint8_t)(1 + rng.range(m.site_dims[i]));
    for (int64_t p = 0; p < w.L; ++p)
        w.ops[p] = 0u;
    w.n = 0;
    for (int k = 0; k < w.warmup; ++k) {
        if (!lane_grow(w, pool, 2 * m.nsb)) {
            flag_error(pool, w.error);
            break;
        }
        lane_diagonal(w, m, rng);
    }
    w.ctr = rng.ctr;
    w.pending = 0;
    w.active = 1;
    W[wi] = w;
}

// Carlo.measure! for one walker, sequentially (sse.jl:70-87,305-376).
__device__ void measure_walker_serial(const DevWalker &w, const DevModel &m, double *out, uint8_t *st)
{
    const int nsb = m.nsb;
    int neg = 0;
    for (int64_t p = 0; p < w.L; ++p) {
        uint32_t op = w.ops[p];
        if (op != 0u)
            neg += m.types[m.bond_type[op_bond(op)]].signs[op_vcode(op) >> 1] < 0;
    }
    double sign = (neg & 1) ? -1.0 : 1.0;
    double nn = (double)w.n;
    double Nn = m.norm;
    out[0] = sign;
    out[1] = nn;
    out[2] = sign * nn;
    out[3] = sign * (nn * nn);
    out[4] = ((-sign) * (nn * w.T + m.energy_offset)) / Nn;
    const int N = m.n_sites, D = m.est_max_dim;
    for (int e = 0; e < m.n_est; ++e) {
        const double *coef = m.est_coef + (size_t)e * N;
        const double *mv = m.est_mval + (size_t)e * N * D;
        for (int i = 0; i < N; ++i)
            st[i] = w.spins[i];
        double tmp = 0.0;
        for (int i = 0; i < N; ++i)
            tmp += coef[i] * mv[i * D + st[i] - 1];
        double mag = tmp, absmag = fabs(tmp), t2 = tmp * tmp, mag2 = t2, mag4 = t2 * t2, en = 1.0;
        for (int64_t p = 0; p < w.L; ++p) {
            uint32_t op = w.ops[p];
            if (op == 0u)
                continue;
            uint32_t vc = op_vcode(op);
            if (!(vc & 1u)) {
                const uint32_t b = op_bond(op);
                const uint8_t *ls = m.types[m.bond_type[b]].leg_states + 2 * nsb * (vc >> 1);
                const int32_t *bs = m.bond_sites + (size_t)nsb * b;
                for (int l = 0; l < nsb; ++l) {
                    int site = bs[l];
                    tmp += coef[site] * (mv[site * D + ls[nsb + l] - 1] - mv[site * D + ls[l] - 1]);
                    st[site] = ls[nsb + l];
                }
            }
            mag += tmp;
            absmag += fabs(tmp);
            t2 = tmp * tmp;
            mag2 += t2;
            mag4 += t2 * t2;
            en += 1.0;
        }
        double norm = 1.0 / Nn;
        mag *= norm;
        absmag *= norm;
        mag2 *= norm * norm;
        mag4 *= (norm * norm) * (norm * norm);
        double *o = out + 5 + 5 * e;
        o[0] = sign * mag / en;
        o[1] = sign * absmag / en;
        o[2] = sign * mag2 / en;
        o[3] = sign * mag4 / en;
        double chi = 1.0 / w.T / (en + 1.0) / en * (mag * mag + mag2) * Nn;
        o[4] = sign * chi;
    }
}

// Accumulate one measurement into the per-walker sums.
__device__ __forceinline__ void accumulate(double *sum, double *sumsq, const double *x, int width)
{
    for (int j = 0; j < width; ++j) {
        sum[j] += x[j];
        sumsq[j] += x[j] * x[j];
    }
}

// mode 0: measure every walker (explicit Carlo.measure! call)
// mode 1: measure thermalized walkers (Carlo measures after a sweep once
//         sweeps > thermalization) and accumulate into sums
// budget/step as in lane_sweep_kernel: a budgeted call measures a walker
// only after the sweeps it took part in
__global__ void lane_measure_kernel(DevWalker *W, const DevModel *models, double *out, double *sum,
                                    double *sumsq, int64_t *count, double *wlf,
                                    uint8_t *scratch_spins, int max_sites, int n_walkers, int mode,
                                    int32_t *done, const int32_t *budget, int step)
{
    int wi = blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= n_walkers)
        return;
    if (budget && step >= budget[wi])
        return;
    const DevWalker &w = W[wi];
    // a pending slot's opstring and spins are not initialised yet
    if (w.error || w.pending || (mode == 1 && !w.active))
        return;
    const DevModel &m = models[w.model];
    double *o = out + w.meas_offset;
    if (mode == 0 || w.sweeps > w.thermalization) {
        measure_walker_serial(w, m, o, scratch_spins + (size_t)wi * max_sites);
        if (mode == 1) {
            accumulate(sum + w.meas_offset, sumsq + w.meas_offset, o, w.meas_width);
            count[w.acc_slot] += 1;
            worm_fraction_accumulate(w, wlf);
        }
    }
    if (mode == 1 && w.retire) {
        W[wi].retire = 0;
        W[wi].active = 0;
        done[wi] = 1;
    }
}

// zero the per-walker work counters (L, n, W sums and sweep count)
__global__ void reset_work_counters_kernel(DevWalker *W, int n_walkers)
{
    int wi = blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= n_walkers)
        return;
    W[wi].acc_L = 0;
    W[wi].acc_n = 0;
    W[wi].acc_W = 0;
    W[wi].acc_sweeps = 0;
}

// retire walkers whose sweep budget ran out (sweeps without measurement)
__global__ void retire_kernel(DevWalker *W, int n_walkers, int32_t *done)
{
    int wi = blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= n_walkers)
        return;
    if (W[wi].active && W[wi].retire) {
        W[wi].retire = 0;
        W[wi].active = 0;
        done[wi] = 1;
    }
}

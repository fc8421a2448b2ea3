/*
 * sse_oracle.c -- CPU restatement of the StochasticSeriesExpansion.jl (v0.1.2)
 * SSE Markov chain for one walker, in the reference's order of operations
 * and random draws.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * library never links it.  Build with -O2 -ffp-contract=off (no fast-math).
 *
 * Random numbers: one counter-based Philox4x32-10 stream per walker
 * (Salmon et al., SC'11).  64-bit word i of the stream with key k is
 *   block b = i/2 : ctr = (lo32 b, hi32 b, 0, 0), key = (lo32 k, hi32 k)
 *   word  = out0 | out1<<32 (i even)   or   out2 | out3<<32 (i odd).
 * rand() Float64 = (word >> 11) * 2^-53; rand(1:s) = Lemire's nearly
 * divisionless method on 64-bit words (Lemire 2019).  In validation mode the
 * reference consumes the same words through its rng argument.
 *
 * What pins each part (tests/): Philox against the Random123 known-answer
 * vectors and rocRAND's engine; the range sampler against its closed form
 * and a chi-square test; so_tanh / so_log within 4 ulp of libm; the vertex
 * list against the reference's golden fixture (test/test_vertex_list.jl);
 * the chain (init, diagonal update, worm update, growth, controller) by
 * worldline consistency after every sweep, checkpoint/resume equality and
 * exact diagonalisation of the reference's three ED jobs (|z| <= 4); the
 * measurement and estimators by the same ED comparison and the reference's
 * published bani2v2o8 results; the parallel-tempering exchange by closed
 * forms, an acceptance-rate check and ED per temperature.
 * Parity unpinned: the draw-by-draw trace of the Julia reference under its
 * own RNG (no Julia here; the PhiloxRNG shim contract is unverified), the
 * FP64 bits of Julia's tanh in the controller, and Carlo's exchange
 * schedule and its RNG use for parallel tempering.
 */
#include "sse_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ RNG */
void so_philox4x32_10(const uint32_t in[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint64_t so_philox_word(uint64_t key, uint64_t index)
{
    uint64_t b = index >> 1;
    uint32_t ctr[4] = {(uint32_t)b, (uint32_t)(b >> 32), 0u, 0u};
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    uint32_t o[4];
    so_philox4x32_10(ctr, k, o);
    if (index & 1)
        return (uint64_t)o[2] | ((uint64_t)o[3] << 32);
    return (uint64_t)o[0] | ((uint64_t)o[1] << 32);
}

double so_word_to_float(uint64_t word) { return (double)(word >> 11) * 0x1.0p-53; }

static uint64_t next_word(uint64_t key, uint64_t *ctr) { return so_philox_word(key, (*ctr)++); }

/* rand(1:s) - 1: Lemire's nearly divisionless range sampler on 64-bit words. */
uint64_t so_range(uint64_t key, uint64_t *ctr, uint64_t s)
{
    uint64_t x = next_word(key, ctr);
    unsigned __int128 m = (unsigned __int128)x * s;
    uint64_t l = (uint64_t)m;
    if (l < s) {
        uint64_t t = (0 - s) % s;
        while (l < t) {
            x = next_word(key, ctr);
            m = (unsigned __int128)x * s;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

static double next_float(uint64_t key, uint64_t *ctr) { return so_word_to_float(next_word(key, ctr)); }

/* -------------------------------------------------------------- tanh
 * The worm-number controller needs tanh (sse.jl:213).  GPU and oracle must
 * agree bit for bit, so both evaluate this same sequence of IEEE operations:
 * expm1 by Cody-Waite reduction x = k ln2 + r and the degree-16 Taylor
 * polynomial of expm1(r), |r| <= ln2/2; tanh(x) = e/(e+2), e = expm1(2|x|). */
static const double EXPM1_C[16] = {
    0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
    0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
    0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
    0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45};
static const double LN2_HI = 0x1.62e42fee00000p-1;
static const double LN2_LO = 0x1.a39ef35793c76p-33;
static const double INV_LN2 = 0x1.71547652b82fep+0;

double so_expm1(double x)
{
    if (x == 0.0)
        return x;
    double kf = rint(x * INV_LN2);
    double r = (x - kf * LN2_HI) - kf * LN2_LO;
    double p = EXPM1_C[15];
    for (int j = 14; j >= 0; --j)
        p = p * r + EXPM1_C[j];
    double em1 = p * r;
    int k = (int)kf;
    if (k == 0)
        return em1;
    double two_k = ldexp(1.0, k);
    return two_k * em1 + (two_k - 1.0);
}

double so_tanh(double x)
{
    double ax = fabs(x);
    if (!(ax < 22.0))
        return copysign(1.0, x);
    if (ax < 0x1p-28)
        return x;
    double e = so_expm1(2.0 * ax);
    double t = e / (e + 2.0);
    return copysign(t, x);
}

/* Natural logarithm for the exchange test of parallel tempering, restated
 * from the textbook series log(m) = 2 atanh(s) = 2 sum s^(2j+1)/(2j+1),
 * s = (m-1)/(m+1), after splitting x = m 2^e with m in (sqrt(1/2), sqrt(2)].
 * |s| <= 0.1716, so twelve terms leave < 1e-19 relative.  The GPU evaluates
 * the same operation sequence, so both sides decide identically.  x must be
 * positive, finite and normal.                                             */
double so_log(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, sizeof bits);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    uint64_t mb = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &mb, sizeof m);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 1.0 / 23.0;
    for (int k = 21; k >= 1; k -= 2)
        p = p * z + 1.0 / (double)k;
    double lm = 2.0 * s * p;
    double ef = (double)e;
    return ef * LN2_HI + (lm + ef * LN2_LO);
}

/* ------------------------------------------------------------ walker */
struct so_walker {
    const sse_model_desc *m;
    double T, target_wlf, atten, avg_worm_length, num_worms;
    int64_t n, L, cap;
    uint32_t *ops;
    uint8_t *state;
    int64_t *vert;          /* [legs*cap] linked leg id legs*p+leg, or -1 */
    int64_t *v_first, *v_last;
    uint64_t key, ctr;
    int64_t thermalization, sweeps, init_cutoff;
    int warmup;
    int64_t last_scatter;
    double last_total, last_wlf;
    int error;
    int64_t acc_L, acc_n, acc_W, acc_sweeps;
};

static inline uint32_t op_bond(uint32_t op) { return op >> SSE_OP_BOND_SHIFT; }
static inline uint32_t op_vcode(uint32_t op) { return (op >> 1) & ((1u << SSE_OP_VCODE_BITS) - 1); }
static inline uint32_t op_make(uint32_t bond, uint32_t vcode) { return 1u | (vcode << 1) | (bond << SSE_OP_BOND_SHIFT); }
static inline const sse_vertex_data_desc *bond_vd(const sse_model_desc *m, uint32_t b) { return &m->types[m->bond_type[b]]; }
/* Bonds act on NS = bond_n_sites sites (SSEData{NSites}); an operator has
 * leg_count = 2 NS legs, leg ids leg_count * p + leg, legs 0..NS-1 lower and
 * NS..2NS-1 upper; site_of_leg(l) = l mod NS (vertex_data.jl:190).          */
static inline int nsb_of(const sse_model_desc *m) { return m->bond_n_sites; }
static inline const int32_t *bond_sites_of(const sse_model_desc *m, uint32_t b)
{
    return m->bond_sites + (size_t)m->bond_n_sites * b;
}
/* join_idx over the bond's sites, the first site fastest (util.jl:3-13) */
static int bond_join(const sse_model_desc *m, const sse_vertex_data_desc *vd, uint32_t b,
                     const uint8_t *state)
{
    const int32_t *bs = bond_sites_of(m, b);
    int idx = 0, stride = 1;
    for (int i = 0; i < nsb_of(m); ++i) {
        idx += (state[bs[i]] - 1) * stride;
        stride *= vd->dims[i];
    }
    return idx;
}

so_walker *so_walker_create(const sse_model_desc *model, const sse_walker_desc *wd)
{
    so_walker *w = (so_walker *)calloc(1, sizeof(so_walker));
    w->m = model;
    w->T = wd->T;
    w->target_wlf = wd->target_worm_length_fraction;
    w->atten = wd->num_worms_attenuation_factor;
    w->avg_worm_length = 1.0;           /* sse.jl:36 */
    w->num_worms = wd->init_num_worms;  /* sse.jl:37 */
    w->key = wd->seed;
    w->thermalization = wd->thermalization;
    w->warmup = wd->diagonal_warmup_sweeps;
    w->init_cutoff = wd->init_opstring_cutoff;
    w->state = (uint8_t *)calloc((size_t)model->n_sites, 1);
    w->v_first = (int64_t *)malloc(sizeof(int64_t) * (size_t)model->n_sites);
    w->v_last = (int64_t *)malloc(sizeof(int64_t) * (size_t)model->n_sites);
    w->last_wlf = NAN;
    return w;
}

void so_walker_destroy(so_walker *w)
{
    if (!w)
        return;
    free(w->ops);
    free(w->vert);
    free(w->state);
    free(w->v_first);
    free(w->v_last);
    free(w);
}

static void ensure_cap(so_walker *w, int64_t L)
{
    if (L <= w->cap)
        return;
    int64_t cap = L;
    w->ops = (uint32_t *)realloc(w->ops, sizeof(uint32_t) * (size_t)cap);
    w->vert = (int64_t *)realloc(w->vert, sizeof(int64_t) * 2 * (size_t)nsb_of(w->m) * (size_t)cap);
    for (int64_t p = w->cap; p < cap; ++p)
        w->ops[p] = 0;
    w->cap = cap;
}

/* diagonal_update (sse.jl:137-191) */
static void diagonal_update(so_walker *w)
{
    const sse_model_desc *m = w->m;
    /* growth check, no draws (sse.jl:138-145) */
    if ((double)w->n >= (double)w->L * 0.5) {
        int64_t old = w->L;
        int64_t nl = (int64_t)floor((double)old * 1.5 + 100.0);
        ensure_cap(w, nl);
        for (int64_t p = old; p < nl; ++p)
            w->ops[p] = 0;
        w->L = nl;
    }
    double pm_raw = (double)m->n_bonds / w->T;   /* sse.jl:147 */
    double pr_raw = w->T / (double)m->n_bonds;   /* sse.jl:148 */
    for (int64_t p = 0; p < w->L; ++p) {
        uint32_t op = w->ops[p];
        if (op == 0) {
            uint32_t b = (uint32_t)so_range(w->key, &w->ctr, (uint64_t)m->n_bonds);  /* :152 */
            const sse_vertex_data_desc *vd = bond_vd(m, b);
            int32_t dv = vd->diagonal_vertices[bond_join(m, vd, b, w->state)]; /* :156-158 */
            double weight = dv < 0 ? 0.0 : vd->weights[dv >> 1];            /* :160-162 */
            double p_make = pm_raw / (double)(w->L - w->n);                   /* :164 */
            if (next_float(w->key, &w->ctr) < p_make * weight) {              /* :166 */
                w->ops[p] = op_make(b, (uint32_t)dv);
                w->n += 1;
            }
        } else {
            uint32_t b = op_bond(op);
            const sse_vertex_data_desc *vd = bond_vd(m, b);
            uint32_t vc = op_vcode(op);
            if (vc & 1) {
                double weight = vd->weights[vc >> 1];                          /* :175 */
                double p_remove = (double)(w->L - w->n + 1) * pr_raw;          /* :176-177 */
                if (next_float(w->key, &w->ctr) * weight < p_remove) {         /* :178 */
                    w->ops[p] = 0;
                    w->n -= 1;
                }
            } else {
                const int ns = nsb_of(m);                                      /* :183-187 */
                const uint8_t *ls = vd->leg_states + 2 * ns * (vc >> 1);
                const int32_t *bs = bond_sites_of(m, b);
                for (int i = 0; i < ns; ++i)
                    w->state[bs[i]] = ls[ns + i];
            }
        }
    }
}

/* make_vertex_list! (vertex_list.jl:15-54) for bonds of ns sites: legs
 * 0..ns-1 lower, ns..2ns-1 upper, leg id 2ns*p + leg */
void so_make_vertex_list(int32_t n_sites, int32_t ns, const int32_t *bond_sites,
                         const uint32_t *ops, int64_t L, int64_t *vert, int64_t *v_first,
                         int64_t *v_last)
{
    const int64_t legs = 2 * (int64_t)ns;
    for (int64_t k = 0; k < legs * L; ++k)
        vert[k] = -1;
    for (int i = 0; i < n_sites; ++i)
        v_first[i] = v_last[i] = -1;
    for (int64_t p = 0; p < L; ++p) {
        uint32_t op = ops[p];
        if (op == 0)
            continue;
        uint32_t b = op_bond(op);
        for (int s = 0; s < ns; ++s) {
            int site = bond_sites[(size_t)ns * b + s];
            int64_t last = v_last[site];
            if (last != -1) {
                vert[last] = legs * p + s;
                vert[legs * p + s] = last;
            } else {
                v_first[site] = legs * p + s;
            }
            v_last[site] = legs * p + ns + s;
        }
    }
    for (int i = 0; i < n_sites; ++i) {
        if (v_first[i] != -1) {
            vert[v_first[i]] = v_last[i];
            vert[v_last[i]] = v_first[i];
        }
    }
}

/* scatter (vertex_data.jl:106-125) */
static int scatter(const sse_vertex_data_desc *vd, uint32_t vcode, int leg_in, int worm_in,
                   double r, int *leg_out, int *worm_out, uint32_t *new_vc)
{
    int vi = (int)(vcode >> 1);
    int k = (vi * vd->max_worm_count + (worm_in - 1)) * 2 * vd->n_sites + leg_in;
    int32_t off = vd->transition_offset[k];
    int32_t cnt = vd->transition_count[k];
    if (off < 0)
        return SSE_ERR_SCATTER;
    for (int e = off; e < off + cnt; ++e) {
        if (r < vd->exit_cumprob[e]) {
            *leg_out = vd->exit_leg[e];
            *worm_out = vd->exit_worm[e];
            *new_vc = (uint32_t)vd->exit_target[e];
            return SSE_OK;
        }
    }
    return SSE_ERR_SCATTER;
}

/* worm_traverse! (sse.jl:233-303); returns worm_length, counts scatter calls */
static int64_t traverse_from(so_walker *w, int64_t p0, int l0, int wf0, int64_t *scatters);

static int64_t worm_traverse(so_walker *w, int64_t *scatters)
{
    const sse_model_desc *m = w->m;
    if (w->n == 0)
        return 0;
    const int ns = nsb_of(m), legs = 2 * ns;
    int64_t p0;
    int l0;
    for (;;) { /* sse.jl:241-247 */
        p0 = (int64_t)so_range(w->key, &w->ctr, (uint64_t)w->L);
        l0 = (int)so_range(w->key, &w->ctr, (uint64_t)legs);
        if (w->vert[legs * p0 + l0] >= 0)
            break;
    }
    uint32_t b0 = op_bond(w->ops[p0]);
    int site0 = bond_sites_of(m, b0)[l0 % ns];
    int wf0 = 1 + (int)so_range(w->key, &w->ctr, (uint64_t)(m->site_dims[site0] - 1)); /* :251 */
    return traverse_from(w, p0, l0, wf0, scatters);
}

/* worm_traverse! from the start (leg l0, worm wf0, operator p0), 0-based leg
 * and slot (sse.jl:253-303): one scatter per step until the worm closes. */
static int64_t traverse_from(so_walker *w, int64_t p0, int l0, int wf0, int64_t *scatters)
{
    const sse_model_desc *m = w->m;
    const int ns = nsb_of(m), legs = 2 * ns;
    int leg_in = l0, wf = wf0;
    int64_t p = p0;
    int64_t len = 1;
    for (;;) { /* sse.jl:274-300 */
        uint32_t op = w->ops[p];
        uint32_t b = op_bond(op);
        const sse_vertex_data_desc *vd = bond_vd(m, b);
        int leg_out, wf_out;
        uint32_t nv;
        double r = next_float(w->key, &w->ctr);
        *scatters += 1;
        if (scatter(vd, op_vcode(op), leg_in, wf, r, &leg_out, &wf_out, &nv) != SSE_OK) {
            w->error = SSE_ERR_SCATTER;
            return len;
        }
        w->ops[p] = op_make(b, nv);
        int dim_out = m->site_dims[bond_sites_of(m, b)[leg_out % ns]];
        if (p == p0 && leg_out == l0 && wf_out == dim_out - wf0)
            break;
        len += 1;
        wf = wf_out;
        int64_t nx = w->vert[legs * p + leg_out];
        leg_in = (int)(nx % legs);
        p = nx / legs;
        if (p == p0 && leg_in == l0 && wf == wf0)
            break;
    }
    return len;
}

/* worm_update (sse.jl:193-231) */
static void worm_update(so_walker *w, int thermalized)
{
    const sse_model_desc *m = w->m;
    double total = 1.0;
    int64_t nworms = (int64_t)ceil(w->num_worms);
    int64_t scat = 0;
    for (int64_t k = 0; k < nworms && !w->error; ++k)
        total += (double)worm_traverse(w, &scat);
    w->last_scatter = scat;
    w->last_total = total;
    w->last_wlf = NAN;
    if (thermalized && w->n != 0)
        w->last_wlf = total / (double)w->n;                                  /* :200-202 */
    double avg = total / ceil(w->num_worms);                                  /* :204 */
    if (!thermalized) {
        w->avg_worm_length += w->atten * (avg - w->avg_worm_length);          /* :206-207 */
        double target = w->target_wlf * (double)w->n / w->avg_worm_length;    /* :208-209 */
        w->num_worms += w->atten * (target - w->num_worms + 100.0 * so_tanh(target - w->num_worms));
        if (w->atten != 0.0) {                                                /* :214-216 */
            double hi = 1.0 + (double)w->n / 2.0;
            double x = w->num_worms;
            w->num_worms = x > hi ? hi : (x < 1.0 ? 1.0 : x);
        }
    }
    const int64_t legs = 2 * (int64_t)nsb_of(m);
    for (int i = 0; i < m->n_sites; ++i) {                                    /* :219-228 */
        int64_t f = w->v_first[i];
        if (f < 0) {
            w->state[i] = (uint8_t)(1 + so_range(w->key, &w->ctr, m->site_dims[i]));
        } else {
            uint32_t op = w->ops[f / legs];
            const sse_vertex_data_desc *vd = bond_vd(m, op_bond(op));
            w->state[i] = vd->leg_states[legs * (op_vcode(op) >> 1) + (f % legs)];
        }
    }
}

void so_init(so_walker *w)
{
    const sse_model_desc *m = w->m;
    for (int i = 0; i < m->n_sites; ++i)                                      /* sse.jl:48 */
        w->state[i] = (uint8_t)(1 + so_range(w->key, &w->ctr, m->site_dims[i]));
    int64_t L0 = w->init_cutoff;
    if (L0 < 0)
        L0 = (int64_t)nearbyint((double)m->n_sites * w->T);                  /* :50-51 */
    ensure_cap(w, L0 > 0 ? L0 : 1);
    for (int64_t p = 0; p < L0; ++p)
        w->ops[p] = 0;
    w->L = L0;
    w->n = 0;
    for (int k = 0; k < w->warmup; ++k)                                       /* :54-57 */
        diagonal_update(w);
}

int so_sweep_phases(so_walker *w, int phases)
{
    int thermalized = w->sweeps > w->thermalization;
    if (phases & 1)
        diagonal_update(w);
    if (phases & 2) {
        ensure_cap(w, w->L);
        so_make_vertex_list(w->m->n_sites, nsb_of(w->m), w->m->bond_sites, w->ops, w->L, w->vert,
                            w->v_first, w->v_last);
    }
    if (phases & 4)
        worm_update(w, thermalized);
    if (phases == 7) {
        w->sweeps += 1;
        w->acc_L += w->L;
        w->acc_n += w->n;
        w->acc_W += w->last_scatter;
        w->acc_sweeps += 1;
    }
    return w->error;
}

int so_sweep(so_walker *w) { return so_sweep_phases(w, 7); }

/* The reference's worm_traverse! test (test/test_sse.jl:30-60): build the
 * vertex list of the current operator string, then run one worm from (leg
 * l0, worm wf0, slot p0), scatters drawn from the walker's stream.  Returns
 * the worm length, or -1 for a start on an identity slot or an error. */
int64_t so_worm_traverse_from(so_walker *w, int64_t p0, int l0, int wf0)
{
    const int legs = 2 * nsb_of(w->m);
    if (p0 < 0 || p0 >= w->L || l0 < 0 || l0 >= legs || w->ops[p0] == 0)
        return -1;
    so_make_vertex_list(w->m->n_sites, nsb_of(w->m), w->m->bond_sites, w->ops, w->L, w->vert,
                        w->v_first, w->v_last);
    int64_t scat = 0;
    int64_t len = traverse_from(w, p0, l0, wf0, &scat);
    return w->error ? -1 : len;
}

int so_measure_width(const so_walker *w) { return 5 + 5 * w->m->n_estimators; }

/* Carlo.measure! (sse.jl:70-87), measure_sign (:305-314), fused estimator
 * loop (:321-376) with MagnetizationEstimator init/measure/result
 * (magnetization_estimator.jl:96-165,205-230). */
void so_measure(so_walker *w, double *out)
{
    const sse_model_desc *m = w->m;
    int neg = 0;
    for (int64_t p = 0; p < w->L; ++p) {
        uint32_t op = w->ops[p];
        if (op != 0)
            neg += bond_vd(m, op_bond(op))->signs[op_vcode(op) >> 1] < 0;
    }
    double sign = (neg & 1) ? -1.0 : 1.0;
    double nn = (double)w->n;
    double Nn = m->normalization_site_count;
    out[0] = sign;
    out[1] = nn;
    out[2] = sign * nn;
    out[3] = sign * (nn * nn);
    out[4] = ((-sign) * (nn * w->T + m->energy_offset)) / Nn;
    int E = m->n_estimators;
    if (E == 0)
        return;
    int N = m->n_sites, D = m->est_max_dim;
    uint8_t *st = (uint8_t *)malloc((size_t)N);
    for (int e = 0; e < E; ++e) {
        const double *coef = m->est_coef + (size_t)e * N;
        const double *mv = m->est_mval + (size_t)e * N * D;
        memcpy(st, w->state, (size_t)N);
        double tmp = 0.0;
        for (int i = 0; i < N; ++i)
            tmp += coef[i] * mv[i * D + st[i] - 1];
        double mag = tmp, absmag = fabs(tmp), t2 = tmp * tmp, mag2 = t2, mag4 = t2 * t2, en = 1.0;
        for (int64_t p = 0; p < w->L; ++p) {
            uint32_t op = w->ops[p];
            if (op == 0)
                continue;
            uint32_t b = op_bond(op), vc = op_vcode(op);
            if (!(vc & 1)) {
                const int ns = nsb_of(m);
                const uint8_t *ls = bond_vd(m, b)->leg_states + 2 * ns * (vc >> 1);
                const int32_t *bs = bond_sites_of(m, b);
                for (int l = 0; l < ns; ++l) {
                    int site = bs[l];
                    tmp += coef[site] * (mv[site * D + ls[ns + l] - 1] - mv[site * D + ls[l] - 1]);
                    st[site] = ls[ns + l];
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
        double chi = 1.0 / w->T / (en + 1.0) / en * (mag * mag + mag2) * Nn;
        o[4] = sign * chi;
    }
    free(st);
}

void so_get_scalars(const so_walker *w, sse_walker_scalars *s)
{
    s->cutoff = w->L;
    s->num_operators = w->n;
    s->avg_worm_length = w->avg_worm_length;
    s->num_worms = w->num_worms;
    s->rng_counter = w->ctr;
    s->sweeps = w->sweeps;
    s->last_scatter_calls = w->last_scatter;
    s->last_total_worm_length = w->last_total;
    s->last_worm_length_fraction = w->last_wlf;
}

int64_t so_get_ops(const so_walker *w, uint32_t *ops, int64_t cap)
{
    int64_t k = w->L < cap ? w->L : cap;
    memcpy(ops, w->ops, sizeof(uint32_t) * (size_t)k);
    return w->L;
}

void so_get_spins(const so_walker *w, uint8_t *spins) { memcpy(spins, w->state, (size_t)w->m->n_sites); }

int so_set_state(so_walker *w, const sse_walker_scalars *s, const uint32_t *ops, const uint8_t *spins)
{
    ensure_cap(w, s->cutoff > 0 ? s->cutoff : 1);
    w->L = s->cutoff;
    memcpy(w->ops, ops, sizeof(uint32_t) * (size_t)w->L);
    memcpy(w->state, spins, (size_t)w->m->n_sites);
    int64_t n = 0;
    for (int64_t p = 0; p < w->L; ++p)
        n += ops[p] != 0;
    w->n = n;
    w->avg_worm_length = s->avg_worm_length;
    w->num_worms = s->num_worms;
    w->ctr = s->rng_counter;
    w->sweeps = s->sweeps;
    return n == s->num_operators ? 0 : SSE_ERR_INVALID;
}

void so_get_vertex_list(const so_walker *w, int64_t *links, int64_t *first, int64_t *last)
{
    memcpy(links, w->vert, sizeof(int64_t) * 2 * (size_t)nsb_of(w->m) * (size_t)w->L);
    memcpy(first, w->v_first, sizeof(int64_t) * (size_t)w->m->n_sites);
    memcpy(last, w->v_last, sizeof(int64_t) * (size_t)w->m->n_sites);
}

double so_run_timed(so_walker *w, int64_t n_sweeps, int measure, double *sums)
{
    int width = so_measure_width(w);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)width);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int64_t k = 0; k < n_sweeps; ++k) {
        so_sweep(w);
        if (measure && w->sweeps > w->thermalization) {
            so_measure(w, tmp);
            if (sums)
                for (int j = 0; j < width; ++j)
                    sums[j] += tmp[j];
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(tmp);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* Run n sweeps; after each sweep of a thermalized walker (Carlo rule
 * sweeps > thermalization) write one measurement row into out[k*width..].
 * Returns the number of rows written.                                   */
int64_t so_run_series(so_walker *w, int64_t n_sweeps, double *out)
{
    int width = so_measure_width(w);
    int64_t rows = 0;
    for (int64_t k = 0; k < n_sweeps; ++k) {
        if (so_sweep(w))
            return -1;
        if (w->sweeps > w->thermalization) {
            so_measure(w, out + rows * width);
            rows += 1;
        }
    }
    return rows;
}

/* Sums of cutoff L, operator count n and scatter calls W over the sweeps
 * since the last call (counters from the chain itself); resets them.     */
void so_work_counters(so_walker *w, int64_t *out4)
{
    out4[0] = w->acc_L;
    out4[1] = w->acc_n;
    out4[2] = w->acc_W;
    out4[3] = w->acc_sweeps;
    w->acc_L = w->acc_n = w->acc_W = w->acc_sweeps = 0;
}

/* ------------------------------------------------------------ parallel tempering */
/* Carlo.parallel_tempering_log_weight_ratio (sse.jl:390-396): moving this
 * walker's configuration from T to new_T changes its log weight by
 * -n log(new_T / T).                                                       */
double so_pt_log_weight_ratio(const so_walker *w, double new_T)
{
    return -(double)w->n * so_log(new_T / w->T);
}

/* Carlo.parallel_tempering_change_parameter! (sse.jl:398-405): only T
 * changes; the configuration and the worm controller stay.                */
void so_pt_change_parameter(so_walker *w, double new_T) { w->T = new_T; }

double so_pt_get_T(const so_walker *w) { return w->T; }

/* Uniform of the exchange test of pair (slot, slot+1) of one ladder in one
 * round: Philox4x32-10 block (round lo, round hi, slot | 2^31, ladder) under
 * the ladder seed; words 0|1 -> (w >> 11) 2^-53.  The set bit 31 keeps the
 * block apart from every walker stream (whose third word is 0).            */
double so_pt_uniform(uint64_t seed, int32_t ladder, int64_t round, int32_t slot)
{
    uint32_t ctr[4] = {(uint32_t)round, (uint32_t)((uint64_t)round >> 32),
                       (uint32_t)slot | 0x80000000u, (uint32_t)ladder};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    so_philox4x32_10(ctr, k, o);
    return so_word_to_float((uint64_t)o[0] | ((uint64_t)o[1] << 32));
}

/* One replica-exchange round (Hukushima-Nemoto, neighbour pairs of one
 * parity) over one temperature ladder, the exchange step Carlo drives
 * through the two hooks above.  slot_walker[k] is the walker at temperature
 * slot k.  Round r proposes the pairs (k, k+1) with k % 2 == r % 2; the
 * swap is accepted iff log(u) < lwr(a -> T_b) + lwr(b -> T_a), i.e. with
 * probability min(1, exp(sum)); an accepted swap exchanges the walkers'
 * temperatures (change_parameter on both) and their slots.  accepted and
 * proposed count per lower slot.  Walkers in an error state are skipped.   */
void so_pt_exchange(so_walker *const *walkers, int32_t *slot_walker, int32_t len, uint64_t seed,
                    int32_t ladder, int64_t round, int64_t *accepted, int64_t *proposed)
{
    for (int32_t k = (int32_t)(round & 1); k + 1 < len; k += 2) {
        so_walker *a = walkers[slot_walker[k]];
        so_walker *b = walkers[slot_walker[k + 1]];
        if (a->error || b->error)
            continue;
        double Ta = a->T, Tb = b->T;
        double d = so_pt_log_weight_ratio(a, Tb) + so_pt_log_weight_ratio(b, Ta);
        double u = so_pt_uniform(seed, ladder, round, k);
        int acc = d >= 0.0 || u == 0.0 || so_log(u) < d;
        proposed[k] += 1;
        if (acc) {
            accepted[k] += 1;
            so_pt_change_parameter(a, Tb);
            so_pt_change_parameter(b, Ta);
            int32_t t = slot_walker[k];
            slot_walker[k] = slot_walker[k + 1];
            slot_walker[k + 1] = t;
        }
    }
}

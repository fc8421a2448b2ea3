// sse_wave_n.h -- the wave kernels (one wavefront per walker) for bonds of
// any site count NS = 1..4 (SSEData{NSites}, sse_data.jl:15-42; vertices of
// 2*NS legs, vertex_list.jl:24-62).
//
// The two-site kernels of sse_wave.h pack four legs into their records and
// chase tables; these kernels take the site count as a template parameter and
// keep the same structure where it carries over:
//   K1n  diagonal update + vertex list, 64 slots per step (as K1: the same
//        draws, acceptance windows and slot-by-slot fallback), with one
//        per-site lane mask per bond position;
//   K2n  worm update, a wave-uniform chase: the chase codes in LDS (2 bytes
//        per operator), one chase-table entry per step (DevModel::gtab,
//        integer thresholds; the exact exit search of vertex_data.jl:106-125
//        where the entry says so), the operator's links read from global
//        memory by one vector load while the exit is picked; walkers whose
//        codes do not fit the launch's LDS chase with every table read from
//        global memory;
//   K3n  measurement (sign and the magnetization estimators), 64 slots per step.
// Compact records of a walker with PL = 2*NS rounded up to a power of two:
//   vt[cap]      vertex code | type << 10 of the operator of each rank
//   link[PL*cap] leg id PL*rank + leg of the leg each leg is linked to
//   mask, base   per 64-slot chunk: non-identity bits, rank of the first
// (for NS = 2 this is the layout of sse_wave.h).
// Included by sse_wave.h after the two-site kernels it builds on (K1's
// stream and acceptance code, the worm controller, the measurement sums) and
// before the launch helpers that dispatch to it.
#pragma once

template <int NS>
struct LegsOf {
    static constexpr int LEGS = 2 * NS;
    static constexpr int LG = NS == 1 ? 1 : NS == 2 ? 2 : 3;  // log2 of the padded leg count
    static constexpr int PL = 1 << LG;
};

__device__ __forceinline__ uint32_t *recn_vt(const DevWalker &w) { return (uint32_t *)w.links; }
__device__ __forceinline__ uint32_t *recn_link(const DevWalker &w) { return (uint32_t *)w.links + w.cap; }
template <int PL>
__device__ __forceinline__ uint64_t *recn_mask(const DevWalker &w)
{
    return (uint64_t *)((uint32_t *)w.links + (1 + PL) * w.cap);
}
template <int PL>
__device__ __forceinline__ uint32_t *recn_base(const DevWalker &w)
{
    return (uint32_t *)(recn_mask<PL>(w) + (w.cap / 64 + 1));
}

// LDS of K1n (one walker): Philox words u64[256], lane masks u64[NS][N],
// first/last leg i32[N] each, upper leg states of the chunk's operators
// u32[64] (one byte per bond position), spins u8[N]
__host__ __device__ inline size_t k1n_lds_bytes(int N, int NS)
{
    size_t b = 2048 + 8 * (size_t)NS * N + 8 * (size_t)N + 256 + (size_t)N;
    return (b + 15) & ~(size_t)15;
}

// upper (after the operator) states of vertex v, one byte per bond position
template <int NS>
__device__ __forceinline__ uint32_t k1n_upper(const DevVD &vd, uint32_t v)
{
    const uint8_t *ls = vd.leg_states + 2 * NS * v + NS;
    uint32_t u = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
        u |= (uint32_t)ls[k] << (8 * k);
    return u;
}

// Slot-by-slot processing of one chunk in reference order (Lemire rejection
// in a bond draw, or forced); emits the same records as the parallel path.
template <int NS>
__device__ void k1n_serial_chunk(DevWalker &w, const DevModel &m, WaveStream &rs, int64_t base,
                                 int64_t &n, int64_t &nr, double pm_raw, double pr_raw,
                                 bool do_diag, bool do_vl, uint8_t *sp, int32_t *vfirst,
                                 int32_t *vlast)
{
    constexpr int PL = LegsOf<NS>::PL;
    const int lane = threadIdx.x;
    const int64_t L = w.L;
    uint32_t *gvt = recn_vt(w);
    uint32_t *glk = recn_link(w);
    const int64_t nr0 = nr;
    uint64_t cmask = 0ull;
    for (int i = 0; i < 64 && base + i < L; ++i) {
        const int64_t p = base + i;
        uint32_t op = w.ops[p];
        uint32_t nop = op;
        if (do_diag) {
            if (op == 0u) {
                uint32_t b = (uint32_t)rs.range((uint64_t)m.n_bonds);
                const int32_t *bs = m.bond_sites + (size_t)NS * b;
                const DevVD &vd = m.types[m.bond_type[b]];
                int idx = 0, stride = 1;
                for (int k = 0; k < NS; ++k) {
                    idx += (sp[bs[k]] - 1) * stride;
                    stride *= vd.dims[k];
                }
                int32_t dv = vd.diag_vertex[idx];
                double wgt = dv < 0 ? 0.0 : vd.weights[dv >> 1];
                double p_make = pm_raw / (double)(L - n);
                if (rs.next_float() < p_make * wgt) {
                    nop = op_make(b, (uint32_t)dv);
                    n += 1;
                }
            } else {
                uint32_t vc = op_vcode(op);
                uint32_t b = op_bond(op);
                const DevVD &vd = m.types[m.bond_type[b]];
                if (vc & 1u) {
                    double wgt = vd.weights[vc >> 1];
                    double p_remove = (double)(L - n + 1) * pr_raw;
                    if (rs.next_float() * wgt < p_remove) {
                        nop = 0u;
                        n -= 1;
                    }
                } else {
                    const uint32_t up = k1n_upper<NS>(vd, vc >> 1);
                    __syncthreads();
                    if (lane == 0)
                        for (int k = 0; k < NS; ++k)
                            sp[m.bond_sites[(size_t)NS * b + k]] = (uint8_t)(up >> (8 * k));
                    __syncthreads();
                }
            }
        }
        if (do_vl && nop != 0u) {
            const uint32_t b = op_bond(nop);
            const int32_t rk = (int32_t)nr;
            for (int k = 0; k < NS; ++k) {
                const int s = m.bond_sites[(size_t)NS * b + k];
                const int32_t me = PL * rk + k;
                const int32_t lst = vlast[s];
                __syncthreads();
                if (lane == 0) {
                    if (lst >= 0) {
                        glk[lst] = (uint32_t)me;
                        glk[me] = (uint32_t)lst;
                    } else {
                        vfirst[s] = me;
                    }
                    vlast[s] = PL * rk + NS + k;
                }
                __syncthreads();
            }
            if (lane == 0)
                gvt[rk] = op_vcode(nop) | ((uint32_t)m.bond_type[b] << 10);
            cmask |= 1ull << i;
            nr += 1;
        }
        if (lane == 0 && nop != op)
            w.ops[p] = nop;
    }
    if (do_vl && lane == 0) {
        recn_mask<PL>(w)[base >> 6] = cmask;
        recn_base<PL>(w)[base >> 6] = (uint32_t)nr0;
    }
    __syncthreads();
}

// Model lookups of K1n: from the LDS copy of DevModel::k1ntab (N1Staged) or
// from the model's global arrays (N1Global).  Same values either way.
template <int NS>
struct N1Global {
    const DevModel *m;
    __device__ __forceinline__ void stage(char *) const {}
    __device__ __forceinline__ int type(uint32_t b) const { return m->bond_type[b]; }
    __device__ __forceinline__ int site(uint32_t b, int k) const { return m->bond_sites[(size_t)NS * b + k]; }
    __device__ __forceinline__ int nv(int t) const { return m->types[t].nv; }
    __device__ __forceinline__ int dim(int t, int j) const { return m->types[t].dims[j]; }
    __device__ __forceinline__ int diag(int t, int i) const { return m->types[t].diag_vertex[i]; }
    __device__ __forceinline__ double weight(int t, int v) const { return m->types[t].weights[v]; }
    __device__ __forceinline__ uint32_t upper(int t, int v) const { return k1n_upper<NS>(m->types[t], (uint32_t)v); }
};
template <int NS>
struct N1Staged {
    const DevModel *m;
    const char *t;  // the LDS copy of m->k1ntab
    // copy the table into LDS (whole wave; the caller's barrier follows)
    __device__ __forceinline__ void stage(char *dst) const
    {
        const uint32_t *g = (const uint32_t *)m->k1ntab;
        uint32_t *d = (uint32_t *)dst;
        for (int i = threadIdx.x; i < m->k1ntab_bytes / 4; i += 64)
            d[i] = g[i];
    }
    __device__ __forceinline__ const uint32_t *hdr(int ty) const { return (const uint32_t *)(t + 32 * ty); }
    __device__ __forceinline__ const int32_t *bond(uint32_t b) const
    {
        return (const int32_t *)(t + m->k1n_bond_off) + (size_t)(NS + 1) * b;
    }
    __device__ __forceinline__ int type(uint32_t b) const { return bond(b)[0]; }
    __device__ __forceinline__ int site(uint32_t b, int k) const { return bond(b)[1 + k]; }
    __device__ __forceinline__ int nv(int ty) const { return (int)hdr(ty)[3]; }
    __device__ __forceinline__ int dim(int ty, int j) const { return (int)((hdr(ty)[4] >> (8 * j)) & 255u); }
    __device__ __forceinline__ int diag(int ty, int i) const { return ((const int32_t *)(t + hdr(ty)[1]))[i]; }
    __device__ __forceinline__ double weight(int ty, int v) const
    {
        return ((const double *)(t + hdr(ty)[0]))[v];
    }
    __device__ __forceinline__ uint32_t upper(int ty, int v) const
    {
        return ((const uint32_t *)(t + hdr(ty)[2]))[v];
    }
};

// LDS bytes of K1n with the model's lookups staged
__host__ __device__ inline size_t k1n_lds_staged(int N, int NS, int k1ntab_bytes)
{
    return k1n_lds_bytes(N, NS) + (((size_t)k1ntab_bytes + 15) & ~(size_t)15);
}

// K1n of walker wi: diagonal update (sse.jl:137-191) and vertex list
// (vertex_list.jl:15-54), 64 slots per step.  phases: 1 diagonal update,
// 2 vertex list, 8 force the slot-by-slot path; warm >= 0: init warm-up
// round r (diagonal only, pending walkers with more warm-up sweeps).
template <int NS, class Tab>
__device__ void k1n_walker(DevWalker *W, const DevModel *models, DevPool *pool, int wi, int phases,
                           int warm, char *lds, int grp, Tab tab)
{
    constexpr int LEGS = LegsOf<NS>::LEGS, PL = LegsOf<NS>::PL;
    const int lane = threadIdx.x;
    DevWalker w = W[wi];
    if (w.error)
        return;
    if (warm >= 0 ? (warm >= w.warmup || !w.pending) : (!w.active || w.retire))
        return;
    const DevModel &m = models[w.model];
    const int N = m.n_sites;
    uint64_t *rbuf = (uint64_t *)lds;
    uint64_t *tm = (uint64_t *)(lds + 2048);  // tm[k * N + site]
    int32_t *vfirst = (int32_t *)(tm + (size_t)NS * N);
    int32_t *vlast = vfirst + N;
    uint32_t *cups = (uint32_t *)(vlast + N);
    uint8_t *sp = (uint8_t *)(cups + 64);
    uint32_t *tmw = (uint32_t *)tm;

    const bool do_diag = phases & 1;
    const bool do_vl = (phases & 2) && warm < 0;
    const bool force_serial = phases & 8;

    tab.stage(lds + k1n_lds_bytes(N, NS));
    for (int i = lane; i < NS * N; i += 64)
        tm[i] = 0ull;
    for (int i = lane; i < N; i += 64) {
        vfirst[i] = -1;
        vlast[i] = -1;
        sp[i] = w.spins[i];
    }
    // growth check, no draws (sse.jl:138-145)
    if (do_diag && (double)w.n >= (double)w.L * 0.5) {
        int64_t nl = grown_cutoff(w.L);
        if (nl > w.cap) {
            int64_t cap = slab_capacity(nl);
            unsigned long long sp_addr = 0;
            if (lane == 0)
                sp_addr = (unsigned long long)pool_alloc(pool, slab_bytes(cap, LEGS));
            uint32_t lo = __shfl((uint32_t)sp_addr, 0), hi = __shfl((uint32_t)(sp_addr >> 32), 0);
            char *s = (char *)(((unsigned long long)hi << 32) | lo);
            if (!s) {
                if (lane == 0) {
                    W[wi].error = SSE_ERR_OOM;
                    flag_error(pool, SSE_ERR_OOM);
                }
                return;
            }
            uint32_t *nops = (uint32_t *)s;
            for (int64_t p = lane; p < w.L; p += 64)
                nops[p] = w.ops[p];
            w.ops = nops;
            w.links = (int32_t *)(s + (size_t)cap * 4);
            w.cap = cap;
        }
        for (int64_t p = w.L + lane; p < nl; p += 64)
            w.ops[p] = 0u;
        w.L = nl;
    }
    __syncthreads();

    WaveStream rs;
    rs.init(rbuf, w.key, w.ctr);
    const double pm_raw = (double)m.n_bonds / w.T;
    const double pr_raw = w.T / (double)m.n_bonds;
    const uint64_t Nb = (uint64_t)m.n_bonds;
    const int64_t L = w.L;
    int64_t n0 = w.n;
    int64_t nr = 0;
    const uint64_t below = below_mask(lane), above = above_mask(lane);
    gmem<uint32_t> *gvt = as_global(recn_vt(w));
    gmem<uint32_t> *glk = as_global(recn_link(w));
    gmem<uint64_t> *gmask = as_global(recn_mask<PL>(w));
    gmem<uint32_t> *gbase = as_global(recn_base<PL>(w));
    gmem<uint32_t> *ops = as_global(w.ops);

    // read once: a read of the model inside the loop would wait for the
    // chunk's global stores
    const uint32_t nbonds = (uint32_t)m.n_bonds;
    uint32_t op_next = (int64_t)lane < L ? ops[lane] : 0u;
    for (int64_t base = 0; base < L; base += 64) {
        const int64_t p = base + lane;
        const bool valid = p < L;
        const uint32_t op = op_next;
        op_next = p + 64 < L ? ops[p + 64] : 0u;
        const uint32_t vc = op_vcode(op);
        const bool ident = valid && op == 0u;
        const bool isop = valid && op != 0u;
        const bool dia = isop && (vc & 1u);
        const bool offd = isop && !(vc & 1u);
        const uint64_t identm = __ballot(ident), diam = __ballot(dia), offm = __ballot(offd);
        uint32_t b = op_bond(op);
        double u = 0.0;
        {
            bool bad = false;
            if (isop) {
                bad = b >= nbonds;
                if (!bad)
                    bad = (int)(vc >> 1) >= tab.nv(tab.type(b));
            }
            if (__ballot(bad)) {
                if (lane == 0) {
                    W[wi].error = SSE_ERR_STATE;
                    flag_error(pool, SSE_ERR_STATE);
                }
                return;
            }
        }

        if (do_diag) {
            const int D = 2 * popc(identm) + popc(diam);
            const int off = 2 * popc(identm & below) + popc(diam & below);
            bool reject = force_serial;
            if (D > 0) {
                rs.ensure((uint32_t)D);
                if (ident) {
                    uint64_t x = rs.peek((uint32_t)off);
                    uint64_t lo = x * Nb;
                    b = (uint32_t)__umul64hi(x, Nb);
                    if (lo < Nb && lo < (0ull - Nb) % Nb)
                        reject = true;
                    u = word_to_float(rs.peek((uint32_t)off + 1));
                } else if (dia) {
                    u = word_to_float(rs.peek((uint32_t)off));
                }
            }
            if (__ballot(reject)) {
                k1n_serial_chunk<NS>(w, m, rs, base, n0, nr, pm_raw, pr_raw, do_diag, do_vl, sp,
                                     vfirst, vlast);
                continue;
            }
            rs.advance((uint32_t)D);
        } else if (force_serial) {
            k1n_serial_chunk<NS>(w, m, rs, base, n0, nr, pm_raw, pr_raw, do_diag, do_vl, sp, vfirst,
                                 vlast);
            continue;
        }

        const bool has = isop || (ident && do_diag);
        int st[NS];
        int type = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k)
            st[k] = 0;
        if (has) {
            type = tab.type(b);
#pragma unroll
            for (int k = 0; k < NS; ++k)
                st[k] = tab.site(b, k);
        }
        // per-site lane masks, one per bond position
        if (has) {
#pragma unroll
            for (int k = 0; k < NS; ++k)
                atomicOr(&tmw[2 * ((size_t)k * N + st[k]) + (lane >> 5)], 1u << (lane & 31));
        }
        if (offd)
            cups[lane] = tab.upper(type, (int)(vc >> 1));
        wave_lds_sync();
        // M[j][k]: lanes whose bond has the site of this lane's position j at
        // position k; A[j]: lanes whose bond touches that site at all
        uint64_t M[NS][NS], A[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            A[j] = 0ull;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                M[j][k] = has ? tm[(size_t)k * N + st[j]] : 0ull;
                A[j] |= M[j][k];
            }
        }
        // position of this lane's site j in the bond of lane q
        auto pos_in = [&](int j, int q) -> int {
            int r = 0;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                if ((M[j][k] >> q) & 1ull)
                    r = k;
            return r;
        };

        uint64_t plus = 0, minus = 0;
        uint32_t fop = op;
        if (do_diag) {
            double wgt = 0.0;
            int32_t dv = -1;
            if (ident) {
                int idx = 0, stride = 1;
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    int v;
                    const uint64_t cand = A[j] & offm & below;
                    if (cand) {
                        const int q = hi_bit(cand);
                        v = (int)((cups[q] >> (8 * pos_in(j, q))) & 255u);
                    } else {
                        v = sp[st[j]];
                    }
                    idx += (v - 1) * stride;
                    stride *= tab.dim(type, j);
                }
                dv = tab.diag(type, idx);
                wgt = dv < 0 ? 0.0 : tab.weight(type, dv >> 1);
            } else if (dia) {
                wgt = tab.weight(type, (int)(vc >> 1));
            }
            bool amb = false, dec = false;
            if (ident) {
                int64_t lo = n0 - lane < 0 ? 0 : n0 - lane;
                int64_t hi = n0 + lane > L - 1 ? L - 1 : n0 + lane;
                bool d_lo = u < (pm_raw / (double)(L - lo)) * wgt;
                bool d_hi = u < (pm_raw / (double)(L - hi)) * wgt;
                amb = d_lo != d_hi;
                dec = d_lo;
            } else if (dia) {
                int64_t lo = n0 - lane < 1 ? 1 : n0 - lane;
                int64_t hi = n0 + lane > L ? L : n0 + lane;
                double uw = u * wgt;
                bool d_lo = uw < (double)(L - lo + 1) * pr_raw;
                bool d_hi = uw < (double)(L - hi + 1) * pr_raw;
                amb = d_lo != d_hi;
                dec = d_lo;
            }
            plus = __ballot(ident && !amb && dec);
            minus = __ballot(dia && !amb && dec);
            uint64_t ambm = __ballot(amb);
            while (ambm) {
                const int a = __ffsll((long long)ambm) - 1;
                const uint64_t ba = below_mask(a);
                const int64_t na = n0 + popc(plus & ba) - popc(minus & ba);
                bool d = false;
                if (lane == a) {
                    if (ident)
                        d = u < (pm_raw / (double)(L - na)) * wgt;
                    else
                        d = u * wgt < (double)(L - na + 1) * pr_raw;
                }
                const uint64_t dm = __ballot(d);
                if ((identm >> a) & 1ull)
                    plus |= dm;
                else
                    minus |= dm;
                ambm &= ambm - 1;
            }
            n0 += popc(plus) - popc(minus);
            if (ident && ((plus >> lane) & 1ull))
                fop = op_make(b, (uint32_t)dv);
            if (dia && ((minus >> lane) & 1ull))
                fop = 0u;
            if (valid && fop != op)
                ops[p] = fop;
        }

        // vertex list over the chunk's final non-identity operators, compact
        // form: operator rank r = nr + (non-identity lanes below)
        const uint64_t NI = do_diag ? ((identm & plus) | (diam & ~minus) | offm) : (diam | offm);
        const bool isni = (NI >> lane) & 1ull;
        const int32_t rank = (int32_t)(nr + popc(NI & below));
        int32_t pred[NS];
        bool last[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            pred[j] = -1;
            last[j] = false;
        }
        if (do_vl && isni) {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const uint64_t mm = A[j] & NI;
                const uint64_t cand = mm & below;
                if (cand) {
                    const int q = hi_bit(cand);
                    pred[j] = PL * (int32_t)(nr + popc(NI & below_mask(q))) + NS + pos_in(j, q);
                } else {
                    pred[j] = vlast[st[j]];
                    if (pred[j] < 0)
                        pred[j] = -2;  // first occurrence of the site
                }
                last[j] = (mm & above) == 0ull;
            }
        }
        wave_lds_sync();
        if (do_vl && isni) {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const int32_t me = PL * rank + j;
                if (pred[j] >= 0) {
                    glk[me] = (uint32_t)pred[j];
                    glk[pred[j]] = (uint32_t)me;
                } else {
                    vfirst[st[j]] = me;
                }
                if (last[j])
                    vlast[st[j]] = PL * rank + NS + j;
            }
            gvt[rank] = op_vcode(fop) | ((uint32_t)type << 10);
        }
        if (do_vl && lane == 0) {
            gmask[base >> 6] = NI;
            gbase[base >> 6] = (uint32_t)nr;
        }
        nr += popc(NI);
        // spins after the chunk: the last off-diagonal operator on a site wins
        if (do_diag && offd) {
#pragma unroll
            for (int j = 0; j < NS; ++j)
                if (((A[j] & offm) & above) == 0ull)
                    sp[st[j]] = (uint8_t)(cups[lane] >> (8 * j));
        }
        if (has) {
#pragma unroll
            for (int k = 0; k < NS; ++k)
                tm[(size_t)k * N + st[k]] = 0ull;
        }
        wave_lds_sync();
    }

    // periodic closure (vertex_list.jl:46-51) and write-back
    if (do_vl) {
        for (int s = lane; s < N; s += 64) {
            int32_t f = vfirst[s];
            if (f >= 0) {
                int32_t l = vlast[s];
                glk[f] = (uint32_t)l;
                glk[l] = (uint32_t)f;
            }
            w.first[s] = f;
            w.last[s] = vlast[s];
        }
    }
    if (do_diag) {
        for (int s = lane; s < N; s += 64)
            w.spins[s] = sp[s];
    }
    if (lane == 0) {
        DevWalker &o = W[wi];
        o.n = n0;
        o.L = L;
        o.cap = w.cap;
        o.ops = w.ops;
        o.links = w.links;
        o.ctr = rs.ctr();
        atomicMax(&pool->max_n, (unsigned long long)n0);
        atomicMax(&pool->group_max_n[grp & 7], (unsigned long long)n0);
        atomicMax(&pool->max_L, (unsigned long long)L);
    }
}

// One worm (sse.jl:241-303) of walker w from a random start, wave-uniform,
// over the compact records in global memory.  Returns the worm length (the
// reference's return value: scatters, + 1 when it closes on re-entry), 0 for
// n == 0, or -(error code).
template <int NS>
__device__ __forceinline__ int64_t k2n_worm(const DevWalker &w, const DevModel &m, K2Stream &rs,
                                            gmem<uint32_t> *gvt, const gmem<uint32_t> *glk,
                                            const gmem<uint64_t> *gmask,
                                            const gmem<uint32_t> *gbase, int64_t &scat);

// The same worm with the chase codes (ecode = type << lgp | vcode) in LDS and
// one chase-table entry per step (DevModel::gtab): the exit is picked by
// comparing the step's 53-bit mantissa with the entry's integer thresholds
// (r < cum exactly when (word >> 11) < ceil(cum * 2^53)); entries marked slow
// (a fourth exit with a threshold below 2^53, more than four exits) take the
// exact search.
template <int NS>
__device__ __forceinline__ int64_t k2n_worm_lds(const DevWalker &w, const DevModel &m, K2Stream &rs,
                                                lds_u16 *lec, const gmem<uint32_t> *glk,
                                                const gmem<uint64_t> *gmask,
                                                const gmem<uint32_t> *gbase, int64_t &scat)
{
    constexpr int LEGS = LegsOf<NS>::LEGS, PL = LegsOf<NS>::PL, LG = LegsOf<NS>::LG;
    if (w.n == 0)
        return 0;
    const int64_t L = w.L;
    const uint32_t lgp = uni((uint32_t)m.dt_lgp), dmw = uni((uint32_t)m.dt_mw);
    const_char *GT = (const_char *)(const void *)uni64((uint64_t)(const void *)m.gtab);
    int64_t p0;
    uint32_t l0, r0;
    for (;;) {  // sse.jl:241-247
        p0 = (int64_t)uni64(rs.range((uint64_t)L));
        l0 = uni((uint32_t)rs.range((uint64_t)LEGS));
        const uint64_t mk = uni64(gmask[p0 >> 6]);
        const int bit = (int)(p0 & 63);
        if ((mk >> bit) & 1ull) {
            r0 = uni(gbase[p0 >> 6] + (uint32_t)popc(mk & below_mask(bit)));
            break;
        }
    }
    const uint32_t b0 = uni(op_bond(w.ops[p0]));
    const int site0 = (int)uni((uint32_t)m.bond_sites[(size_t)NS * b0 + l0 % NS]);
    const uint32_t dim0 = uni((uint32_t)m.site_dims[site0]);
    const uint32_t wf0 = 1u + uni((uint32_t)rs.range((uint64_t)(dim0 - 1)));  // sse.jl:251
    uint32_t r = r0, leg_in = l0, wf = wf0;
    int64_t len = 1;
    for (;;) {  // sse.jl:274-300
        // the links of all legs of this operator, lane j holding leg j's
        // (one vector load, in flight while the exit is picked)
        const uint32_t lkv = threadIdx.x < (unsigned)LEGS ? glk[PL * r + threadIdx.x] : 0u;
        const uint32_t e = uni((uint32_t)lec[r]);
        const uint32_t eix = ((e * dmw + (wf - 1u)) << LG) + leg_in;
        ConstEntry *E = (ConstEntry *)(GT + (size_t)eix * sizeof(DirectEntry));
        const uint64_t mm = rs.next() >> 11;
        scat += 1;
        uint32_t eo;
        if (mm < E->K[0])
            eo = E->out[0];
        else if (mm < E->K[1])
            eo = E->out[1];
        else if (mm < E->K[2])
            eo = E->out[2];
        else
            eo = E->out[3];
        eo = uni(eo);
        if (eo & kGenSlowBit) {
            const uint32_t cnt = uni(E->count);
            if (cnt == 0u)
                return -SSE_ERR_SCATTER;
            if (cnt <= 4u) {
                uint32_t j = 0;
                for (; j < cnt; ++j)
                    if (mm < uni64(E->K[j]))
                        break;
                if (j >= cnt)
                    return -SSE_ERR_SCATTER;
                eo = uni(j < 3u ? E->out[j] : E->pad[0]);
            } else {
                const uint32_t type = e >> lgp;
                const DevVD &vd = m.types[type];
                const double rr = (double)mm * 0x1.0p-53;
                const int32_t off = (int32_t)uni(E->trans);
                int32_t x = off;
                for (; x < off + (int32_t)cnt; ++x) {
                    const double c = __longlong_as_double((long long)uni64(
                        (uint64_t)__double_as_longlong(vd.cum[x])));
                    if (rr < c)
                        break;
                }
                if (x >= off + (int32_t)cnt)
                    return -SSE_ERR_SCATTER;
                eo = gen_out((type << lgp) | uni((uint32_t)vd.tgt[x]), uni((uint32_t)vd.eleg[x]),
                             uni((uint32_t)vd.ewrm[x]));
            }
        }
        const uint32_t leg_out = (eo >> 12) & 7u;
        const uint32_t wf_out = (eo & 0xFFFu) + 1u;
        lec[r] = (uint16_t)(eo >> 16);  // every lane writes the same value
        if (r == r0 && leg_out == l0 && wf_out == dim0 - wf0)
            break;
        len += 1;
        wf = wf_out;
        const uint32_t nx = (uint32_t)__builtin_amdgcn_readlane((int)lkv, (int)leg_out);
        leg_in = nx & (uint32_t)(PL - 1);
        r = nx >> LG;
        if (r >= (uint32_t)w.n || leg_in >= (uint32_t)LEGS)
            return -SSE_ERR_STATE;
        if (r == r0 && leg_in == l0 && wf == wf0)
            break;
    }
    return len;
}

template <int NS>
__device__ __forceinline__ int64_t k2n_worm(const DevWalker &w, const DevModel &m, K2Stream &rs,
                                            gmem<uint32_t> *gvt, const gmem<uint32_t> *glk,
                                            const gmem<uint64_t> *gmask,
                                            const gmem<uint32_t> *gbase, int64_t &scat)
{
    constexpr int LEGS = LegsOf<NS>::LEGS, PL = LegsOf<NS>::PL, LG = LegsOf<NS>::LG;
    if (w.n == 0)
        return 0;  // worm_traverse! returns 0 without draws (sse.jl:234-236)
    const int64_t L = w.L;
    int64_t p0;
    uint32_t l0, r0;
    for (;;) {  // sse.jl:241-247
        p0 = (int64_t)uni64(rs.range((uint64_t)L));
        l0 = uni((uint32_t)rs.range((uint64_t)LEGS));
        const uint64_t mk = uni64(gmask[p0 >> 6]);
        const int bit = (int)(p0 & 63);
        if ((mk >> bit) & 1ull) {
            r0 = uni(gbase[p0 >> 6] + (uint32_t)popc(mk & below_mask(bit)));
            break;
        }
    }
    const uint32_t b0 = uni(op_bond(w.ops[p0]));
    const int site0 = (int)uni((uint32_t)m.bond_sites[(size_t)NS * b0 + l0 % NS]);
    const uint32_t dim0 = uni((uint32_t)m.site_dims[site0]);
    const uint32_t wf0 = 1u + uni((uint32_t)rs.range((uint64_t)(dim0 - 1)));  // sse.jl:251
    uint32_t r = r0, leg_in = l0, wf = wf0;
    int64_t len = 1;
    for (;;) {  // sse.jl:274-300
        const uint32_t vt = uni(gvt[r]);
        const uint32_t type = vt >> 10, vc = vt & 1023u;
        const DevVD &vd = m.types[type];
        const double rr = rs.next_float();
        scat += 1;
        // scatter (vertex_data.jl:106-125): first exit with rr < cum
        const int tix = ((int)(vc >> 1) * vd.max_worm + (int)(wf - 1)) * LEGS + (int)leg_in;
        const int32_t off = uni((uint32_t)vd.toff[tix]);
        const int32_t cnt = uni((uint32_t)vd.tcnt[tix]);
        if (off < 0)
            return -SSE_ERR_SCATTER;
        int32_t e = off;
        for (; e < off + cnt; ++e) {
            const double c = __longlong_as_double((long long)uni64(
                (uint64_t)__double_as_longlong(vd.cum[e])));
            if (rr < c)
                break;
        }
        if (e >= off + cnt)
            return -SSE_ERR_SCATTER;
        const uint32_t leg_out = uni((uint32_t)vd.eleg[e]);
        const uint32_t wf_out = uni((uint32_t)vd.ewrm[e]);
        const uint32_t nvc = uni((uint32_t)vd.tgt[e]);
        if (threadIdx.x == 0)
            gvt[r] = nvc | (type << 10);
        if (r == r0 && leg_out == l0 && wf_out == dim0 - wf0)
            break;  // leaves through its start leg with the inverse worm
        len += 1;
        wf = wf_out;
        const uint32_t nx = uni(glk[PL * r + leg_out]);
        leg_in = nx & (uint32_t)(PL - 1);
        r = nx >> LG;
        if (r >= (uint32_t)w.n || leg_in >= (uint32_t)LEGS)
            return -SSE_ERR_STATE;
        if (r == r0 && leg_in == l0 && wf == wf0)
            break;  // re-enters its start leg
    }
    return len;
}

// Worm update of walker wi (sse.jl:193-231): the worms, the state refresh of
// sites without operators and with them (sse.jl:219-228), the operator
// write-back and the controller.
template <int NS>
__device__ void k2n_walker(DevWalker *W, const DevModel *models, DevPool *pool, int wi,
                           int full_sweep, char *lds, int lds_bytes)
{
    constexpr int LEGS = LegsOf<NS>::LEGS, PL = LegsOf<NS>::PL, LG = LegsOf<NS>::LG;
    const int lane = threadIdx.x;
    DevWalker w = W[wi];
    if (w.error || !w.active || w.retire)
        return;
    const DevModel &m = models[w.model];
    const bool therm = w.sweeps > w.thermalization;
    K2Stream rs;
    rs.init(w.key, w.ctr);
    gmem<uint32_t> *gvt = as_global(recn_vt(w));
    const gmem<uint32_t> *glk = as_global((const uint32_t *)recn_link(w));
    const gmem<uint64_t> *gmask = as_global((const uint64_t *)recn_mask<PL>(w));
    const gmem<uint32_t> *gbase = as_global((const uint32_t *)recn_base<PL>(w));
    double total = 1.0;
    int64_t scat = 0;
    int err = 0;
    const int64_t nworms = (int64_t)ceil(w.num_worms);
    // chase codes in LDS when they fit and the codes fit 16 bits
    const bool in_lds = (int64_t)2 * w.n <= (int64_t)lds_bytes &&
                        ((int64_t)m.n_types << m.dt_lgp) <= 65536;
    if (in_lds) {
        lds_u16 *lec = (lds_u16 *)(uint16_t *)lds;
        const uint32_t lgp = (uint32_t)m.dt_lgp;
        for (int64_t r = lane; r < w.n; r += 64)
            lec[r] = (uint16_t)ecode_of_vt(gvt[r], lgp);
        __syncthreads();
        for (int64_t k = 0; k < nworms; ++k) {
            const int64_t len = k2n_worm_lds<NS>(w, m, rs, lec, glk, gmask, gbase, scat);
            if (len < 0) {
                err = (int)-len;
                break;
            }
            total += (double)len;
        }
        __syncthreads();
        if (!err)
            for (int64_t r = lane; r < w.n; r += 64)
                gvt[r] = vt_of_ecode(lec[r], lgp);
    } else {
        for (int64_t k = 0; k < nworms; ++k) {
            const int64_t len = k2n_worm<NS>(w, m, rs, gvt, glk, gmask, gbase, scat);
            if (len < 0) {
                err = (int)-len;
                break;
            }
            total += (double)len;
        }
    }
    __syncthreads();
    if (!err) {
        // state refresh (sse.jl:219-228): a site without operators draws a new
        // state (in site order), one with operators takes the lower leg state
        // of its first vertex
        const int N = m.n_sites;
        for (int sb = 0; sb < N; sb += 32) {
            const int s = sb + lane;
            const bool vs = lane < 32 && s < N;
            const int32_t f = vs ? w.first[s] : 0;
            const bool fr = vs && f < 0;
            const uint64_t frm = __ballot(fr);
            const int D = popc(frm);
            const uint64_t dim = vs ? (uint64_t)m.site_dims[s] : 1;
            bool reject = false;
            uint8_t val = 0;
            if (D > 0) {
                rs.ensure((uint32_t)D);
                const uint64_t x = rs.peek(fr ? (uint32_t)popc(frm & below_mask(lane)) : 0u);
                if (fr) {
                    uint64_t lo = x * dim;
                    if (lo < dim && lo < (0ull - dim) % dim)
                        reject = true;
                    val = (uint8_t)(1 + __umul64hi(x, dim));
                }
            }
            if (vs && !fr) {
                const uint32_t rk = (uint32_t)f >> LG, lg = (uint32_t)f & (uint32_t)(PL - 1);
                if ((int64_t)rk < w.n && lg < (uint32_t)NS) {
                    const uint32_t vt = gvt[rk];
                    const DevVD &vd = m.types[vt >> 10];
                    val = vd.leg_states[(size_t)LEGS * ((vt & 1023u) >> 1) + lg];
                } else {
                    err = SSE_ERR_STATE;
                }
            }
            if (__ballot(reject)) {
                for (int i = 0; i < 32 && sb + i < N; ++i) {
                    const bool fi = (frm >> i) & 1ull;
                    uint8_t vi = 0;
                    if (fi)
                        vi = (uint8_t)(1 + rs.range((uint64_t)m.site_dims[sb + i]));
                    if (lane == i && fi)
                        val = vi;
                }
            } else {
                rs.advance((uint32_t)D);
            }
            if (vs)
                w.spins[s] = val;
        }
        err = __ballot(err != 0) ? SSE_ERR_STATE : 0;
        // the worm-updated vertex codes back into the operator string
        const int64_t L = w.L;
        gmem<uint32_t> *ops = as_global(w.ops);
        for (int64_t base = 0; base < L && !err; base += 64) {
            const uint64_t mk = gmask[base >> 6];
            if ((mk >> lane) & 1ull) {
                const uint32_t rk = gbase[base >> 6] + (uint32_t)popc(mk & below_mask(lane));
                const uint32_t op = ops[base + lane];
                const uint32_t vc = gvt[rk] & 1023u;
                const uint32_t nop = (op & ~(1023u << 1)) | (vc << 1);
                if (nop != op)
                    ops[base + lane] = nop;
            }
        }
    }
    if (err) {
        if (lane == 0) {
            W[wi].error = err;
            flag_error(pool, err);
        }
        return;
    }
    w.last_scatter = scat;
    worm_controller(w, total, therm);
    if (lane == 0) {
        DevWalker &o = W[wi];
        o.ctr = rs.ctr();
        o.avg_worm_length = w.avg_worm_length;
        o.num_worms = w.num_worms;
        o.last_scatter = w.last_scatter;
        o.last_total = w.last_total;
        o.last_wlf = w.last_wlf;
        if (full_sweep) {
            o.sweeps = w.sweeps + 1;
            o.acc_L = w.acc_L + w.L;
            o.acc_n = w.acc_n + w.n;
            o.acc_W = w.acc_W + scat;
            o.acc_sweeps = w.acc_sweeps + 1;
            if (w.remaining > 0) {
                o.remaining = w.remaining - 1;
                if (w.remaining == 1)
                    o.retire = 1;
            }
        }
    }
}

// Measurement (sse.jl:70-87, 305-376) of walker wi, 64 slots per step.  With
// the model's per-(bond, vertex) tables it is k3_walker's (their values do not
// depend on the bond size); otherwise the sign and operator count in one pass
// over the operator string, then one pass per magnetization estimator.  mode
// as in k3_walker.
template <int NS>
__device__ void k3n_walker(DevWalker *W, const DevModel *models, double *out, double *sum,
                           double *sumsq, int64_t *count, double *wlf, int wi, int mode,
                           int32_t *done)
{
    constexpr int LEGS = LegsOf<NS>::LEGS;
    const int lane = threadIdx.x;
    if (models[W[wi].model].est_dtab) {
        k3_walker<true>(W, models, out, sum, sumsq, count, wlf, wi, mode, done);
        return;
    }
    const DevWalker w = W[wi];
    if (w.error || w.pending || (mode == 1 && !w.active))
        return;
    if (mode == 1 && !(w.sweeps > w.thermalization)) {
        if (w.retire && lane == 0) {
            W[wi].retire = 0;
            W[wi].active = 0;
            done[wi] = 1;
        }
        return;
    }
    const DevModel &m = models[w.model];
    double *o = out + w.meas_offset;
    const int N = m.n_sites, D = m.est_max_dim;
    const int64_t L = w.L;
    const gmem<uint32_t> *ops = as_global((const uint32_t *)w.ops);
    int neg = 0;
    int64_t cnt = 0;
    double sign = 1.0;
    // the sign and the operator count
    for (int64_t base = 0; base < L; base += 64) {
        const int64_t p = base + lane;
        const uint32_t op = p < L ? ops[p] : 0u;
        bool ng = false;
        if (op != 0u) {
            const DevVD &vd = m.types[m.bond_type[op_bond(op)]];
            ng = vd.signs[op_vcode(op) >> 1] < 0;
        }
        neg += popc(__ballot(ng));
        cnt += popc(__ballot(op != 0u));
    }
    sign = (neg & 1) ? -1.0 : 1.0;
    for (int e = 0; e < m.n_est; ++e) {
        const gmem<double> *cf = as_global(m.est_coef + (size_t)e * N);
        const gmem<double> *mk = as_global(m.est_mval + (size_t)e * N * D);
        double t0 = 0.0;
        for (int i = lane; i < N; i += 64)
            t0 += cf[i] * mk[(size_t)i * D + w.spins[i] - 1];
        const double tmp0 = wave_sum(t0);
        double run = tmp0, mag = 0.0, absmag = 0.0, mag2 = 0.0, mag4 = 0.0;
        for (int64_t base = 0; base < L; base += 64) {
            const int64_t p = base + lane;
            const uint32_t op = p < L ? ops[p] : 0u;
            double delta = 0.0;
            if (op != 0u && !(op_vcode(op) & 1u)) {
                const uint32_t b = op_bond(op);
                const DevVD &vd = m.types[m.bond_type[b]];
                const uint8_t *ls = vd.leg_states + (size_t)LEGS * (op_vcode(op) >> 1);
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const int s = m.bond_sites[(size_t)NS * b + k];
                    delta += cf[s] * (mk[(size_t)s * D + ls[NS + k] - 1] - mk[(size_t)s * D + ls[k] - 1]);
                }
            }
            const double inc = wave_incl_scan(delta);
            if (op != 0u) {
                const double t = run + inc;
                const double t2 = t * t;
                mag += t;
                absmag += fabs(t);
                mag2 += t2;
                mag4 += t2 * t2;
            }
            run += __shfl(inc, 63);
        }
        // as k3_pass: the estimator's five values (magnetization_estimator.jl:125-165,205-230)
        const double Nn = m.norm;
        const double en = 1.0 + (double)cnt;
        const double norm = 1.0 / Nn;
        double smag = tmp0 + wave_sum(mag);
        double sabs = fabs(tmp0) + wave_sum(absmag);
        double s2 = tmp0 * tmp0;
        double smag2 = s2 + wave_sum(mag2);
        double smag4 = s2 * s2 + wave_sum(mag4);
        smag *= norm;
        sabs *= norm;
        smag2 *= norm * norm;
        smag4 *= (norm * norm) * (norm * norm);
        if (lane == 0) {
            double *oe = o + 5 + 5 * e;
            oe[0] = sign * smag / en;
            oe[1] = sign * sabs / en;
            oe[2] = sign * smag2 / en;
            oe[3] = sign * smag4 / en;
            const double chi = 1.0 / w.T / (en + 1.0) / en * (smag * smag + smag2) * Nn;
            oe[4] = sign * chi;
        }
    }
    const double nn = (double)w.n;
    if (lane == 0) {
        o[0] = sign;
        o[1] = nn;
        o[2] = sign * nn;
        o[3] = sign * (nn * nn);
        o[4] = ((-sign) * (nn * w.T + m.energy_offset)) / m.norm;
    }
    if (mode == 1) {
        __syncthreads();
        for (int j = lane; j < w.meas_width; j += 64) {
            const double x = o[j];
            sum[w.meas_offset + j] += x;
            sumsq[w.meas_offset + j] += x * x;
        }
        if (lane == 0) {
            count[w.acc_slot] += 1;
            worm_fraction_accumulate(w, wlf);
            if (w.retire) {
                W[wi].retire = 0;
                W[wi].active = 0;
                done[wi] = 1;
            }
        }
    }
}

// ------------------------------------------------------------------ kernels
// The site count of a walker's model selects the instantiation (all bonds of
// a model have the same count).
#define SSE_NS_DISPATCH(ns, CALL) \
    switch (ns) {                 \
    case 1: CALL(1); break;       \
    case 2: CALL(2); break;       \
    case 3: CALL(3); break;       \
    default: CALL(4); break;      \
    }

// budget non-null: sweep `step` of a budgeted call; walker wi takes part
// only while step < budget[wi]
// K1n of walker wi with the model's lookups in LDS when the launch's LDS
// (lds_bytes) holds them
template <int NS>
__device__ __forceinline__ void k1n_walker_sel(DevWalker *W, const DevModel *models, DevPool *pool, int wi,
                                               int phases, int warm, char *lds, int grp, int lds_bytes)
{
    const DevModel *m = &models[W[wi].model];
    if (k1n_lds_staged(m->n_sites, NS, m->k1ntab_bytes) <= (size_t)lds_bytes)
        k1n_walker<NS>(W, models, pool, wi, phases, warm, lds, grp,
                       N1Staged<NS>{m, lds + k1n_lds_bytes(m->n_sites, NS)});
    else
        k1n_walker<NS>(W, models, pool, wi, phases, warm, lds, grp, N1Global<NS>{m});
}

__global__ __launch_bounds__(64) void waven_diag_vlist_kernel(DevWalker *W, const DevModel *models,
                                                              DevPool *pool, int w_begin, int w_end,
                                                              int phases, int warm,
                                                              const int32_t *order, int grp,
                                                              const int32_t *budget, int step,
                                                              int lds_bytes)
{
    extern __shared__ __align__(16) char lds[];
    const int pos = w_begin + (int)blockIdx.x;
    if (pos >= w_end)
        return;
    const int wi = walker_at(order, pos);
    if (budget && step >= budget[wi])
        return;
    const int ns = models[W[wi].model].nsb;
#define K1N_CALL(K) k1n_walker_sel<K>(W, models, pool, wi, phases, warm, lds, grp, lds_bytes)
    SSE_NS_DISPATCH(ns, K1N_CALL)
#undef K1N_CALL
}

__global__ __launch_bounds__(64) void waven_worm_kernel(DevWalker *W, const DevModel *models,
                                                        DevPool *pool, int w_begin, int w_end,
                                                        int full_sweep, const int32_t *order,
                                                        const int32_t *budget, int step,
                                                        int lds_bytes)
{
    extern __shared__ __align__(16) char lds[];
    const int pos = w_begin + (int)blockIdx.x;
    if (pos >= w_end)
        return;
    const int wi = walker_at(order, pos);
    if (budget && step >= budget[wi])
        return;
    const int ns = models[W[wi].model].nsb;
#define K2N_CALL(K) k2n_walker<K>(W, models, pool, wi, full_sweep, lds, lds_bytes)
    SSE_NS_DISPATCH(ns, K2N_CALL)
#undef K2N_CALL
}

__global__ __launch_bounds__(64) void waven_measure_kernel(DevWalker *W, const DevModel *models,
                                                           double *out, double *sum, double *sumsq,
                                                           int64_t *count, double *wlf, int w_begin,
                                                           int w_end, int mode, int32_t *done,
                                                           const int32_t *order,
                                                           const int32_t *budget, int step)
{
    const int pos = w_begin + (int)blockIdx.x;
    if (pos >= w_end)
        return;
    const int wi = walker_at(order, pos);
    if (budget && step >= budget[wi])
        return;
    const int ns = models[W[wi].model].nsb;
#define K3N_CALL(K) k3n_walker<K>(W, models, out, sum, sumsq, count, wlf, wi, mode, done)
    SSE_NS_DISPATCH(ns, K3N_CALL)
#undef K3N_CALL
}

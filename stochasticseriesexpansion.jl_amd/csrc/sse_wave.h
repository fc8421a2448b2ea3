// sse_wave.h -- one wavefront (64 lanes) per walker.
//
// K1 wave_diag_vlist_kernel: diagonal update (sse.jl:137-191) and vertex-list
//   construction (vertex_list.jl:15-54) fused into one streaming pass over the
//   operator string, 64 slots per step:
//   * draw offsets: slot types (identity 2 draws, diagonal 1, off-diagonal 0)
//     are fixed before the update, so every lane's Philox words follow from a
//     ballot prefix count; the words come from a 256-word LDS buffer that all
//     lanes refill cooperatively.
//   * spins at identity slots and vertex-list predecessors: every lane ORs its
//     bit into per-site LDS lane masks (one for bond position 0, one for 1);
//     "latest earlier operator on site s" is the highest lower bit of the mask
//     filtered by the off-diagonal (spins) or final non-identity (links) set.
//   * the acceptance tests depend on the running operator count n, but each
//     is monotone in n and n moves by at most one per slot: a lane's decision
//     is fixed if it agrees at both ends of its window [n0-lane, n0+lane];
//     the rare ambiguous lanes are resolved in slot order with the exact n.
//   * a Lemire rejection of the bond draw (probability Nb/2^64) would shift
//     later draw offsets; that chunk is then processed slot by slot.
// K2 wave_worm_kernel: worm_update (sse.jl:193-303) -- the traversal is a
//   sequential pointer chase executed uniformly by the wave; the lanes refill
//   the Philox buffer together (64 blocks per refill).
// K3 wave_measure_kernel: measure (sse.jl:70-87,305-376) with the estimator
//   running sums as wave prefix sums.
#pragma once
#include <type_traits>
#include <vector>

#include "sse_common.h"
#include "sse_lane.h"

struct WaveConfig {
    int max_sites = 0;
    int max_warmup = 0;
    int max_wtab = 0;
    size_t lds_k1 = 0, lds_k2 = 0, lds_k3 = 0;
    size_t lds_limit = 65536;  // device limit on dynamic LDS per workgroup
    bool small_codes = true;   // every model has <= 256 chase codes (Rec8 records)
    int rec_mode = 0;          // worm-kernel record layout: 0 auto, 1 no Rec8, 2 global
    int chase = 0;             // LDS-record chase: 0 plain, 1 with record lookahead
    size_t lds_k2_fixed = 0;   // worm-kernel LDS per walker set by the caller (0 = sized here)
    size_t lds_k2_group[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // per stream group (0 = lds_k2)
    int fuse_measure = 0;      // measuring sweeps: worm update + measurement in one kernel
    int generic = 0;           // some model has bonds of other than two sites: sse_wave_n.h kernels
    size_t lds_k1n = 0;        // LDS of the generic diagonal-update kernel
    size_t lds_k2n = 0;        // LDS of the generic worm kernel (chase codes, 2 bytes per operator)
};

__device__ __forceinline__ uint64_t below_mask(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }
__device__ __forceinline__ uint64_t above_mask(int lane) { return lane >= 63 ? 0ull : (~0ull << (lane + 1)); }
__device__ __forceinline__ int hi_bit(uint64_t x) { return 63 - __clzll((long long)x); }
__device__ __forceinline__ int popc(uint64_t x) { return __popcll(x); }

// Philox words of one walker staged in LDS: buf holds words
// [bbase, bbase+256) and the next word to draw is bbase+off.  All members are
// wave-uniform; every call must be made by the whole wave.
template <int NW>
__device__ __noinline__ void wave_refill(uint64_t *buf, uint64_t key, uint64_t base)
{
    const int l = threadIdx.x;
    __syncthreads();
#pragma unroll
    for (int j = l; j < NW / 2; j += 64) {
        uint32_t o0, o1, o2, o3;
        philox_block((base >> 1) + (uint64_t)j, key, o0, o1, o2, o3);
        buf[2 * j] = (uint64_t)o0 | ((uint64_t)o1 << 32);
        buf[2 * j + 1] = (uint64_t)o2 | ((uint64_t)o3 << 32);
    }
    __syncthreads();
}

template <int NW>
struct WaveStreamT {
    static constexpr uint32_t kNW = NW;
    uint64_t *buf;
    uint64_t key, bbase;
    uint32_t off;  // next word is bbase + off; off > 256 marks an empty buffer
    __device__ void init(uint64_t *b, uint64_t k, uint64_t c)
    {
        buf = b;
        key = k;
        bbase = c;
        off = 1024;  // empty: the first ensure() refills at c
    }
    __device__ __forceinline__ uint64_t ctr() const { return bbase + off - (off >= 1024 ? 1024 : 0); }
    // make words [ctr, ctr+need) resident (need <= 255)
    __device__ __forceinline__ void ensure(uint32_t need)
    {
        if (off >= 1024) {  // empty: start at the pending counter bbase
            const uint64_t c = bbase;
            bbase = c & ~1ull;
            off = (uint32_t)(c - bbase);
            wave_refill<NW>(buf, key, bbase);
        } else if (off + need > NW) {
            const uint64_t c = bbase + off;
            bbase = c & ~1ull;
            off = (uint32_t)(c - bbase);
            wave_refill<NW>(buf, key, bbase);
        }
    }
    // the buffer is used up (off == NW): load the next NW words
    __device__ __forceinline__ void refill_next()
    {
        bbase += off;
        off = 0;
        wave_refill<NW>(buf, key, bbase);
    }
    // word k positions after the next one (after ensure)
    __device__ __forceinline__ uint64_t peek(uint32_t k) const { return buf[off + k]; }
    __device__ __forceinline__ void advance(uint32_t k) { off += k; }
    __device__ __forceinline__ uint64_t next()
    {
        ensure(1);
        uint64_t w = buf[off];
        ++off;
        return w;
    }
    __device__ __forceinline__ double next_float() { return word_to_float(next()); }
    __device__ __forceinline__ uint64_t range(uint64_t s)
    {
        uint64_t x = next();
        uint64_t lo = x * s, hi = __umul64hi(x, s);
        if (lo < s) {
            uint64_t t = (0ull - s) % s;
            while (lo < t) {
                x = next();
                lo = x * s;
                hi = __umul64hi(x, s);
            }
        }
        return hi;
    }
};

typedef WaveStreamT<256> WaveStream;
typedef WaveStreamT<128> WaveStream128;
// The worm kernel's Philox words in registers: lane l holds word bbase + l of
// the walker's stream (bbase even) and that word's 53-bit mantissa
// (word >> 11, the integer behind rand()).  A wave-uniform draw is two
// v_readlane with a scalar index; lanes that draw different words read them
// with ds_bpermute.  No LDS and no barrier; all members except the four
// per-lane words are wave-uniform and every call is made by the whole wave.
struct K2RegStream {
    static constexpr uint32_t kNW = 64;
    uint32_t wlo, whi, mlo, mhi;
    uint64_t key, bbase;
    uint32_t off;  // next word is bbase + off; off >= 1024 marks an empty buffer
    __device__ void init(uint64_t k, uint64_t c)
    {
        key = k;
        bbase = c;
        off = 1024;
    }
    __device__ __forceinline__ void fill()
    {
        const uint32_t lane = threadIdx.x;
        uint32_t o0, o1, o2, o3;
        philox_block((bbase >> 1) + (lane >> 1), key, o0, o1, o2, o3);
        wlo = (lane & 1u) ? o2 : o0;
        whi = (lane & 1u) ? o3 : o1;
        const uint64_t mm = (((uint64_t)whi << 32) | wlo) >> 11;
        mlo = (uint32_t)mm;
        mhi = (uint32_t)(mm >> 32);
    }
    __device__ __forceinline__ uint64_t ctr() const { return bbase + off - (off >= 1024 ? 1024 : 0); }
    // make words [ctr, ctr+need) resident (need <= 63)
    __device__ __forceinline__ void ensure(uint32_t need)
    {
        if (off >= 1024) {
            const uint64_t c = bbase;
            bbase = c & ~1ull;
            off = (uint32_t)(c - bbase);
            fill();
        } else if (off + need > kNW) {
            const uint64_t c = bbase + off;
            bbase = c & ~1ull;
            off = (uint32_t)(c - bbase);
            fill();
        }
    }
    // the buffer is used up (off == kNW): load the next kNW words
    __device__ __forceinline__ void refill_next()
    {
        bbase += off;
        off = 0;
        fill();
    }
    // word k (wave-uniform, < kNW) of the buffer and its mantissa
    __device__ __forceinline__ uint64_t word_at(uint32_t k) const
    {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)wlo, (int)k);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)whi, (int)k);
        return ((uint64_t)hi << 32) | lo;
    }
    // word k positions after the next one, k per lane (after ensure)
    __device__ __forceinline__ uint64_t peek(uint32_t k) const
    {
        const int a = (int)((off + k) << 2);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)wlo);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)whi);
        return ((uint64_t)hi << 32) | lo;
    }
    __device__ __forceinline__ void advance(uint32_t k) { off += k; }
    __device__ __forceinline__ uint64_t next()
    {
        ensure(1);
        const uint64_t x = word_at(off);
        ++off;
        return x;
    }
    __device__ __forceinline__ double next_float() { return word_to_float(next()); }
    __device__ __forceinline__ uint64_t range(uint64_t s)
    {
        uint64_t x = next();
        uint64_t lo = x * s, hi = __umul64hi(x, s);
        if (lo < s) {
            uint64_t t = (0ull - s) % s;
            while (lo < t) {
                x = next();
                lo = x * s;
                hi = __umul64hi(x, s);
            }
        }
        return hi;
    }
};
typedef K2RegStream K2Stream;

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x)
{
    return (uint64_t)uni((uint32_t)x) | ((uint64_t)uni((uint32_t)(x >> 32)) << 32);
}

// LDS layout of K1 (per workgroup = one walker):
//   rbuf  u64[256]        Philox words
//   tm0   u64[N]          per-site lane mask, bond position 0
//   tm1   u64[N]          per-site lane mask, bond position 1
//   vfirst, vlast i32[N]  vertex-list first/last leg per site
//   cups  u16[64]         upper-leg states of this chunk's operators
//   sp    u8[N]           spins at the current slot
__host__ __device__ inline size_t k1_lds_bytes(int N)
{
    size_t b = 2048 + 16 * (size_t)N + 8 * (size_t)N + 128 + (size_t)N;
    return (b + 15) & ~(size_t)15;
}

constexpr int K1_BLOCK = 16;  // 64-slot chunks per block of operator-string loads

// Orders the LDS accesses of the lanes of a one-wave workgroup.  The LDS runs
// a wave's instructions in order, so only code motion must be prevented; unlike
// __syncthreads this does not wait for global loads still in flight.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Model lookups of the diagonal update / vertex list: from the LDS copies of
// the bond records and of DevModel::k1tab (K1Staged), or from global memory.
struct K1Global {
    const DevModel *m;
    __device__ __forceinline__ void stage(const DevModel &) const {}
    __device__ __forceinline__ int2 bond(uint32_t b) const { return m->bond_rec[b]; }
    __device__ __forceinline__ int nv(int t) const { return m->types[t].nv; }
    __device__ __forceinline__ int dims0(int t) const { return m->types[t].dims0; }
    __device__ __forceinline__ int diag(int t, int i) const { return m->types[t].diag_vertex[i]; }
    __device__ __forceinline__ double weight(int t, int v) const { return m->types[t].weights[v]; }
    __device__ __forceinline__ uint32_t upper(int t, int v) const
    {
        const uint8_t *ls = m->types[t].leg_states + 4 * v;
        return ls[2] | ((uint32_t)ls[3] << 8);
    }
};
struct K1Staged {
    const int2 *br;
    const uint8_t *kt;
    // copy the bond records and the model's k1tab into LDS (whole wave)
    __device__ __forceinline__ void stage(const DevModel &m) const
    {
        int2 *b = (int2 *)br;
        for (int i = threadIdx.x; i < m.n_bonds; i += 64)
            b[i] = m.bond_rec[i];
        uint32_t *k = (uint32_t *)kt;
        const uint32_t *g = (const uint32_t *)m.k1tab;
        for (int i = threadIdx.x; i < m.k1tab_bytes / 4; i += 64)
            k[i] = g[i];
    }
    __device__ __forceinline__ int2 bond(uint32_t b) const { return br[b]; }
    __device__ __forceinline__ uint4 hdr(int t) const { return ((const uint4 *)kt)[t]; }
    __device__ __forceinline__ int nv(int t) const { return (int)(hdr(t).w & 0xFFFFu); }
    __device__ __forceinline__ int dims0(int t) const { return (int)(hdr(t).w >> 16); }
    __device__ __forceinline__ int diag(int t, int i) const { return ((const int16_t *)(kt + hdr(t).y))[i]; }
    __device__ __forceinline__ double weight(int t, int v) const
    {
        return ((const double *)(kt + hdr(t).x))[v];
    }
    __device__ __forceinline__ uint32_t upper(int t, int v) const
    {
        const uint8_t *ls = kt + hdr(t).z + 4 * v;
        return ls[2] | ((uint32_t)ls[3] << 8);
    }
};

// LDS bytes of K1 with the model's lookups staged: the base layout, then the
// bond records and the k1tab of the model
__host__ __device__ inline size_t k1_lds_staged(int N, int n_bonds, int k1tab_bytes)
{
    return k1_lds_bytes(N) + ((8 * (size_t)n_bonds + 15) & ~(size_t)15) + (size_t)k1tab_bytes;
}

// Slot-by-slot processing of one chunk (reference order).  Used when a bond
// draw of the chunk hits a Lemire rejection, and when forced for validation.
// Emits the same compact records as the parallel path.
__device__ void k1_serial_chunk(DevWalker &w, const DevModel &m, WaveStream &rs, int64_t base,
                                int64_t &n, int64_t &nr, double pm_raw, double pr_raw,
                                bool do_diag, bool do_vl, uint8_t *sp, int32_t *vfirst,
                                int32_t *vlast)
{
    const int lane = threadIdx.x;
    const int64_t L = w.L;
    uint32_t *gvt = rec_vt(w);
    uint32_t *glk = rec_link(w);
    const int64_t nr0 = nr;
    uint64_t cmask = 0ull;
    for (int i = 0; i < 64 && base + i < L; ++i) {
        const int64_t p = base + i;
        uint32_t op = w.ops[p];
        uint32_t nop = op;
        if (do_diag) {
            if (op == 0u) {
                uint32_t b = (uint32_t)rs.range((uint64_t)m.n_bonds);
                int2 br = m.bond_rec[b];
                const DevVD &vd = m.types[br.x >> 20];
                int s0 = br.x & 0xFFFFF, s1 = br.y;
                int idx = (sp[s0] - 1) + vd.dims0 * (sp[s1] - 1);
                int32_t dv = vd.diag_vertex[idx];
                double wgt = dv < 0 ? 0.0 : vd.weights[dv >> 1];
                double p_make = pm_raw / (double)(L - n);
                if (rs.next_float() < p_make * wgt) {
                    nop = op_make(b, (uint32_t)dv);
                    n += 1;
                }
            } else {
                uint32_t vc = op_vcode(op);
                int2 br = m.bond_rec[op_bond(op)];
                const DevVD &vd = m.types[br.x >> 20];
                if (vc & 1u) {
                    double wgt = vd.weights[vc >> 1];
                    double p_remove = (double)(L - n + 1) * pr_raw;
                    if (rs.next_float() * wgt < p_remove) {
                        nop = 0u;
                        n -= 1;
                    }
                } else {
                    const uint8_t *ls = vd.leg_states + 4 * (vc >> 1);
                    __syncthreads();
                    if (lane == 0) {
                        sp[br.x & 0xFFFFF] = ls[2];
                        sp[br.y] = ls[3];
                    }
                    __syncthreads();
                }
            }
        }
        if (do_vl && nop != 0u) {
            int2 br = m.bond_rec[op_bond(nop)];
            int sites[2] = {br.x & 0xFFFFF, br.y};
            const int32_t rk = (int32_t)nr;
            for (int k = 0; k < 2; ++k) {
                int s = sites[k];
                int32_t me = 4 * rk + k;
                int32_t lst = vlast[s];
                __syncthreads();
                if (lane == 0) {
                    if (lst >= 0) {
                        glk[lst] = (uint32_t)me;
                        glk[me] = (uint32_t)lst;
                    } else {
                        vfirst[s] = me;
                    }
                    vlast[s] = 4 * rk + 2 + k;
                }
                __syncthreads();
            }
            if (lane == 0)
                gvt[rk] = op_vcode(nop) | ((uint32_t)(br.x >> 20) << 10);
            cmask |= 1ull << i;
            nr += 1;
        }
        if (lane == 0 && nop != op)
            w.ops[p] = nop;
    }
    if (do_vl && lane == 0) {
        rec_mask(w)[base >> 6] = cmask;
        rec_base(w)[base >> 6] = (uint32_t)nr0;
    }
    __syncthreads();
}

// k1_serial_chunk out of line, everything by value (the caller's walker copy
// and stream stay in registers; the rare path gets a call instead of inlined
// code in the chunk loop).  Returns the counts and the stream position.
struct K1SerialOut {
    int64_t n, nr;
    uint64_t bbase;
    uint32_t off;
};
__device__ __noinline__ K1SerialOut k1_serial_chunk_ool(DevWalker w, const DevModel *m, WaveStream rs,
                                                        int64_t base, int64_t n, int64_t nr, double pm_raw,
                                                        double pr_raw, bool do_diag, bool do_vl, uint8_t *sp,
                                                        int32_t *vfirst, int32_t *vlast)
{
    k1_serial_chunk(w, *m, rs, base, n, nr, pm_raw, pr_raw, do_diag, do_vl, sp, vfirst, vlast);
    return K1SerialOut{n, nr, rs.bbase, rs.off};
}

// phases: 1 diagonal update, 2 vertex list, 8 force the slot-by-slot path.
// warm >= 0: init warm-up round (diagonal only; walkers with fewer warm-up
// sweeps are skipped).
template <class Tab>
__device__ __forceinline__ void k1_walker_t(DevWalker *W, const DevModel *models, DevPool *pool, int wi,
                                            int phases, int warm, char *lds, Tab tab, int grp)
{
    const int lane = threadIdx.x;
    DevWalker w = W[wi];
    if (w.error)
        return;
    // warm-up rounds touch pending walkers only; a sweep skips inactive
    // walkers and slots whose sweep budget ran out (retire set)
    if (warm >= 0 ? (warm >= w.warmup || !w.pending) : (!w.active || w.retire))
        return;
    const DevModel &m = models[w.model];
    tab.stage(m);
    const int N = m.n_sites;
    uint64_t *rbuf = (uint64_t *)lds;
    uint64_t *tm0 = (uint64_t *)(lds + 2048);
    uint64_t *tm1 = tm0 + N;
    int32_t *vfirst = (int32_t *)(tm1 + N);
    int32_t *vlast = vfirst + N;
    uint16_t *cups = (uint16_t *)(vlast + N);
    uint8_t *sp = (uint8_t *)(cups + 64);
    uint32_t *tm0w = (uint32_t *)tm0;
    uint32_t *tm1w = (uint32_t *)tm1;

    const bool do_diag = phases & 1;
    const bool do_vl = (phases & 2) && warm < 0;
    const bool force_serial = phases & 8;

    for (int i = lane; i < N; i += 64) {
        tm0[i] = 0ull;
        tm1[i] = 0ull;
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
                sp_addr = (unsigned long long)pool_alloc(pool, slab_bytes(cap));
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
    int64_t nr = 0;  // operators (ranks) emitted so far
    const uint64_t below = below_mask(lane), above = above_mask(lane);
    gmem<uint32_t> *gvt = as_global(rec_vt(w));
    gmem<uint32_t> *glk = as_global(rec_link(w));
    gmem<uint64_t> *gmask = as_global(rec_mask(w));
    gmem<uint32_t> *gbase = as_global(rec_base(w));
    gmem<uint32_t> *ops = as_global(w.ops);

    // The operator string reaches the chunks in blocks of K1_BLOCK chunks held
    // in registers (lane l: slot l of each chunk): a block's loads are issued a
    // whole block ahead and moved into place when the block starts, so the loop
    // waits on global memory once per block.  (A load consumed in the next
    // chunk would be waited on together with the chunk's global stores: the
    // memory counter does not order loads against stores.)  K1 writes only the
    // current chunk's slots, so the loaded values of later chunks stay current.
    // The chunk index within the block is wave-uniform, so cur[slot] is a
    // register-indexed move.  Model scalars are read once, before the loop,
    // for the same reason (a read inside it would wait for the stores).
    const uint32_t nbonds = (uint32_t)m.n_bonds;
    uint32_t cur[K1_BLOCK], pf[K1_BLOCK];
#pragma unroll
    for (int u = 0; u < K1_BLOCK; ++u) {
        const int64_t q = 64 * u + lane;
        cur[u] = q < L ? ops[q] : 0u;
    }
#pragma unroll
    for (int u = 0; u < K1_BLOCK; ++u) {
        const int64_t q = 64 * (K1_BLOCK + u) + lane;
        pf[u] = q < L ? ops[q] : 0u;
    }
    for (int64_t base = 0; base < L; base += 64) {
        const int64_t p = base + lane;
        const bool valid = p < L;
        const int64_t chunk = base >> 6;
        const int slot = (int)uni((uint32_t)(chunk % K1_BLOCK));
        if (slot == 0 && chunk > 0) {
#pragma unroll
            for (int u = 0; u < K1_BLOCK; ++u)
                cur[u] = pf[u];
#pragma unroll
            for (int u = 0; u < K1_BLOCK; ++u) {
                const int64_t q = base + 64 * (K1_BLOCK + u) + lane;
                pf[u] = q < L ? ops[q] : 0u;
            }
        }
        const uint32_t op = cur[slot];
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
                    bad = (int)(vc >> 1) >= tab.nv(tab.bond(b).x >> 20);
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
                const K1SerialOut so = k1_serial_chunk_ool(w, &m, rs, base, n0, nr, pm_raw, pr_raw, do_diag,
                                                           do_vl, sp, vfirst, vlast);
                n0 = so.n;
                nr = so.nr;
                rs.bbase = so.bbase;
                rs.off = so.off;
                continue;
            }
            rs.advance((uint32_t)D);
        } else if (force_serial) {
            const K1SerialOut so = k1_serial_chunk_ool(w, &m, rs, base, n0, nr, pm_raw, pr_raw, do_diag,
                                                       do_vl, sp, vfirst, vlast);
            n0 = so.n;
            nr = so.nr;
            rs.bbase = so.bbase;
            rs.off = so.off;
            continue;
        }

        const bool has = isop || (ident && do_diag);
        int s0 = 0, s1 = 0, type = 0;
        if (has) {
            int2 br = tab.bond(b);
            s0 = br.x & 0xFFFFF;
            s1 = br.y;
            type = br.x >> 20;
        }
        // per-site lane masks
        if (has) {
            atomicOr(&tm0w[2 * s0 + (lane >> 5)], 1u << (lane & 31));
            atomicOr(&tm1w[2 * s1 + (lane >> 5)], 1u << (lane & 31));
        }
        if (offd) {
            cups[lane] = (uint16_t)tab.upper(type, vc >> 1);
        }
        wave_lds_sync();
        uint64_t a0 = 0, a1 = 0, c0 = 0, c1 = 0;  // site0: pos-0/pos-1 masks; site1: same
        if (has) {
            a0 = tm0[s0];
            a1 = tm1[s0];
            c0 = tm0[s1];
            c1 = tm1[s1];
        }

        uint64_t plus = 0, minus = 0;
        uint32_t fop = op;
        if (do_diag) {
            double wgt = 0.0;
            int32_t dv = -1;
            if (ident) {
                int v0, v1;
                uint64_t cand = (a0 | a1) & offm & below;
                if (cand) {
                    int q = hi_bit(cand);
                    v0 = (cups[q] >> (((a0 >> q) & 1ull) ? 0 : 8)) & 255;
                } else {
                    v0 = sp[s0];
                }
                cand = (c0 | c1) & offm & below;
                if (cand) {
                    int q = hi_bit(cand);
                    v1 = (cups[q] >> (((c0 >> q) & 1ull) ? 0 : 8)) & 255;
                } else {
                    v1 = sp[s1];
                }
                dv = tab.diag(type, (v0 - 1) + tab.dims0(type) * (v1 - 1));
                wgt = dv < 0 ? 0.0 : tab.weight(type, dv >> 1);
            } else if (dia) {
                wgt = tab.weight(type, vc >> 1);
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

        // vertex list over the final non-identity operators of the chunk, in
        // compact form: operator rank r = nr + (non-identity lanes below)
        const uint64_t NI = do_diag ? ((identm & plus) | (diam & ~minus) | offm) : (diam | offm);
        const bool isni = (NI >> lane) & 1ull;
        const int32_t rank = (int32_t)(nr + popc(NI & below));
        int32_t pred0 = -1, pred1 = -1;
        bool last0 = false, last1 = false;
        if (do_vl && isni) {
            uint64_t m0 = (a0 | a1) & NI, m1 = (c0 | c1) & NI;
            uint64_t cand = m0 & below;
            if (cand) {
                int q = hi_bit(cand);
                pred0 = 4 * (int32_t)(nr + popc(NI & below_mask(q))) + 2 + (((a0 >> q) & 1ull) ? 0 : 1);
            } else {
                pred0 = vlast[s0];
                if (pred0 < 0)
                    pred0 = -2;  // first occurrence of the site
            }
            cand = m1 & below;
            if (cand) {
                int q = hi_bit(cand);
                pred1 = 4 * (int32_t)(nr + popc(NI & below_mask(q))) + 2 + (((c0 >> q) & 1ull) ? 0 : 1);
            } else {
                pred1 = vlast[s1];
                if (pred1 < 0)
                    pred1 = -2;
            }
            last0 = (m0 & above) == 0ull;
            last1 = (m1 & above) == 0ull;
        }
        wave_lds_sync();
        if (do_vl && isni) {
            const int32_t me0 = 4 * rank, me1 = 4 * rank + 1;
            if (pred0 >= 0) {
                glk[me0] = (uint32_t)pred0;
                glk[pred0] = (uint32_t)me0;
            } else {
                vfirst[s0] = me0;
            }
            if (pred1 >= 0) {
                glk[me1] = (uint32_t)pred1;
                glk[pred1] = (uint32_t)me1;
            } else {
                vfirst[s1] = me1;
            }
            if (last0)
                vlast[s0] = 4 * rank + 2;
            if (last1)
                vlast[s1] = 4 * rank + 3;
            gvt[rank] = op_vcode(fop) | ((uint32_t)type << 10);
        }
        if (do_vl && lane == 0) {
            gmask[base >> 6] = NI;
            gbase[base >> 6] = (uint32_t)nr;
        }
        nr += popc(NI);
        // spins after the chunk: the last off-diagonal operator on a site wins
        if (do_diag && offd) {
            if ((((a0 | a1) & offm) & above) == 0ull)
                sp[s0] = (uint8_t)(cups[lane] & 255);
            if ((((c0 | c1) & offm) & above) == 0ull)
                sp[s1] = (uint8_t)(cups[lane] >> 8);
        }
        if (has) {
            tm0[s0] = 0ull;
            tm1[s1] = 0ull;
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

// K1 of walker wi; the model's lookups come from LDS when the launch's LDS
// (lds_bytes) holds them
__device__ __forceinline__ void k1_walker(DevWalker *W, const DevModel *models, DevPool *pool, int wi,
                                          int phases, int warm, char *lds, int lds_bytes, int grp)
{
    const DevModel &m = models[W[wi].model];
    const size_t base = k1_lds_bytes(m.n_sites);
    if (k1_lds_staged(m.n_sites, m.n_bonds, m.k1tab_bytes) <= (size_t)lds_bytes) {
        K1Staged tab{(const int2 *)(lds + base),
                     (const uint8_t *)(lds + base + ((8 * (size_t)m.n_bonds + 15) & ~(size_t)15))};
        k1_walker_t(W, models, pool, wi, phases, warm, lds, tab, grp);
    } else {
        K1Global tab{&m};
        k1_walker_t(W, models, pool, wi, phases, warm, lds, tab, grp);
    }
}

// kernels of the per-sweep path take a walker range [w_begin, w_end), one
// workgroup (one wave) per walker
// Walker of workgroup b of a launch over positions [w_begin, w_end): with a
// schedule (order: walker indices, largest expected operator count first) the
// position indexes it, else it is the walker index itself.
__device__ __forceinline__ int walker_at(const int32_t *order, int pos)
{
    return order ? order[pos] : pos;
}

__global__ __launch_bounds__(64) void wave_diag_vlist_kernel(DevWalker *W, const DevModel *models,
                                                             DevPool *pool, int w_begin, int w_end,
                                                             int phases, int warm, int lds_bytes,
                                                             const int32_t *order, int grp)
{
    extern __shared__ __align__(16) char lds[];
    const int pos = w_begin + (int)blockIdx.x;
    if (pos < w_end)
        k1_walker(W, models, pool, walker_at(order, pos), phases, warm, lds, lds_bytes, grp);
}

// ------------------------------------------------------------------ K2
// Exit lists of the packed worm tables (DevModel::wtab) for transitions with
// more than four exits.
struct WormTab {
    const uint32_t *hdr, *trans, *eout;
    const double *cum;
};

__device__ __forceinline__ WormTab worm_tab(const char *base, const DevModel &m)
{
    WormTab t;
    t.hdr = (const uint32_t *)base;
    t.trans = (const uint32_t *)(base + m.wtab_trans_off);
    t.eout = (const uint32_t *)(base + m.wtab_eout_off);
    t.cum = (const double *)(base + m.wtab_cum_off);
    return t;
}

// Worm records of one walker during K2, indexed by operator rank r: the chase
// code ecode = type << dt_lgp | vcode and the four links (leg ids 4r'+leg').
// K1 leaves them in global memory (u32 vt = vcode | type << 10, uint4 links);
// K2 stages them in LDS when they fit:
//   Rec8   one u64 per rank: links in 14-bit fields, ecode in bits 56..63
//          (n <= 4096, at most 256 ecodes)
//   Rec10  u64 links in 16-bit fields + a u16 ecode array (n <= 16384)
//   RecG   global memory: ecode rewritten in place over vt, links read through
//          the scalar cache (K2 never writes them)
// load() returns wave-uniform words; put_ecode() is called by the whole wave
// with one value (all lanes store it).
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(3))) uint16_t lds_u16;
typedef __attribute__((address_space(3))) uint8_t lds_u8;
struct U4 {
    uint32_t x, y, z, w;
};
typedef const __attribute__((address_space(4))) U4 const_u4;
typedef const __attribute__((address_space(4))) char const_char;
typedef const __attribute__((address_space(4))) DirectEntry ConstEntry;

struct Rec8 {
    lds_u64 *rec;
    static constexpr bool kChecked = false;
    typedef uint32_t Key;  // leg id < 2^16: (leg id << 14 | worm-1) fits 32 bits
    struct Word {
        uint32_t lo, hi;
    };
    __device__ __forceinline__ Word load(uint32_t r) const
    {
        const uint64_t x = rec[r];
        return Word{uni((uint32_t)x), uni((uint32_t)(x >> 32))};
    }
    __device__ __forceinline__ static uint32_t ecode(const Word &x) { return x.hi >> 24; }
    __device__ __forceinline__ static uint32_t link(const Word &x, uint32_t leg)
    {
        const uint64_t v = (uint64_t)x.lo | ((uint64_t)x.hi << 32);
        return (uint32_t)(v >> (14u * leg)) & 0x3FFFu;
    }
    __device__ __forceinline__ static void set_link0(Word &x, uint32_t v) { x.lo = (x.lo & ~0x3FFFu) | v; }
    // link of the exit leg when the exit word carries 14 * leg in its low six
    // bits (one-worm tables): one 64-bit shift by the word itself
    __device__ __forceinline__ static uint32_t link_sh(const Word &x, uint32_t eo)
    {
        const uint64_t v = (uint64_t)x.lo | ((uint64_t)x.hi << 32);
        return (uint32_t)(v >> (eo & 63u)) & 0x3FFFu;
    }
    // every lane writes the same byte (no exec-mask switch in the chase loop)
    __device__ __forceinline__ void put_ecode(uint32_t r, uint32_t e) const
    {
        ((lds_u8 *)rec)[8u * r + 7u] = (uint8_t)e;
    }
    __device__ __forceinline__ uint32_t get_ecode(uint32_t r) const { return (uint32_t)(rec[r] >> 56); }
    // lookahead chase: each lane reads one record (lane-varying rank); the
    // wave then takes the word of one lane with v_readlane
    struct Cand {
        uint32_t lo, hi;
    };
    __device__ __forceinline__ Cand load_lane(uint32_t r) const
    {
        const uint64_t x = rec[r];
        return Cand{(uint32_t)x, (uint32_t)(x >> 32)};
    }
    __device__ __forceinline__ static Word pick(const Cand &c, uint32_t lane)
    {
        return Word{(uint32_t)__builtin_amdgcn_readlane((int)c.lo, (int)lane),
                    (uint32_t)__builtin_amdgcn_readlane((int)c.hi, (int)lane)};
    }
    __device__ __forceinline__ static void set_ecode(Word &x, uint32_t e)
    {
        x.hi = (x.hi & 0x00FFFFFFu) | (e << 24);
    }
};

struct Rec10 {
    lds_u64 *lk;
    lds_u16 *ec;
    static constexpr bool kChecked = false;
    typedef uint32_t Key;
    struct Word {
        uint32_t lo, hi, e;
    };
    __device__ __forceinline__ Word load(uint32_t r) const
    {
        const uint64_t x = lk[r];
        const uint32_t e = ec[r];
        return Word{uni((uint32_t)x), uni((uint32_t)(x >> 32)), uni(e)};
    }
    __device__ __forceinline__ static uint32_t ecode(const Word &x) { return x.e; }
    __device__ __forceinline__ static uint32_t link(const Word &x, uint32_t leg)
    {
        const uint64_t v = (uint64_t)x.lo | ((uint64_t)x.hi << 32);
        return (uint32_t)(v >> (16u * leg)) & 0xFFFFu;
    }
    __device__ __forceinline__ static void set_link0(Word &x, uint32_t v) { x.lo = (x.lo & ~0xFFFFu) | v; }
    __device__ __forceinline__ void put_ecode(uint32_t r, uint32_t e) const { ec[r] = (uint16_t)e; }
    __device__ __forceinline__ uint32_t get_ecode(uint32_t r) const { return ec[r]; }
    struct Cand {
        uint32_t lo, hi, e;
    };
    __device__ __forceinline__ Cand load_lane(uint32_t r) const
    {
        const uint64_t x = lk[r];
        return Cand{(uint32_t)x, (uint32_t)(x >> 32), (uint32_t)ec[r]};
    }
    __device__ __forceinline__ static Word pick(const Cand &c, uint32_t lane)
    {
        return Word{(uint32_t)__builtin_amdgcn_readlane((int)c.lo, (int)lane),
                    (uint32_t)__builtin_amdgcn_readlane((int)c.hi, (int)lane),
                    (uint32_t)__builtin_amdgcn_readlane((int)c.e, (int)lane)};
    }
    __device__ __forceinline__ static void set_ecode(Word &x, uint32_t e) { x.e = e; }
};

struct RecG {
    uint32_t *ec;
    const_u4 *lk;
    static constexpr bool kChecked = true;
    typedef uint64_t Key;
    struct Word {
        U4 l;
        uint32_t e;
    };
    __device__ __forceinline__ Word load(uint32_t r) const
    {
        const_u4 &q = lk[r];
        return Word{U4{q.x, q.y, q.z, q.w}, uni(ec[r])};
    }
    __device__ __forceinline__ static uint32_t ecode(const Word &x) { return x.e; }
    __device__ __forceinline__ static uint32_t link(const Word &x, uint32_t leg)
    {
        return leg == 0 ? x.l.x : leg == 1 ? x.l.y : leg == 2 ? x.l.z : x.l.w;
    }
    __device__ __forceinline__ static void set_link0(Word &x, uint32_t v) { x.l.x = v; }
    // every lane stores, so each lane's own later load is ordered after it
    __device__ __forceinline__ void put_ecode(uint32_t r, uint32_t e) const { ec[r] = e; }
    __device__ __forceinline__ uint32_t get_ecode(uint32_t r) const { return ec[r]; }
};

// Hybrid for walkers too large for Rec8/Rec10: chase codes in LDS (u16), links
// in global memory through the scalar cache (K2 never writes them)
struct RecH {
    lds_u16 *ec;
    const_u4 *lk;
    static constexpr bool kChecked = true;
    typedef uint64_t Key;
    struct Word {
        U4 l;
        uint32_t e;
    };
    __device__ __forceinline__ Word load(uint32_t r) const
    {
        const_u4 &q = lk[r];
        return Word{U4{q.x, q.y, q.z, q.w}, uni(ec[r])};
    }
    __device__ __forceinline__ static uint32_t ecode(const Word &x) { return x.e; }
    __device__ __forceinline__ static uint32_t link(const Word &x, uint32_t leg)
    {
        return leg == 0 ? x.l.x : leg == 1 ? x.l.y : leg == 2 ? x.l.z : x.l.w;
    }
    __device__ __forceinline__ static void set_link0(Word &x, uint32_t v) { x.l.x = v; }
    __device__ __forceinline__ void put_ecode(uint32_t r, uint32_t e) const { ec[r] = (uint16_t)e; }
    __device__ __forceinline__ uint32_t get_ecode(uint32_t r) const { return ec[r]; }
};

// The four links of record r read by one vector load (lane j reads link
// j & 3; lane-varying addresses are never turned into scalar loads) and made
// wave-uniform with v_readlane.  Used inside the persistent launch, where the
// diagonal/vertex-list phase rewrites the records between worm phases: the
// scalar cache is not kept coherent with those vector stores within a launch.
__device__ __forceinline__ U4 links_vec(const uint32_t *lk4, uint32_t r)
{
    const uint32_t v = as_global(lk4)[4u * r + (threadIdx.x & 3u)];
    return U4{(uint32_t)__builtin_amdgcn_readlane((int)v, 0),
              (uint32_t)__builtin_amdgcn_readlane((int)v, 1),
              (uint32_t)__builtin_amdgcn_readlane((int)v, 2),
              (uint32_t)__builtin_amdgcn_readlane((int)v, 3)};
}

// RecG with vector link loads (persistent launch)
struct RecGV {
    uint32_t *ec;
    const uint32_t *lk;
    static constexpr bool kChecked = true;
    typedef uint64_t Key;
    typedef RecG::Word Word;
    __device__ __forceinline__ Word load(uint32_t r) const
    {
        const U4 l = links_vec(lk, r);
        // the code: lane j reads code r + (j & 1) (r + 1 stays inside the
        // slab: the link region follows the code array), so the load is a
        // vector load as well
        const uint32_t e = as_global(ec)[r + (threadIdx.x & 1u)];
        return Word{l, (uint32_t)__builtin_amdgcn_readlane((int)e, 0)};
    }
    __device__ __forceinline__ static uint32_t ecode(const Word &x) { return x.e; }
    __device__ __forceinline__ static uint32_t link(const Word &x, uint32_t leg) { return RecG::link(x, leg); }
    __device__ __forceinline__ static void set_link0(Word &x, uint32_t v) { x.l.x = v; }
    __device__ __forceinline__ void put_ecode(uint32_t r, uint32_t e) const { ec[r] = e; }
    __device__ __forceinline__ uint32_t get_ecode(uint32_t r) const { return ec[r]; }
};

// RecH with vector link loads (persistent launch)
struct RecHV {
    lds_u16 *ec;
    const uint32_t *lk;
    static constexpr bool kChecked = true;
    typedef uint64_t Key;
    typedef RecH::Word Word;
    __device__ __forceinline__ Word load(uint32_t r) const { return Word{links_vec(lk, r), uni(ec[r])}; }
    __device__ __forceinline__ static uint32_t ecode(const Word &x) { return x.e; }
    __device__ __forceinline__ static uint32_t link(const Word &x, uint32_t leg) { return RecH::link(x, leg); }
    __device__ __forceinline__ static void set_link0(Word &x, uint32_t v) { x.l.x = v; }
    __device__ __forceinline__ void put_ecode(uint32_t r, uint32_t e) const { ec[r] = (uint16_t)e; }
    __device__ __forceinline__ uint32_t get_ecode(uint32_t r) const { return ec[r]; }
};

// Record index of the leg id held shifted in lidS (lid << S).  For the 8-byte
// LDS records it is produced on the vector unit (v_bfe), so the record's LDS
// address takes no scalar instruction and no scalar-to-vector move: the scalar
// unit is what the chase runs out of when several walkers share a compute unit.
template <class R, uint32_t S>
struct RecIndex {
    __device__ __forceinline__ static uint32_t of(typename R::Key lidS) { return (uint32_t)(lidS >> (S + 2)); }
};
template <uint32_t S>
struct RecIndex<Rec8, S> {
    __device__ __forceinline__ static uint32_t of(uint32_t lidS)
    {
        uint32_t r;
        asm("v_bfe_u32 %0, %1, %2, 16" : "=v"(r) : "s"(lidS), "i"(S + 2));
        return r;
    }
};

// Exit selection of one scatter on the scalar unit: m = word >> 11 against the
// entry's first three thresholds by 64-bit borrow chains (s_sub/s_subb leave
// the borrow, m < K, in SCC), from the last to the first: out[j] for the first
// j < 3 with m < K[j], else out[3].  The table puts the chase's slow marker
// (kChaseSlow) wherever the answer is not simply out[3] (a fourth exit whose
// threshold is below 2^53, more than four exits, a fall-through), so the
// caller tests one bit of the result.
__device__ __forceinline__ uint32_t chase_select(uint32_t mlo, uint32_t mhi, ConstEntry *E)
{
    // K[3] is loaded too (unused by the selection) so that the thresholds
    // arrive in one 32-byte scalar load
    const uint64_t K0 = E->K[0], K1 = E->K[1], K2 = E->K[2], K3 = E->K[3];
    const uint32_t o0 = E->out[0], o1 = E->out[1], o2 = E->out[2], o3 = E->out[3];
    uint32_t eo, t;
    asm("s_sub_u32 %1, %2, %8\n\t"
        "s_subb_u32 %1, %3, %9\n\t"
        "s_cselect_b32 %0, %12, %13\n\t"
        "s_sub_u32 %1, %2, %6\n\t"
        "s_subb_u32 %1, %3, %7\n\t"
        "s_cselect_b32 %0, %11, %0\n\t"
        "s_sub_u32 %1, %2, %4\n\t"
        "s_subb_u32 %1, %3, %5\n\t"
        "s_cselect_b32 %0, %10, %0"
        : "=&s"(eo), "=&s"(t)
        : "s"(uni(mlo)), "s"(uni(mhi)), "s"((uint32_t)K0), "s"((uint32_t)(K0 >> 32)), "s"((uint32_t)K1),
          "s"((uint32_t)(K1 >> 32)), "s"((uint32_t)K2), "s"((uint32_t)(K2 >> 32)), "s"(o0), "s"(o1),
          "s"(o2), "s"(o3), "s"((uint32_t)K3), "s"((uint32_t)(K3 >> 32))
        : "scc");
    return eo;
}

// Worm start (sse.jl:241-251): draw (slot, leg) pairs until the leg belongs
// to an operator (an operator's four legs are all linked), then the worm
// function.  Returns the operator's rank r0, leg l0, the site's dimension and
// the worm function wf0 (1-based); false when wf0 exceeds the chase table.
__device__ __forceinline__ bool worm_start(const DevWalker &w, const DevModel &m, K2Stream &rs,
                                           const gmem<uint64_t> *MASK, const gmem<uint32_t> *BASE,
                                           uint32_t dmw, uint32_t &r0, uint32_t &l0, uint32_t &dim0,
                                           uint32_t &wf0)
{
    const int64_t L = w.L;
    int64_t p0;
    for (;;) {
        p0 = (int64_t)uni64(rs.range((uint64_t)L));
        l0 = uni((uint32_t)rs.range(4));
        const uint64_t mk = uni64(MASK[p0 >> 6]);
        const int bit = (int)(p0 & 63);
        if ((mk >> bit) & 1ull) {
            r0 = uni(BASE[p0 >> 6] + (uint32_t)popc(mk & below_mask(bit)));
            break;
        }
    }
    const int2 b0 = m.bond_rec[op_bond(w.ops[p0])];
    const uint32_t site0 = uni((uint32_t)((l0 & 1) ? b0.y : (b0.x & 0xFFFFF)));
    dim0 = uni(m.site_dims[site0]);
    wf0 = 1u + uni((uint32_t)rs.range((uint64_t)(dim0 - 1)));
    return wf0 - 1u < dmw;
}

// The exact scatter (vertex_data.jl:113-125) of a step whose chase entry
// carries the slow marker: the entry's own thresholds for up to four exits,
// the exit list for more.  A fall-through or an invalid transition sets err
// and returns efail (the code rewritten unchanged, exit leg 0, the worm field
// of the closing test), which the caller turns into a loop exit.
template <bool kOneWorm>
__device__ __forceinline__ uint32_t chase_exact(const DevModel &m, const_char *CT, uint32_t eoff,
                                                uint32_t mlo, uint32_t mhi, uint32_t e, uint32_t lgp,
                                                uint32_t efail, uint32_t &err)
{
    ConstEntry *ES = (ConstEntry *)(CT + eoff);
    const uint32_t cnt = uni(ES->count);
    if (cnt <= 4u) {
        const uint64_t mm = ((uint64_t)mhi << 32) | mlo;
        uint32_t j = 0;
        for (; j < cnt; ++j)
            if (mm < uni64(ES->K[j]))
                break;
        if (j >= cnt) {
            err = SSE_ERR_SCATTER;
            return efail;
        }
        return uni(j < 3u ? ES->out[j] : ES->pad[0]);
    }
    const uint32_t trw = uni(ES->trans);
    const WormTab tab = worm_tab((const char *)m.wtab, m);
    const double xr = (double)(((uint64_t)mhi << 32) | mlo) * 0x1.0p-53;
    uint32_t ex = trw & 0xFFFFFu;
    const uint32_t eend = ex + (trw >> 20);
    for (; ex < eend; ++ex) {
        const uint64_t cb = uni64((uint64_t)__double_as_longlong(tab.cum[ex]));
        if (xr < __longlong_as_double((long long)cb))
            break;
    }
    if (ex >= eend) {
        err = SSE_ERR_SCATTER;
        return efail;
    }
    const uint32_t vc = uni(tab.eout[ex]);
    const uint32_t ecn = (e >> lgp << lgp) | (vc & 1023u), lg = (vc >> 10) & 3u;
    return kOneWorm ? chase_out1(ecn, lg) : chase_out(ecn, lg, vc >> 12);
}

// The worm update of one walker (sse.jl:193-303), executed uniformly by the
// whole wave.  The chase is scalar work: the record and the next Philox word
// leave LDS together and become wave-uniform at once, the transition entry
// arrives in one scalar load (constant address space, uniform offset), the exit
// is selected on the scalar unit, and the closing tests compare packed keys
// (leg id << 14 | worm-1).  Returns an error code (0 = ok).
template <class R, bool kOneWorm>
__device__ __forceinline__ int worm_chase(DevWalker &w, const DevModel &m, K2Stream &rs, const R &rec,
                          const gmem<uint64_t> *MASK, const gmem<uint32_t> *BASE, double &total,
                          int64_t &scat)
{
    const uint32_t n = (uint32_t)w.n;
    const int64_t nworms = (int64_t)ceil(w.num_worms);
    const uint32_t dmw = uni((uint32_t)m.dt_mw);
    const uint32_t lgp = uni((uint32_t)m.dt_lgp);
    const uint32_t estride = dmw << 8;  // bytes per ecode
    const_char *CT = (const_char *)(const void *)uni64((uint64_t)(const void *)m.dtab);
    total = 1.0;
    uint32_t sc = 0;
    for (int64_t k = 0; k < nworms; ++k) {
        if (n == 0)
            continue;  // worm_traverse! returns 0 without draws (sse.jl:234-236)
        uint32_t r0, l0, dim0, wf0;
        if (!worm_start(w, m, rs, MASK, BASE, dmw, r0, l0, dim0, wf0))
            return SSE_ERR_STATE;
        const uint32_t lid0 = 4u * r0 + l0;
        // closing tests (sse.jl:290-297), both on the key of the next leg: the
        // worm leaves through (r0, l0) with worm dim0 - wf0 exactly when the hop
        // lands on the leg linked to (r0, l0) (the links are an involution
        // without fixed points, so at most one test holds), and it re-enters
        // (r0, l0) with worm wf0.  Key = lid << S | worm-1; with one worm type
        // the worm field is always 0 and S = 6 puts the leg at bits 6..7, the
        // entry's leg offset in the transition table.
        constexpr uint32_t S = kOneWorm ? 6u : 14u;
        typedef typename R::Key Key;
        const uint32_t lnk0 = uni(R::link(rec.load(r0), l0));
        const Key keyA = ((Key)lnk0 << S) | (kOneWorm ? 0u : dim0 - wf0 - 1u);
        const Key keyB = ((Key)lid0 << S) | (kOneWorm ? 0u : wf0 - 1u);
        Key lidS = (Key)lid0 << S;
        uint32_t wfm = wf0 - 1u;
        // one Philox word per step: the step count is the stream advance
        const uint64_t pos0 = rs.bbase + rs.off;
        uint32_t err = 0;
        Key kB = 0;
        do {  // sse.jl:274-300
            // the rare paths (refill, exact scatter, worm end) are marked
            // unlikely so the step runs straight through with one taken branch
            if (__builtin_expect(rs.off >= K2Stream::kNW, 0)) {
                if (rs.off == K2Stream::kNW)
                    rs.refill_next();
                else
                    rs.ensure(1);
            }
            const uint32_t off = uni(rs.off);
            const uint32_t r = RecIndex<R, S>::of(lidS);
            typename R::Word x = rec.load(r);
            // rand() = mm * 2^-53: the mantissa of this step's word
            const uint32_t mlo = (uint32_t)__builtin_amdgcn_readlane((int)rs.mlo, (int)off);
            const uint32_t mhi = (uint32_t)__builtin_amdgcn_readlane((int)rs.mhi, (int)off);
            rs.off = off + 1u;
            const uint32_t e = R::ecode(x);
            const uint32_t legbits = kOneWorm ? (uint32_t)lidS & 0xC0u          // leg_in << 6
                                              : (uint32_t)(lidS >> 8) & 0xC0u;
            const uint32_t eoff = kOneWorm ? (e << 8) | legbits : e * estride + (wfm << 8) + legbits;
            ConstEntry *E = (ConstEntry *)(CT + eoff);
            uint32_t eo = chase_select(mlo, mhi, E);
            if (__builtin_expect((eo & kChaseSlowBit) != 0, 0)) {
                // the entry's address is formed here, not hoisted into the fast path
                uint32_t eoff_s = eoff;
                asm volatile("" : "+s"(eoff_s));
                const uint32_t efail = (e << 16) | (kOneWorm ? 0u : dim0 - wf0 - 1u);
                eo = chase_exact<kOneWorm>(m, CT, eoff_s, mlo, mhi, e, lgp, efail, err);
                if (err)  // exit leg 0 of the record copy becomes the leg linked to (r0, l0)
                    R::set_link0(x, lnk0);
            }
            rec.put_ecode(r, eo >> 16);
            uint32_t nx;
            if constexpr (kOneWorm && std::is_same<R, Rec8>::value)
                nx = Rec8::link_sh(x, eo);
            else
                nx = R::link(x, (eo >> 14) & 3u);
            if (R::kChecked && (nx >> 2) >= n) {
                err = SSE_ERR_STATE;
                break;
            }
            const uint32_t wfn = kOneWorm ? 0u : eo & 0x3FFFu;
            const Key nxS = (Key)nx << S;
            kB = kOneWorm ? nxS : nxS | wfn;
            lidS = nxS;
            wfm = wfn;
        } while (__builtin_expect(min(kB ^ keyA, kB ^ keyB) != 0, 1));
        const uint32_t endB = kB != keyA;  // closed on re-entry
        if (err)
            return (int)err;
        const uint32_t steps = (uint32_t)(rs.bbase + rs.off - pos0);
        sc += steps;
        // worm length: scatters + 1 when the worm closed on re-entry
        total += (double)(steps + endB);
    }
    scat = sc;
    return 0;
}

// The worm update with record lookahead, for the LDS record layouts (Rec8,
// Rec10).  Same steps, draws and results as worm_chase; what changes is when
// the next record is read.  At the top of a step the current record x is
// wave-uniform, so lane j (j = lane % 4) reads, with one LDS instruction, the
// record of the operator linked to leg j of x.  While those loads are in
// flight the scalar unit selects the exit; the next record is then the word of
// lane exit-leg (v_readlane), and the LDS latency leaves the step's
// dependency chain.  The loads were issued before this step rewrote the
// current operator's chase code, so a hop onto the same operator (a site
// with one operator, linked to itself through the periodic closure) patches
// the code into the picked record.
template <class R, bool kOneWorm>
__device__ __forceinline__ int worm_chase_la(DevWalker &w, const DevModel &m, K2Stream &rs, const R &rec,
                             const gmem<uint64_t> *MASK, const gmem<uint32_t> *BASE, double &total,
                             int64_t &scat)
{
    const uint32_t n = (uint32_t)w.n;
    const int64_t nworms = (int64_t)ceil(w.num_worms);
    const uint32_t dmw = uni((uint32_t)m.dt_mw);
    const uint32_t lgp = uni((uint32_t)m.dt_lgp);
    const uint32_t estride = dmw << 8;  // bytes per ecode
    const_char *CT = (const_char *)(const void *)uni64((uint64_t)(const void *)m.dtab);
    const uint32_t myleg = threadIdx.x & 3u;
    total = 1.0;
    uint32_t sc = 0;
    for (int64_t k = 0; k < nworms; ++k) {
        if (n == 0)
            continue;  // worm_traverse! returns 0 without draws (sse.jl:234-236)
        uint32_t r0, l0, dim0, wf0;
        if (!worm_start(w, m, rs, MASK, BASE, dmw, r0, l0, dim0, wf0))
            return SSE_ERR_STATE;
        const uint32_t lid0 = 4u * r0 + l0;
        constexpr uint32_t S = kOneWorm ? 6u : 14u;
        typedef typename R::Key Key;
        typename R::Word x = rec.load(r0);
        const uint32_t lnk0 = uni(R::link(x, l0));
        // closing tests as in worm_chase
        const Key keyA = ((Key)lnk0 << S) | (kOneWorm ? 0u : dim0 - wf0 - 1u);
        const Key keyB = ((Key)lid0 << S) | (kOneWorm ? 0u : wf0 - 1u);
        Key lidS = (Key)lid0 << S;
        uint32_t wfm = wf0 - 1u;
        const uint64_t pos0 = rs.bbase + rs.off;
        uint32_t err = 0;
        Key kB = 0;
        do {  // sse.jl:274-300
            // the records linked to the current operator's four legs
            const typename R::Cand cand = rec.load_lane(R::link(x, myleg) >> 2);
            if (__builtin_expect(rs.off >= K2Stream::kNW, 0)) {
                if (rs.off == K2Stream::kNW)
                    rs.refill_next();
                else
                    rs.ensure(1);
            }
            const uint32_t off = uni(rs.off);
            const uint32_t r = RecIndex<R, S>::of(lidS);
            const uint32_t mlo = (uint32_t)__builtin_amdgcn_readlane((int)rs.mlo, (int)off);
            const uint32_t mhi = (uint32_t)__builtin_amdgcn_readlane((int)rs.mhi, (int)off);
            rs.off = off + 1u;
            const uint32_t e = R::ecode(x);
            const uint32_t legbits = kOneWorm ? (uint32_t)lidS & 0xC0u
                                              : (uint32_t)(lidS >> 8) & 0xC0u;
            const uint32_t eoff = kOneWorm ? (e << 8) | legbits : e * estride + (wfm << 8) + legbits;
            ConstEntry *E = (ConstEntry *)(CT + eoff);
            uint32_t eo = chase_select(mlo, mhi, E);
            if (__builtin_expect((eo & kChaseSlowBit) != 0, 0)) {
                uint32_t eoff_s = eoff;
                asm volatile("" : "+s"(eoff_s));
                const uint32_t efail = (e << 16) | (kOneWorm ? 0u : dim0 - wf0 - 1u);
                eo = chase_exact<kOneWorm>(m, CT, eoff_s, mlo, mhi, e, lgp, efail, err);
                if (err)
                    R::set_link0(x, lnk0);
            }
            const uint32_t enew = eo >> 16;
            rec.put_ecode(r, enew);
            const uint32_t leg = (eo >> 14) & 3u;
            uint32_t nx;
            if constexpr (kOneWorm && std::is_same<R, Rec8>::value)
                nx = Rec8::link_sh(x, eo);
            else
                nx = R::link(x, leg);
            typename R::Word xn = R::pick(cand, leg);
            // (the record index on the scalar unit: r itself is a vector
            // register, the LDS address of the code store)
            if ((nx >> 2) == (uint32_t)(lidS >> (S + 2)))
                R::set_ecode(xn, enew);
            const uint32_t wfn = kOneWorm ? 0u : eo & 0x3FFFu;
            const Key nxS = (Key)nx << S;
            kB = kOneWorm ? nxS : nxS | wfn;
            lidS = nxS;
            wfm = wfn;
            x = xn;
        } while (__builtin_expect(min(kB ^ keyA, kB ^ keyB) != 0, 1));
        const uint32_t endB = kB != keyA;
        if (err)
            return (int)err;
        const uint32_t steps = (uint32_t)(rs.bbase + rs.off - pos0);
        sc += steps;
        total += (double)(steps + endB);
    }
    scat = sc;
    return 0;
}

// state refresh (sse.jl:219-228) and operator write-back after the worms
template <class R>
__device__ __noinline__ void worm_writeback(R rec, uint32_t lgp, uint32_t *ops_p, int64_t L,
                                            const gmem<uint64_t> *MASK, const gmem<uint32_t> *BASE);

template <class R, class RS>
__device__ __forceinline__ void worm_finish(DevWalker &w, const DevModel &m, RS &rs, const R &rec,
                            const gmem<uint64_t> *MASK, const gmem<uint32_t> *BASE)
{
    const int lane = threadIdx.x;
    const int N = m.n_sites;
    const uint32_t lgp = (uint32_t)m.dt_lgp;
    // sites in blocks of 32 so that one block draws at most 32 words (the
    // stream holds 64)
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
            // every lane takes part in the cross-lane read; lane s of the
            // refreshed sites uses word off + (its rank among them)
            const uint64_t x = rs.peek(fr ? (uint32_t)popc(frm & below_mask(lane)) : 0u);
            if (fr) {
                uint64_t lo = x * dim;
                if (lo < dim && lo < (0ull - dim) % dim)
                    reject = true;
                val = (uint8_t)(1 + __umul64hi(x, dim));
            }
        }
        if (vs && !fr && (int64_t)(f >> 2) < w.n) {
            const uint32_t vt = vt_of_ecode(rec.get_ecode((uint32_t)f >> 2), lgp);
            const DevVD &vd = m.types[vt >> 10];
            val = vd.leg_states[4 * ((vt & 1023u) >> 1) + (f & 3)];
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
    worm_writeback<R>(rec, lgp, w.ops, w.L, MASK, BASE);
}

// Write the worm-updated vertex codes back into the operator string (the
// last step of the worm update), K 64-slot chunks per pass so their loads are
// in flight together: lane u loads chunk u's mask and base (vector loads, read
// back by v_readlane), every lane its slot of each chunk.  A function of its
// own: inlined, its registers would crowd the chase loop's scalar operands.
template <class R>
__device__ __noinline__ void worm_writeback(R rec, uint32_t lgp, uint32_t *ops_p, int64_t L,
                                            const gmem<uint64_t> *MASK, const gmem<uint32_t> *BASE)
{
    const int lane = threadIdx.x;
    constexpr int K = 16;
    gmem<uint32_t> *ops = as_global(ops_p);
    for (int64_t base = 0; base < L; base += 64 * K) {
        const int64_t cl = (base >> 6) + lane;
        const bool lin = lane < K && base + 64 * lane < L;
        const uint64_t mkl = lin ? MASK[cl] : 0ull;
        const uint32_t bsl = lin ? BASE[cl] : 0u;
        uint32_t op[K];
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int64_t p = base + 64 * u + lane;
            op[u] = p < L ? ops[p] : 0u;
        }
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const uint64_t mk = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(mkl >> 32), u) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mkl, u);
            if ((mk >> lane) & 1ull) {
                const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)bsl, u) +
                                   (uint32_t)popc(mk & below_mask(lane));
                const uint32_t vc = vt_of_ecode(rec.get_ecode(r), lgp) & 1023u;
                const uint32_t nop = (op[u] & ~(1023u << 1)) | (vc << 1);
                if (nop != op[u])
                    ops[base + 64 * u + lane] = nop;
            }
        }
    }
}

// LDS bytes the worm kernel needs for a walker with n operators, cutoff L, in
// the Rec8 (fmt 8) or Rec10 (fmt 10) layout
__host__ __device__ inline size_t k2_lds_need(int64_t n, int64_t, int fmt)
{
    const size_t b = 0;  // the Philox words live in registers
    if (fmt == 2)                        // RecH: chase codes only
        return b + (((size_t)n * 2 + 15) & ~(size_t)15);
    return b + (size_t)n * 8 + (fmt == 10 ? (((size_t)n * 2 + 15) & ~(size_t)15) : 0);
}

// LDS per walker the worm kernel is sized for at most: beyond it a walker uses
// the hybrid or global records instead of lowering every walker's residency
constexpr size_t K2_LDS_CAP = 65536;

// Stage the compact records of K1 (global) into the LDS layout R, converting
// vt to ecodes and checking every ecode and link; returns false on a bad record.
// K 64-record chunks per pass, their loads in flight together; a function of
// its own (arguments by value) so that its registers stay out of the chase's.
template <int FMT>
__device__ __noinline__ bool k2_stage_records(uint32_t n, uint32_t lgp, uint32_t nt, const uint32_t *vt_p,
                                              const U4 *lk_p, lds_u64 *lk, lds_u16 *ec)
{
    constexpr int K = 16;
    const int lane = threadIdx.x;
    const gmem<uint32_t> *gvt = as_global(vt_p);
    const gmem<U4> *glk = as_global(lk_p);
    const uint32_t lmax = 4u * n;
    bool bad = false;
    for (uint32_t i0 = 0; i0 < n; i0 += 64 * K) {
        uint32_t vt[K];
        uint4 l4[K];
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const uint32_t i = i0 + 64u * u + lane;
            if (i < n) {
                vt[u] = gvt[i];
                const gmem<U4> &q = glk[i];
                l4[u] = make_uint4(q.x, q.y, q.z, q.w);
            }
        }
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const uint32_t i = i0 + 64u * u + lane;
            if (i < n) {
                const uint32_t e = ecode_of_vt(vt[u], lgp);
                bad |= (vt[u] & 1023u) >= (1u << lgp) || (vt[u] >> 10) >= nt || l4[u].x >= lmax ||
                       l4[u].y >= lmax || l4[u].z >= lmax || l4[u].w >= lmax;
                if (FMT == 8) {
                    lk[i] = (uint64_t)l4[u].x | ((uint64_t)l4[u].y << 14) | ((uint64_t)l4[u].z << 28) |
                            ((uint64_t)l4[u].w << 42) | ((uint64_t)e << 56);
                } else {
                    lk[i] = (uint64_t)l4[u].x | ((uint64_t)l4[u].y << 16) | ((uint64_t)l4[u].z << 32) |
                            ((uint64_t)l4[u].w << 48);
                    ec[i] = (uint16_t)e;
                }
            }
        }
    }
    return __ballot(bad) == 0;
}

template <int FMT>
__device__ __forceinline__ bool k2_stage(const DevWalker &w, const DevModel &m, lds_u64 *lk, lds_u16 *ec)
{
    return k2_stage_records<FMT>((uint32_t)w.n, (uint32_t)m.dt_lgp, (uint32_t)m.n_types,
                                 (const uint32_t *)rec_vt(w), (const U4 *)rec_link(w), lk, ec);
}

// vt -> ecode into an LDS u16 array (RecH path), checking every code.
__device__ bool k2_stage_codes(const DevWalker &w, const DevModel &m, lds_u16 *ec)
{
    const int lane = threadIdx.x;
    const uint32_t n = (uint32_t)w.n, lgp = (uint32_t)m.dt_lgp, nt = (uint32_t)m.n_types;
    const gmem<uint32_t> *gvt = as_global((const uint32_t *)rec_vt(w));
    bool bad = false;
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t vt = gvt[i];
        bad |= (vt & 1023u) >= (1u << lgp) || (vt >> 10) >= nt;
        ec[i] = (uint16_t)ecode_of_vt(vt, lgp);
    }
    return __ballot(bad) == 0;
}

// In-place vt -> ecode conversion of the global records (RecG path).
__device__ bool k2_stage_global(const DevWalker &w, const DevModel &m)
{
    const int lane = threadIdx.x;
    const uint32_t n = (uint32_t)w.n, lgp = (uint32_t)m.dt_lgp, nt = (uint32_t)m.n_types;
    uint32_t *gvt = rec_vt(w);
    bool bad = false;
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t vt = gvt[i];
        bad |= (vt & 1023u) >= (1u << lgp) || (vt >> 10) >= nt;
        gvt[i] = ecode_of_vt(vt, lgp);
    }
    return __ballot(bad) == 0;
}

// Worm update of walker wi.  LDS holds the records only (the Philox words are
// in registers).  RecG reads links through the scalar cache, which is
// valid because K1 wrote them in an earlier kernel launch.
// kFused: called inside the persistent launch (global link records read by
// vector loads, RecHV / RecGV)
template <bool kFused>
__device__ __forceinline__ void k2_walker(DevWalker *W, const DevModel *models, DevPool *pool, int wi,
                                          int full_sweep, int lds_bytes, int rec_mode, char *lds,
                                          int chase)
{
    const int lane = threadIdx.x;
    DevWalker w = W[wi];
    if (w.error || !w.active || w.retire)
        return;
    const DevModel &m = models[w.model];
    const bool therm = w.sweeps > w.thermalization;
    K2Stream rs;
    rs.init(w.key, w.ctr);
    const bool small_codes = ((int64_t)m.n_types << m.dt_lgp) <= 256;
    const bool u16_codes = ((int64_t)m.n_types << m.dt_lgp) <= 65536;
    int fmt = 0;
    if (rec_mode == 0 && w.n <= 4096 && small_codes && k2_lds_need(w.n, w.L, 8) <= (size_t)lds_bytes)
        fmt = 8;
    else if (rec_mode <= 1 && w.n <= 16384 && u16_codes && k2_lds_need(w.n, w.L, 10) <= (size_t)lds_bytes)
        fmt = 10;
    else if ((rec_mode <= 1 || rec_mode == 3) && u16_codes && k2_lds_need(w.n, w.L, 2) <= (size_t)lds_bytes)
        fmt = 2;
    // chunk masks and bases stay in global memory (worm starts and write-back)
    const gmem<uint64_t> *gmask = as_global((const uint64_t *)rec_mask(w));
    const gmem<uint32_t> *gbase = as_global((const uint32_t *)rec_base(w));
    double total = 1.0;
    int64_t scat = 0;
    int err = 0;
    // records at the start of the walker's LDS
    char *q = lds;
    if (fmt == 2) {
        lds_u16 *lec = (lds_u16 *)(uint16_t *)q;
        const bool ok = k2_stage_codes(w, m, lec);
        __syncthreads();
        if (!ok) {
            err = SSE_ERR_STATE;
        } else {
            if constexpr (kFused) {
                RecHV rec{lec, (const uint32_t *)rec_link(w)};
                err = m.dt_mw == 1 ? worm_chase<RecHV, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase<RecHV, false>(w, m, rs, rec, gmask, gbase, total, scat);
                __syncthreads();
                if (!err)
                    worm_finish(w, m, rs, rec, gmask, gbase);
            } else {
                RecH rec{lec, (const_u4 *)(const void *)rec_link(w)};
                err = m.dt_mw == 1 ? worm_chase<RecH, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase<RecH, false>(w, m, rs, rec, gmask, gbase, total, scat);
                __syncthreads();
                if (!err)
                    worm_finish(w, m, rs, rec, gmask, gbase);
            }
        }
    } else if (fmt) {
        lds_u64 *llk = (lds_u64 *)(uint64_t *)q;
        lds_u16 *lec = (lds_u16 *)(uint16_t *)(q + w.n * 8);
        const bool ok = fmt == 8 ? k2_stage<8>(w, m, llk, lec) : k2_stage<10>(w, m, llk, lec);
        __syncthreads();
        if (!ok) {
            err = SSE_ERR_STATE;
        } else if (fmt == 8) {
            Rec8 rec{llk};
            if (chase == 1)
                err = m.dt_mw == 1 ? worm_chase_la<Rec8, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase_la<Rec8, false>(w, m, rs, rec, gmask, gbase, total, scat);
            else
                err = m.dt_mw == 1 ? worm_chase<Rec8, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase<Rec8, false>(w, m, rs, rec, gmask, gbase, total, scat);
            __syncthreads();
            if (!err)
                worm_finish(w, m, rs, rec, gmask, gbase);
        } else {
            Rec10 rec{llk, lec};
            if (chase == 1)
                err = m.dt_mw == 1 ? worm_chase_la<Rec10, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase_la<Rec10, false>(w, m, rs, rec, gmask, gbase, total, scat);
            else
                err = m.dt_mw == 1 ? worm_chase<Rec10, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase<Rec10, false>(w, m, rs, rec, gmask, gbase, total, scat);
            __syncthreads();
            if (!err)
                worm_finish(w, m, rs, rec, gmask, gbase);
        }
    } else {
        if (!k2_stage_global(w, m)) {
            err = SSE_ERR_STATE;
        } else {
            // one-worm tables carry 14 * leg in the worm field: the one-worm chase
            if constexpr (kFused) {
                RecGV rec{rec_vt(w), (const uint32_t *)rec_link(w)};
                err = m.dt_mw == 1 ? worm_chase<RecGV, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase<RecGV, false>(w, m, rs, rec, gmask, gbase, total, scat);
                __syncthreads();
                if (!err)
                    worm_finish(w, m, rs, rec, gmask, gbase);
            } else {
                RecG rec{rec_vt(w), (const_u4 *)(const void *)rec_link(w)};
                err = m.dt_mw == 1 ? worm_chase<RecG, true>(w, m, rs, rec, gmask, gbase, total, scat)
                                   : worm_chase<RecG, false>(w, m, rs, rec, gmask, gbase, total, scat);
                __syncthreads();
                if (!err)
                    worm_finish(w, m, rs, rec, gmask, gbase);
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

// K2.  rec_mode 0 picks the smallest record layout that fits (Rec8, Rec10,
// RecH, then RecG); 1 skips Rec8; 2 forces RecG; 3 forces RecH.  chase 1
// runs the LDS layouts with record lookahead, 0 without.
__global__ __launch_bounds__(64) void wave_worm_kernel(DevWalker *W, const DevModel *models,
                                                       DevPool *pool, int w_begin, int w_end,
                                                       int full_sweep, int lds_bytes, int rec_mode,
                                                       const int32_t *order, int chase)
{
    extern __shared__ __align__(16) char lds[];
    const int pos = w_begin + (int)blockIdx.x;
    if (pos < w_end)
        k2_walker<false>(W, models, pool, walker_at(order, pos), full_sweep, lds_bytes, rec_mode, lds,
                         chase);
}

// ------------------------------------------------------------------ init
__global__ __launch_bounds__(64) void wave_init_kernel(DevWalker *W, const DevModel *models, int n_walkers)
{
    __shared__ uint64_t rbuf[256];
    const int wi = blockIdx.x;
    if (wi >= n_walkers)
        return;
    const int lane = threadIdx.x;
    DevWalker w = W[wi];
    if (!w.pending)
        return;
    const DevModel &m = models[w.model];
    WaveStream rs;
    rs.init(rbuf, w.key, 0);
    // random state per site (sse.jl:48), one int draw each in site order
    for (int i = 0; i < m.n_sites; ++i) {
        uint8_t v = (uint8_t)(1 + rs.range((uint64_t)m.site_dims[i]));
        if (lane == 0)
            w.spins[i] = v;
    }
    for (int64_t p = lane; p < w.L; p += 64)
        w.ops[p] = 0u;
    if (lane == 0) {
        W[wi].n = 0;
        W[wi].ctr = rs.ctr();
    }
}

// ------------------------------------------------------------------ K3
__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        x += __shfl_xor(x, o);
    return x;
}

// x moved across lanes by one DPP pattern, both 32-bit halves; lanes without
// a source (or outside the row mask) read 0
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_f64(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, kCtrl, kRowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), kCtrl, kRowMask, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Inclusive prefix sum over the wave on the VALU's cross-lane (DPP) paths:
// shifts by 1, 2, 4 and 8 within each row of 16 lanes, then lane 15 of rows
// 0 and 2 added to rows 1 and 3, then lane 31 added to rows 2 and 3.
__device__ __forceinline__ double wave_incl_scan(double x)
{
    x += dpp_f64<0x111, 0xf>(x);  // row_shr:1
    x += dpp_f64<0x112, 0xf>(x);  // row_shr:2
    x += dpp_f64<0x114, 0xf>(x);  // row_shr:4
    x += dpp_f64<0x118, 0xf>(x);  // row_shr:8
    x += dpp_f64<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
    x += dpp_f64<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
    return x;
}

// mode 0: explicit measure of every walker; mode 1: thermalized walkers only,
// accumulated into the per-walker sums (fused sweep+measure loop).
// One pass over the operator string: the sign (sse.jl:305-314) when
// want_sign, and estimators e0 .. e0+ne-1 (ne <= K3_MAXE) of the
// MagnetizationEstimator family (magnetization_estimator.jl:96-230), whose
// running sums over the operator positions are wave prefix sums.  Writes the
// five outputs of each of its estimators.
constexpr int K3_MAXE = 4;
__device__ __forceinline__ void k3_finish(const DevWalker &w, const DevModel &m, int e0, int ne,
                                          bool want_sign, int neg, int64_t cnt, const double *tmp0,
                                          const double *mag, const double *absmag, const double *mag2,
                                          const double *mag4, double &sign, double *o);
template <bool kTable = true>
__device__ __forceinline__ void k3_pass(const DevWalker &w, const DevModel &m, int e0, int ne,
                                        bool want_sign, double &sign, double *o)
{
    const int lane = threadIdx.x;
    const int64_t L = w.L;
    const int N = m.n_sites, D = m.est_max_dim;
    const double *coef[K3_MAXE], *mv[K3_MAXE];
    double tmp0[K3_MAXE], run[K3_MAXE], mag[K3_MAXE], absmag[K3_MAXE], mag2[K3_MAXE], mag4[K3_MAXE];
#pragma unroll
    for (int k = 0; k < K3_MAXE; ++k) {
        coef[k] = m.est_coef + (size_t)(e0 + k) * N;
        mv[k] = m.est_mval + (size_t)(e0 + k) * N * D;
        double t0 = 0.0;
        if (k < ne)
            for (int i = lane; i < N; i += 64)
                t0 += coef[k][i] * mv[k][i * D + w.spins[i] - 1];
        tmp0[k] = k < ne ? wave_sum(t0) : 0.0;
        run[k] = tmp0[k];
        mag[k] = absmag[k] = mag2[k] = mag4[k] = 0.0;
    }
    int neg = 0;
    int64_t cnt = 0;
    const gmem<uint32_t> *ops = as_global((const uint32_t *)w.ops);
    if (kTable && m.est_dtab) {
        // per-(bond, vertex) tables: an operator's estimator changes and sign
        // are one gather each; the operator is loaded two chunks ahead, its
        // table values one ahead
        const gmem<double> *DT = as_global(m.est_dtab);
        const gmem<int8_t> *SG = as_global(m.est_dsgn);
        const uint32_t dnv = (uint32_t)m.est_dnv, NE = (uint32_t)m.n_est;
        const bool sgn = want_sign && m.any_neg;
        auto op_at = [&](int64_t q) -> uint32_t { return q < L ? ops[q] : 0u; };
        auto vals_of = [&](uint32_t op, double *d, int8_t &sg) {
            const uint32_t ix = op_bond(op) * dnv + (op_vcode(op) >> 1);
            const bool offd = op != 0u && !(op_vcode(op) & 1u);
#pragma unroll
            for (int k = 0; k < K3_MAXE; ++k)
                d[k] = k < ne && offd ? DT[(size_t)ix * NE + e0 + k] : 0.0;
            sg = sgn && op != 0u ? SG[ix] : (int8_t)0;
        };
        uint32_t op0 = op_at(lane), op1 = op_at(64 + lane);
        double d0[K3_MAXE];
        int8_t sg0;
        vals_of(op0, d0, sg0);
        for (int64_t base = 0; base < L; base += 64) {
            const uint32_t op2 = op_at(base + 128 + lane);
            double d1[K3_MAXE];
            int8_t sg1;
            vals_of(op1, d1, sg1);
            if (sgn)
                neg += popc(__ballot(op0 != 0u && sg0 < 0));
            cnt += popc(__ballot(op0 != 0u));
#pragma unroll
            for (int k = 0; k < K3_MAXE; ++k) {
                if (k < ne) {
                    const double inc = wave_incl_scan(d0[k]);
                    if (op0 != 0u) {
                        const double t = run[k] + inc;
                        const double t2 = t * t;
                        mag[k] += t;
                        absmag[k] += fabs(t);
                        mag2[k] += t2;
                        mag4[k] += t2 * t2;
                    }
                    run[k] += __shfl(inc, 63);
                }
            }
            op0 = op1;
            op1 = op2;
#pragma unroll
            for (int k = 0; k < K3_MAXE; ++k)
                d0[k] = d1[k];
            sg0 = sg1;
        }
        k3_finish(w, m, e0, ne, want_sign, neg, cnt, tmp0, mag, absmag, mag2, mag4, sign, o);
        return;
    }
    const gmem<uint64_t> *brec = as_global((const uint64_t *)m.bond_rec);  // int2 pairs
    // Software pipeline over the 64-slot chunks: the operator is loaded three
    // chunks ahead, its bond record two ahead, the vertex's sign and leg states
    // one ahead, so a chunk waits only for its estimator table values.  The
    // vertex tables of up to four bond types are addressed from registers.
    const int nt = m.n_types;
    const int8_t *sgt[4];
    const uint8_t *lst[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        sgt[t] = t < nt ? m.types[t].signs : nullptr;
        lst[t] = t < nt ? m.types[t].leg_states : nullptr;
    }
    auto op_at = [&](int64_t q) -> uint32_t { return q < L ? ops[q] : 0u; };
    auto bond_of = [&](uint32_t op) -> uint64_t { return op != 0u ? brec[op_bond(op)] : 0ull; };
    // sign of a vertex and, for an off-diagonal one, its four leg states (one u32)
    auto vertex_of = [&](uint32_t op, uint64_t bv, int8_t &sg, uint32_t &legs) {
        sg = 0;
        legs = 0u;
        if (op != 0u) {
            const int type = (int)((uint32_t)bv >> 20);
            const uint32_t v = op_vcode(op) >> 1;
            const int8_t *sp = nt <= 4 ? (type == 0 ? sgt[0] : type == 1 ? sgt[1] : type == 2 ? sgt[2] : sgt[3])
                                       : m.types[type].signs;
            const uint8_t *lp = nt <= 4 ? (type == 0 ? lst[0] : type == 1 ? lst[1] : type == 2 ? lst[2] : lst[3])
                                        : m.types[type].leg_states;
            sg = as_global(sp)[v];
            if (!(op_vcode(op) & 1u))  // off-diagonal: the leg states matter
                legs = *(const gmem<uint32_t> *)as_global((const uint32_t *)(lp + 4 * v));
        }
    };
    // prime the pipeline
    uint32_t op0 = op_at(lane), op1 = op_at(64 + lane), op2 = op_at(128 + lane);
    uint64_t bv0 = bond_of(op0), bv1 = bond_of(op1);
    int8_t sg0;
    uint32_t lg0;
    vertex_of(op0, bv0, sg0, lg0);
    for (int64_t base = 0; base < L; base += 64) {
        // issue the later stages' loads for the chunks ahead
        const uint32_t op3 = op_at(base + 192 + lane);
        const uint64_t bv2 = bond_of(op2);
        int8_t sg1;
        uint32_t lg1;
        vertex_of(op1, bv1, sg1, lg1);
        // this chunk
        const uint32_t op = op0;
        const uint32_t vc = op_vcode(op);
        if (want_sign)
            neg += popc(__ballot(op != 0u && sg0 < 0));
        cnt += popc(__ballot(op != 0u));
        const bool offd = op != 0u && !(vc & 1u);
        const uint32_t l0 = lg0 & 255u, l1 = (lg0 >> 8) & 255u, l2 = (lg0 >> 16) & 255u, l3 = lg0 >> 24;
        const int sa = (int)((uint32_t)bv0 & 0xFFFFFu), sb = (int)(uint32_t)(bv0 >> 32);
#pragma unroll
        for (int k = 0; k < K3_MAXE; ++k) {
            if (k < ne) {
                double delta = 0.0;
                if (offd) {
                    const gmem<double> *cf = as_global(coef[k]), *mk = as_global(mv[k]);
                    delta = cf[sa] * (mk[sa * D + l2 - 1] - mk[sa * D + l0 - 1]);
                    delta += cf[sb] * (mk[sb * D + l3 - 1] - mk[sb * D + l1 - 1]);
                }
                const double inc = wave_incl_scan(delta);
                if (op != 0u) {
                    const double t = run[k] + inc;
                    const double t2 = t * t;
                    mag[k] += t;
                    absmag[k] += fabs(t);
                    mag2[k] += t2;
                    mag4[k] += t2 * t2;
                }
                run[k] += __shfl(inc, 63);
            }
        }
        // rotate the pipeline
        op0 = op1;
        op1 = op2;
        op2 = op3;
        bv0 = bv1;
        bv1 = bv2;
        sg0 = sg1;
        lg0 = lg1;
    }
    k3_finish(w, m, e0, ne, want_sign, neg, cnt, tmp0, mag, absmag, mag2, mag4, sign, o);
}

// The five outputs of each estimator of a pass (and the sign of the first).
__device__ __forceinline__ void k3_finish(const DevWalker &w, const DevModel &m, int e0, int ne,
                                          bool want_sign, int neg, int64_t cnt, const double *tmp0,
                                          const double *mag, const double *absmag, const double *mag2,
                                          const double *mag4, double &sign, double *o)
{
    const int lane = threadIdx.x;
    if (want_sign)
        sign = (neg & 1) ? -1.0 : 1.0;
    const double Nn = m.norm;
    const double en = 1.0 + (double)cnt;
    const double norm = 1.0 / Nn;
#pragma unroll
    for (int k = 0; k < K3_MAXE; ++k) {
        if (k < ne) {
            double smag = tmp0[k] + wave_sum(mag[k]);
            double sabs = fabs(tmp0[k]) + wave_sum(absmag[k]);
            double s2 = tmp0[k] * tmp0[k];
            double smag2 = s2 + wave_sum(mag2[k]);
            double smag4 = s2 * s2 + wave_sum(mag4[k]);
            smag *= norm;
            sabs *= norm;
            smag2 *= norm * norm;
            smag4 *= (norm * norm) * (norm * norm);
            if (lane == 0) {
                double *oe = o + 5 + 5 * (e0 + k);
                oe[0] = sign * smag / en;
                oe[1] = sign * sabs / en;
                oe[2] = sign * smag2 / en;
                oe[3] = sign * smag4 / en;
                const double chi = 1.0 / w.T / (en + 1.0) / en * (smag * smag + smag2) * Nn;
                oe[4] = sign * chi;
            }
        }
    }
}

// measure (sse.jl:70-87, 305-376) of walker wi.  mode 0: explicit measure;
// mode 1: thermalized walkers only, accumulated into the per-walker sums
// (fused sweep+measure loop).  Estimators are taken K3_MAXE per pass over the
// operator string; the first pass also computes the sign.
template <bool kTable = true>
__device__ __forceinline__ void k3_walker(DevWalker *W, const DevModel *models, double *out, double *sum,
                                          double *sumsq, int64_t *count, double *wlf, int wi, int mode,
                                          int32_t *done)
{
    const int lane = threadIdx.x;
    const DevWalker w = W[wi];
    // a pending slot's opstring and spins are not initialised yet
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
    double sign = 1.0;
    int e0 = 0;
    do {
        const int ne = m.n_est - e0 < K3_MAXE ? m.n_est - e0 : K3_MAXE;
        k3_pass<kTable>(w, m, e0, ne, e0 == 0, sign, o);
        e0 += K3_MAXE;
    } while (e0 < m.n_est);
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

__global__ __launch_bounds__(64) void wave_measure_kernel(DevWalker *W, const DevModel *models,
                                                          double *out, double *sum, double *sumsq,
                                                          int64_t *count, double *wlf, int w_begin,
                                                          int w_end, int mode, int32_t *done,
                                                          const int32_t *order)
{
    const int pos = w_begin + (int)blockIdx.x;
    if (pos < w_end)
        k3_walker(W, models, out, sum, sumsq, count, wlf, walker_at(order, pos), mode, done);
}

// K2 and the accumulated measurement of the same sweep in one kernel: each
// walker's wave measures as soon as its own worm update is written back, so the
// measurement runs inside the launch's worm tail instead of after it.  The
// measurement reads the operator string, spins and DevWalker that the worm
// phase of the same wave has just stored, through vector loads after the
// barrier (nothing of it goes through the scalar cache).  Its measurement
// gathers the estimator values (no per-(bond, vertex) tables): with those the
// compiler cannot keep the chase's operands in scalar registers.
__global__ __launch_bounds__(64) void wave_worm_measure_kernel(
    DevWalker *W, const DevModel *models, DevPool *pool, int w_begin, int w_end, int lds_bytes,
    int rec_mode, const int32_t *order, int chase, double *out, double *sum, double *sumsq,
    int64_t *count, double *wlf, int32_t *done)
{
    extern __shared__ __align__(16) char lds[];
    const int pos = w_begin + (int)blockIdx.x;
    if (pos < w_end) {
        const int wi = walker_at(order, pos);
        k2_walker<false>(W, models, pool, wi, 1, lds_bytes, rec_mode, lds, chase);
        __syncthreads();
        k3_walker<false>(W, models, out, sum, sumsq, count, wlf, wi, 1, done);
    }
}

// pending walkers become active once initialised
__global__ void activate_pending_kernel(DevWalker *W, int n_walkers)
{
    int wi = blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= n_walkers)
        return;
    if (W[wi].pending) {
        W[wi].pending = 0;
        W[wi].active = 1;
    }
}

// ------------------------------------------------------------------ host side
#include "sse_wave_n.h"

// LDS of the generic diagonal-update kernel: the base layout for the largest
// site count and bond size, raised to stage a model's lookups in LDS where that
// keeps it within 32 KiB (other models read them from global memory)
static inline size_t k1n_lds_for(const std::vector<DevModel> &models, int max_sites, int max_ns)
{
    size_t b = k1n_lds_bytes(max_sites, max_ns);
    for (const DevModel &m : models) {
        const size_t st = k1n_lds_staged(m.n_sites, m.nsb, m.k1ntab_bytes);
        if (st <= 32768 && st > b)
            b = st;
    }
    return b;
}

static inline int wave_configure(WaveConfig &c, const std::vector<DevModel> &models, int max_sites)
{
    c.max_sites = max_sites;
    c.lds_k1 = k1_lds_bytes(max_sites);
    int max_ns = 2;
    c.generic = 0;
    for (const DevModel &m : models) {
        max_ns = std::max(max_ns, (int)m.nsb);
        if (m.nsb != 2)
            c.generic = 1;
    }
    c.lds_k1n = k1n_lds_for(models, max_sites, max_ns);
    if (c.generic && c.lds_k1n > 65536)
        (void)hipFuncSetAttribute((const void *)waven_diag_vlist_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds_k1n);
    // stage each model's lookups in LDS if that keeps K1 within 32 KiB (high
    // occupancy matters more than the lookups of larger models)
    for (const DevModel &m : models) {
        const size_t st = (k1_lds_staged(m.n_sites, m.n_bonds, m.k1tab_bytes) + 15) & ~(size_t)15;
        if (st <= 32768 && st > c.lds_k1)
            c.lds_k1 = st;
    }
    int dev = 0, lim = 65536;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        lim > 0)
        c.lds_limit = (size_t)lim;
    if (c.lds_k1 > c.lds_limit || (c.generic && c.lds_k1n > c.lds_limit))
        return SSE_ERR_INVALID;
    c.small_codes = true;
    for (const DevModel &m : models) {
        c.max_wtab = std::max(c.max_wtab, (int)m.dtab_bytes);
        if (((int64_t)m.n_types << m.dt_lgp) > 256)
            c.small_codes = false;
    }
    c.lds_k2 = 4096;
    c.lds_k2n = 4096;
    if (c.lds_limit > 65536) {
        (void)hipFuncSetAttribute((const void *)wave_worm_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds_limit);
        (void)hipFuncSetAttribute((const void *)wave_diag_vlist_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds_limit);
    }
    return SSE_OK;
}

// Worm-kernel LDS per walker for the largest walker of the last call (max n,
// max L) with a little headroom; walkers that do not fit take a slower record
// layout for that sweep.  The allocation is then widened to the largest size
// that keeps the same number of walkers per CU (more headroom at no cost in
// residency).
static inline size_t wave_k2_lds_for(const WaveConfig &c, unsigned long long max_n,
                                     unsigned long long max_L)
{
    if (c.lds_k2_fixed)
        return c.lds_k2_fixed;
    const int64_t nh = (int64_t)(max_n + 16);
    size_t need = k2_lds_need(nh, (int64_t)max_L, nh <= 4096 && c.small_codes && c.rec_mode == 0 ? 8
                                                  : c.rec_mode == 3                           ? 2
                                                                                              : 10);
    if (need > K2_LDS_CAP)  // larger walkers take the hybrid or global records
        need = K2_LDS_CAP;
    need = (need + 255) & ~(size_t)255;
    if (need > c.lds_limit)
        need = c.lds_limit;
    // Residency measured on MI355X: a workgroup takes round512(bytes + 256) of
    // the CU's LDS and about 162816 bytes are usable per CU (26880 B runs 6
    // walkers per CU, 27136-30720 B run 5, 32256 B and up only 4, whatever the
    // occupancy query says).
    const size_t per_cu = std::min(c.lds_limit, (size_t)162816);
    auto alloc = [](size_t b) { return (b + 256 + 511) & ~(size_t)511; };
    const size_t k = std::max((size_t)1, per_cu / alloc(need));  // walkers per CU
    size_t wide = ((per_cu / k) & ~(size_t)511) - 256;
    if (wide < need)
        wide = need;
    return wide;
}

static inline void wave_resize_lds(WaveConfig &c, unsigned long long max_n, unsigned long long max_L)
{
    c.lds_k2 = wave_k2_lds_for(c, max_n, max_L);
    // generic worm kernel: the chase codes of the largest walker and some
    // headroom, at most 64 KiB (larger walkers chase the global records)
    size_t g = (2 * ((size_t)max_n + 64) + 255) & ~(size_t)255;
    c.lds_k2n = g > 65536 ? 65536 : g;
}

static inline int wave_launch_init(WaveConfig &c, hipStream_t s, DevWalker *W, const DevModel *M,
                                   DevPool *pool, int n, int max_warmup)
{
    hipLaunchKernelGGL(wave_init_kernel, dim3(n), dim3(64), 0, s, W, M, n);
    for (int r = 0; r < max_warmup; ++r) {
        if (c.generic)
            hipLaunchKernelGGL(waven_diag_vlist_kernel, dim3(n), dim3(64), c.lds_k1n, s, W, M, pool, 0,
                               n, 1, r, (const int32_t *)nullptr, 0, (const int32_t *)nullptr, 0,
                               (int)c.lds_k1n);
        else
            hipLaunchKernelGGL(wave_diag_vlist_kernel, dim3(n), dim3(64), c.lds_k1, s, W, M, pool, 0, n,
                               1, r, (int)c.lds_k1, (const int32_t *)nullptr, 0);
    }
    hipLaunchKernelGGL(activate_pending_kernel, dim3((n + 63) / 64), dim3(64), 0, s, W, n);
    return hipGetLastError() == hipSuccess ? SSE_OK : SSE_ERR_HIP;
}

// one sweep (phase mask) of the walkers at schedule positions [w0, w1) on
// stream s; grp is their stream group (0..6: its own worm-kernel LDS size
// once sized; -1: one stream for all walkers)
static inline int wave_launch_sweep(WaveConfig &c, hipStream_t s, DevWalker *W, const DevModel *M,
                                    DevPool *pool, int w0, int w1, int phases, hipEvent_t *ev,
                                    const int32_t *order = nullptr, int grp = -1)
{
    const int n = w1 - w0;
    const int k1 = phases & (1 | 2 | 8);
    const int g = grp >= 0 && grp < 7 ? grp : 7;  // slot 7 collects the rest
    if (c.generic) {
        if (k1 & 3)
            hipLaunchKernelGGL(waven_diag_vlist_kernel, dim3(n), dim3(64), c.lds_k1n, s, W, M, pool, w0,
                               w1, k1, -1, order, g, (const int32_t *)nullptr, 0, (int)c.lds_k1n);
        if (ev)
            (void)hipEventRecord(ev[1], s);
        if (phases & 4)
            hipLaunchKernelGGL(waven_worm_kernel, dim3(n), dim3(64), c.lds_k2n, s, W, M, pool, w0, w1,
                               (phases & 7) == 7 ? 1 : 0, order, (const int32_t *)nullptr, 0,
                               (int)c.lds_k2n);
        if (ev)
            (void)hipEventRecord(ev[2], s);
        return hipGetLastError() == hipSuccess ? SSE_OK : SSE_ERR_HIP;
    }
    const size_t lds2 = g < 7 && c.lds_k2_group[g] ? c.lds_k2_group[g] : c.lds_k2;
    if (k1 & 3) {
        hipLaunchKernelGGL(wave_diag_vlist_kernel, dim3(n), dim3(64), c.lds_k1, s, W, M, pool, w0, w1, k1,
                           -1, (int)c.lds_k1, order, g);
    }
    if (ev)
        (void)hipEventRecord(ev[1], s);
    if (phases & 4)
        hipLaunchKernelGGL(wave_worm_kernel, dim3(n), dim3(64), lds2, s, W, M, pool, w0, w1,
                           (phases & 7) == 7 ? 1 : 0, (int)lds2, c.rec_mode, order, c.chase);
    if (ev)
        (void)hipEventRecord(ev[2], s);
    return hipGetLastError() == hipSuccess ? SSE_OK : SSE_ERR_HIP;
}

// one measuring sweep of the walkers at schedule positions [w0, w1): K1, then
// the worm update and measurement in one kernel (ev[1] after K1, ev[2] after
// the fused kernel)
static inline int wave_launch_sweep_measure(WaveConfig &c, hipStream_t s, DevWalker *W,
                                            const DevModel *M, DevPool *pool, int w0, int w1,
                                            hipEvent_t *ev, const int32_t *order, int grp,
                                            double *out, double *sum, double *sumsq,
                                            int64_t *count, double *wlf, int32_t *done)
{
    const int n = w1 - w0;
    const int g = grp >= 0 && grp < 7 ? grp : 7;
    if (c.generic) {  // no fused form: sweep, then measure
        int rc = wave_launch_sweep(c, s, W, M, pool, w0, w1, 7, ev, order, grp);
        if (!rc)
            hipLaunchKernelGGL(waven_measure_kernel, dim3(n), dim3(64), 0, s, W, M, out, sum, sumsq,
                               count, wlf, w0, w1, 1, done, order, (const int32_t *)nullptr, 0);
        return rc ? rc : (hipGetLastError() == hipSuccess ? SSE_OK : SSE_ERR_HIP);
    }
    const size_t lds2 = g < 7 && c.lds_k2_group[g] ? c.lds_k2_group[g] : c.lds_k2;
    hipLaunchKernelGGL(wave_diag_vlist_kernel, dim3(n), dim3(64), c.lds_k1, s, W, M, pool, w0, w1, 3,
                       -1, (int)c.lds_k1, order, g);
    if (ev)
        (void)hipEventRecord(ev[1], s);
    hipLaunchKernelGGL(wave_worm_measure_kernel, dim3(n), dim3(64), lds2, s, W, M, pool, w0, w1,
                       (int)lds2, c.rec_mode, order, c.chase, out, sum, sumsq, count, wlf, done);
    if (ev)
        (void)hipEventRecord(ev[2], s);
    return hipGetLastError() == hipSuccess ? SSE_OK : SSE_ERR_HIP;
}

static inline int wave_launch_measure(WaveConfig &c, hipStream_t s, DevWalker *W, const DevModel *M,
                                      double *out, double *sum, double *sumsq, int64_t *count,
                                      double *wlf, int w0, int w1, int mode, int32_t *done,
                                      const int32_t *order = nullptr)
{
    if (c.generic)
        hipLaunchKernelGGL(waven_measure_kernel, dim3(w1 - w0), dim3(64), 0, s, W, M, out, sum, sumsq,
                           count, wlf, w0, w1, mode, done, order, (const int32_t *)nullptr, 0);
    else
        hipLaunchKernelGGL(wave_measure_kernel, dim3(w1 - w0), dim3(64), 0, s, W, M, out, sum, sumsq,
                           count, wlf, w0, w1, mode, done, order);
    return hipGetLastError() == hipSuccess ? SSE_OK : SSE_ERR_HIP;
}


// sse_common.h -- device-side data layout and helpers of the SSE sweep.
//
// Layout in HBM (one allocation per batch unless noted):
//   DevModel[n_models]      model graph + per-bond-type VertexData tables
//   DevWalker[n_walkers]    per-walker scalars and slab pointers
//   slab pool               per walker one slab: ops[cap] (u32), then a record
//                           region used as dense links[legs*cap] (i32, legs =
//                           2 * sites per bond) by the lane kernels, or by the
//                           wave kernels (two-site bonds) as compact records
//                           indexed by operator rank r (r-th non-identity slot):
//                             vt[cap]      vcode | bond type << 10
//                             link[4*cap]  linked leg id 4r'+leg'
//                             mask[cap/64+1], base[cap/64+1]  per 64-slot chunk:
//                                          non-identity bit mask, rank of its
//                                          first operator
//                           carved from a device bump allocator, so the cutoff
//                           L grows inside the kernels
//   spins/first/last        [n_walkers][n_sites] u8 / i32 / i32
//
// Operator code (u32): 0 = identity, else 1 | vcode<<1 | bond<<11,
// vcode = diag | vidx<<1 (vidx 0-based).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SSE_WAVE 64

struct DevVD {
    int32_t dims0, dims1, nv, max_worm, n_exits, nsites;
    int32_t dims[4];           // local dimension of each of the nsites bond sites
    double energy_offset;
    const double *weights;     // [nv]
    const int8_t *signs;       // [nv]
    const uint8_t *leg_states; // [nv][2*nsites]
    const int32_t *diag_vertex;// [prod(dims)]
    const int32_t *toff;       // [nv][max_worm][2*nsites]
    const int32_t *tcnt;       // [nv][max_worm][2*nsites]
    const double *cum;         // [n_exits]
    const int32_t *tgt;        // [n_exits]
    const uint8_t *eleg;       // [n_exits]
    const uint8_t *ewrm;       // [n_exits]
};

struct DevModel {
    int32_t n_sites, n_bonds, n_types, n_est, est_max_dim, max_dim;
    int32_t nsb, pad_nsb;      // sites per bond (all bonds alike), leg count 2*nsb
    double energy_offset, norm;
    const uint8_t *site_dims;  // [n_sites]
    const int32_t *bond_type;  // [n_bonds]
    const int32_t *bond_sites; // [n_bonds][nsb]
    const DevVD *types;        // [n_types]
    const double *est_coef;    // [n_est][n_sites]
    const double *est_mval;    // [n_est][n_sites][est_max_dim]
    // packed per-bond record for the kernels: sites and type in one int2
    const int2 *bond_rec;      // [n_bonds] (site0 | type<<20, site1)
    // packed worm tables (staged in LDS by the worm kernel):
    //   u32 hdr[3*n_types]   (trans offset, nv, max_worm) per type
    //   u32 trans[...]       exit index | count << 20, or 0xFFFFFFFF
    //   u32 eout[n_exits]    target vcode | leg << 10 | worm << 12
    //   f64 cum[n_exits]     at byte offset wtab_cum_off
    const uint32_t *wtab;
    int32_t wtab_bytes, wtab_cum_off, wtab_eout_off, wtab_trans_off;
    // chase table: one 64-byte entry (DirectEntry) per (ecode, worm-1, leg_in)
    // at index (ecode*dt_mw + worm-1)*4 + leg_in, ecode = type << dt_lgp | vcode
    // (2^dt_lgp >= 2*dt_nv); up to 4 exits inline, more are looked up in wtab
    const struct DirectEntry *dtab;
    int32_t dt_nv, dt_mw, dtab_bytes, dt_lgp;
    // diagonal-update table, copied to LDS by K1 next to the bond records:
    //   per type a 16-byte header {weights off, diag-vertex off, leg-state off,
    //   nv | dims0 << 16}, then f64 weights[nv], i16 diagonal vertex codes
    //   [dims0*dims1] (or -1), u8 leg states [nv][4]; offsets from the table start
    const uint8_t *k1tab;
    int32_t k1tab_bytes, pad_k1;
    // lookups of the generic diagonal-update kernel, copied to LDS when they
    // fit (any bond size): per type a 32-byte header {weights off, diagonal-
    // vertex off, upper-states off, nv, dims packed one byte each}, then per
    // bond {type, sites[nsb]} (i32 each, from k1n_bond_off), then per type f64
    // weights[nv], i32 diagonal vertex codes [prod dims] (or -1) and u32 upper
    // leg states [nv] (one byte per bond position); offsets from the start
    const uint8_t *k1ntab;
    int32_t k1ntab_bytes, k1n_bond_off;
    // chase table of the generic wave kernels (any bond size): one entry per
    // (ecode, worm-1, leg_in) at ((ecode*dt_mw + worm-1) << g_lg) + leg_in,
    // 2^g_lg = leg count rounded up to a power of two; exit words gen_out()
    const struct DirectEntry *gtab;
    int32_t gtab_bytes, g_lg;
    // measurement tables (null when too large): per
    // (bond, vertex) at bond * est_dnv + vertex, the change of each estimator
    // across the vertex, f64 [n_est] (the measurement adds it for off-diagonal
    // vertices), and the vertex's sign, i8
    const double *est_dtab;
    const int8_t *est_dsgn;
    int32_t est_dnv, any_neg;  // vertex stride; some vertex has a negative sign
};

// Exit j is taken by the first j with r < cum[j], r = (word >> 11) * 2^-53.
// With m = word >> 11 (an integer < 2^53) and K = ceil(cum * 2^53) this is
// exactly m < K, so the worm kernel compares integers.  Layout for the chase:
// out[0] is taken for m < K[0], out[j] for K[j-1] <= m < K[j]; m >= K[3] falls
// through.  With count < 4 exits the last threshold is repeated and the unused
// outs are 0xFFFFFFFF, so m >= K[count-1] reaches the fall-through test.  An
// invalid transition (count 0) or one with more than 4 exits has all K = 0:
// every word falls through and the count decides.
// The chase compares m with K[0..2] only and takes out[3] otherwise, so
// out[j] holds kChaseSlow (send the step to the exact path) wherever that
// would be wrong: slots past the last exit (K = 2^53 there), the fourth exit
// when its threshold is below 2^53 (its word is then kept in pad[0]), and
// out[3] of an entry with more than four exits (the first three are inline).
struct __align__(16) DirectEntry {
    uint64_t K[4];     // integer thresholds ceil(cum[j] * 2^53)
    uint32_t out[4];   // chase_out(target ecode, exit leg, exit worm) or kChaseSlow
    uint32_t count;    // number of exits
    uint32_t trans;    // wtab trans word (exit index | count << 20)
    uint32_t pad[2];   // pad[0]: the fourth exit's word when count == 4
};

// Marker exit word: worm-1 has bit 13 set, which no real exit word has (the
// model check keeps every worm count below 8192).
#define kChaseSlow 0xFFFFFFFFu
#define kChaseSlowBit 0x2000u

// Exit word of the generic chase table: worm-1 in bits 0..11, exit leg in
// bits 12..14, target ecode in bits 16..31; bit 15 is set only in the slow
// marker kChaseSlow (worm counts stay below 4096, legs below 8).
__host__ __device__ inline uint32_t gen_out(uint32_t ecode, uint32_t leg, uint32_t worm)
{
    return (worm - 1u) | (leg << 12) | (ecode << 16);
}
#define kGenSlowBit 0x8000u

// exit word of the chase: worm-1 in bits 0..13, exit leg in 14..15, target
// ecode (type << dt_lgp | vcode) in 16..31
__host__ __device__ inline uint32_t chase_out(uint32_t ecode, uint32_t leg, uint32_t worm)
{
    return (worm - 1u) | (leg << 14) | (ecode << 16);
}
// exit word of a model with one worm type: worm-1 is always 0, so bits 0..5
// carry 14 * exit leg instead, the shift that takes the exit leg's link out of
// an 8-byte record
__host__ __device__ inline uint32_t chase_out1(uint32_t ecode, uint32_t leg)
{
    return (14u * leg) | (leg << 14) | (ecode << 16);
}
__host__ __device__ inline uint32_t ecode_of_vt(uint32_t vt, uint32_t lgp)
{
    return ((vt >> 10) << lgp) | (vt & 1023u);
}
__host__ __device__ inline uint32_t vt_of_ecode(uint32_t e, uint32_t lgp)
{
    return (e & ((1u << lgp) - 1u)) | ((e >> lgp) << 10);
}

// Pointers known to address global memory.  Loads and stores through them are
// global (not flat) instructions, which count only against the vector-memory
// counter: an LDS or scalar-load wait then does not also wait for them.
template <class T>
using gmem = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ gmem<T> *as_global(T *p)
{
    return (gmem<T> *)p;
}
template <class T>
__device__ __forceinline__ const gmem<T> *as_global(const T *p)
{
    return (const gmem<T> *)p;
}

struct DevWalker {
    // parameters
    int32_t model, warmup;
    double T, target_wlf, atten;
    uint64_t key;
    int64_t thermalization, init_cutoff;
    // chain state
    double avg_worm_length, num_worms;
    int64_t L, n, cap;
    uint64_t ctr;
    int64_t sweeps, measured;
    // diagnostics of the last sweep
    int64_t last_scatter;
    double last_total, last_wlf;
    int32_t error, n_sites;
    // buffers
    uint32_t *ops;    // [cap]
    int32_t *links;   // record region (dense links [legs*cap] or compact records)
    uint8_t *spins;   // [n_sites]
    int32_t *first;   // [n_sites]
    int32_t *last;    // [n_sites]
    // work counters (sum over sweeps since reset)
    int64_t acc_L, acc_n, acc_W, acc_sweeps;
    int32_t meas_offset, meas_width;
    // serving slots: active walkers are swept; a finite budget of sweeps
    // (remaining >= 0) retires the walker after its last sweep + measure
    int32_t active, pending;
    int64_t remaining;
    int32_t retire;
    int32_t acc_slot;  // accumulator row (count index) of the walker's temperature slot
};

struct DevPool {
    unsigned long long next;  // bytes used
    unsigned long long size;
    char *base;
    int32_t overflow;
    int32_t status;           // bit (1 << code) per walker error code seen
    unsigned long long max_n; // largest operator count after a diagonal update
    unsigned long long max_L; // largest cutoff
    unsigned long long group_max_n[8];  // the same per walker group (stream group)
};

__device__ __forceinline__ void flag_error(DevPool *pool, int code)
{
    atomicOr(&pool->status, 1 << code);
}

#define SSE_OP_BOND_SHIFT_D 11
__device__ __forceinline__ uint32_t op_bond(uint32_t op) { return op >> SSE_OP_BOND_SHIFT_D; }
__device__ __forceinline__ uint32_t op_vcode(uint32_t op) { return (op >> 1) & 1023u; }
__device__ __forceinline__ uint32_t op_make(uint32_t bond, uint32_t vcode)
{
    return 1u | (vcode << 1) | (bond << SSE_OP_BOND_SHIFT_D);
}

// ---------------------------------------------------------------- Philox4x32-10
// Counter-based generator of Salmon et al. (SC'11).  64-bit word i of a
// walker's stream with key k: block b = i/2 with counter (lo32 b, hi32 b, 0, 0)
// and key (lo32 k, hi32 k); word = out0|out1<<32 for even i, out2|out3<<32 odd.
__device__ __forceinline__ void philox_block(uint64_t b, uint64_t key, uint32_t &o0, uint32_t &o1,
                                             uint32_t &o2, uint32_t &o3)
{
    uint32_t c0 = (uint32_t)b, c1 = (uint32_t)(b >> 32), c2 = 0u, c3 = 0u;
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

__device__ __forceinline__ double word_to_float(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }

// Per-lane sequential stream with a one-block cache.
struct LaneRng {
    uint64_t key, ctr, blk;
    uint32_t o0, o1, o2, o3;
    __device__ void init(uint64_t k, uint64_t c)
    {
        key = k;
        ctr = c;
        blk = ~0ull;
    }
    __device__ __forceinline__ uint64_t next()
    {
        uint64_t b = ctr >> 1;
        if (b != blk) {
            philox_block(b, key, o0, o1, o2, o3);
            blk = b;
        }
        uint64_t w = (ctr & 1) ? ((uint64_t)o2 | ((uint64_t)o3 << 32))
                               : ((uint64_t)o0 | ((uint64_t)o1 << 32));
        ++ctr;
        return w;
    }
    __device__ __forceinline__ double next_float() { return word_to_float(next()); }
    // rand(1:s) - 1, Lemire's nearly divisionless method on 64-bit words
    __device__ __forceinline__ uint64_t range(uint64_t s)
    {
        uint64_t x = next();
        uint64_t l = x * s;
        uint64_t h = __umul64hi(x, s);
        if (l < s) {
            uint64_t t = (0ull - s) % s;
            while (l < t) {
                x = next();
                l = x * s;
                h = __umul64hi(x, s);
            }
        }
        return h;
    }
};

// ---------------------------------------------------------------- tanh
// Worm-number controller tanh.  Evaluated by exactly this sequence of IEEE
// operations (no contraction): Cody-Waite reduction x = k ln2 + r and the
// degree-16 Taylor polynomial of expm1(r); tanh(x) = e/(e+2), e = expm1(2|x|).
__device__ __forceinline__ double sse_expm1(double x)
{
    const double C[16] = {
        0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
        0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
        0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
        0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45};
    if (x == 0.0)
        return x;
    double kf = rint(x * 0x1.71547652b82fep+0);
    double r = (x - kf * 0x1.62e42fee00000p-1) - kf * 0x1.a39ef35793c76p-33;
    double p = C[15];
#pragma unroll
    for (int j = 14; j >= 0; --j)
        p = p * r + C[j];
    double em1 = p * r;
    int k = (int)kf;
    if (k == 0)
        return em1;
    double two_k = ldexp(1.0, k);
    return two_k * em1 + (two_k - 1.0);
}

__device__ __forceinline__ double sse_tanh(double x)
{
    double ax = fabs(x);
    if (!(ax < 22.0))
        return copysign(1.0, x);
    if (ax < 0x1p-28)
        return x;
    double e = sse_expm1(2.0 * ax);
    double t = e / (e + 2.0);
    return copysign(t, x);
}

// Natural logarithm of the parallel-tempering exchange test: x = m 2^e with
// m in (sqrt(1/2), sqrt(2)], log(m) = 2 atanh(s) = 2 sum_j s^(2j+1)/(2j+1),
// s = (m-1)/(m+1), twelve terms (|s| <= 0.1716).  Exactly this sequence of
// IEEE operations, so the host-side checker decides every exchange the same
// way.  x positive, finite, normal.
__device__ __forceinline__ double sse_log(double x)
{
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 1.0 / 23.0;
#pragma unroll
    for (int k = 21; k >= 1; k -= 2)
        p = p * z + 1.0 / (double)k;
    const double lm = 2.0 * s * p;
    const double ef = (double)e;
    return ef * 0x1.62e42fee00000p-1 + (lm + ef * 0x1.a39ef35793c76p-33);
}

// Philox4x32-10 of a full 128-bit counter block.
__device__ __forceinline__ void philox_block4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint64_t key, uint32_t &o0, uint32_t &o1)
{
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    o0 = c0;
    o1 = c1;
}

// Worm-number controller and WormLengthFraction (sse.jl:200-217).
__device__ __forceinline__ void worm_controller(DevWalker &w, double total, bool thermalized)
{
    w.last_total = total;
    w.last_wlf = __longlong_as_double(0x7ff8000000000000ll);
    if (thermalized && w.n != 0)
        w.last_wlf = total / (double)w.n;
    double avg = total / ceil(w.num_worms);
    if (!thermalized) {
        w.avg_worm_length += w.atten * (avg - w.avg_worm_length);
        double target = w.target_wlf * (double)w.n / w.avg_worm_length;
        w.num_worms += w.atten * (target - w.num_worms + 100.0 * sse_tanh(target - w.num_worms));
        if (w.atten != 0.0) {
            double hi = 1.0 + (double)w.n / 2.0;
            double x = w.num_worms;
            w.num_worms = x > hi ? hi : (x < 1.0 ? 1.0 : x);
        }
    }
}

// WormLengthFraction (sse.jl:200-202) of the last sweep into the slot's
// accumulator {sum, sum of squares, count}: the reference measures it in
// every thermalized sweep with n != 0 (last_wlf is NaN otherwise).  The slot
// follows the temperature under parallel tempering, like the other sums.
__device__ __forceinline__ void worm_fraction_accumulate(const DevWalker &w, double *wlf)
{
    const double v = w.last_wlf;
    if (v == v) {
        double *q = wlf + 3 * (size_t)w.acc_slot;
        q[0] += v;
        q[1] += v * v;
        q[2] += 1.0;
    }
}

// Growth rule of the cutoff (sse.jl:138-145): n >= L/2 -> L = floor(1.5 L + 100).
__device__ __forceinline__ int64_t grown_cutoff(int64_t L) { return (3 * L) / 2 + 100; }

// Capacity of a fresh slab for cutoff L: two further growth steps of headroom.
__device__ __host__ __forceinline__ int64_t slab_capacity(int64_t L)
{
    int64_t c = L < 64 ? 64 : L;
    c = (3 * c) / 2 + 100;
    c = (3 * c) / 2 + 100;
    return (c + 63) & ~(int64_t)63;
}

// Slab of a walker with leg_count legs per operator: ops[cap], then a record
// region that holds either the compact records of the wave kernels (4 bytes
// of vertex code and 4 * PL bytes of links per slot, plus the chunk masks and
// bases; 20 bytes per slot for two-site bonds) or the dense links of the lane
// kernels (4 * legs bytes per slot), whichever is larger.
__device__ __host__ __forceinline__ uint64_t slab_bytes(int64_t cap, int legs = 4)
{
    // wave layout: vt[cap] and links[PL*cap] (PL = legs rounded up to a power
    // of two), then the chunk masks and bases
    const uint64_t pl = legs <= 2 ? 2u : legs <= 4 ? 4u : 8u;
    const uint64_t wave = (uint64_t)cap * 4u * (1u + pl) + ((uint64_t)cap / 64u + 1u) * 16u;
    const uint64_t lane = (uint64_t)cap * 4u * (uint64_t)legs;
    uint64_t b = (uint64_t)cap * 4u + (wave > lane ? wave : lane);
    return (b + 255u) & ~(uint64_t)255u;
}

// compact-record views of a walker's record region (wave kernels)
__device__ __forceinline__ uint32_t *rec_vt(const DevWalker &w) { return (uint32_t *)w.links; }
__device__ __forceinline__ uint32_t *rec_link(const DevWalker &w) { return (uint32_t *)w.links + w.cap; }
__device__ __forceinline__ uint64_t *rec_mask(const DevWalker &w)
{
    return (uint64_t *)((uint32_t *)w.links + 5 * w.cap);
}
__device__ __forceinline__ uint32_t *rec_base(const DevWalker &w)
{
    return (uint32_t *)(rec_mask(w) + (w.cap / 64 + 1));
}

// Bump allocation from the batch's slab pool; returns nullptr on exhaustion.
__device__ __forceinline__ char *pool_alloc(DevPool *pool, uint64_t bytes)
{
    unsigned long long off = atomicAdd(&pool->next, (unsigned long long)bytes);
    if (off + bytes > pool->size) {
        atomicExch(&pool->overflow, 1);
        return nullptr;
    }
    return pool->base + off;
}

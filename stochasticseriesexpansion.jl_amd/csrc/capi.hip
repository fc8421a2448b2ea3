// capi.hip -- C ABI of libssegpu.so (declared in include/ssegpu.h).
//
// A batch owns: the flattened models (uploaded once), the walker array, a
// slab pool from which every walker's opstring/vertex-list slab is carved
// (device-side bump allocation lets the cutoff grow inside the kernels), the
// spin/first/last arrays, measurement buffers and one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ssegpu.h"
#include "sse_common.h"
#include "vertex_lp.h"
#include "sse_lane.h"
#include "sse_wave.h"
#include "sse_pt.h"
#include "sse_persist.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHECK(expr)                                                                     \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return fail(SSE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
    } while (0)

// What a captured launch sequence (HIP graph) bakes in: a replay is valid
// while every field is unchanged.  Compared bytewise (constructed zeroed).
struct GraphKey {
    int32_t kind;       // 1 one stream, 2 stream groups, 3 persistent launch
    int32_t sweeps, measure, groups, rec_mode, chase, use_order, n_tasks, fuse, lane_width, generic;
    size_t lds[9];      // worm-kernel LDS (one stream / per group), persistent LDS
    const void *models, *budgets, *order;
    GraphKey() { memset(this, 0, sizeof *this); }
    bool same(const GraphKey &o) const { return memcmp(this, &o, sizeof *this) == 0; }
};

// What sse_set_walker checks a restored state against, per model: the vertex
// count of each bond's type and each site's local dimension.
struct HostModelCheck {
    std::vector<int16_t> bond_nv;
    std::vector<uint8_t> site_dims;
};

}  // namespace

struct sse_batch {
    int device = 0;
    int mode = SSE_MODE_WAVE;
    hipStream_t stream = nullptr;
    int n_models = 0, n_walkers = 0;
    std::vector<sse_walker_desc> wdesc;
    std::vector<int32_t> widths, offsets, nsites;
    std::vector<DevModel> h_models;
    std::vector<int> model_max_dim;
    std::vector<HostModelCheck> h_check;  // [model], in upload order
    DevModel *d_models = nullptr;
    std::vector<void *> allocs;
    DevWalker *d_walkers = nullptr;
    std::vector<DevWalker> h_walkers;
    DevPool *d_pool = nullptr;
    char *pool_base = nullptr;
    size_t pool_size = 0;
    size_t pool_host_next = 0;
    uint8_t *d_spins = nullptr;
    int32_t *d_first = nullptr, *d_last = nullptr;
    double *d_out = nullptr, *d_sum = nullptr, *d_sumsq = nullptr;
    int64_t *d_count = nullptr;
    double *d_wlf = nullptr;  // [slot][3]: WormLengthFraction sum, sum of squares, count
    int32_t *d_done = nullptr;
    uint8_t *d_scratch = nullptr;
    bool any_pending = true;
    bool any_budget = false;
    // warm-up rounds the pending walkers need: the largest
    // diagonal_warmup_sweeps among the walkers created or restarted
    // (sse_slot_start) since the last initialisation
    int pending_warmup = 0;
    int max_sites = 0, total_width = 0;
    int64_t max_records = 0;
    bool inited = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> evpool;  // 4 per sweep of the last timed call
    // captured sweep(+measure) sequence, replayed by sse_run/sse_sweep
    bool use_graph = true, use_timing = true;
    // stream groups replayed from one captured multi-stream graph (SSE_OPT_GRAPH
    // 2): measured slower than direct launches on this runtime, so opt-in
    bool graph_groups = false;
    // SSE_OPT_TIMING 2: events around every kernel launch of a call on the
    // stream it runs on (stream groups and the persistent launch included;
    // direct launches), summed per kernel by sse_kernel_times
    bool kernel_timing = false;
    double kern_ms[4] = {0, 0, 0, 0};
    int64_t kern_launches[4] = {0, 0, 0, 0};
    struct KernEv {
        int kind;          // 0 K1 + K2 + K3 of one group-sweep (1 without K3), 3 persistent
        size_t e;          // first of 4 events in kevpool
    };
    std::vector<KernEv> kern_ev;
    std::vector<hipEvent_t> kevpool;
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey;
    // persistent work-queue launch (SSE_OPT_PERSISTENT, sse_run_budgets):
    // task counter, per-walker budgets and task order of a budgeted call
    bool persistent = false;
    int n_cu = 0;
    int32_t *d_queue = nullptr, *d_budget = nullptr, *d_task_order = nullptr;
    // expected operator count per sweep of each walker's current job (bonds /
    // T of its model; a scheduling estimate kept by create and slot_start)
    std::vector<double> work_key;
    // autotuner (SSE_OPT_TUNE): launch forms measured on the caller's own
    // sweeps; all of them give the same chains and accumulators
    struct TuneCfg {
        int persistent, groups, chase;
        size_t k2_lds;  // SSE_OPT_K2_LDS value (0 = sized from the operator counts)
    };
    struct Tuner {
        std::vector<TuneCfg> cands;
        std::vector<double> rate;  // best vertex updates per second seen per candidate
        std::vector<int> runs;
        int best = -1;             // chosen candidate (-1: still measuring)
        bool twin = true;          // add the lookahead-chase twin of the fastest
    };
    bool tune = true;         // tuning enabled (an explicit launch option turns it off)
    int lane_width = 64;      // lane mode: walkers per wavefront (threads per workgroup)
    Tuner tune_run, tune_bud; // sse_run/sse_sweep calls; sse_run_budgets calls
    unsigned long long *d_work = nullptr;  // vertex-update sum of a tuning measurement
    double phase_ms[3] = {0, 0, 0};
    int64_t phase_launches = 0;
    // wave mode: walkers split into groups, each swept on its own stream, so
    // one group's sweeps overlap the slow tail of another's (0: automatic,
    // see group_count)
    int n_groups = 0;
    unsigned long long seen_max_n = 0, seen_max_L = 0;  // operator counts behind the LDS size
    std::vector<hipStream_t> gstreams;
    std::vector<hipEvent_t> gjoin;
    hipEvent_t gfork = nullptr;
    // device time of the last sse_sweep/sse_run call (HIP events, timing on)
    hipEvent_t call_ev[2] = {nullptr, nullptr};
    double call_ms = 0;
    int64_t call_sweeps = 0;
    WaveConfig wave;
    // replica exchange in T: ladders of walkers in slot order, one exchange
    // round after every pt_interval-th sweep (0: only on sse_pt_exchange)
    int32_t pt_ladders = 0, pt_len = 0, pt_interval = 0, pt_since = 0;
    uint64_t pt_seed = 0;
    int64_t pt_round = 0;
    size_t pt_cap = 0;
    int32_t *d_pt_slot = nullptr;
    // schedule: walker indices by expected operator count, largest first (the
    // longest chains start first; stream groups then hold walkers of similar
    // size, each group's worm kernel sized for its own largest walker)
    int32_t *d_order = nullptr;
    bool use_order = true;
    int n_big = 0;  // leading walkers of the schedule too large for 8-byte LDS records
    int64_t *d_pt_acc = nullptr, *d_pt_prop = nullptr;
    std::vector<uint8_t> pt_member;
};

// Copies and fills of the host-side entry points go through the batch's own
// non-blocking stream and wait for it.  The legacy stream is never used, so a
// batch driven by one host thread never waits on, or invalidates, another
// thread's batch (one that is capturing a graph at that moment, say).
static hipError_t bmemcpy(sse_batch *b, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, b->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(b->stream);
}
static hipError_t bmemset(sse_batch *b, void *dst, int value, size_t bytes)
{
    hipError_t e = hipMemsetAsync(dst, value, bytes, b->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(b->stream);
}

template <class T>
static int dev_alloc(sse_batch *b, T **p, size_t count)
{
    void *q = nullptr;
    size_t bytes = count * sizeof(T);
    if (bytes == 0)
        bytes = 16;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(SSE_ERR_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    b->allocs.push_back(q);
    *p = (T *)q;
    return SSE_OK;
}

static int validate_model(const sse_model_desc &m, int k)
{
    char buf[256];
    if (m.n_sites <= 0 || m.n_bonds <= 0 || m.n_types <= 0 || !m.site_dims || !m.bond_type ||
        !m.bond_sites || !m.types) {
        snprintf(buf, sizeof buf, "model %d: empty graph or missing arrays", k);
        return fail(SSE_ERR_INVALID, buf);
    }
    const int ns = m.bond_n_sites;
    if (ns < 1 || ns > SSE_MAX_BOND_SITES) {
        snprintf(buf, sizeof buf, "model %d: bond_n_sites must be 1..%d", k, SSE_MAX_BOND_SITES);
        return fail(SSE_ERR_INVALID, buf);
    }
    if (m.n_sites >= (1 << 20) || m.n_bonds >= (1 << 21) || m.n_types >= (1 << 11)) {
        snprintf(buf, sizeof buf, "model %d: too many sites/bonds/types", k);
        return fail(SSE_ERR_INVALID, buf);
    }
    for (int i = 0; i < m.n_sites; ++i)
        if (m.site_dims[i] < 1) {
            snprintf(buf, sizeof buf, "model %d: site %d has dimension 0 (not on any bond)", k, i);
            return fail(SSE_ERR_INVALID, buf);
        }
    for (int t = 0; t < m.n_types; ++t) {
        const sse_vertex_data_desc &vd = m.types[t];
        bool ok = vd.n_sites == ns && vd.n_vertices > 0 && vd.n_vertices <= 512 &&
                  vd.max_worm_count >= 1 && vd.max_worm_count <= 4096;
        int64_t states = 1;
        for (int i = 0; ok && i < ns; ++i) {
            ok = vd.dims[i] >= 1 && vd.dims[i] <= 255;
            states *= vd.dims[i];
        }
        if (!ok || states > (1 << 20)) {
            snprintf(buf, sizeof buf, "model %d type %d: unsupported table shape", k, t);
            return fail(SSE_ERR_INVALID, buf);
        }
    }
    for (int b = 0; b < m.n_bonds; ++b) {
        const int t = m.bond_type[b];
        bool ok = t >= 0 && t < m.n_types;
        for (int i = 0; ok && i < ns; ++i) {
            const int si = m.bond_sites[ns * b + i];
            ok = si >= 0 && si < m.n_sites;
            for (int j = 0; ok && j < i; ++j)
                ok = m.bond_sites[ns * b + j] != si;
        }
        if (!ok) {
            snprintf(buf, sizeof buf, "model %d: bond %d invalid", k, b);
            return fail(SSE_ERR_INVALID, buf);
        }
        const sse_vertex_data_desc &vd = m.types[t];
        for (int i = 0; i < ns; ++i)
            if (vd.dims[i] != m.site_dims[m.bond_sites[ns * b + i]]) {
                snprintf(buf, sizeof buf, "model %d: bond %d site dimensions inconsistent", k, b);
                return fail(SSE_ERR_INVALID, buf);
            }
    }
    for (int t = 0; t < m.n_types; ++t) {
        const sse_vertex_data_desc &vd = m.types[t];
        int64_t states = 1;
        for (int i = 0; i < ns; ++i)
            states *= vd.dims[i];
        for (int64_t i = 0; i < states; ++i) {
            int dv = vd.diagonal_vertices[i];
            if (dv >= 0 && ((dv >> 1) >= vd.n_vertices || !(dv & 1))) {
                snprintf(buf, sizeof buf, "model %d type %d: bad diagonal vertex", k, t);
                return fail(SSE_ERR_INVALID, buf);
            }
        }
        // leg states are 1-based local states of the leg's site (leg j sits on
        // bond site j % ns); the kernels index per-state tables with them
        for (int v = 0; v < vd.n_vertices; ++v)
            for (int j = 0; j < 2 * ns; ++j) {
                const int st = vd.leg_states[(size_t)2 * ns * v + j];
                if (st < 1 || st > vd.dims[j % ns] || (m.n_estimators > 0 && st > m.est_max_dim)) {
                    snprintf(buf, sizeof buf, "model %d type %d: vertex %d leg %d state out of range",
                             k, t, v, j);
                    return fail(SSE_ERR_INVALID, buf);
                }
            }
        int nt = vd.n_vertices * vd.max_worm_count * 2 * ns;
        for (int i = 0; i < nt; ++i) {
            int off = vd.transition_offset[i], cnt = vd.transition_count[i];
            if (off >= 0 && (cnt < 0 || off + cnt > vd.n_exits)) {
                snprintf(buf, sizeof buf, "model %d type %d: transition out of range", k, t);
                return fail(SSE_ERR_INVALID, buf);
            }
        }
        for (int e = 0; e < vd.n_exits; ++e) {
            if (vd.exit_leg[e] >= 2 * ns || (vd.exit_target[e] >> 1) >= vd.n_vertices ||
                vd.exit_target[e] < 0 || vd.exit_worm[e] < 1 || vd.exit_worm[e] > vd.max_worm_count) {
                snprintf(buf, sizeof buf, "model %d type %d: exit %d out of range", k, t, e);
                return fail(SSE_ERR_INVALID, buf);
            }
        }
    }
    if (m.n_estimators < 0 || (m.n_estimators > 0 && (!m.est_coef || !m.est_mval)))
        return fail(SSE_ERR_INVALID, "estimator tables missing");
    return SSE_OK;
}

// Upload one model: every array into one device buffer, pointers fixed up.
// The model's HostModelCheck goes into hc; the caller commits both hc and dm
// only once every step of the upload has succeeded.
static int upload_model(sse_batch *b, const sse_model_desc &m, DevModel &dm, HostModelCheck &hc)
{
    hc.site_dims.assign(m.site_dims, m.site_dims + m.n_sites);
    hc.bond_nv.resize(m.n_bonds);
    for (int k = 0; k < m.n_bonds; ++k)
        hc.bond_nv[k] = (int16_t)m.types[m.bond_type[k]].n_vertices;
    std::vector<char> blob;
    auto put = [&](const void *src, size_t bytes) {
        size_t off = (blob.size() + 15) & ~(size_t)15;
        blob.resize(off + bytes + 16);
        if (bytes)
            memcpy(blob.data() + off, src, bytes);
        blob.resize(off + bytes);
        return off;
    };
    struct VOff {
        size_t w, s, ls, dv, to, tc, cum, tgt, el, ew;
    };
    std::vector<VOff> voff(m.n_types);
    const int ns = m.bond_n_sites, legs = 2 * ns;
    int max_dim = 1;
    for (int i = 0; i < m.n_sites; ++i)
        max_dim = m.site_dims[i] > max_dim ? m.site_dims[i] : max_dim;
    for (int t = 0; t < m.n_types; ++t) {
        const sse_vertex_data_desc &vd = m.types[t];
        size_t nt = (size_t)vd.n_vertices * vd.max_worm_count * legs;
        size_t states = 1;
        for (int i = 0; i < ns; ++i)
            states *= (size_t)vd.dims[i];
        voff[t].w = put(vd.weights, sizeof(double) * vd.n_vertices);
        voff[t].s = put(vd.signs, vd.n_vertices);
        voff[t].ls = put(vd.leg_states, (size_t)legs * vd.n_vertices);
        voff[t].dv = put(vd.diagonal_vertices, sizeof(int32_t) * states);
        voff[t].to = put(vd.transition_offset, sizeof(int32_t) * nt);
        voff[t].tc = put(vd.transition_count, sizeof(int32_t) * nt);
        voff[t].cum = put(vd.exit_cumprob, sizeof(double) * vd.n_exits);
        voff[t].tgt = put(vd.exit_target, sizeof(int32_t) * vd.n_exits);
        voff[t].el = put(vd.exit_leg, vd.n_exits);
        voff[t].ew = put(vd.exit_worm, vd.n_exits);
    }
    // packed bond records: the first two sites and the type (one site:
    // site1 = -1; wider bonds read bond_sites)
    std::vector<int2> brec(m.n_bonds);
    for (int k = 0; k < m.n_bonds; ++k) {
        brec[k].x = m.bond_sites[ns * k] | (m.bond_type[k] << 20);
        brec[k].y = ns >= 2 ? m.bond_sites[ns * k + 1] : -1;
    }
    size_t o_sd = put(m.site_dims, m.n_sites);
    size_t o_bt = put(m.bond_type, sizeof(int32_t) * m.n_bonds);
    size_t o_bs = put(m.bond_sites, sizeof(int32_t) * ns * m.n_bonds);
    size_t o_br = put(brec.data(), sizeof(int2) * m.n_bonds);
    size_t ne = (size_t)m.n_estimators;
    size_t o_ec = put(m.est_coef, sizeof(double) * ne * m.n_sites);
    size_t o_em = put(m.est_mval, sizeof(double) * ne * m.n_sites * (m.n_estimators ? m.est_max_dim : 0));
    size_t o_types = put(nullptr, sizeof(DevVD) * m.n_types);
    // packed worm tables: hdr | trans | eout | cum
    std::vector<uint32_t> hdr, trans, eout;
    std::vector<double> cum;
    int total_exits = 0;
    for (int t = 0; t < m.n_types; ++t) {
        const sse_vertex_data_desc &vd = m.types[t];
        hdr.push_back((uint32_t)trans.size());
        hdr.push_back((uint32_t)vd.n_vertices);
        hdr.push_back((uint32_t)vd.max_worm_count);
        int nt = vd.n_vertices * vd.max_worm_count * legs;
        for (int i = 0; i < nt; ++i) {
            int off = vd.transition_offset[i], cnt = vd.transition_count[i];
            if (off < 0 || cnt <= 0)
                trans.push_back(0xFFFFFFFFu);
            else
                trans.push_back((uint32_t)(total_exits + off) | ((uint32_t)cnt << 20));
            if (cnt >= 4096)
                return fail(SSE_ERR_INVALID, "transition with more than 4095 exits");
        }
        for (int e = 0; e < vd.n_exits; ++e) {
            eout.push_back((uint32_t)vd.exit_target[e] | ((uint32_t)vd.exit_leg[e] << 10) |
                           ((uint32_t)vd.exit_worm[e] << 12));
            cum.push_back(vd.exit_cumprob[e]);
        }
        total_exits += vd.n_exits;
    }
    if (total_exits >= (1 << 20))
        return fail(SSE_ERR_INVALID, "too many transition exits");
    std::vector<char> wt;
    size_t w_trans = hdr.size() * 4;
    size_t w_eout = w_trans + trans.size() * 4;
    size_t w_cum = (w_eout + eout.size() * 4 + 7) & ~(size_t)7;
    size_t w_bytes = (w_cum + cum.size() * 8 + 15) & ~(size_t)15;
    wt.assign(w_bytes, 0);
    memcpy(wt.data(), hdr.data(), hdr.size() * 4);
    if (!trans.empty())
        memcpy(wt.data() + w_trans, trans.data(), trans.size() * 4);
    if (!eout.empty())
        memcpy(wt.data() + w_eout, eout.data(), eout.size() * 4);
    if (!cum.empty())
        memcpy(wt.data() + w_cum, cum.data(), cum.size() * 8);
    size_t o_wtab = put(wt.data(), wt.size());
    // chase table: one entry per (ecode, worm-1, leg_in), ecode = type << lgp | vcode
    // with P = 2^lgp >= 2 * max vertex count, so both diagonal-flag variants of a
    // vertex index address the same transitions (the flag rides in the vcode)
    int dnv = 1, dmw = 1;
    for (int t = 0; t < m.n_types; ++t) {
        dnv = std::max(dnv, (int)m.types[t].n_vertices);
        dmw = std::max(dmw, (int)m.types[t].max_worm_count);
    }
    int lgp = 1;
    while ((1 << lgp) < 2 * dnv)
        ++lgp;
    const size_t P = (size_t)1 << lgp;
    // (the wave kernels' tables; built for two-site bonds only, the bonds
    // the wave kernels take)
    std::vector<DirectEntry> dt(ns == 2 ? (size_t)m.n_types * P * dmw * 4 : 1);
    for (auto &x : dt) {  // invalid transition: every word takes the exact path, which fails
        for (int j = 0; j < 4; ++j) {
            x.K[j] = 0;
            x.out[j] = kChaseSlow;
        }
        x.count = 0;
        x.trans = 0xFFFFFFFFu;
        x.pad[0] = x.pad[1] = 0;
    }
    if (ns == 2) {
        int ebase = 0;
        for (int t = 0; t < m.n_types; ++t) {
            const sse_vertex_data_desc &vd = m.types[t];
            for (int v = 0; v < vd.n_vertices; ++v)
                for (int wv = 0; wv < vd.max_worm_count; ++wv)
                    for (int l = 0; l < 4; ++l) {
                        int src = (v * vd.max_worm_count + wv) * 4 + l;
                        int off = vd.transition_offset[src], cnt = vd.transition_count[src];
                        if (off < 0 || cnt <= 0)
                            continue;
                        DirectEntry x;
                        x.count = (uint32_t)cnt;
                        x.trans = (uint32_t)(ebase + off) | ((uint32_t)cnt << 20);
                        x.pad[0] = x.pad[1] = 0;
                        // r < c  <=>  (word >> 11) < ceil(c * 2^53)
                        auto thr = [](double c) {
                            return c <= 0.0 ? 0ull
                                   : c >= 1.0 ? (1ull << 53)
                                              : (uint64_t)std::ceil(std::ldexp(c, 53));
                        };
                        auto word = [&](int j) {
                            const uint32_t ec = (uint32_t)t << lgp | (uint32_t)vd.exit_target[off + j];
                            return dmw == 1 ? chase_out1(ec, vd.exit_leg[off + j])
                                            : chase_out(ec, vd.exit_leg[off + j], vd.exit_worm[off + j]);
                        };
                        // slots past the last exit: K = 2^53 (always taken), slow marker
                        for (int j = 0; j < 4; ++j) {
                            x.K[j] = j < cnt ? thr(vd.exit_cumprob[off + j]) : (1ull << 53);
                            x.out[j] = j < cnt ? word(j) : kChaseSlow;
                        }
                        if (cnt > 4) {  // first three inline, the rest from the exit list
                            x.K[3] = 1ull << 53;
                            x.out[3] = kChaseSlow;
                        } else if (cnt == 4) {
                            x.pad[0] = x.out[3];
                            if (x.K[3] < (1ull << 53))  // the fourth exit needs its compare
                                x.out[3] = kChaseSlow;
                        }
                        for (int dg = 0; dg < 2; ++dg)
                            dt[((((size_t)t << lgp) + (2 * v + dg)) * dmw + wv) * 4 + l] = x;
                    }
            ebase += vd.n_exits;
        }
    }
    size_t o_dtab = put(dt.data(), sizeof(DirectEntry) * dt.size());
    // the generic wave kernels' chase table (every bond size): the same
    // entries with PL = 2^glg legs per (ecode, worm) and gen_out exit words
    const int glg = legs <= 2 ? 1 : legs <= 4 ? 2 : 3;
    std::vector<DirectEntry> gt((size_t)m.n_types * P * dmw << glg);
    for (auto &x : gt) {
        for (int j = 0; j < 4; ++j) {
            x.K[j] = 0;
            x.out[j] = kChaseSlow;
        }
        x.count = 0;
        x.trans = 0xFFFFFFFFu;
        x.pad[0] = x.pad[1] = 0;
    }
    {
        auto thr = [](double c) {
            return c <= 0.0 ? 0ull : c >= 1.0 ? (1ull << 53) : (uint64_t)std::ceil(std::ldexp(c, 53));
        };
        for (int t = 0; t < m.n_types; ++t) {
            const sse_vertex_data_desc &vd = m.types[t];
            for (int v = 0; v < vd.n_vertices; ++v)
                for (int wv = 0; wv < vd.max_worm_count; ++wv)
                    for (int l = 0; l < legs; ++l) {
                        int src = (v * vd.max_worm_count + wv) * legs + l;
                        int off = vd.transition_offset[src], cnt = vd.transition_count[src];
                        if (off < 0 || cnt <= 0)
                            continue;
                        DirectEntry x;
                        x.count = (uint32_t)cnt;
                        x.trans = (uint32_t)off;  // first exit in the type's own exit list
                        x.pad[0] = x.pad[1] = 0;
                        auto word = [&](int j) {
                            const uint32_t ec = (uint32_t)t << lgp | (uint32_t)vd.exit_target[off + j];
                            return gen_out(ec, vd.exit_leg[off + j], vd.exit_worm[off + j]);
                        };
                        for (int j = 0; j < 4; ++j) {
                            x.K[j] = j < cnt ? thr(vd.exit_cumprob[off + j]) : (1ull << 53);
                            x.out[j] = j < cnt ? word(j) : kChaseSlow;
                        }
                        if (cnt > 4) {
                            x.K[3] = 1ull << 53;
                            x.out[3] = kChaseSlow;
                        } else if (cnt == 4) {
                            x.pad[0] = x.out[3];
                            if (x.K[3] < (1ull << 53))
                                x.out[3] = kChaseSlow;
                        }
                        for (int dg = 0; dg < 2; ++dg)
                            gt[(((((size_t)t << lgp) + (2 * v + dg)) * dmw + wv) << glg) + l] = x;
                    }
        }
    }
    size_t o_gtab = put(gt.data(), sizeof(DirectEntry) * gt.size());
    // measurement tables: per (bond, vertex) the change of each estimator
    // across the vertex, the same expression in the same order as the
    // measurement kernels' own (two sites: cf[a] (mv[a][u0] - mv[a][l0]) +
    // cf[b] (mv[b][u1] - mv[b][l1]); other sizes: 0 + the site terms in bond
    // order), and the vertex's sign
    int any_neg = 0;
    for (int t = 0; t < m.n_types; ++t)
        for (int v = 0; v < m.types[t].n_vertices; ++v)
            any_neg |= m.types[t].signs[v] < 0;
    const size_t edt_n = (size_t)m.n_bonds * dnv;
    const bool edt = edt_n * (ne + 1) * sizeof(double) <= ((size_t)16 << 20);
    size_t o_edt = 0, o_esg = 0;
    if (edt) {
        std::vector<double> dtb(edt_n * ne + 1, 0.0);
        std::vector<int8_t> sgb(edt_n, 0);
        const int D = m.n_estimators ? m.est_max_dim : 1;
        for (int bd = 0; bd < m.n_bonds; ++bd) {
            const sse_vertex_data_desc &vd = m.types[m.bond_type[bd]];
            const int32_t *bs = m.bond_sites + (size_t)ns * bd;
            for (int v = 0; v < vd.n_vertices; ++v) {
                const size_t ix = (size_t)bd * dnv + v;
                sgb[ix] = vd.signs[v];
                const uint8_t *ls = vd.leg_states + (size_t)legs * v;
                for (size_t k = 0; k < ne; ++k) {
                    const double *cf = m.est_coef + k * m.n_sites;
                    const double *mk = m.est_mval + k * m.n_sites * D;
                    auto term = [&](int j) {
                        const size_t st = (size_t)bs[j] * D;
                        return cf[bs[j]] * (mk[st + ls[ns + j] - 1] - mk[st + ls[j] - 1]);
                    };
                    double delta;
                    if (ns == 2) {
                        delta = term(0);
                        delta += term(1);
                    } else {
                        delta = 0.0;
                        for (int j = 0; j < ns; ++j)
                            delta += term(j);
                    }
                    dtb[ix * ne + k] = delta;
                }
            }
        }
        o_edt = put(dtb.data(), sizeof(double) * dtb.size());
        o_esg = put(sgb.data(), sgb.size());
    }
    // diagonal-update table (layout: DevModel::k1tab)
    std::vector<char> k1((size_t)16 * m.n_types);
    for (int t = 0; t < m.n_types && ns == 2; ++t) {
        const sse_vertex_data_desc &vd = m.types[t];
        auto align8 = [&]() { k1.resize((k1.size() + 7) & ~(size_t)7); };
        align8();
        const uint32_t w_off = (uint32_t)k1.size();
        k1.resize(k1.size() + 8 * (size_t)vd.n_vertices);
        memcpy(k1.data() + w_off, vd.weights, 8 * (size_t)vd.n_vertices);
        const uint32_t dv_off = (uint32_t)k1.size();
        const int ndv = vd.dims[0] * vd.dims[1];
        k1.resize(k1.size() + 2 * (size_t)ndv);
        for (int i = 0; i < ndv; ++i) {
            const int16_t c = (int16_t)vd.diagonal_vertices[i];
            memcpy(k1.data() + dv_off + 2 * i, &c, 2);
        }
        const uint32_t ls_off = (uint32_t)k1.size();
        k1.resize(k1.size() + 4 * (size_t)vd.n_vertices);
        memcpy(k1.data() + ls_off, vd.leg_states, 4 * (size_t)vd.n_vertices);
        const uint32_t h[4] = {w_off, dv_off, ls_off,
                               (uint32_t)vd.n_vertices | ((uint32_t)vd.dims[0] << 16)};
        memcpy(k1.data() + 16 * (size_t)t, h, 16);
    }
    k1.resize((k1.size() + 15) & ~(size_t)15);
    size_t o_k1 = put(k1.data(), k1.size());
    // the generic diagonal-update kernel's lookups (layout: DevModel::k1ntab)
    std::vector<char> kn((size_t)32 * m.n_types + (size_t)4 * (ns + 1) * m.n_bonds);
    const uint32_t kn_bond_off = (uint32_t)(32 * (size_t)m.n_types);
    for (int bd = 0; bd < m.n_bonds; ++bd) {
        int32_t rec[5] = {m.bond_type[bd], 0, 0, 0, 0};
        for (int k = 0; k < ns; ++k)
            rec[1 + k] = m.bond_sites[(size_t)ns * bd + k];
        memcpy(kn.data() + kn_bond_off + (size_t)4 * (ns + 1) * bd, rec, (size_t)4 * (ns + 1));
    }
    for (int t = 0; t < m.n_types; ++t) {
        const sse_vertex_data_desc &vd = m.types[t];
        size_t states = 1;
        uint32_t dimspack = 0;
        for (int i = 0; i < ns; ++i) {
            states *= (size_t)vd.dims[i];
            dimspack |= (uint32_t)vd.dims[i] << (8 * i);
        }
        kn.resize((kn.size() + 7) & ~(size_t)7);
        const uint32_t w_off = (uint32_t)kn.size();
        kn.resize(kn.size() + 8 * (size_t)vd.n_vertices);
        memcpy(kn.data() + w_off, vd.weights, 8 * (size_t)vd.n_vertices);
        const uint32_t dv_off = (uint32_t)kn.size();
        kn.resize(kn.size() + 4 * states);
        memcpy(kn.data() + dv_off, vd.diagonal_vertices, 4 * states);
        const uint32_t up_off = (uint32_t)kn.size();
        kn.resize(kn.size() + 4 * (size_t)vd.n_vertices);
        for (int v = 0; v < vd.n_vertices; ++v) {
            uint32_t u = 0;
            for (int k = 0; k < ns; ++k)
                u |= (uint32_t)vd.leg_states[(size_t)2 * ns * v + ns + k] << (8 * k);
            memcpy(kn.data() + up_off + 4 * (size_t)v, &u, 4);
        }
        const uint32_t h[8] = {w_off, dv_off, up_off, (uint32_t)vd.n_vertices, dimspack, 0, 0, 0};
        memcpy(kn.data() + 32 * (size_t)t, h, 32);
    }
    kn.resize((kn.size() + 15) & ~(size_t)15);
    size_t o_kn = put(kn.data(), kn.size());

    char *d = nullptr;
    int rc = dev_alloc(b, &d, blob.size() + 16);
    if (rc)
        return rc;
    std::vector<DevVD> types(m.n_types);
    for (int t = 0; t < m.n_types; ++t) {
        const sse_vertex_data_desc &vd = m.types[t];
        DevVD &x = types[t];
        x.dims0 = vd.dims[0];
        x.dims1 = ns >= 2 ? vd.dims[1] : 1;
        x.nsites = ns;
        for (int i = 0; i < 4; ++i)
            x.dims[i] = i < ns ? vd.dims[i] : 1;
        x.nv = vd.n_vertices;
        x.max_worm = vd.max_worm_count;
        x.n_exits = vd.n_exits;
        x.energy_offset = vd.energy_offset;
        x.weights = (const double *)(d + voff[t].w);
        x.signs = (const int8_t *)(d + voff[t].s);
        x.leg_states = (const uint8_t *)(d + voff[t].ls);
        x.diag_vertex = (const int32_t *)(d + voff[t].dv);
        x.toff = (const int32_t *)(d + voff[t].to);
        x.tcnt = (const int32_t *)(d + voff[t].tc);
        x.cum = (const double *)(d + voff[t].cum);
        x.tgt = (const int32_t *)(d + voff[t].tgt);
        x.eleg = (const uint8_t *)(d + voff[t].el);
        x.ewrm = (const uint8_t *)(d + voff[t].ew);
    }
    memcpy(blob.data() + o_types, types.data(), sizeof(DevVD) * m.n_types);
    HIPCHECK(bmemcpy(b, d, blob.data(), blob.size(), hipMemcpyHostToDevice));
    dm.n_sites = m.n_sites;
    dm.n_bonds = m.n_bonds;
    dm.nsb = ns;
    dm.pad_nsb = 0;
    dm.n_types = m.n_types;
    dm.n_est = m.n_estimators;
    dm.est_max_dim = m.n_estimators ? m.est_max_dim : 1;
    dm.max_dim = max_dim;
    dm.energy_offset = m.energy_offset;
    dm.norm = m.normalization_site_count;
    dm.site_dims = (const uint8_t *)(d + o_sd);
    dm.bond_type = (const int32_t *)(d + o_bt);
    dm.bond_sites = (const int32_t *)(d + o_bs);
    dm.bond_rec = (const int2 *)(d + o_br);
    dm.est_coef = (const double *)(d + o_ec);
    dm.est_mval = (const double *)(d + o_em);
    dm.types = (const DevVD *)(d + o_types);
    dm.wtab = (const uint32_t *)(d + o_wtab);
    dm.wtab_bytes = (int32_t)w_bytes;
    dm.wtab_trans_off = (int32_t)w_trans;
    dm.wtab_eout_off = (int32_t)w_eout;
    dm.wtab_cum_off = (int32_t)w_cum;
    dm.dtab = (const DirectEntry *)(d + o_dtab);
    dm.dt_nv = dnv;
    dm.dt_mw = dmw;
    dm.dt_lgp = lgp;
    dm.dtab_bytes = (int32_t)(sizeof(DirectEntry) * dt.size());
    dm.k1tab = (const uint8_t *)(d + o_k1);
    dm.k1tab_bytes = (int32_t)k1.size();
    dm.pad_k1 = 0;
    dm.k1ntab = (const uint8_t *)(d + o_kn);
    dm.k1ntab_bytes = (int32_t)kn.size();
    dm.k1n_bond_off = (int32_t)kn_bond_off;
    dm.gtab = (const DirectEntry *)(d + o_gtab);
    dm.gtab_bytes = (int32_t)(sizeof(DirectEntry) * gt.size());
    dm.g_lg = glg;
    dm.est_dtab = edt ? (const double *)(d + o_edt) : nullptr;
    dm.est_dsgn = edt ? (const int8_t *)(d + o_esg) : nullptr;
    dm.est_dnv = dnv;
    dm.any_neg = any_neg;
    return SSE_OK;
}

static size_t env_size(const char *name, size_t dflt)
{
    const char *s = getenv(name);
    if (!s || !*s)
        return dflt;
    return (size_t)strtoull(s, nullptr, 10);
}

// host-side bump allocation of a slab (only while no kernel runs)
static char *host_slab(sse_batch *b, int64_t cap, int legs)
{
    size_t bytes = slab_bytes(cap, legs);
    if (b->pool_host_next + bytes > b->pool_size)
        return nullptr;
    char *p = b->pool_base + b->pool_host_next;
    b->pool_host_next += bytes;
    return p;
}

static int sync_pool_to_device(sse_batch *b)
{
    DevPool p;
    p.next = b->pool_host_next;
    p.size = b->pool_size;
    p.base = b->pool_base;
    p.overflow = 0;
    p.status = 0;
    p.max_n = 0;
    p.max_L = 0;
    HIPCHECK(bmemcpy(b, b->d_pool, &p, sizeof p, hipMemcpyHostToDevice));
    return SSE_OK;
}

static int sync_pool_from_device(sse_batch *b)
{
    DevPool p;
    HIPCHECK(bmemcpy(b, &p, b->d_pool, sizeof p, hipMemcpyDeviceToHost));
    b->pool_host_next = p.next > p.size ? p.size : p.next;
    if (p.overflow)
        return fail(SSE_ERR_OOM, "slab pool exhausted; raise SSE_POOL_MB");
    return SSE_OK;
}

static int check_walker_errors(sse_batch *b)
{
    DevPool p;
    HIPCHECK(hipStreamSynchronize(b->stream));
    HIPCHECK(bmemcpy(b, &p, b->d_pool, sizeof p, hipMemcpyDeviceToHost));
    if (p.overflow)
        return fail(SSE_ERR_OOM, "slab pool exhausted; raise SSE_POOL_MB");
    if (b->mode == SSE_MODE_WAVE && (p.max_n || p.max_L)) {
        // size the worm kernel's LDS from the operator counts of this call's
        // sweeps (the device maxima restart at zero for the next call): one
        // size for single-stream launches, one per stream group
        b->seen_max_n = p.max_n;
        b->seen_max_L = p.max_L;
        const unsigned long long zero[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        static_assert(sizeof zero == sizeof p.max_n + sizeof p.max_L + sizeof p.group_max_n,
                      "pool maxima layout");
        HIPCHECK(bmemcpy(b, &b->d_pool->max_n, zero, sizeof zero, hipMemcpyHostToDevice));
        wave_resize_lds(b->wave, b->seen_max_n, b->seen_max_L);
        for (int g = 0; g < 7; ++g)
            if (p.group_max_n[g])
                b->wave.lds_k2_group[g] = wave_k2_lds_for(b->wave, p.group_max_n[g], p.max_L);
    }
    if (p.status) {
        if (p.status & (1 << SSE_ERR_SCATTER))
            return fail(SSE_ERR_SCATTER, "scatter fell through (r >= last cumprob) or hit an "
                                         "invalid transition");
        return fail(SSE_ERR_STATE, "walker error flag set");
    }
    return SSE_OK;
}

extern "C" {

int32_t sse_abi_version(void) { return SSEGPU_ABI_VERSION; }

const char *sse_last_error(void) { return g_err.c_str(); }

uint64_t sse_opcode_to_reference(uint32_t op)
{
    if (op == 0)
        return 0;
    uint64_t vc = (op >> 1) & 1023u;
    uint64_t bond = op >> SSE_OP_BOND_SHIFT;
    uint64_t vref = (vc & 1) | (((vc >> 1) + 1) << 1);
    return 1ull | (vref << 1) | ((bond + 1) << 26);
}

uint32_t sse_opcode_from_reference(uint64_t op)
{
    if (op == 0)
        return 0;
    uint64_t vref = (op & ((1ull << 25) - 1)) >> 1;
    uint64_t bond = (op >> 26) - 1;
    uint64_t vc = (vref & 1) | (((vref >> 1) - 1) << 1);
    return (uint32_t)(1u | (vc << 1) | (bond << SSE_OP_BOND_SHIFT));
}

static int max_dim_of(const int32_t *dims, int32_t n)
{
    int m = 1;
    for (int32_t i = 0; i < n; ++i)
        m = dims[i] > m ? dims[i] : m;
    return m;
}

int sse_vertex_transitions(int32_t n_sites, const int32_t *dims, int32_t n_vertices,
                           const uint8_t *leg_states, const double *weights, double tolerance,
                           double lp_tolerance, int32_t *trans_offset, int32_t *trans_length,
                           int64_t exit_capacity, double *cumprobs, int32_t *targets,
                           int32_t *out_leg, int32_t *out_worm, int64_t *n_exits, int32_t *flags)
{
    if (!dims || !leg_states || !weights || !trans_offset || !trans_length || !n_exits ||
        !flags || n_sites < 1 || n_sites > 8 || n_vertices < 0 || exit_capacity < 0 ||
        (exit_capacity > 0 && (!cumprobs || !targets || !out_leg || !out_worm)))
        return fail(SSE_ERR_INVALID, "sse_vertex_transitions: invalid arguments");
    ssevd::Transitions t;
    int rc;
    try {
        rc = ssevd::build_transitions(n_sites, dims, n_vertices, leg_states, weights, tolerance,
                                      lp_tolerance, t);
    } catch (...) {
        return fail(SSE_ERR_OOM, "sse_vertex_transitions: out of host memory");
    }
    static const char *const what[] = {
        "", "sse_vertex_transitions: bad dimensions or vertex count",
        "sse_vertex_transitions: a transition LP has no optimum",
        "sse_vertex_transitions: LP solution below its bound",
        "sse_vertex_transitions: LP solution above its bound",
        "sse_vertex_transitions: transition to an invalid vertex"};
    if (rc)
        return fail(SSE_ERR_INVALID, what[rc]);
    *n_exits = (int64_t)t.cumprobs.size();
    *flags = (t.inaccurate ? 1 : 0) | (t.norm_warning ? 2 : 0) | (t.alt_optima ? 4 : 0);
    if ((int64_t)t.cumprobs.size() > exit_capacity)
        return fail(SSE_ERR_INVALID, "sse_vertex_transitions: exit_capacity too small");
    const size_t nt = (size_t)n_vertices * (size_t)(max_dim_of(dims, n_sites) - 1) * 2 * n_sites;
    if (t.offset.empty()) {  // no entries: the caller's table is all zeros
        memset(trans_offset, 0, sizeof(int32_t) * nt);
        memset(trans_length, 0, sizeof(int32_t) * nt);
    } else {
        memcpy(trans_offset, t.offset.data(), sizeof(int32_t) * t.offset.size());
        memcpy(trans_length, t.length.data(), sizeof(int32_t) * t.length.size());
    }
    if (!t.cumprobs.empty()) {
        memcpy(cumprobs, t.cumprobs.data(), sizeof(double) * t.cumprobs.size());
        memcpy(targets, t.targets.data(), sizeof(int32_t) * t.targets.size());
        memcpy(out_leg, t.out_leg.data(), sizeof(int32_t) * t.out_leg.size());
        memcpy(out_worm, t.out_worm.data(), sizeof(int32_t) * t.out_worm.size());
    }
    return SSE_OK;
}

int sse_batch_create(const sse_model_desc *models, int32_t n_models, const sse_walker_desc *walkers,
                     int32_t n_walkers, int32_t mode, sse_batch **out)
{
    if (!out || !models || !walkers || n_models <= 0 || n_walkers <= 0)
        return fail(SSE_ERR_INVALID, "sse_batch_create: invalid arguments");
    if (mode != SSE_MODE_WAVE && mode != SSE_MODE_LANE)
        return fail(SSE_ERR_INVALID, "sse_batch_create: unknown mode");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(SSE_ERR_NO_DEVICE, "no HIP device");
    for (int k = 0; k < n_models; ++k) {
        int rc = validate_model(models[k], k);
        if (rc)
            return rc;
    }
    for (int w = 0; w < n_walkers; ++w) {
        const sse_walker_desc &wd = walkers[w];
        if (wd.model < 0 || wd.model >= n_models || !(wd.T > 0) || wd.init_num_worms < 1 ||
            wd.diagonal_warmup_sweeps < 0)
            return fail(SSE_ERR_INVALID, "sse_batch_create: invalid walker parameters");
    }
    sse_batch *b = new sse_batch();
    b->mode = mode;
    (void)hipGetDevice(&b->device);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(SSE_ERR_HIP, "hipStreamCreate failed");
    }
    for (int i = 0; i < 4; ++i)
        (void)hipEventCreate(&b->ev[i]);
    b->n_models = n_models;
    b->n_walkers = n_walkers;
    b->wdesc.assign(walkers, walkers + n_walkers);
    for (int w = 0; w < n_walkers; ++w)
        b->pending_warmup = std::max(b->pending_warmup, (int)walkers[w].diagonal_warmup_sweeps);
    b->h_models.resize(n_models);
    int rc = SSE_OK;
    for (int k = 0; k < n_models && !rc; ++k) {
        HostModelCheck hc;
        rc = upload_model(b, models[k], b->h_models[k], hc);
        if (!rc)
            b->h_check.push_back(std::move(hc));
        b->max_sites = std::max(b->max_sites, (int)models[k].n_sites);
    }
    if (!rc)
        rc = dev_alloc(b, &b->d_models, n_models);
    if (!rc && bmemcpy(b, b->d_models, b->h_models.data(), sizeof(DevModel) * n_models,
                         hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SSE_ERR_HIP, "model upload failed");
    // initial cutoffs and slab capacities
    size_t need = 0;
    std::vector<int64_t> L0(n_walkers), cap0(n_walkers);
    b->widths.resize(n_walkers);
    b->offsets.resize(n_walkers);
    b->nsites.resize(n_walkers);
    int off = 0;
    for (int w = 0; w < n_walkers && !rc; ++w) {
        const sse_walker_desc &wd = walkers[w];
        const sse_model_desc &m = models[wd.model];
        L0[w] = wd.init_opstring_cutoff >= 0 ? wd.init_opstring_cutoff
                                             : (int64_t)std::nearbyint((double)m.n_sites * wd.T);
        cap0[w] = slab_capacity(L0[w]);
        need += slab_bytes(cap0[w], 2 * m.bond_n_sites);
        b->widths[w] = 5 + 5 * m.n_estimators;
        b->offsets[w] = off;
        b->nsites[w] = m.n_sites;
        off += b->widths[w];
    }
    b->total_width = off;
    if (!rc) {
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        size_t dflt = std::max<size_t>((size_t)2 << 30, need * 64);
        dflt = std::min(dflt, (size_t)((double)free_b * 0.4));
        dflt = std::max(dflt, need * 4 + ((size_t)64 << 20));
        b->pool_size = env_size("SSE_POOL_MB", 0) ? env_size("SSE_POOL_MB", 0) << 20 : dflt;
        if (b->pool_size < need)
            rc = fail(SSE_ERR_OOM, "SSE_POOL_MB smaller than the initial slabs");
    }
    if (!rc)
        rc = dev_alloc(b, &b->pool_base, b->pool_size);
    if (!rc)
        rc = dev_alloc(b, &b->d_pool, 1);
    size_t nsp = (size_t)n_walkers * b->max_sites;
    if (!rc)
        rc = dev_alloc(b, &b->d_spins, nsp);
    if (!rc)
        rc = dev_alloc(b, &b->d_first, nsp);
    if (!rc)
        rc = dev_alloc(b, &b->d_last, nsp);
    if (!rc)
        rc = dev_alloc(b, &b->d_scratch, nsp);
    if (!rc)
        rc = dev_alloc(b, &b->d_out, (size_t)b->total_width);
    if (!rc)
        rc = dev_alloc(b, &b->d_sum, (size_t)b->total_width);
    if (!rc)
        rc = dev_alloc(b, &b->d_sumsq, (size_t)b->total_width);
    if (!rc)
        rc = dev_alloc(b, &b->d_count, (size_t)n_walkers);
    if (!rc)
        rc = dev_alloc(b, &b->d_wlf, 3 * (size_t)n_walkers);
    if (!rc)
        rc = dev_alloc(b, &b->d_done, (size_t)n_walkers);
    if (!rc)
        rc = dev_alloc(b, &b->d_walkers, (size_t)n_walkers);
    if (!rc)
        rc = dev_alloc(b, &b->d_queue, 1);
    if (!rc)
        rc = dev_alloc(b, &b->d_work, 1);
    if (!rc)
        rc = dev_alloc(b, &b->d_budget, (size_t)n_walkers);
    if (!rc)
        rc = dev_alloc(b, &b->d_task_order, (size_t)n_walkers);
    if (!rc) {
        b->h_walkers.resize(n_walkers);
        for (int w = 0; w < n_walkers && !rc; ++w) {
            const sse_walker_desc &wd = walkers[w];
            DevWalker &x = b->h_walkers[w];
            memset(&x, 0, sizeof x);
            x.model = wd.model;
            x.warmup = wd.diagonal_warmup_sweeps;
            x.T = wd.T;
            x.target_wlf = wd.target_worm_length_fraction;
            x.atten = wd.num_worms_attenuation_factor;
            x.key = wd.seed;
            x.thermalization = wd.thermalization;
            x.init_cutoff = L0[w];
            x.avg_worm_length = 1.0;
            x.num_worms = wd.init_num_worms;
            x.L = L0[w];
            x.n = 0;
            x.cap = cap0[w];
            x.ctr = 0;
            x.n_sites = b->nsites[w];
            x.last_wlf = NAN;
            char *s = host_slab(b, cap0[w], 2 * b->h_models[wd.model].nsb);
            if (!s) {
                rc = fail(SSE_ERR_OOM, "slab pool too small");
                break;
            }
            x.ops = (uint32_t *)s;
            x.links = (int32_t *)(s + (size_t)x.cap * 4);
            x.spins = b->d_spins + (size_t)w * b->max_sites;
            x.first = b->d_first + (size_t)w * b->max_sites;
            x.last = b->d_last + (size_t)w * b->max_sites;
            x.meas_offset = b->offsets[w];
            x.meas_width = b->widths[w];
            x.acc_slot = w;
            x.active = 0;
            x.pending = 1;
            x.remaining = -1;
        }
    }
    if (!rc && bmemcpy(b, b->d_walkers, b->h_walkers.data(), sizeof(DevWalker) * n_walkers,
                         hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SSE_ERR_HIP, "walker upload failed");
    if (!rc)
        rc = sync_pool_to_device(b);
    if (!rc) {
        // schedule: expected operator count n ~ beta * bonds, largest first
        // (stable, so equal walkers keep their index order)
        std::vector<int32_t> order(n_walkers);
        std::vector<double> &key = b->work_key;
        key.resize(n_walkers);
        for (int w = 0; w < n_walkers; ++w) {
            order[w] = w;
            key[w] = (double)b->h_models[walkers[w].model].n_bonds / walkers[w].T;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t c) { return key[a] > key[c]; });
        // n ~ 0.7 beta * bonds for the magnets here; beyond ~3600 operators a
        // walker outgrows the 8-byte records, and with stream groups these
        // walkers get a group (and a worm-kernel LDS size) of their own
        b->n_big = 0;
        while (b->n_big < n_walkers && 0.7 * key[order[b->n_big]] > 3600.0)
            ++b->n_big;
        rc = dev_alloc(b, &b->d_order, (size_t)n_walkers);
        if (!rc && bmemcpy(b, b->d_order, order.data(), sizeof(int32_t) * n_walkers,
                           hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(SSE_ERR_HIP, "schedule upload failed");
    }
    if (!rc) {
        (void)bmemset(b, b->d_sum, 0, sizeof(double) * b->total_width);
        (void)bmemset(b, b->d_sumsq, 0, sizeof(double) * b->total_width);
        (void)bmemset(b, b->d_count, 0, sizeof(int64_t) * n_walkers);
        (void)bmemset(b, b->d_wlf, 0, sizeof(double) * 3 * n_walkers);
        (void)bmemset(b, b->d_done, 0, sizeof(int32_t) * n_walkers);
        rc = wave_configure(b->wave, b->h_models, b->max_sites);
        if (!rc && b->wave.lds_limit > 65536)
            (void)hipFuncSetAttribute((const void *)wave_persist_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->wave.lds_limit);
        if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device) !=
                hipSuccess ||
            b->n_cu <= 0)
            b->n_cu = 256;
        // lane mode: one walker per wavefront until the batch fills 16
        // wavefronts per compute unit, then more per wavefront (powers of two
        // up to 64).  On 4096 walkers one per wavefront measured 1.5-1.7x the
        // rate of 64 per wavefront: a lane's chain waits on its own memory
        // loads, and more wavefronts per unit hide more of that wait.
        int lw = 1;
        while (lw < 64 && (int64_t)n_walkers >= (int64_t)2 * lw * 16 * b->n_cu)
            lw *= 2;
        b->lane_width = lw;
    }
    if (rc) {
        std::string keep = g_err;
        sse_batch_destroy(b);
        g_err = keep;
        return rc;
    }
    *out = b;
    return SSE_OK;
}

int sse_batch_destroy(sse_batch *b)
{
    if (!b)
        return SSE_OK;
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    for (void *p : b->allocs)
        (void)hipFree(p);
    for (int i = 0; i < 4; ++i)
        if (b->ev[i])
            (void)hipEventDestroy(b->ev[i]);
    for (hipEvent_t e : b->evpool)
        (void)hipEventDestroy(e);
    for (hipEvent_t e : b->kevpool)
        (void)hipEventDestroy(e);
    if (b->gexec)
        (void)hipGraphExecDestroy(b->gexec);
    for (hipStream_t gs : b->gstreams)
        (void)hipStreamDestroy(gs);
    for (hipEvent_t e : b->gjoin)
        (void)hipEventDestroy(e);
    if (b->gfork)
        (void)hipEventDestroy(b->gfork);
    for (int i = 0; i < 2; ++i)
        if (b->call_ev[i])
            (void)hipEventDestroy(b->call_ev[i]);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
    return SSE_OK;
}

// ev (optional): 3 events recorded after the diagonal/vertex-list kernel,
// after the worm kernel (lane mode: both after the single sweep kernel)
static int launch_sweep(sse_batch *b, int phases, hipEvent_t *ev = nullptr)
{
    int n = b->n_walkers;
    if (b->mode == SSE_MODE_LANE) {
        const int threads = b->lane_width;
        const int blocks = (n + threads - 1) / threads;
        hipLaunchKernelGGL(lane_sweep_kernel, dim3(blocks), dim3(threads), 0, b->stream,
                           b->d_walkers, b->d_models, b->d_pool, 0, n, phases,
                           (const int32_t *)nullptr, 0);
        if (ev) {
            HIPCHECK(hipEventRecord(ev[1], b->stream));
            HIPCHECK(hipEventRecord(ev[2], b->stream));
        }
    } else {
        int rc = wave_launch_sweep(b->wave, b->stream, b->d_walkers, b->d_models, b->d_pool, 0, n,
                                   phases, ev, b->use_order ? b->d_order : nullptr, -1);
        if (rc)
            return rc;
    }
    HIPCHECK(hipGetLastError());
    return SSE_OK;
}

static int ensure_kernel_events(sse_batch *b, size_t count)
{
    while (b->kevpool.size() < count) {
        hipEvent_t e;
        HIPCHECK(hipEventCreate(&e));
        b->kevpool.push_back(e);
    }
    return SSE_OK;
}

static int ensure_events(sse_batch *b, size_t count)
{
    while (b->evpool.size() < count) {
        hipEvent_t e;
        HIPCHECK(hipEventCreate(&e));
        b->evpool.push_back(e);
    }
    return SSE_OK;
}

// per-phase device time summed over the sweeps of the last call
static void collect_phase_times(sse_batch *b, int n_sweeps, bool with_measure)
{
    double t[3] = {0, 0, 0};
    for (int k = 0; k < n_sweeps; ++k) {
        hipEvent_t *e = &b->evpool[4 * (size_t)k];
        float a = 0, c = 0, d = 0;
        (void)hipEventElapsedTime(&a, e[0], e[1]);
        (void)hipEventElapsedTime(&c, e[1], e[2]);
        if (with_measure)
            (void)hipEventElapsedTime(&d, e[2], e[3]);
        t[0] += a;
        t[1] += c;
        t[2] += d;
    }
    for (int i = 0; i < 3; ++i)
        b->phase_ms[i] = t[i];
    b->phase_launches = n_sweeps;
}

static void clear_phase_times(sse_batch *b)
{
    for (int i = 0; i < 3; ++i)
        b->phase_ms[i] = 0;
    b->phase_launches = 0;
}

// Per-kernel durations of the last call's timed launches (SSE_OPT_TIMING 2),
// read after the call has synchronized.
static void collect_kernel_times(sse_batch *b)
{
    for (int i = 0; i < 4; ++i) {
        b->kern_ms[i] = 0;
        b->kern_launches[i] = 0;
    }
    for (const auto &k : b->kern_ev) {
        float a = 0, c = 0, d = 0;
        const hipEvent_t *e = &b->kevpool[k.e];
        if (k.kind == 3) {
            (void)hipEventElapsedTime(&a, e[0], e[1]);
            b->kern_ms[3] += a;
            b->kern_launches[3] += 1;
            continue;
        }
        // kind 0: K1, K2, K3; 1: K1, K2 (no measure); 4: K1, fused K2 + K3
        // (counted as the worm kernel)
        (void)hipEventElapsedTime(&a, e[0], e[1]);
        (void)hipEventElapsedTime(&c, e[1], e[2]);
        b->kern_ms[0] += a;
        b->kern_ms[1] += c;
        b->kern_launches[0] += 1;
        b->kern_launches[1] += 1;
        if (k.kind == 0) {
            (void)hipEventElapsedTime(&d, e[2], e[3]);
            b->kern_ms[2] += d;
            b->kern_launches[2] += 1;
        }
    }
    b->kern_ev.clear();
}

static int launch_measure(sse_batch *b, int mode);
static int launch_sweep_measure(sse_batch *b, int measure, hipEvent_t *ev = nullptr);

// Replay the captured launch sequence of key k, capturing issue() into a
// HIP graph first when the cached one was made for another key.  No memset
// nodes: accumulators and counters are cleared by kernels.  Launches issued
// on other streams during the capture join it through events (the stream
// groups' fork and join), so a multi-stream pipeline becomes one graph.
}  // extern "C"
template <class F>
static int graph_replay(sse_batch *b, const GraphKey &k, F &&issue)
{
    if (!b->gexec || !b->gkey.same(k)) {
        if (b->gexec) {
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = issue();
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess)
            return fail(SSE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->gexec = nullptr;
            return fail(SSE_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        b->gkey = k;
    }
    HIPCHECK(hipGraphLaunch(b->gexec, b->stream));
    return SSE_OK;
}

static GraphKey graph_key(const sse_batch *b, int kind, int sweeps, int measure)
{
    GraphKey k;
    k.kind = kind;
    k.sweeps = sweeps;
    k.measure = measure;
    k.rec_mode = b->wave.rec_mode;
    k.chase = b->wave.chase;
    k.fuse = b->wave.fuse_measure;
    k.lane_width = b->lane_width;
    k.generic = b->wave.generic;
    k.use_order = b->use_order;
    k.models = b->d_models;
    k.lds[0] = b->wave.lds_k2;
    for (int g = 0; g < 7; ++g)
        k.lds[1 + g] = b->wave.lds_k2_group[g];
    k.lds[8] = b->wave.lds_k1;
    return k;
}

extern "C" {

// Capture `sweeps` x (sweep [+ fused measure]) on the batch stream.
static int graph_run(sse_batch *b, int sweeps, int measure)
{
    return graph_replay(b, graph_key(b, 1, sweeps, measure), [&]() {
        int rc = SSE_OK;
        for (int k = 0; k < sweeps && !rc; ++k)
            rc = launch_sweep_measure(b, measure);
        return rc;
    });
}

// Device time of a sweep/run call: stage 0 records the start event, 1 the
// end event, 2 (after the call synchronized) reads the elapsed time.
static int call_timer(sse_batch *b, int stage, int64_t sweeps)
{
    if (stage == 0) {
        b->call_ms = 0;
        b->call_sweeps = 0;
    }
    if (!b->use_timing)
        return SSE_OK;
    for (int i = 0; i < 2; ++i)
        if (!b->call_ev[i])
            HIPCHECK(hipEventCreate(&b->call_ev[i]));
    if (stage < 2) {
        HIPCHECK(hipEventRecord(b->call_ev[stage], b->stream));
        if (stage == 1)
            b->call_sweeps = sweeps;
        return SSE_OK;
    }
    float t = 0;
    HIPCHECK(hipEventElapsedTime(&t, b->call_ev[0], b->call_ev[1]));
    b->call_ms = t;
    return SSE_OK;
}

// Wave mode with n_groups > 1: walkers [g*n/G, (g+1)*n/G) are swept on stream
// g.  The group streams fork from and join back into the batch stream, so
// everything else (copies, syncs, the next call) stays ordered on it.  Sweeps
// are issued round-robin over the groups; a group whose slowest walker is
// still in its worm update leaves the rest of the device to the others.
// Hardware queues the HIP runtime gives this process: GPU_MAX_HW_QUEUES, 4
// when unset (HIP's default).  Streams beyond it share queues and serialize.
static int hw_queues()
{
    static const int q = [] {
        const char *e = getenv("GPU_MAX_HW_QUEUES");
        const int v = e ? atoi(e) : 0;
        return v > 0 ? v : 4;
    }();
    return q;
}

// Walker groups of a wave-mode sweep: the SSE_OPT_STREAMS value, or
// automatically one group per 512 walkers, at least 2 and at most one per
// hardware queue (smaller groups would leave CUs idle while they drain).
static int group_count(const sse_batch *b)
{
    const int n = b->n_walkers;
    const int g = b->n_groups > 0 ? b->n_groups : std::min(hw_queues(), std::max(2, n / 512));
    return std::max(1, std::min(g, n));
}

// Streams and events of G stream groups (created before any capture).
static int groups_prepare(sse_batch *b, int G)
{
    while ((int)b->gstreams.size() < G - 1) {
        hipStream_t gs;
        HIPCHECK(hipStreamCreateWithFlags(&gs, hipStreamNonBlocking));
        b->gstreams.push_back(gs);
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->gjoin.push_back(e);
    }
    if (!b->gfork)
        HIPCHECK(hipEventCreateWithFlags(&b->gfork, hipEventDisableTiming));
    return SSE_OK;
}

static int groups_run(sse_batch *b, int n_sweeps, int measure, bool ktime = false)
{
    const int n = b->n_walkers;
    const int G = group_count(b);
    size_t ev0 = 0;
    if (ktime) {
        ev0 = b->kern_ev.size() * 4;
        int rc = ensure_kernel_events(b, ev0 + 4 * (size_t)G * n_sweeps);
        if (rc)
            return rc;
    }
    // group 0 runs on the batch stream itself: G groups use G streams (the
    // device maps streams onto a few hardware queues; streams sharing one
    // queue would serialize behind each other)
    int rc0 = groups_prepare(b, G);
    if (rc0)
        return rc0;
    HIPCHECK(hipEventRecord(b->gfork, b->stream));
    for (int g = 1; g < G; ++g)
        HIPCHECK(hipStreamWaitEvent(b->gstreams[g - 1], b->gfork, 0));
    // group bounds in schedule positions: equal shares, or, when the schedule
    // leads with large walkers, those in group 0 and the rest shared equally
    std::vector<int> bound(G + 1);
    const int big = b->use_order && G > 1 && b->n_big > 0 && b->n_big < n ? b->n_big : 0;
    for (int g = 0; g <= G; ++g)
        bound[g] = big ? (g == 0 ? 0 : big + (int)((int64_t)(n - big) * (g - 1) / (G - 1)))
                       : (int)((int64_t)n * g / G);
    for (int k = 0; k < n_sweeps; ++k) {
        for (int g = 0; g < G; ++g) {
            hipStream_t gs = g == 0 ? b->stream : b->gstreams[g - 1];
            const int w0 = bound[g], w1 = bound[g + 1];
            if (w0 >= w1)
                continue;
            const int32_t *ord = b->use_order ? b->d_order : nullptr;
            hipEvent_t *e = nullptr;
            const bool fused = measure && b->wave.fuse_measure;
            if (ktime) {
                const size_t i = ev0 + 4 * ((size_t)k * G + g);
                e = &b->kevpool[i];
                b->kern_ev.push_back({fused ? 4 : measure ? 0 : 1, i});
                HIPCHECK(hipEventRecord(e[0], gs));
            }
            int rc;
            if (fused) {
                rc = wave_launch_sweep_measure(b->wave, gs, b->d_walkers, b->d_models, b->d_pool, w0,
                                               w1, e, ord, g, b->d_out, b->d_sum, b->d_sumsq,
                                               b->d_count, b->d_wlf, b->d_done);
            } else {
                rc = wave_launch_sweep(b->wave, gs, b->d_walkers, b->d_models, b->d_pool, w0, w1, 7,
                                       e, ord, g);
                if (!rc && measure)
                    rc = wave_launch_measure(b->wave, gs, b->d_walkers, b->d_models, b->d_out,
                                             b->d_sum, b->d_sumsq, b->d_count, b->d_wlf, w0, w1, 1,
                                             b->d_done, ord);
            }
            if (!rc && ktime)
                rc = hipEventRecord(e[3], gs) == hipSuccess ? SSE_OK : SSE_ERR_HIP;
            if (rc)
                return fail(rc, "group sweep launch failed");
        }
    }
    for (int g = 1; g < G; ++g) {
        HIPCHECK(hipEventRecord(b->gjoin[g - 1], b->gstreams[g - 1]));
        HIPCHECK(hipStreamWaitEvent(b->stream, b->gjoin[g - 1], 0));
    }
    return SSE_OK;
}

// Stream groups through graph replays: the fork/join pipeline of a chunk of
// kGroupChunk sweeps is captured once (the group streams join the capture of
// the batch stream through the fork event) and replayed; a remainder shorter
// than a chunk is launched directly.
constexpr int kGroupChunk = 10;
static int groups_graph_many(sse_batch *b, int n_sweeps, int measure)
{
    const int G = group_count(b);
    int rc = groups_prepare(b, G);
    if (rc)
        return rc;
    int done = 0;
    for (; done + kGroupChunk <= n_sweeps; done += kGroupChunk) {
        GraphKey k = graph_key(b, 2, kGroupChunk, measure);
        k.groups = G;
        rc = graph_replay(b, k, [&]() { return groups_run(b, kGroupChunk, measure); });
        if (rc)
            return rc;
    }
    return done < n_sweeps ? groups_run(b, n_sweeps - done, measure) : SSE_OK;
}

// The persistent work-queue launch: n_tasks tasks in the order `order`,
// budgets per walker or n_sweeps each (sse_persist.h), captured as a graph
// when graphs are on.
static int persist_issue(sse_batch *b, int n_sweeps, int measure, const int32_t *budgets,
                         int n_tasks, const int32_t *order)
{
    auto issue = [&]() {
        int rc = wave_launch_persist(b->wave, b->stream, b->d_walkers, b->d_models, b->d_pool, b->d_out,
                                     b->d_sum, b->d_sumsq, b->d_count, b->d_wlf, b->d_done, order,
                                     n_tasks, budgets, n_sweeps, measure, b->d_queue, b->n_cu);
        return rc ? fail(rc, "persistent launch failed") : SSE_OK;
    };
    if (!b->use_graph)
        return issue();
    GraphKey k = graph_key(b, 3, n_sweeps, measure);
    k.n_tasks = n_tasks;
    k.budgets = budgets;
    k.order = order;
    k.lds[0] = persist_lds_bytes(b->wave);
    return graph_replay(b, k, issue);
}

// n_sweeps through replays of a captured chunk of up to 32 sweeps
static int graph_run_many(sse_batch *b, int n_sweeps, int measure)
{
    const int chunk = n_sweeps < 32 ? n_sweeps : 32;
    int done = 0;
    while (done + chunk <= n_sweeps) {
        int rc = graph_run(b, chunk, measure);
        if (rc)
            return rc;
        done += chunk;
    }
    for (; done < n_sweeps; ++done) {
        int rc = launch_sweep_measure(b, measure);
        if (rc)
            return rc;
    }
    return SSE_OK;
}

// Carlo.init! of every pending walker (all walkers after creation, or the
// slots restarted by sse_slot_start)
static int launch_init_pending(sse_batch *b)
{
    int n = b->n_walkers;
    if (b->mode == SSE_MODE_LANE) {
        const int lw = b->lane_width;
        hipLaunchKernelGGL(lane_init_kernel, dim3((n + lw - 1) / lw), dim3(lw), 0, b->stream,
                           b->d_walkers, b->d_models, b->d_pool, 0, n, 1);
        HIPCHECK(hipGetLastError());
    } else {
        // K1's warm-up round r skips walkers with fewer warm-up sweeps, so the
        // largest count among the pending walkers sets the number of rounds
        int rc = wave_launch_init(b->wave, b->stream, b->d_walkers, b->d_models, b->d_pool, n,
                                  b->pending_warmup);
        if (rc)
            return rc;
    }
    b->any_pending = false;
    b->pending_warmup = 0;
    return SSE_OK;
}

int sse_init(sse_batch *b)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    int rc = launch_init_pending(b);
    if (!rc)
        rc = check_walker_errors(b);
    if (rc)
        return rc;
    b->inited = true;
    return SSE_OK;
}

static int launch_pt_exchange(sse_batch *b);

// The sweeps of one call between exchanges: grouped streams, per-sweep timing
// events (event slots from k0), graph replays or plain launches.
static int issue_sweeps(sse_batch *b, int n_sweeps, int measure, bool grouped, bool timed, int k0)
{
    if (b->mode == SSE_MODE_WAVE && b->persistent) {
        if (!b->kernel_timing)
            return persist_issue(b, n_sweeps, measure, nullptr, b->n_walkers,
                                 b->use_order ? b->d_order : nullptr);
        const size_t ev0 = b->kern_ev.size() * 4;
        int rc = ensure_kernel_events(b, ev0 + 4);
        if (rc)
            return rc;
        hipEvent_t *e = &b->kevpool[ev0];
        b->kern_ev.push_back({3, ev0});
        HIPCHECK(hipEventRecord(e[0], b->stream));
        const bool graph = b->use_graph;
        b->use_graph = false;
        rc = persist_issue(b, n_sweeps, measure, nullptr, b->n_walkers,
                           b->use_order ? b->d_order : nullptr);
        b->use_graph = graph;
        HIPCHECK(hipEventRecord(e[1], b->stream));
        return rc;
    }
    if (grouped) {
        if (b->kernel_timing)
            return groups_run(b, n_sweeps, measure, true);
        return b->use_graph && b->graph_groups ? groups_graph_many(b, n_sweeps, measure)
                                               : groups_run(b, n_sweeps, measure);
    }
    if (timed) {
        for (int k = 0; k < n_sweeps; ++k) {
            hipEvent_t *e = &b->evpool[4 * (size_t)(k0 + k)];
            HIPCHECK(hipEventRecord(e[0], b->stream));
            int rc = launch_sweep_measure(b, measure, e);
            if (rc)
                return rc;
            HIPCHECK(hipEventRecord(e[3], b->stream));
        }
        return SSE_OK;
    }
    if (b->use_graph)
        return graph_run_many(b, n_sweeps, measure);
    for (int k = 0; k < n_sweeps; ++k) {
        int rc = launch_sweep_measure(b, measure);
        if (rc)
            return rc;
    }
    return SSE_OK;
}

// ---------------------------------------------------------------- autotuner
// Sum over walkers of acc_L + acc_W (the vertex updates counted by the chains).
__global__ void work_sum_kernel(const DevWalker *W, int n, unsigned long long *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        atomicAdd(out, (unsigned long long)(W[i].acc_L + W[i].acc_W));
}
__global__ void work_zero_kernel(unsigned long long *out) { *out = 0ull; }

static int work_sum(sse_batch *b, unsigned long long *v)
{
    hipLaunchKernelGGL(work_zero_kernel, dim3(1), dim3(1), 0, b->stream, b->d_work);
    hipLaunchKernelGGL(work_sum_kernel, dim3((b->n_walkers + 255) / 256), dim3(256), 0, b->stream,
                       b->d_walkers, b->n_walkers, b->d_work);
    HIPCHECK(hipGetLastError());
    HIPCHECK(bmemcpy(b, v, b->d_work, sizeof *v, hipMemcpyDeviceToHost));
    return SSE_OK;
}

static void tune_apply(sse_batch *b, const sse_batch::TuneCfg &c)
{
    b->persistent = c.persistent != 0 && !b->wave.generic;
    b->n_groups = c.groups;
    b->wave.chase = c.chase;
    b->wave.lds_k2_fixed = c.k2_lds;
    wave_resize_lds(b->wave, b->seen_max_n, b->seen_max_L);
    for (size_t &g : b->wave.lds_k2_group)
        g = 0;
}

// Candidates: the stream-group pipeline at 2, 4 and (queue count) groups
// (not for budgeted calls, which always use the work-queue launch), and the
// persistent work-queue launch with the worm-kernel LDS sized from the
// largest walker or capped at 32 KiB (more workgroups per CU; larger walkers
// then take the global record layouts).  The record-lookahead chase is tried
// on the winner afterwards.
static void tune_init(sse_batch *b, sse_batch::Tuner &t, bool budgets)
{
    t.cands.clear();
    const int n = b->n_walkers;
    if (!budgets) {
        std::vector<int> gs = {2, 4, hw_queues()};
        if (hw_queues() > 8)
            gs.push_back(8);
        std::sort(gs.begin(), gs.end());
        gs.erase(std::unique(gs.begin(), gs.end()), gs.end());
        for (int g : gs)
            if (g <= std::max(1, n / 32))
                t.cands.push_back({0, g, 0, 0});
        if (t.cands.empty())
            t.cands.push_back({0, 1, 0, 0});
    }
    if (!b->wave.generic) {  // the generic kernels have no work-queue launch
        t.cands.push_back({1, 0, 0, 0});
        t.cands.push_back({1, 0, 0, 32768});
    }
    if (t.cands.empty())
        t.cands.push_back({0, 1, 0, 0});
    t.twin = !b->wave.generic;  // the generic worm kernel has one chase form
    t.rate.assign(t.cands.size(), 0.0);
    t.runs.assign(t.cands.size(), 0);
    t.best = -1;
}

// Next candidate to measure (each twice, round-robin), or -1 when done.
static int tune_next(const sse_batch::Tuner &t)
{
    for (int round = 0; round < 2; ++round)
        for (size_t i = 0; i < t.cands.size(); ++i)
            if (t.runs[i] == round)
                return (int)i;
    return -1;
}

// Record one measurement; when every candidate ran twice, keep the fastest
// (after adding its lookahead-chase twin once).
static void tune_record(sse_batch::Tuner &t, int i, double updates, double seconds)
{
    const double rate = seconds > 0 ? updates / seconds : 0.0;
    t.rate[i] = std::max(t.rate[i], rate);
    t.runs[i] += 1;
    if (tune_next(t) >= 0)
        return;
    int best = 0;
    for (size_t k = 1; k < t.cands.size(); ++k)
        if (t.rate[k] > t.rate[best])
            best = (int)k;
    sse_batch::TuneCfg tw = t.cands[best];
    if (tw.chase == 0 && t.twin) {
        tw.chase = 1;
        bool have = false;
        for (auto &c : t.cands)
            have |= c.persistent == tw.persistent && c.groups == tw.groups && c.chase == 1 &&
                    c.k2_lds == tw.k2_lds;
        if (!have) {
            t.cands.push_back(tw);
            t.rate.push_back(0.0);
            t.runs.push_back(1);  // one measurement suffices for the twin
            return;
        }
    }
    t.best = best;
}

static int issue_sweeps(sse_batch *b, int n_sweeps, int measure, bool grouped, bool timed, int k0);
static int sweep_call_rest(sse_batch *b, int32_t n_sweeps, int32_t measure, int tuned_sweeps);

// Tuning inside an sse_run call: candidate segments of four sweeps each (fewer
// when fewer are left; direct launches: no graph capture is timed), the
// worm-kernel LDS resized from the operator counts between segments, until a
// choice is made or the call's sweeps run out.  Returns the sweeps it ran, or
// -(error code).
static int tune_in_call(sse_batch *b, int n_sweeps, int measure)
{
    sse_batch::Tuner &t = b->tune_run;
    if (t.cands.empty())
        tune_init(b, t, false);
    const bool graph = b->use_graph;
    b->use_graph = false;
    int used = 0;
    int rc = SSE_OK;
    while (used < n_sweeps && t.best < 0) {
        const int i = tune_next(t);
        tune_apply(b, t.cands[i]);
        const bool grouped = !b->persistent && group_count(b) > 1;
        unsigned long long w0 = 0, w1 = 0;
        rc = work_sum(b, &w0);
        if (rc)
            break;
        // four sweeps per segment: a stream-group pipeline needs a few sweeps
        // past its fill before a segment shows its steady-state rate
        const int k = std::min(4, n_sweeps - used);
        const auto t0 = std::chrono::steady_clock::now();
        rc = issue_sweeps(b, k, measure, grouped, false, 0);
        if (!rc && hipStreamSynchronize(b->stream) != hipSuccess)
            rc = fail(SSE_ERR_HIP, "tuning: stream synchronize failed");
        const auto t1 = std::chrono::steady_clock::now();
        if (!rc)
            rc = work_sum(b, &w1);
        if (!rc)
            rc = check_walker_errors(b);  // also resizes the worm-kernel LDS
        if (rc)
            break;
        used += k;
        tune_record(t, i, (double)(w1 - w0), std::chrono::duration<double>(t1 - t0).count());
    }
    b->use_graph = graph;
    return rc ? -rc : used;
}

// Body of sse_sweep (measure 0) and sse_run: n_sweeps sweeps of every active
// walker, each followed by the measurement of the thermalized ones when
// measure is set, and an exchange round after every pt_interval-th sweep
// (counted across calls) when temperature ladders are set.
static int sweep_call(sse_batch *b, int32_t n_sweeps, int32_t measure)
{
    b->kern_ev.clear();
    if (b->any_pending) {
        int rc = launch_init_pending(b);
        if (rc)
            return rc;
    }
    // the autotuner measures on the first measuring calls of a wave-mode
    // batch (thermalization from random states is no steady state), not
    // between the exchanges of temperature ladders
    if (b->mode == SSE_MODE_WAVE && b->tune && !(b->pt_ladders > 0 && b->pt_interval > 0)) {
        if (b->tune_run.best >= 0) {
            tune_apply(b, b->tune_run.cands[b->tune_run.best]);
        } else if (measure && n_sweeps >= 1) {
            int rc0 = call_timer(b, 0, 0);
            if (rc0)
                return rc0;
            const int used = tune_in_call(b, n_sweeps, 1);
            if (used < 0)
                return -used;
            if (b->tune_run.best >= 0)
                tune_apply(b, b->tune_run.cands[b->tune_run.best]);
            return sweep_call_rest(b, n_sweeps - used, measure, used);
        }
    }
    return sweep_call_rest(b, n_sweeps, measure, 0);
}

// sweep_call after tuning; `tuned_sweeps` were already run (and timed from
// the call's start event when > 0)
static int sweep_call_rest(sse_batch *b, int32_t n_sweeps, int32_t measure, int tuned_sweeps)
{
    const bool grouped = b->mode == SSE_MODE_WAVE && !b->persistent && group_count(b) > 1;
    const bool timed = !grouped && !(b->mode == SSE_MODE_WAVE && b->persistent) && b->use_timing &&
                       n_sweeps <= 256;
    const bool pt = b->pt_ladders > 0 && b->pt_interval > 0;
    if (!tuned_sweeps) {
        int rc0 = call_timer(b, 0, 0);
        if (rc0)
            return rc0;
    }
    if (timed) {
        int rc = ensure_events(b, 4 * (size_t)n_sweeps);
        if (rc)
            return rc;
    }
    for (int done = 0; done < n_sweeps;) {
        int seg = n_sweeps - done;
        if (pt)
            seg = std::min(seg, b->pt_interval - b->pt_since);
        int rc = issue_sweeps(b, seg, measure ? 1 : 0, grouped, timed, done);
        if (rc)
            return rc;
        done += seg;
        if (pt) {
            b->pt_since += seg;
            if (b->pt_since >= b->pt_interval) {
                b->pt_since = 0;
                rc = launch_pt_exchange(b);
                if (rc)
                    return rc;
            }
        }
    }
    if (!measure && b->any_budget) {
        hipLaunchKernelGGL(retire_kernel, dim3((b->n_walkers + 63) / 64), dim3(64), 0, b->stream,
                           b->d_walkers, b->n_walkers, b->d_done);
        HIPCHECK(hipGetLastError());
    }
    int rc1 = call_timer(b, 1, n_sweeps + tuned_sweeps);
    if (rc1)
        return rc1;
    int rc = check_walker_errors(b);
    if (b->kernel_timing)
        collect_kernel_times(b);
    if (!rc)
        rc = call_timer(b, 2, 0);
    if (timed && !tuned_sweeps)
        collect_phase_times(b, n_sweeps, measure != 0);
    else
        clear_phase_times(b);
    return rc;
}

int sse_sweep(sse_batch *b, int32_t n_sweeps)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (!b->inited)
        return fail(SSE_ERR_STATE, "sse_sweep before sse_init");
    if (n_sweeps < 0)
        return fail(SSE_ERR_INVALID, "negative sweep count");
    return sweep_call(b, n_sweeps, 0);
}

int sse_sweep_phases(sse_batch *b, int32_t phases)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (!b->inited)
        return fail(SSE_ERR_STATE, "sweep before init");
    if (phases <= 0 || phases > 15 || (phases & 7) == 0)
        return fail(SSE_ERR_INVALID, "phases must be a mask of 1|2|4 (+8)");
    int rc = launch_sweep(b, phases);
    if (rc)
        return rc;
    return check_walker_errors(b);
}

static int launch_measure(sse_batch *b, int mode)
{
    int n = b->n_walkers;
    if (b->mode == SSE_MODE_LANE) {
        const int lw = b->lane_width;
        hipLaunchKernelGGL(lane_measure_kernel, dim3((n + lw - 1) / lw), dim3(lw), 0, b->stream,
                           b->d_walkers, b->d_models, b->d_out, b->d_sum, b->d_sumsq, b->d_count,
                           b->d_wlf, b->d_scratch, b->max_sites, n, mode, b->d_done,
                           (const int32_t *)nullptr, 0);
    } else {
        int rc = wave_launch_measure(b->wave, b->stream, b->d_walkers, b->d_models, b->d_out,
                                     b->d_sum, b->d_sumsq, b->d_count, b->d_wlf, 0, n, mode,
                                     b->d_done, b->use_order ? b->d_order : nullptr);
        if (rc)
            return rc;
    }
    HIPCHECK(hipGetLastError());
    return SSE_OK;
}

// One sweep and, with measure, the accumulated measurement of every walker on
// the batch stream: wave mode with fuse_measure runs K1 and the fused worm +
// measurement kernel; otherwise the sweep kernels, then the measurement
// kernel.  ev (optional) as in launch_sweep.
static int launch_sweep_measure(sse_batch *b, int measure, hipEvent_t *ev)
{
    if (measure && b->mode == SSE_MODE_WAVE && b->wave.fuse_measure) {
        int rc = wave_launch_sweep_measure(b->wave, b->stream, b->d_walkers, b->d_models, b->d_pool, 0,
                                           b->n_walkers, ev, b->use_order ? b->d_order : nullptr, -1,
                                           b->d_out, b->d_sum, b->d_sumsq, b->d_count, b->d_wlf,
                                           b->d_done);
        HIPCHECK(hipGetLastError());
        return rc;
    }
    int rc = launch_sweep(b, 7, ev);
    if (!rc && measure)
        rc = launch_measure(b, 1);
    return rc;
}

int32_t sse_measure_width(const sse_batch *b, int32_t walker)
{
    if (!b || walker < 0 || walker >= b->n_walkers)
        return -1;
    return b->widths[walker];
}

int sse_measure(sse_batch *b, double *out)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (!b->inited)
        return fail(SSE_ERR_STATE, "measure before init");
    // slots restarted since the last sweep are initialised first (Carlo.init!
    // precedes any measure!), so every measured walker has a valid state
    if (b->any_pending) {
        int rc = launch_init_pending(b);
        if (rc)
            return rc;
    }
    int rc = launch_measure(b, 0);
    if (rc)
        return rc;
    if (out) {
        HIPCHECK(hipMemcpyAsync(out, b->d_out, sizeof(double) * b->total_width,
                                hipMemcpyDeviceToHost, b->stream));
    }
    return check_walker_errors(b);
}

int sse_run(sse_batch *b, int32_t n_sweeps, int32_t measure)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (!b->inited)
        return fail(SSE_ERR_STATE, "run before init");
    if (n_sweeps < 0)
        return fail(SSE_ERR_INVALID, "negative sweep count");
    return sweep_call(b, n_sweeps, measure);
}

// sse_run_budgets in lane mode, and in wave mode with bonds of other than two
// sites (no work-queue launch there): a sequence of sweep (+ measure)
// launches; launch i runs the walkers whose budget exceeds i.
static int lane_run_budgets(sse_batch *b, const int64_t *sweeps, int32_t measure)
{
    const int n = b->n_walkers;
    std::vector<int32_t> bud(n);
    int64_t most = 0;
    for (int w = 0; w < n; ++w) {
        if (sweeps[w] < 0 || sweeps[w] > INT32_MAX)
            return fail(SSE_ERR_INVALID, "sse_run_budgets: sweep counts must be in 0..2^31-1");
        bud[w] = (int32_t)sweeps[w];
        most = std::max(most, sweeps[w]);
    }
    if (b->any_pending) {
        int rc = launch_init_pending(b);
        if (rc)
            return rc;
    }
    int rc = call_timer(b, 0, 0);
    if (rc)
        return rc;
    HIPCHECK(bmemcpy(b, b->d_budget, bud.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    const int lw = b->lane_width;
    const int blocks = (n + lw - 1) / lw;
    const int32_t *bp = (const int32_t *)b->d_budget;
    const int32_t *ord = b->use_order ? b->d_order : nullptr;
    for (int64_t i = 0; i < most; ++i) {
        if (b->mode == SSE_MODE_LANE) {
            hipLaunchKernelGGL(lane_sweep_kernel, dim3(blocks), dim3(lw), 0, b->stream, b->d_walkers,
                               b->d_models, b->d_pool, 0, n, 7, bp, (int)i);
            if (measure)
                hipLaunchKernelGGL(lane_measure_kernel, dim3(blocks), dim3(lw), 0, b->stream,
                                   b->d_walkers, b->d_models, b->d_out, b->d_sum, b->d_sumsq,
                                   b->d_count, b->d_wlf, b->d_scratch, b->max_sites, n, 1,
                                   b->d_done, bp, (int)i);
        } else {  // the generic wave kernels, one walker per workgroup
            hipLaunchKernelGGL(waven_diag_vlist_kernel, dim3(n), dim3(64), b->wave.lds_k1n, b->stream,
                               b->d_walkers, b->d_models, b->d_pool, 0, n, 3, -1, ord, 7, bp, (int)i,
                               (int)b->wave.lds_k1n);
            hipLaunchKernelGGL(waven_worm_kernel, dim3(n), dim3(64), b->wave.lds_k2n, b->stream,
                               b->d_walkers, b->d_models, b->d_pool, 0, n, 1, ord, bp, (int)i,
                               (int)b->wave.lds_k2n);
            if (measure)
                hipLaunchKernelGGL(waven_measure_kernel, dim3(n), dim3(64), 0, b->stream,
                                   b->d_walkers, b->d_models, b->d_out, b->d_sum, b->d_sumsq,
                                   b->d_count, b->d_wlf, 0, n, 1, b->d_done, ord, bp, (int)i);
        }
        HIPCHECK(hipGetLastError());
    }
    if (!measure && b->any_budget) {
        hipLaunchKernelGGL(retire_kernel, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_walkers, n,
                           b->d_done);
        HIPCHECK(hipGetLastError());
    }
    rc = call_timer(b, 1, most);
    if (rc)
        return rc;
    rc = check_walker_errors(b);
    if (!rc)
        rc = call_timer(b, 2, 0);
    clear_phase_times(b);
    return rc;
}

int sse_run_budgets(sse_batch *b, const int64_t *sweeps, int32_t measure)
{
    if (!b || !sweeps)
        return fail(SSE_ERR_INVALID, "sse_run_budgets: bad arguments");
    if (!b->inited)
        return fail(SSE_ERR_STATE, "run before init");
    if (b->pt_ladders > 0 && b->pt_interval > 0)
        return fail(SSE_ERR_INVALID, "sse_run_budgets: walkers of a temperature ladder advance "
                                     "together (sse_run)");
    if (b->mode == SSE_MODE_LANE || b->wave.generic)
        return lane_run_budgets(b, sweeps, measure);
    const int n = b->n_walkers;
    std::vector<int32_t> bud(n);
    int64_t most = 0;
    for (int w = 0; w < n; ++w) {
        if (sweeps[w] < 0 || sweeps[w] > INT32_MAX)
            return fail(SSE_ERR_INVALID, "sse_run_budgets: sweep counts must be in 0..2^31-1");
        bud[w] = (int32_t)sweeps[w];
        most = std::max(most, sweeps[w]);
    }
    // tasks: walkers with a budget, largest expected work (budget x bonds / T) first
    std::vector<int32_t> order;
    std::vector<double> work(n);
    for (int w = 0; w < n; ++w)
        if (bud[w] > 0) {
            order.push_back(w);
            work[w] = (double)bud[w] * b->work_key[w];
        }
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) { return work[a] > work[c]; });
    if (b->any_pending) {
        int rc = launch_init_pending(b);
        if (rc)
            return rc;
    }
    // autotuner: one candidate per call (each twice), then the fastest
    sse_batch::Tuner &t = b->tune_bud;
    int cand = -1;
    const bool graph = b->use_graph;
    const bool persistent = b->persistent;
    if (b->tune && measure) {
        if (t.cands.empty())
            tune_init(b, t, true);
        if (t.best >= 0) {
            tune_apply(b, t.cands[t.best]);
        } else {
            cand = tune_next(t);
            tune_apply(b, t.cands[cand]);
            b->use_graph = false;
        }
    }
    unsigned long long w0 = 0, w1 = 0;
    if (cand >= 0) {
        int rw = work_sum(b, &w0);
        if (rw)
            return rw;
    }
    int rc = call_timer(b, 0, 0);
    if (rc)
        return rc;
    const auto t0 = std::chrono::steady_clock::now();
    b->kern_ev.clear();
    if (!order.empty()) {
        HIPCHECK(bmemcpy(b, b->d_budget, bud.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        HIPCHECK(bmemcpy(b, b->d_task_order, order.data(), sizeof(int32_t) * order.size(),
                         hipMemcpyHostToDevice));
        // SSE_OPT_TIMING 2: events around the persistent launch (a direct
        // launch, as for sse_run), summed by sse_kernel_times
        hipEvent_t *e = nullptr;
        const bool graph_k = b->use_graph;
        if (b->kernel_timing) {
            const size_t ev0 = b->kern_ev.size() * 4;
            rc = ensure_kernel_events(b, ev0 + 4);
            if (rc)
                return rc;
            e = &b->kevpool[ev0];
            b->kern_ev.push_back({3, ev0});
            b->use_graph = false;
            HIPCHECK(hipEventRecord(e[0], b->stream));
        }
        rc = persist_issue(b, 0, measure ? 1 : 0, b->d_budget, (int)order.size(), b->d_task_order);
        b->use_graph = graph_k;
        if (rc)
            return rc;
        if (e)
            HIPCHECK(hipEventRecord(e[1], b->stream));
    }
    if (cand >= 0) {
        HIPCHECK(hipStreamSynchronize(b->stream));
        const auto t1 = std::chrono::steady_clock::now();
        int rw = work_sum(b, &w1);
        if (rw)
            return rw;
        tune_record(t, cand, (double)(w1 - w0), std::chrono::duration<double>(t1 - t0).count());
        b->use_graph = graph;
    }
    if (b->tune && measure)  // sse_run keeps its own launch form
        b->persistent = persistent;
    if (!measure && b->any_budget) {
        hipLaunchKernelGGL(retire_kernel, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_walkers,
                           n, b->d_done);
        HIPCHECK(hipGetLastError());
    }
    rc = call_timer(b, 1, most);
    if (rc)
        return rc;
    rc = check_walker_errors(b);
    if (b->kernel_timing)
        collect_kernel_times(b);
    if (!rc)
        rc = call_timer(b, 2, 0);
    clear_phase_times(b);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------- no-alloc
// Device scratch of sse_run_budgets_noalloc, one copy per GPU for the life of
// the process: [0, 256) the task counter, then the caller's budgets as int64
// [SSE_NOALLOC_MAX_WALKERS], then the same narrowed to int32.  Host code uses
// only the per-device address from hipGetSymbolAddress, never the name.
constexpr size_t kNoAllocBudget64 = 256;
constexpr size_t kNoAllocBudget32 = kNoAllocBudget64 + 8 * (size_t)SSE_NOALLOC_MAX_WALKERS;
constexpr size_t kNoAllocBytes = kNoAllocBudget32 + 4 * (size_t)SSE_NOALLOC_MAX_WALKERS;
static __device__ __align__(256) char scratch[kNoAllocBytes];

__global__ void budget_narrow_kernel(const int64_t *in, int32_t *out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = (int32_t)in[i];
}

// scratch address per device, taken once per device with that device current
static char *noalloc_scratch(int device)
{
    static std::mutex mu;
    static char *ptr[64] = {};
    if (device < 0 || device >= 64)
        return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess)
            return nullptr;
        if (cur != device && hipSetDevice(device) != hipSuccess)
            return nullptr;
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device)
            (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return nullptr;
        ptr[device] = (char *)p;
    }
    return ptr[device];
}

extern "C" {

int sse_run_budgets_noalloc(sse_batch *b, const int64_t *sweeps, int32_t measure)
{
    if (!b || !sweeps)
        return fail(SSE_ERR_INVALID, "sse_run_budgets_noalloc: bad arguments");
    if (!b->inited)
        return fail(SSE_ERR_STATE, "run before init");
    if (b->mode != SSE_MODE_WAVE || b->wave.generic)
        return fail(SSE_ERR_INVALID, "sse_run_budgets_noalloc: wave mode with two-site bonds only");
    if (b->n_walkers > SSE_NOALLOC_MAX_WALKERS)
        return fail(SSE_ERR_INVALID, "sse_run_budgets_noalloc: the batch has more walkers than the "
                                     "static scratch holds (SSE_NOALLOC_MAX_WALKERS)");
    if (b->pt_ladders > 0 && b->pt_interval > 0)
        return fail(SSE_ERR_INVALID, "sse_run_budgets_noalloc: walkers of a temperature ladder "
                                     "advance together (sse_run)");
    const int n = b->n_walkers;
    for (int w = 0; w < n; ++w)
        if (sweeps[w] < 0 || sweeps[w] > INT32_MAX)
            return fail(SSE_ERR_INVALID, "sse_run_budgets_noalloc: sweep counts must be in 0..2^31-1");
    char *sc = noalloc_scratch(b->device);
    if (!sc)
        return fail(SSE_ERR_HIP, "sse_run_budgets_noalloc: no address for the device scratch");
    if (b->any_pending) {
        int rc = launch_init_pending(b);
        if (rc)
            return rc;
    }
    int64_t *bud64 = (int64_t *)(sc + kNoAllocBudget64);
    int32_t *bud32 = (int32_t *)(sc + kNoAllocBudget32);
    HIPCHECK(hipMemcpyAsync(bud64, sweeps, sizeof(int64_t) * n, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(budget_narrow_kernel, dim3((n + 255) / 256), dim3(256), 0, b->stream, bud64,
                       bud32, n);
    HIPCHECK(hipGetLastError());
    // tasks in the batch's schedule order (largest expected sweep first);
    // walkers with budget 0 end their task at once.  The counter is cleared
    // by the launch's reset kernel.
    int rc = wave_launch_persist(b->wave, b->stream, b->d_walkers, b->d_models, b->d_pool, b->d_out,
                                 b->d_sum, b->d_sumsq, b->d_count, b->d_wlf, b->d_done,
                                 b->use_order ? b->d_order : nullptr, n, bud32, 0, measure ? 1 : 0,
                                 (int32_t *)sc, b->n_cu);
    if (rc)
        return fail(rc, "persistent launch failed");
    if (!measure && b->any_budget) {
        hipLaunchKernelGGL(retire_kernel, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_walkers,
                           n, b->d_done);
        HIPCHECK(hipGetLastError());
    }
    return check_walker_errors(b);
}

int sse_read_accumulators(sse_batch *b, double *sums, double *sumsq, int64_t *counts)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    HIPCHECK(hipStreamSynchronize(b->stream));
    if (sums)
        HIPCHECK(bmemcpy(b, sums, b->d_sum, sizeof(double) * b->total_width, hipMemcpyDeviceToHost));
    if (sumsq)
        HIPCHECK(bmemcpy(b, sumsq, b->d_sumsq, sizeof(double) * b->total_width,
                           hipMemcpyDeviceToHost));
    if (counts)
        HIPCHECK(bmemcpy(b, counts, b->d_count, sizeof(int64_t) * b->n_walkers,
                           hipMemcpyDeviceToHost));
    return SSE_OK;
}

int sse_read_worm_fraction(sse_batch *b, double *sums, double *sumsq, int64_t *counts)
{
    if (!b || !sums || !sumsq || !counts)
        return fail(SSE_ERR_INVALID, "sse_read_worm_fraction: bad arguments");
    HIPCHECK(hipStreamSynchronize(b->stream));
    std::vector<double> h(3 * (size_t)b->n_walkers);
    HIPCHECK(bmemcpy(b, h.data(), b->d_wlf, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int w = 0; w < b->n_walkers; ++w) {
        sums[w] = h[3 * (size_t)w];
        sumsq[w] = h[3 * (size_t)w + 1];
        counts[w] = (int64_t)h[3 * (size_t)w + 2];
    }
    return SSE_OK;
}

int sse_reset_accumulators(sse_batch *b)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    HIPCHECK(hipMemsetAsync(b->d_sum, 0, sizeof(double) * b->total_width, b->stream));
    HIPCHECK(hipMemsetAsync(b->d_sumsq, 0, sizeof(double) * b->total_width, b->stream));
    HIPCHECK(hipMemsetAsync(b->d_count, 0, sizeof(int64_t) * b->n_walkers, b->stream));
    HIPCHECK(hipMemsetAsync(b->d_wlf, 0, sizeof(double) * 3 * b->n_walkers, b->stream));
    hipLaunchKernelGGL(reset_work_counters_kernel, dim3((b->n_walkers + 63) / 64), dim3(64), 0,
                       b->stream, b->d_walkers, b->n_walkers);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(b->stream));
    return SSE_OK;
}

int sse_get_walker(sse_batch *b, int32_t walker, sse_walker_scalars *s, uint32_t *ops,
                   int64_t ops_capacity, uint8_t *spins)
{
    if (!b || walker < 0 || walker >= b->n_walkers)
        return fail(SSE_ERR_INVALID, "sse_get_walker: bad walker");
    HIPCHECK(hipStreamSynchronize(b->stream));
    DevWalker w;
    HIPCHECK(bmemcpy(b, &w, b->d_walkers + walker, sizeof w, hipMemcpyDeviceToHost));
    if (s) {
        s->cutoff = w.L;
        s->num_operators = w.n;
        s->avg_worm_length = w.avg_worm_length;
        s->num_worms = w.num_worms;
        s->rng_counter = w.ctr;
        s->sweeps = w.sweeps;
        s->last_scatter_calls = w.last_scatter;
        s->last_total_worm_length = w.last_total;
        s->last_worm_length_fraction = w.last_wlf;
    }
    if (ops && ops_capacity > 0 && w.L > 0) {
        int64_t k = w.L < ops_capacity ? w.L : ops_capacity;
        HIPCHECK(bmemcpy(b, ops, w.ops, sizeof(uint32_t) * k, hipMemcpyDeviceToHost));
    }
    if (spins)
        HIPCHECK(bmemcpy(b, spins, w.spins, w.n_sites, hipMemcpyDeviceToHost));
    return SSE_OK;
}

int sse_set_walker(sse_batch *b, int32_t walker, const sse_walker_scalars *s, const uint32_t *ops,
                   const uint8_t *spins)
{
    if (!b || walker < 0 || walker >= b->n_walkers || !s || !ops || !spins || s->cutoff < 0)
        return fail(SSE_ERR_INVALID, "sse_set_walker: bad arguments");
    HIPCHECK(hipStreamSynchronize(b->stream));
    DevWalker w;
    HIPCHECK(bmemcpy(b, &w, b->d_walkers + walker, sizeof w, hipMemcpyDeviceToHost));
    const DevModel &m = b->h_models[w.model];
    const HostModelCheck &hc = b->h_check[w.model];
    int64_t n = 0;
    for (int64_t p = 0; p < s->cutoff; ++p) {
        uint32_t op = ops[p];
        if (op) {
            ++n;
            const uint32_t bond = op >> SSE_OP_BOND_SHIFT;
            if (bond >= (uint32_t)m.n_bonds)
                return fail(SSE_ERR_INVALID, "sse_set_walker: operator bond out of range");
            if (((op >> 2) & 511u) >= (uint32_t)hc.bond_nv[bond])
                return fail(SSE_ERR_INVALID, "sse_set_walker: vertex index out of range for the "
                                             "operator's bond type");
        }
    }
    if (n != s->num_operators)
        return fail(SSE_ERR_INVALID, "sse_set_walker: num_operators does not match ops");
    for (int32_t i = 0; i < w.n_sites; ++i)
        if (spins[i] < 1 || spins[i] > hc.site_dims[i])
            return fail(SSE_ERR_INVALID, "sse_set_walker: local state outside 1..dim");
    int rc = sync_pool_from_device(b);
    if (rc)
        return rc;
    if (s->cutoff > w.cap) {
        int64_t cap = slab_capacity(s->cutoff);
        char *slab = host_slab(b, cap, 2 * m.nsb);
        if (!slab)
            return fail(SSE_ERR_OOM, "slab pool exhausted");
        w.ops = (uint32_t *)slab;
        w.links = (int32_t *)(slab + (size_t)cap * 4);
        w.cap = cap;
        rc = sync_pool_to_device(b);
        if (rc)
            return rc;
    }
    w.L = s->cutoff;
    w.n = n;
    w.avg_worm_length = s->avg_worm_length;
    w.num_worms = s->num_worms;
    w.ctr = s->rng_counter;
    w.sweeps = s->sweeps;
    w.error = 0;
    w.active = 1;
    w.pending = 0;
    w.retire = 0;
    if (w.L > 0)
        HIPCHECK(bmemcpy(b, w.ops, ops, sizeof(uint32_t) * w.L, hipMemcpyHostToDevice));
    HIPCHECK(bmemcpy(b, w.spins, spins, w.n_sites, hipMemcpyHostToDevice));
    HIPCHECK(bmemcpy(b, b->d_walkers + walker, &w, sizeof w, hipMemcpyHostToDevice));
    b->inited = true;
    return SSE_OK;
}

int sse_get_vertex_list(sse_batch *b, int32_t walker, int32_t *links, int64_t links_capacity,
                        int32_t *first, int32_t *last)
{
    if (!b || walker < 0 || walker >= b->n_walkers)
        return fail(SSE_ERR_INVALID, "sse_get_vertex_list: bad walker");
    HIPCHECK(hipStreamSynchronize(b->stream));
    DevWalker w;
    HIPCHECK(bmemcpy(b, &w, b->d_walkers + walker, sizeof w, hipMemcpyDeviceToHost));
    if (b->mode == SSE_MODE_LANE) {
        const int64_t legs = 2 * b->h_models[w.model].nsb;
        if (links && links_capacity >= legs * w.L && w.L > 0)
            HIPCHECK(bmemcpy(b, links, w.links, sizeof(int32_t) * legs * w.L, hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> pos;  // wave path: operator rank -> slot
    if (b->mode != SSE_MODE_LANE) {
        // compact records (rank-indexed, leg ids PL*rank + leg with PL the leg
        // count rounded up to a power of two) -> the reference's dense layout
        const int legs = 2 * b->h_models[w.model].nsb;
        const int lg = legs <= 2 ? 1 : legs <= 4 ? 2 : 3, pl = 1 << lg;
        int64_t nch = (w.L + 63) / 64;
        const uint32_t *vt = (const uint32_t *)w.links;
        std::vector<uint64_t> mask(nch > 0 ? nch : 1);
        std::vector<uint32_t> lk((size_t)pl * (w.n > 0 ? w.n : 1));
        if (nch > 0)
            HIPCHECK(bmemcpy(b, mask.data(), (const uint64_t *)(vt + (1 + pl) * w.cap), 8 * nch,
                               hipMemcpyDeviceToHost));
        if (w.n > 0)
            HIPCHECK(bmemcpy(b, lk.data(), vt + w.cap, 4 * (size_t)pl * w.n, hipMemcpyDeviceToHost));
        for (int64_t c = 0; c < nch; ++c)
            for (int i = 0; i < 64; ++i)
                if ((mask[c] >> i) & 1ull)
                    pos.push_back(64 * c + i);
        if ((int64_t)pos.size() != w.n)
            return fail(SSE_ERR_STATE, "vertex list not built for the current operator string");
        auto dense = [&](int32_t leg) -> int32_t {
            return leg < 0 ? leg : (int32_t)(legs * pos[leg >> lg] + (leg & (pl - 1)));
        };
        if (links && links_capacity >= legs * w.L) {
            for (int64_t k = 0; k < legs * w.L; ++k)
                links[k] = -1;
            for (int64_t r = 0; r < w.n; ++r)
                for (int k = 0; k < legs; ++k)
                    links[legs * pos[r] + k] = dense((int32_t)lk[(size_t)pl * r + k]);
        }
        if (first || last) {
            std::vector<int32_t> f(w.n_sites), l(w.n_sites);
            HIPCHECK(bmemcpy(b, f.data(), w.first, 4 * w.n_sites, hipMemcpyDeviceToHost));
            HIPCHECK(bmemcpy(b, l.data(), w.last, 4 * w.n_sites, hipMemcpyDeviceToHost));
            for (int i = 0; i < w.n_sites; ++i) {
                if (first)
                    first[i] = dense(f[i] < 0 ? -1 : f[i]);
                if (last)
                    last[i] = dense(l[i] < 0 ? -1 : l[i]);
            }
        }
        return SSE_OK;
    }
    if (first)
        HIPCHECK(bmemcpy(b, first, w.first, sizeof(int32_t) * w.n_sites, hipMemcpyDeviceToHost));
    if (last)
        HIPCHECK(bmemcpy(b, last, w.last, sizeof(int32_t) * w.n_sites, hipMemcpyDeviceToHost));
    return SSE_OK;
}

// ---------------------------------------------------------------- serving

// Re-pack every live slab into a fresh pool (the bump allocator never frees).
// Slots that finished their budget keep no slab (their results live in the
// accumulators).  Only called while no kernel runs.
static int compact_pool(sse_batch *b)
{
    HIPCHECK(hipStreamSynchronize(b->stream));
    std::vector<DevWalker> h(b->n_walkers);
    HIPCHECK(bmemcpy(b, h.data(), b->d_walkers, sizeof(DevWalker) * b->n_walkers,
                       hipMemcpyDeviceToHost));
    char *np = nullptr;
    if (hipMalloc((void **)&np, b->pool_size) != hipSuccess)
        return fail(SSE_ERR_OOM, "pool compaction: no memory for the new pool");
    size_t next = 0;
    for (auto &w : h) {
        bool live = w.active || w.pending || w.remaining != 0;
        int64_t cap = live ? w.cap : 64;
        size_t bytes = slab_bytes(cap, 2 * b->h_models[w.model].nsb);
        if (next + bytes > b->pool_size) {
            (void)hipFree(np);
            return fail(SSE_ERR_OOM, "pool compaction: live slabs exceed the pool");
        }
        char *dst = np + next;
        next += bytes;
        if (live && w.L > 0)
            HIPCHECK(hipMemcpyAsync(dst, w.ops, sizeof(uint32_t) * w.L, hipMemcpyDeviceToDevice,
                                    b->stream));
        if (!live)
            w.L = 0;
        w.ops = (uint32_t *)dst;
        w.links = (int32_t *)(dst + (size_t)cap * 4);
        w.cap = cap;
    }
    HIPCHECK(hipStreamSynchronize(b->stream));
    for (auto it = b->allocs.begin(); it != b->allocs.end(); ++it)
        if (*it == (void *)b->pool_base) {
            b->allocs.erase(it);
            break;
        }
    (void)hipFree(b->pool_base);
    b->pool_base = np;
    b->allocs.push_back(np);
    b->pool_host_next = next;
    HIPCHECK(bmemcpy(b, b->d_walkers, h.data(), sizeof(DevWalker) * b->n_walkers,
                       hipMemcpyHostToDevice));
    return sync_pool_to_device(b);
}

int sse_add_model(sse_batch *b, const sse_model_desc *model, int32_t *index)
{
    if (!b || !model || !index)
        return fail(SSE_ERR_INVALID, "sse_add_model: bad arguments");
    int rc = validate_model(*model, b->n_models);
    if (rc)
        return rc;
    if (model->n_sites > b->max_sites)
        return fail(SSE_ERR_INVALID, "sse_add_model: more sites than the batch was created for");
    HIPCHECK(hipStreamSynchronize(b->stream));
    // upload into locals; the batch's model lists (h_models, h_check,
    // d_models, n_models) change together only after every step succeeded
    DevModel dm;
    HostModelCheck hc;
    rc = upload_model(b, *model, dm, hc);
    if (rc)
        return rc;
    std::vector<DevModel> hm = b->h_models;
    hm.push_back(dm);
    DevModel *nd = nullptr;
    rc = dev_alloc(b, &nd, hm.size());
    if (rc)
        return rc;
    if (bmemcpy(b, nd, hm.data(), sizeof(DevModel) * hm.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(SSE_ERR_HIP, "sse_add_model: model table upload failed");
    b->h_models.swap(hm);
    b->h_check.push_back(std::move(hc));
    b->d_models = nd;
    if (((int64_t)dm.n_types << dm.dt_lgp) > 256 && b->wave.small_codes) {
        b->wave.small_codes = false;  // 10-byte worm records from now on: size LDS for them
        wave_resize_lds(b->wave, b->seen_max_n, b->seen_max_L);
    }
    if (dm.nsb != 2 || b->wave.generic) {
        // bonds of other than two sites: the generic wave kernels from now on
        // (they take two-site models too), never the work-queue launch
        b->wave.generic = 1;
        {
            int max_ns = 2;
            for (const DevModel &x : b->h_models)
                max_ns = std::max(max_ns, (int)x.nsb);
            b->wave.lds_k1n = std::max(b->wave.lds_k1n, k1n_lds_for(b->h_models, b->max_sites, max_ns));
        }
        if (b->wave.lds_k1n > 65536)
            (void)hipFuncSetAttribute((const void *)waven_diag_vlist_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->wave.lds_k1n);
        b->persistent = false;
    }
    *index = b->n_models++;
    return SSE_OK;
}

int sse_slot_start(sse_batch *b, int32_t slot, const sse_walker_desc *wd, int64_t n_sweeps)
{
    if (!b || !wd || slot < 0 || slot >= b->n_walkers)
        return fail(SSE_ERR_INVALID, "sse_slot_start: bad slot");
    if (wd->model < 0 || wd->model >= b->n_models || !(wd->T > 0) || wd->init_num_worms < 1 ||
        wd->diagonal_warmup_sweeps < 0 || n_sweeps == 0)
        return fail(SSE_ERR_INVALID, "sse_slot_start: invalid walker parameters");
    const DevModel &m = b->h_models[wd->model];
    int width = 5 + 5 * m.n_est;
    if (width > b->widths[slot])
        return fail(SSE_ERR_INVALID, "sse_slot_start: model has more estimators than the slot");
    if (!b->pt_member.empty() && b->pt_member[slot])
        return fail(SSE_ERR_INVALID, "sse_slot_start: slot is in a temperature ladder");
    HIPCHECK(hipStreamSynchronize(b->stream));
    int rc = sync_pool_from_device(b);
    if (rc)
        return rc;
    DevWalker x;
    HIPCHECK(bmemcpy(b, &x, b->d_walkers + slot, sizeof x, hipMemcpyDeviceToHost));
    int64_t L0 = wd->init_opstring_cutoff >= 0 ? wd->init_opstring_cutoff
                                               : (int64_t)std::nearbyint((double)m.n_sites * wd->T);
    int64_t cap = slab_capacity(L0);
    if (cap > x.cap || !x.ops) {
        char *s = host_slab(b, cap, 2 * m.nsb);
        if (!s) {
            x.active = 0;
            x.pending = 0;
            x.remaining = 0;  // release the old slab during compaction
            HIPCHECK(bmemcpy(b, b->d_walkers + slot, &x, sizeof x, hipMemcpyHostToDevice));
            rc = compact_pool(b);
            if (rc)
                return rc;
            s = host_slab(b, cap, 2 * m.nsb);
            if (!s)
                return fail(SSE_ERR_OOM, "slab pool full; raise SSE_POOL_MB");
        }
        x.ops = (uint32_t *)s;
        x.links = (int32_t *)(s + (size_t)cap * 4);
        x.cap = cap;
        rc = sync_pool_to_device(b);
        if (rc)
            return rc;
    }
    b->work_key[slot] = (double)m.n_bonds / wd->T;
    x.model = wd->model;
    x.warmup = wd->diagonal_warmup_sweeps;
    x.T = wd->T;
    x.target_wlf = wd->target_worm_length_fraction;
    x.atten = wd->num_worms_attenuation_factor;
    x.key = wd->seed;
    x.thermalization = wd->thermalization;
    x.init_cutoff = L0;
    x.avg_worm_length = 1.0;
    x.num_worms = wd->init_num_worms;
    x.L = L0;
    x.n = 0;
    x.ctr = 0;
    x.sweeps = 0;
    x.measured = 0;
    x.last_scatter = 0;
    x.last_total = 0;
    x.last_wlf = NAN;
    x.error = 0;
    x.n_sites = m.n_sites;
    x.acc_L = x.acc_n = x.acc_W = x.acc_sweeps = 0;
    x.meas_offset = b->offsets[slot];
    x.meas_width = width;
    x.acc_slot = slot;
    x.active = 0;
    x.pending = 1;
    x.retire = 0;
    x.remaining = n_sweeps;
    HIPCHECK(bmemcpy(b, b->d_walkers + slot, &x, sizeof x, hipMemcpyHostToDevice));
    HIPCHECK(bmemset(b, b->d_sum + x.meas_offset, 0, sizeof(double) * b->widths[slot]));
    HIPCHECK(bmemset(b, b->d_sumsq + x.meas_offset, 0, sizeof(double) * b->widths[slot]));
    HIPCHECK(bmemset(b, b->d_count + slot, 0, sizeof(int64_t)));
    HIPCHECK(bmemset(b, b->d_wlf + 3 * (size_t)slot, 0, 3 * sizeof(double)));
    HIPCHECK(bmemset(b, b->d_done + slot, 0, sizeof(int32_t)));
    b->any_pending = true;
    b->pending_warmup = std::max(b->pending_warmup, (int)wd->diagonal_warmup_sweeps);
    if (n_sweeps > 0)
        b->any_budget = true;
    b->inited = true;
    return SSE_OK;
}

int sse_slot_stop(sse_batch *b, int32_t slot)
{
    if (!b || slot < 0 || slot >= b->n_walkers)
        return fail(SSE_ERR_INVALID, "sse_slot_stop: bad slot");
    HIPCHECK(hipStreamSynchronize(b->stream));
    DevWalker x;
    HIPCHECK(bmemcpy(b, &x, b->d_walkers + slot, sizeof x, hipMemcpyDeviceToHost));
    x.active = 0;
    x.pending = 0;
    x.retire = 0;
    x.remaining = 0;
    HIPCHECK(bmemcpy(b, b->d_walkers + slot, &x, sizeof x, hipMemcpyHostToDevice));
    b->inited = true;
    return SSE_OK;
}

int sse_slot_poll(sse_batch *b, int32_t *done_slots, int32_t capacity, int32_t *n_done)
{
    if (!b || !n_done)
        return fail(SSE_ERR_INVALID, "sse_slot_poll: bad arguments");
    HIPCHECK(hipStreamSynchronize(b->stream));
    std::vector<int32_t> d(b->n_walkers);
    HIPCHECK(bmemcpy(b, d.data(), b->d_done, sizeof(int32_t) * b->n_walkers, hipMemcpyDeviceToHost));
    int32_t k = 0;
    for (int i = 0; i < b->n_walkers; ++i)
        if (d[i]) {
            if (done_slots && k < capacity)
                done_slots[k] = i;
            ++k;
        }
    if (k)
        HIPCHECK(bmemset(b, b->d_done, 0, sizeof(int32_t) * b->n_walkers));
    *n_done = k;
    return SSE_OK;
}

int sse_pool_usage(sse_batch *b, int64_t *used_bytes, int64_t *total_bytes)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    HIPCHECK(hipStreamSynchronize(b->stream));
    int rc = sync_pool_from_device(b);
    if (rc)
        return rc;
    if (used_bytes)
        *used_bytes = (int64_t)b->pool_host_next;
    if (total_bytes)
        *total_bytes = (int64_t)b->pool_size;
    return SSE_OK;
}

int sse_set_option(sse_batch *b, int32_t option, int64_t value)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    switch (option) {
    case SSE_OPT_GRAPH:
        if (value < 0 || value > 2)
            return fail(SSE_ERR_INVALID, "SSE_OPT_GRAPH takes 0, 1 or 2");
        b->use_graph = value != 0;
        b->graph_groups = value == 2;
        return SSE_OK;
    case SSE_OPT_TIMING:
        if (value < 0 || value > 2)
            return fail(SSE_ERR_INVALID, "SSE_OPT_TIMING takes 0, 1 or 2");
        b->use_timing = value != 0;
        b->kernel_timing = value == 2;
        return SSE_OK;
    case SSE_OPT_PERSISTENT:
        if (value != 0 && value != 1)
            return fail(SSE_ERR_INVALID, "SSE_OPT_PERSISTENT takes 0 or 1");
        if (value && b->mode != SSE_MODE_WAVE)
            return fail(SSE_ERR_INVALID, "SSE_OPT_PERSISTENT needs wave mode");
        if (value && b->wave.generic)
            return fail(SSE_ERR_INVALID, "SSE_OPT_PERSISTENT: the work-queue launch takes two-site "
                                         "bonds only");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->persistent = value != 0;
        b->tune = false;  // the caller chose the launch form
        return SSE_OK;
    case SSE_OPT_TUNE:
        if (value != 0 && value != 1)
            return fail(SSE_ERR_INVALID, "SSE_OPT_TUNE takes 0 or 1");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->tune = value != 0;
        b->tune_run = sse_batch::Tuner();
        b->tune_bud = sse_batch::Tuner();
        return SSE_OK;
    case SSE_OPT_K2_LDS: {
        if (value != 0 && (value < 8 * (int64_t)K2Stream::kNW + 256 || value > (int64_t)b->wave.lds_limit))
            return fail(SSE_ERR_INVALID, "SSE_OPT_K2_LDS: 0 or a size the device allows");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->tune = false;
        b->wave.lds_k2_fixed = (size_t)value & ~(size_t)15;
        wave_resize_lds(b->wave, b->seen_max_n, b->seen_max_L);
        for (size_t &g : b->wave.lds_k2_group)
            g = 0;
        return SSE_OK;
    }
    case SSE_OPT_ORDER:
        if (value != 0 && value != 1)
            return fail(SSE_ERR_INVALID, "SSE_OPT_ORDER takes 0 (walker index order) or 1");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->use_order = value != 0;
        for (size_t &g : b->wave.lds_k2_group)  // the groups hold other walkers now
            g = 0;
        return SSE_OK;
    case SSE_OPT_STREAMS:
        if (value < 0 || value > 64)
            return fail(SSE_ERR_INVALID, "SSE_OPT_STREAMS takes 0 (automatic) or 1..64 walker groups");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->tune = false;
        b->n_groups = (int)value;
        for (size_t &g : b->wave.lds_k2_group)  // the groups hold other walkers now
            g = 0;
        return SSE_OK;
    case SSE_OPT_CHASE:
        if (value != 0 && value != 1)
            return fail(SSE_ERR_INVALID, "SSE_OPT_CHASE takes 0 (plain) or 1 (record lookahead)");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->tune = false;
        b->wave.chase = (int)value;
        return SSE_OK;
    case SSE_OPT_LANE_WIDTH:
        if (value < 1 || value > 64)
            return fail(SSE_ERR_INVALID, "SSE_OPT_LANE_WIDTH takes 1..64 walkers per wavefront");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->lane_width = (int)value;
        return SSE_OK;
    case SSE_OPT_FUSE_MEASURE:
        if (value != 0 && value != 1)
            return fail(SSE_ERR_INVALID, "SSE_OPT_FUSE_MEASURE takes 0 or 1");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->wave.fuse_measure = (int)value;
        return SSE_OK;
    case SSE_OPT_GENERIC: {
        if (value != 0 && value != 1)
            return fail(SSE_ERR_INVALID, "SSE_OPT_GENERIC takes 0 or 1");
        if (b->mode != SSE_MODE_WAVE)
            return fail(SSE_ERR_INVALID, "SSE_OPT_GENERIC needs wave mode");
        int max_ns = 2;
        for (const DevModel &m : b->h_models)
            max_ns = std::max(max_ns, (int)m.nsb);
        if (!value && max_ns != 2)
            return fail(SSE_ERR_INVALID, "SSE_OPT_GENERIC: bonds of other than two sites need the "
                                         "generic kernels");
        HIPCHECK(hipStreamSynchronize(b->stream));
        b->wave.generic = (int)value;
        if (value) {
            b->wave.lds_k1n = std::max(b->wave.lds_k1n, k1n_lds_for(b->h_models, b->max_sites, max_ns));
            if (b->wave.lds_k1n > 65536)
                (void)hipFuncSetAttribute((const void *)waven_diag_vlist_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)b->wave.lds_k1n);
            b->persistent = false;
        }
        b->tune_run = sse_batch::Tuner();  // the candidates depend on the kernel family
        b->tune_bud = sse_batch::Tuner();
        return SSE_OK;
    }
    case SSE_OPT_WORM_RECORDS:
        if (value < 0 || value > 3)
            return fail(SSE_ERR_INVALID, "SSE_OPT_WORM_RECORDS takes 0, 1, 2 or 3");
        b->wave.rec_mode = (int)value;
        HIPCHECK(hipStreamSynchronize(b->stream));
        wave_resize_lds(b->wave, b->seen_max_n, b->seen_max_L);  // layouts differ in size
        return SSE_OK;
    default:
        return fail(SSE_ERR_INVALID, "unknown option");
    }
}

int sse_tuning(const sse_batch *b, int32_t which, int32_t *persistent, int32_t *groups,
               int32_t *chase, int64_t *k2_lds, double *rate, int32_t *candidates)
{
    if (!b || (which != 0 && which != 1))
        return fail(SSE_ERR_INVALID, "sse_tuning: which is 0 (sse_run) or 1 (sse_run_budgets)");
    const sse_batch::Tuner &t = which ? b->tune_bud : b->tune_run;
    int measured = 0;
    for (int r : t.runs)
        measured += r > 0;
    if (candidates)
        *candidates = measured;
    if (t.best < 0)
        return fail(SSE_ERR_STATE, "sse_tuning: no choice made yet");
    const sse_batch::TuneCfg &c = t.cands[t.best];
    if (persistent)
        *persistent = c.persistent;
    if (groups)
        *groups = c.groups;
    if (chase)
        *chase = c.chase;
    if (k2_lds)
        *k2_lds = (int64_t)c.k2_lds;
    if (rate)
        *rate = t.rate[t.best];
    return SSE_OK;
}

int sse_lds_bytes(const sse_batch *b, int64_t *k1, int64_t *k2, int32_t *k2_per_cu)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (k1)
        *k1 = (int64_t)b->wave.lds_k1;
    if (k2)
        *k2 = (int64_t)b->wave.lds_k2;
    if (k2_per_cu) {
        int nb = 0;
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)wave_worm_kernel, 64,
                                                              b->wave.lds_k2));
        *k2_per_cu = nb;
    }
    return SSE_OK;
}

int sse_call_time(const sse_batch *b, double *ms, int64_t *sweeps)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (ms)
        *ms = b->call_ms;
    if (sweeps)
        *sweeps = b->call_sweeps;
    return SSE_OK;
}

int sse_kernel_times(const sse_batch *b, double *ms4, int64_t *launches4)
{
    if (!b || !ms4 || !launches4)
        return fail(SSE_ERR_INVALID, "sse_kernel_times: bad arguments");
    for (int i = 0; i < 4; ++i) {
        ms4[i] = b->kern_ms[i];
        launches4[i] = b->kern_launches[i];
    }
    return SSE_OK;
}

int sse_phase_times(const sse_batch *b, double *ms3)
{
    if (!b || !ms3)
        return fail(SSE_ERR_INVALID, "bad arguments");
    for (int i = 0; i < 3; ++i)
        ms3[i] = b->phase_ms[i];
    return SSE_OK;
}

int sse_work_counters(sse_batch *b, int64_t *L_sum, int64_t *n_sum, int64_t *W_sum,
                      int64_t *walker_sweeps)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    HIPCHECK(hipStreamSynchronize(b->stream));
    std::vector<DevWalker> h(b->n_walkers);
    HIPCHECK(bmemcpy(b, h.data(), b->d_walkers, sizeof(DevWalker) * b->n_walkers,
                       hipMemcpyDeviceToHost));
    int64_t a = 0, c = 0, d = 0, e = 0;
    for (auto &w : h) {
        a += w.acc_L;
        c += w.acc_n;
        d += w.acc_W;
        e += w.acc_sweeps;
    }
    if (L_sum)
        *L_sum = a;
    if (n_sum)
        *n_sum = c;
    if (W_sum)
        *W_sum = d;
    if (walker_sweeps)
        *walker_sweeps = e;
    return SSE_OK;
}

// ------------------------------------------------------------ parallel tempering

static int launch_pt_exchange(sse_batch *b)
{
    const int pairs = b->pt_ladders * (b->pt_len / 2);
    if (pairs <= 0)
        return SSE_OK;
    hipLaunchKernelGGL(pt_exchange_kernel, dim3((pairs + 63) / 64), dim3(64), 0, b->stream,
                       b->d_walkers, b->d_pt_slot, b->d_pt_acc, b->d_pt_prop, b->pt_ladders,
                       b->pt_len, b->pt_seed, b->pt_round);
    HIPCHECK(hipGetLastError());
    b->pt_round += 1;
    return SSE_OK;
}

static void release_alloc(sse_batch *b, void *p)
{
    if (!p)
        return;
    for (size_t i = 0; i < b->allocs.size(); ++i)
        if (b->allocs[i] == p) {
            (void)hipFree(p);
            b->allocs.erase(b->allocs.begin() + (long)i);
            return;
        }
}

int sse_pt_set_ladders(sse_batch *b, const int32_t *slot_walkers, int32_t n_ladders,
                       int32_t ladder_len, uint64_t seed, int32_t interval)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (n_ladders < 0 || interval < 0 || (n_ladders > 0 && (ladder_len < 2 || !slot_walkers)))
        return fail(SSE_ERR_INVALID, "sse_pt_set_ladders: bad ladder shape");
    HIPCHECK(hipStreamSynchronize(b->stream));
    if (n_ladders == 0) {
        b->pt_ladders = 0;
        b->pt_len = 0;
        b->pt_member.assign(b->n_walkers, 0);
        return SSE_OK;
    }
    const size_t total = (size_t)n_ladders * (size_t)ladder_len;
    if (total > (size_t)b->n_walkers)
        return fail(SSE_ERR_INVALID, "sse_pt_set_ladders: more ladder slots than walkers");
    std::vector<DevWalker> h(b->n_walkers);
    HIPCHECK(bmemcpy(b, h.data(), b->d_walkers, sizeof(DevWalker) * b->n_walkers,
                       hipMemcpyDeviceToHost));
    std::vector<uint8_t> member(b->n_walkers, 0);
    for (int32_t l = 0; l < n_ladders; ++l) {
        const int32_t *sw = slot_walkers + (size_t)l * ladder_len;
        for (int32_t k = 0; k < ladder_len; ++k) {
            const int32_t w = sw[k];
            if (w < 0 || w >= b->n_walkers)
                return fail(SSE_ERR_INVALID, "sse_pt_set_ladders: walker index out of range");
            if (member[w])
                return fail(SSE_ERR_INVALID, "sse_pt_set_ladders: walker in two ladder slots");
            member[w] = 1;
            if (h[w].model != h[sw[0]].model || h[w].meas_width != h[sw[0]].meas_width)
                return fail(SSE_ERR_INVALID,
                            "sse_pt_set_ladders: walkers of one ladder must share one model");
            if (b->wdesc[w].model != h[w].model || h[w].remaining >= 0)
                return fail(SSE_ERR_INVALID,
                            "sse_pt_set_ladders: serving slots cannot join a ladder");
        }
    }
    if (total > b->pt_cap) {
        release_alloc(b, b->d_pt_slot);
        release_alloc(b, b->d_pt_acc);
        release_alloc(b, b->d_pt_prop);
        b->d_pt_slot = nullptr;
        b->d_pt_acc = b->d_pt_prop = nullptr;
        b->pt_cap = 0;
        int rc = dev_alloc(b, &b->d_pt_slot, total);
        if (!rc)
            rc = dev_alloc(b, &b->d_pt_acc, total);
        if (!rc)
            rc = dev_alloc(b, &b->d_pt_prop, total);
        if (rc)
            return rc;
        b->pt_cap = total;
    }
    HIPCHECK(bmemcpy(b, b->d_pt_slot, slot_walkers, sizeof(int32_t) * total, hipMemcpyHostToDevice));
    HIPCHECK(bmemset(b, b->d_pt_acc, 0, sizeof(int64_t) * total));
    HIPCHECK(bmemset(b, b->d_pt_prop, 0, sizeof(int64_t) * total));
    b->pt_ladders = n_ladders;
    b->pt_len = ladder_len;
    b->pt_interval = interval;
    b->pt_since = 0;
    b->pt_seed = seed;
    b->pt_round = 0;
    b->pt_member = member;
    return SSE_OK;
}

int sse_pt_exchange(sse_batch *b)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    if (b->pt_ladders <= 0)
        return fail(SSE_ERR_STATE, "sse_pt_exchange: no temperature ladders set");
    if (!b->inited)
        return fail(SSE_ERR_STATE, "sse_pt_exchange before sse_init");
    if (b->any_pending) {
        int rc = launch_init_pending(b);
        if (rc)
            return rc;
    }
    int rc = launch_pt_exchange(b);
    if (rc)
        return rc;
    return check_walker_errors(b);
}

int sse_pt_get_ladders(sse_batch *b, int32_t *slot_walkers, int64_t *accepted, int64_t *proposed,
                       int64_t *rounds)
{
    if (!b)
        return fail(SSE_ERR_INVALID, "null batch");
    HIPCHECK(hipStreamSynchronize(b->stream));
    const size_t total = (size_t)b->pt_ladders * (size_t)b->pt_len;
    if (total) {
        if (slot_walkers)
            HIPCHECK(bmemcpy(b, slot_walkers, b->d_pt_slot, sizeof(int32_t) * total,
                               hipMemcpyDeviceToHost));
        if (accepted)
            HIPCHECK(bmemcpy(b, accepted, b->d_pt_acc, sizeof(int64_t) * total,
                               hipMemcpyDeviceToHost));
        if (proposed)
            HIPCHECK(bmemcpy(b, proposed, b->d_pt_prop, sizeof(int64_t) * total,
                               hipMemcpyDeviceToHost));
    }
    if (rounds)
        *rounds = b->pt_round;
    return SSE_OK;
}

static int read_walker(sse_batch *b, int32_t walker, DevWalker *w)
{
    if (!b || walker < 0 || walker >= b->n_walkers)
        return fail(SSE_ERR_INVALID, "bad walker");
    HIPCHECK(hipStreamSynchronize(b->stream));
    HIPCHECK(bmemcpy(b, w, b->d_walkers + walker, sizeof *w, hipMemcpyDeviceToHost));
    return SSE_OK;
}

int sse_pt_log_weight_ratio(sse_batch *b, int32_t walker, double new_T, double *ratio)
{
    if (!ratio || !(new_T > 0) || !std::isfinite(new_T))
        return fail(SSE_ERR_INVALID, "sse_pt_log_weight_ratio: bad arguments");
    DevWalker w;
    int rc = read_walker(b, walker, &w);
    if (rc)
        return rc;
    *ratio = -(double)w.n * std::log(new_T / w.T);
    return SSE_OK;
}

int sse_pt_change_parameter(sse_batch *b, int32_t walker, double new_T)
{
    if (!(new_T > 0) || !std::isfinite(new_T))
        return fail(SSE_ERR_INVALID, "sse_pt_change_parameter: T must be positive and finite");
    DevWalker w;
    int rc = read_walker(b, walker, &w);
    if (rc)
        return rc;
    w.T = new_T;
    HIPCHECK(bmemcpy(b, b->d_walkers + walker, &w, sizeof w, hipMemcpyHostToDevice));
    return SSE_OK;
}

int sse_pt_get_parameter(sse_batch *b, int32_t walker, double *T)
{
    if (!T)
        return fail(SSE_ERR_INVALID, "sse_pt_get_parameter: null output");
    DevWalker w;
    int rc = read_walker(b, walker, &w);
    if (rc)
        return rc;
    *T = w.T;
    return SSE_OK;
}

}  // extern "C"

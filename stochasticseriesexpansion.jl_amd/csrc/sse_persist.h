// sse_persist.h -- one persistent, work-queue-driven launch for a whole call.
//
// Walkers are independent Markov chains: Carlo.sweep! of one walker
// (sse.jl:62-68) needs nothing from any other walker.  The per-sweep kernels
// hold every walker of a launch to the pace of the launch's slowest worm
// update; here each walker advances on its own.  The launch has as many
// workgroups (one wave each) as the device holds at the launch's LDS size.
// A workgroup takes the next task from an atomic counter -- a task is one
// walker and its whole sweep budget for the call -- and runs that walker's
// sweeps back to back, each the fused diagonal update + vertex list (K1
// phase), worm update (K2 phase) and, once thermalized, measurement (K3 phase,
// accumulated), then takes the next task, until the counter passes the task
// count.
//
// Ownership, and why nothing waits: a walker is one task, taken by exactly one
// workgroup, which alone reads and writes the walker's DevWalker, slab, spins
// and accumulator row during the launch.  No workgroup ever waits for another;
// the only shared state is the task counter and the slab pool's bump pointer
// (both atomics) and the pool's error/maximum words (atomicOr / atomicMax).
// Between the phases of one walker the workgroup passes a barrier (which
// orders its own global and LDS accesses), and every global record a phase
// reads was written by vector stores of the same workgroup: the worm phase
// reads global links with vector loads (RecHV, RecGV), never through the
// scalar cache.  Tasks run largest expected work first (the schedule order),
// so the launch ends with the small tasks.
#pragma once
#include "sse_wave.h"

// Task t of the launch: walker order[t] (or t) with budget sweeps (or the
// call's sweep count when budgets is null).  The counter *queue must be 0 at
// launch (persist_reset_kernel).  measure 1 runs the measurement after every
// thermalized sweep (sse_run), 0 leaves it out (sse_sweep).
__global__ __launch_bounds__(64) void wave_persist_kernel(DevWalker *W, const DevModel *models,
                                                          DevPool *pool, double *out, double *sum,
                                                          double *sumsq, int64_t *count,
                                                          double *wlf, int32_t *done,
                                                          const int32_t *order, int n_tasks,
                                                          const int32_t *budgets, int n_sweeps,
                                                          int measure, int lds_bytes, int rec_mode,
                                                          int chase, int32_t *queue)
{
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x;
    for (;;) {
        int t = 0;
        if (lane == 0)
            t = atomicAdd(queue, 1);
        t = (int)uni((uint32_t)__shfl(t, 0));
        if (t >= n_tasks)
            break;
        const int wi = walker_at(order, t);
        const int ns = budgets ? budgets[wi] : n_sweeps;
        for (int s = 0; s < ns; ++s) {
            // a walker that stopped (error, retired budget, inactive) ends its task
            const DevWalker &cw = W[wi];
            const int stop = cw.error | !cw.active | cw.retire;
            if (uni((uint32_t)stop))
                break;
            k1_walker(W, models, pool, wi, 3, -1, lds, lds_bytes, 7);
            __syncthreads();
            k2_walker<true>(W, models, pool, wi, 1, lds_bytes, rec_mode, lds, chase);
            __syncthreads();
            if (measure) {
                k3_walker(W, models, out, sum, sumsq, count, wlf, wi, 1, done);
                __syncthreads();
            }
        }
    }
}

__global__ void persist_reset_kernel(int32_t *queue) { *queue = 0; }

// LDS bytes of a persistent launch: the diagonal-update phase's (model tables
// staged) and the worm phase's record size, whichever is larger
static inline size_t persist_lds_bytes(const WaveConfig &c)
{
    return std::max(c.lds_k1, c.lds_k2);
}

// Workgroups the device holds at once at lds bytes per workgroup, from the
// residency measured on MI355X (a workgroup takes round512(bytes + 256) of
// about 162816 usable LDS bytes per CU; 64-lane workgroups, registers allow
// at least 8 per CU)
static inline int persist_resident(size_t lds, int n_cu)
{
    const size_t per_cu_lds = 162816;
    const size_t alloc = (lds + 256 + 511) & ~(size_t)511;
    size_t k = per_cu_lds / alloc;
    if (k < 1)
        k = 1;
    if (k > 8)
        k = 8;
    return (int)k * n_cu;
}

static inline int wave_launch_persist(WaveConfig &c, hipStream_t s, DevWalker *W, const DevModel *M,
                                      DevPool *pool, double *out, double *sum, double *sumsq,
                                      int64_t *count, double *wlf, int32_t *done,
                                      const int32_t *order, int n_tasks, const int32_t *budgets,
                                      int n_sweeps, int measure, int32_t *queue, int n_cu)
{
    if (n_tasks <= 0)
        return SSE_OK;
    const size_t lds = persist_lds_bytes(c);
    const int grid = std::min(n_tasks, persist_resident(lds, n_cu));
    hipLaunchKernelGGL(persist_reset_kernel, dim3(1), dim3(1), 0, s, queue);
    hipLaunchKernelGGL(wave_persist_kernel, dim3(grid), dim3(64), lds, s, W, M, pool, out, sum, sumsq,
                       count, wlf, done, order, n_tasks, budgets, n_sweeps, measure, (int)lds,
                       c.rec_mode, c.chase, queue);
    return hipGetLastError() == hipSuccess ? SSE_OK : SSE_ERR_HIP;
}

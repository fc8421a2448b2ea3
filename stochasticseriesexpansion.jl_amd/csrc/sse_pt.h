// Replica exchange in T (parallel tempering) over temperature ladders of the
// batch's walkers.  The reference supplies the two hooks of the exchange,
// Carlo.parallel_tempering_log_weight_ratio (-n log(T'/T), src/sse.jl:390-396)
// and Carlo.parallel_tempering_change_parameter! (T <- T', src/sse.jl:398-405);
// Carlo drives them from its own exchange loop on the host.  Here the exchange
// round runs on the device between sweeps, so a batch with ladders never leaves
// the GPU.
//
// A ladder is a list of walkers in temperature-slot order (slot_walker[k] = the
// walker now at slot k).  Round r proposes the neighbour pairs (k, k+1) with
// k % 2 == r % 2 (pairs of one round are disjoint, one thread each) and accepts
// iff log(u) < lwr(a -> T_b) + lwr(b -> T_a), i.e. with probability
// min(1, exp(sum)).  An accepted swap exchanges the two walkers' temperatures,
// their slots, and their accumulator rows (meas_offset, acc_slot), so the
// measurements of a temperature keep landing in one row.  The uniform of pair
// (ladder l, round r, slot k) is word 0|1 of the Philox4x32-10 block
// (r lo, r hi, k | 2^31, l) under the ladder seed: independent of the thread
// mapping, and apart from every walker stream (third counter word 0).
#pragma once

#include "sse_common.h"

__global__ __launch_bounds__(64) void pt_exchange_kernel(DevWalker *W, int32_t *slot_walker,
                                                         int64_t *accepted, int64_t *proposed,
                                                         int n_ladders, int len, uint64_t seed,
                                                         int64_t round)
{
    const int half = len / 2;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (half <= 0 || t >= n_ladders * half)
        return;
    const int l = t / half;
    const int k = (int)(round & 1) + 2 * (t - l * half);
    if (k + 1 >= len)
        return;
    int32_t *sw = slot_walker + (size_t)l * len;
    const int a = sw[k], b = sw[k + 1];
    DevWalker &A = W[a];
    DevWalker &B = W[b];
    if (A.error || B.error)
        return;
    const double Ta = A.T, Tb = B.T;
    // lwr(a -> Tb) + lwr(b -> Ta), each -n log(new / old) as in sse.jl:395
    const double d = -(double)A.n * sse_log(Tb / Ta) + -(double)B.n * sse_log(Ta / Tb);
    uint32_t o0, o1;
    philox_block4((uint32_t)round, (uint32_t)((uint64_t)round >> 32), (uint32_t)k | 0x80000000u,
                  (uint32_t)l, seed, o0, o1);
    const double u = word_to_float((uint64_t)o0 | ((uint64_t)o1 << 32));
    const bool acc = d >= 0.0 || u == 0.0 || sse_log(u) < d;
    const size_t pk = (size_t)l * len + k;
    proposed[pk] += 1;
    if (acc) {
        accepted[pk] += 1;
        A.T = Tb;
        B.T = Ta;
        const int32_t mo = A.meas_offset, as = A.acc_slot;
        A.meas_offset = B.meas_offset;
        A.acc_slot = B.acc_slot;
        B.meas_offset = mo;
        B.acc_slot = as;
        sw[k] = b;
        sw[k + 1] = a;
    }
}

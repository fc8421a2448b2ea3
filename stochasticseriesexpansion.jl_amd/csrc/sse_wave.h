// sse_wave.h -- one wavefront per walker (placeholder: delegates to the lane kernels)
#pragma once
#include <vector>
#include "sse_common.h"
#include "sse_lane.h"

struct WaveConfig {
    int dummy;
};

static inline int wave_configure(WaveConfig &, const std::vector<DevModel> &, int) { return 0; }

static inline int wave_launch_sweep(WaveConfig &, hipStream_t s, DevWalker *W, const DevModel *M,
                                    DevPool *pool, int n, int phases, hipEvent_t *)
{
    hipLaunchKernelGGL(lane_sweep_kernel, dim3((n + 63) / 64), dim3(64), 0, s, W, M, pool, 0, n, phases);
    return 0;
}
static inline int wave_launch_init(WaveConfig &, hipStream_t s, DevWalker *W, const DevModel *M,
                                   DevPool *pool, int n)
{
    hipLaunchKernelGGL(lane_init_kernel, dim3((n + 63) / 64), dim3(64), 0, s, W, M, pool, 0, n);
    return 0;
}
static inline int wave_launch_measure(WaveConfig &, hipStream_t s, DevWalker *W, const DevModel *M,
                                      double *out, double *sum, double *sumsq, int64_t *count,
                                      uint8_t *scratch, int max_sites, int n, int mode)
{
    hipLaunchKernelGGL(lane_measure_kernel, dim3((n + 63) / 64), dim3(64), 0, s, W, M, out, sum,
                       sumsq, count, scratch, max_sites, n, mode);
    return 0;
}
static inline int wave_dense_vertex_list(WaveConfig &, const DevWalker &w, int32_t *links, int64_t cap)
{
    if (links && cap >= 4 * w.L && w.L > 0)
        return hipMemcpy(links, w.links, sizeof(int32_t) * 4 * w.L, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 2;
    return 0;
}
static inline void wave_fix_first_last(const DevWalker &, int32_t *, int32_t *, int32_t *) {}

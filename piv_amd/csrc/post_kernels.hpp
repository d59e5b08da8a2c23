// Launchers of the post-analysis kernels (post_kernels.hip): flow statistics of validated
// fields and the time-series filters and metrics of a run.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "piv_device.hpp"

namespace piv {

// same layout as piv_flow_stats (include/piv.h)
struct FlowStatsDev {
    float vorticity_mean, vorticity_abs_mean, ke_mean, ke_sum, coverage, noise_px;
    uint32_t n_usable, n_vorticity;
};

// largest temporal window (offsets per output) and spatial kernel width
constexpr int TEMPORAL_MAXW = 63;
constexpr int SPATIAL_MAXR = 32, SPATIAL_MAXK = 2 * SPATIAL_MAXR + 1;

struct TemporalArgs {
    int nt, ny, nx;
    int kernel;  // 0 mean, 1 median
    int need;    // max(min_support, 1)
    int noff;
    int off[TEMPORAL_MAXW];
};

struct SpatialArgs {
    int nt, ny, nx;
    int r;       // kernel radius, weights [(2r+1)^2] row-major (j, i)
    int leave;   // Chebyshev radius left out, < 0: none
    int none;    // the left-out block holds more than half the weight: every output NaN
    float min_full;  // min_weight * (weight of the kept kernel)
};

// dx, dy, flags: [npairs][ny][nx]; k = f32(1 / dt); part: flow_partial_bytes; resid: u32
// [npairs][2][ny*nx]; out: device [npairs]
hipError_t launch_flow_stats(const float *dx, const float *dy, const uint8_t *flags, int npairs,
                             int ny, int nx, float k, float step, void *part, uint32_t *resid,
                             FlowStatsDev *out, hipStream_t st);
size_t flow_partial_bytes(int npairs, int ny, int nx);

hipError_t launch_temporal(const TemporalArgs &args, const float *dx, const float *dy,
                           const uint8_t *usable, const uint8_t *gate, float *odx, float *ody,
                           uint16_t *sup, hipStream_t st);
hipError_t launch_spatial(const SpatialArgs &args, const float *w, const float *dx, const float *dy,
                          const uint8_t *gate, float *odx, float *ody, uint16_t *sup,
                          hipStream_t st);
hipError_t launch_mask(long n, const float *v, const uint8_t *ok, float *out, hipStream_t st);

// cross-validation errors compacted in index order; count_dev: device u64
size_t cv_scratch_bytes(long n);
hipError_t launch_cv_errors(long n, const float *rdx, const float *rdy, const uint8_t *ok,
                            const float *pdx, const float *pdy, void *scratch, float *out,
                            uint64_t *count_dev, hipStream_t st);
// medians (device results); keys: u32 scratch of n (errors), (nt-1)*ny*nx (change) or
// nt*ny*nx (roughness) entries
hipError_t launch_median_errors(long n, const float *e, const uint8_t *keep, uint32_t *keys,
                                float *med, uint64_t *count, hipStream_t st);
hipError_t launch_pair_change(int nt, int ny, int nx, const float *dx, const float *dy,
                              uint32_t *keys, float *med, hipStream_t st);
hipError_t launch_roughness(int nt, int ny, int nx, const float *dx, const float *dy,
                            uint32_t *keys, float *med, hipStream_t st);

}  // namespace piv

// C-ABI of the post-analysis stages (include/piv.h): per-pair flow statistics of validated
// fields (piv-io/src/flow.rs:13-34, piv-core/src/flowstats.rs:69-137) and the time-series
// filters and metrics of a run (piv-core/src/smooth.rs:44-310).  Every entry point runs HIP
// kernels (post_kernels.hip) on device-resident arrays; there is no host fallback.
#include <cmath>
#include <cstdint>
#include <vector>

#include "piv_internal.hpp"
#include "post_kernels.hpp"

using namespace pivi;

static_assert(sizeof(piv_flow_stats) == sizeof(FlowStatsDev), "piv_flow_stats layout");

namespace {

// A stream of the caller's device for the duration of one call.
struct CallStream {
    hipStream_t s = nullptr;
    ~CallStream()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

piv_status open_stream(int device, CallStream &cs)
{
    PIV_TRY(set_device(device));
    HIP_TRY(hipStreamCreate(&cs.s));
    return PIV_OK;
}

piv_status check_dims(uint64_t nt, uint64_t ny, uint64_t nx, uint64_t limit = 0xffffffffull)
{
    if (ny == 0 || nx == 0) return fail(PIV_ERR_INVALID_ARGUMENT, "empty grid");
    if (nt * ny * nx > limit)
        return fail(PIV_ERR_INVALID_ARGUMENT, "series of %llu values exceeds %llu",
                    (unsigned long long)(nt * ny * nx), (unsigned long long)limit);
    return PIV_OK;
}

}  // namespace

extern "C" {

piv_status piv_flow_stats_device(piv_analyzer *a, const piv_grid *grid, uint32_t npairs,
                                 const float *dx, const float *dy, const uint8_t *flags,
                                 double dt_s, piv_flow_stats *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !grid || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (npairs == 0) return PIV_OK;
        if (!dx || !dy || !flags) return fail(PIV_ERR_INVALID_ARGUMENT, "null field");
        if (!(dt_s > 0.0) || !std::isfinite(dt_s))
            return fail(PIV_ERR_INVALID_ARGUMENT, "dt must be positive");
        PIV_TRY(check_dims(npairs, grid->ny, grid->nx, 0x7fffffffull / 2));
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        const int ny = (int)grid->ny, nx = (int)grid->nx;
        const size_t cells = (size_t)ny * nx;
        PIV_TRY(a->post_part.reserve(flow_partial_bytes((int)npairs, ny, nx)));
        PIV_TRY(a->post_keys.reserve(cells * 2 * npairs * sizeof(uint32_t)));
        PIV_TRY(a->post_out.reserve(npairs * sizeof(FlowStatsDev)));
        // flow.rs:27: k = (1 / dt) as f32; flowstats step = grid.step as f32
        const float k = (float)(1.0 / dt_s);
        const double bytes = 9.0 * (double)cells * npairs;  // dx, dy, flags read once
        HIP_TRY(timed(a, "flowstats", 0, bytes, (uint64_t)npairs, [&] {
            return launch_flow_stats(dx, dy, flags, (int)npairs, ny, nx, k, (float)grid->step,
                                     a->post_part.p, a->post_keys.as<uint32_t>(),
                                     a->post_out.as<FlowStatsDev>(), a->stream);
        }));
        HIP_TRY(hipMemcpyAsync(out, a->post_out.p, npairs * sizeof(FlowStatsDev),
                               hipMemcpyDeviceToHost, a->stream));
        HIP_TRY(hipStreamSynchronize(a->stream));
        return PIV_OK;
    });
}

piv_status piv_smooth_temporal_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                      const float *dx, const float *dy, const uint8_t *usable,
                                      const uint8_t *gate, piv_temporal_kernel kernel,
                                      uint32_t window, uint32_t min_support,
                                      const int32_t *leave_out, uint32_t n_leave, float *out_dx,
                                      float *out_dy, uint16_t *support)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!dx || !dy || !usable || !gate || !out_dx || !out_dy || !support || (n_leave && !leave_out))
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (kernel != PIV_TEMPORAL_MEAN && kernel != PIV_TEMPORAL_MEDIAN)
            return fail(PIV_ERR_INVALID_ARGUMENT, "unknown temporal kernel %d", (int)kernel);
        // smooth.rs:54 asserts an odd window
        if (window % 2 != 1) return fail(PIV_ERR_INVALID_ARGUMENT, "temporal window must be odd");
        if (window > (uint32_t)TEMPORAL_MAXW)
            return fail(PIV_ERR_INVALID_ARGUMENT, "temporal window %u exceeds %d", window, TEMPORAL_MAXW);
        PIV_TRY(check_dims(nt, ny, nx));
        if (nt == 0) return PIV_OK;
        // smooth.rs:57-71: the centre unless left out, then half the window on each side, each
        // left-out offset replaced by the next one further out
        TemporalArgs args{};
        args.nt = (int)nt;
        args.ny = (int)ny;
        args.nx = (int)nx;
        args.kernel = kernel == PIV_TEMPORAL_MEDIAN ? 1 : 0;
        args.need = min_support > 1 ? (int)std::min<uint32_t>(min_support, 1u << 30) : 1;
        auto left = [&](long o) {
            for (uint32_t q = 0; q < n_leave; ++q)
                if (leave_out[q] == o) return true;
            return false;
        };
        const int half = (int)window / 2;
        if (!left(0)) args.off[args.noff++] = 0;
        for (int sign = -1; sign <= 1; sign += 2) {
            long kk = 1;
            int taken = 0;
            while (taken < half) {
                if (!left(sign * kk)) {
                    args.off[args.noff++] = (int)(sign * kk);
                    ++taken;
                }
                ++kk;
            }
        }
        CallStream cs;
        PIV_TRY(open_stream(device, cs));
        HIP_TRY(launch_temporal(args, dx, dy, usable, gate, out_dx, out_dy, support, cs.s));
        HIP_TRY(hipStreamSynchronize(cs.s));
        return PIV_OK;
    });
}

piv_status piv_smooth_spatial_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                     const float *dx, const float *dy, const uint8_t *gate,
                                     float sigma, float min_weight, int32_t leave_out,
                                     float *out_dx, float *out_dy, uint16_t *support)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!dx || !dy || !gate || !out_dx || !out_dy || !support)
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (!(sigma > 0.0f) || !std::isfinite(sigma))
            return fail(PIV_ERR_INVALID_ARGUMENT, "sigma must be positive");
        PIV_TRY(check_dims(nt, ny, nx));
        if (nt == 0) return PIV_OK;
        // smooth.rs:142-166: radius, weights, kept and total weight, summed in (j, i) order
        const int r = std::max((int)std::ceil(2.5f * sigma), 1);
        if (r > SPATIAL_MAXR)
            return fail(PIV_ERR_INVALID_ARGUMENT, "sigma %g needs a kernel radius above %d", sigma,
                        SPATIAL_MAXR);
        const int kw = 2 * r + 1;
        std::vector<float> w((size_t)kw * kw);
        for (int j = -r; j <= r; ++j)
            for (int i = -r; i <= r; ++i)
                w[(size_t)(j + r) * kw + (i + r)] =
                    std::exp(-((float)(i * i + j * j)) / (2.0f * sigma * sigma));
        auto excluded = [&](int i, int j) {
            return leave_out >= 0 && std::max(std::abs(i), std::abs(j)) <= leave_out;
        };
        float full = -0.0f, all = -0.0f;
        for (int j = -r; j <= r; ++j)
            for (int i = -r; i <= r; ++i) {
                if (!excluded(i, j)) full += w[(size_t)(j + r) * kw + (i + r)];
                all += w[(size_t)(j + r) * kw + (i + r)];
            }
        SpatialArgs args{};
        args.nt = (int)nt;
        args.ny = (int)ny;
        args.nx = (int)nx;
        args.r = r;
        args.leave = leave_out;
        args.none = full < 0.5f * all ? 1 : 0;
        args.min_full = min_weight * full;
        CallStream cs;
        PIV_TRY(open_stream(device, cs));
        DevBuf dw;
        PIV_TRY(dw.reserve(w.size() * sizeof(float)));
        HIP_TRY(hipMemcpyAsync(dw.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, cs.s));
        HIP_TRY(launch_spatial(args, dw.as<float>(), dx, dy, gate, out_dx, out_dy, support, cs.s));
        HIP_TRY(hipStreamSynchronize(cs.s));
        dw.release();
        return PIV_OK;
    });
}

piv_status piv_mask_device(int device, size_t n, const float *values, const uint8_t *usable,
                           float *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (n && (!values || !usable || !out)) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        CallStream cs;
        PIV_TRY(open_stream(device, cs));
        HIP_TRY(launch_mask((long)n, values, usable, out, cs.s));
        HIP_TRY(hipStreamSynchronize(cs.s));
        return PIV_OK;
    });
}

piv_status piv_cross_validation_errors_device(int device, size_t n, const float *raw_dx,
                                              const float *raw_dy, const uint8_t *usable,
                                              const float *pred_dx, const float *pred_dy,
                                              float *out, uint64_t *count)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!count || (n && (!raw_dx || !raw_dy || !usable || !pred_dx || !pred_dy || !out)))
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (n > 0xffffffffull) return fail(PIV_ERR_INVALID_ARGUMENT, "too many vectors");
        CallStream cs;
        PIV_TRY(open_stream(device, cs));
        DevBuf scratch, cnt;
        PIV_TRY(scratch.reserve(cv_scratch_bytes((long)n)));
        PIV_TRY(cnt.reserve(8));
        HIP_TRY(launch_cv_errors((long)n, raw_dx, raw_dy, usable, pred_dx, pred_dy, scratch.p, out,
                                 cnt.as<uint64_t>(), cs.s));
        HIP_TRY(hipMemcpyAsync(count, cnt.p, 8, hipMemcpyDeviceToHost, cs.s));
        HIP_TRY(hipStreamSynchronize(cs.s));
        return PIV_OK;
    });
}

piv_status piv_median_error_device(int device, const float *errors, size_t n, const uint8_t *keep,
                                   float *median, uint64_t *count)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!median || !count || (n && !errors)) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (n > 0xffffffffull) return fail(PIV_ERR_INVALID_ARGUMENT, "too many errors");
        CallStream cs;
        PIV_TRY(open_stream(device, cs));
        DevBuf keys, res;
        PIV_TRY(keys.reserve(n * 4 + 4));
        PIV_TRY(res.reserve(16));
        HIP_TRY(launch_median_errors((long)n, errors, keep, keys.as<uint32_t>(), res.as<float>(),
                                     reinterpret_cast<uint64_t *>(res.as<char>() + 8), cs.s));
        char h[16];
        HIP_TRY(hipMemcpyAsync(h, res.p, 16, hipMemcpyDeviceToHost, cs.s));
        HIP_TRY(hipStreamSynchronize(cs.s));
        std::memcpy(median, h, 4);
        std::memcpy(count, h + 8, 8);
        return PIV_OK;
    });
}

static piv_status series_median(int device, uint32_t nt, uint32_t ny, uint32_t nx, const float *dx,
                                const float *dy, float *out, bool change)
{
    if (!dx || !dy || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
    PIV_TRY(check_dims(nt, ny, nx));
    CallStream cs;
    PIV_TRY(open_stream(device, cs));
    DevBuf keys, res;
    PIV_TRY(keys.reserve((size_t)nt * ny * nx * 4 + 4));
    PIV_TRY(res.reserve(4));
    if (change)
        HIP_TRY(launch_pair_change((int)nt, (int)ny, (int)nx, dx, dy, keys.as<uint32_t>(),
                                   res.as<float>(), cs.s));
    else
        HIP_TRY(launch_roughness((int)nt, (int)ny, (int)nx, dx, dy, keys.as<uint32_t>(),
                                 res.as<float>(), cs.s));
    HIP_TRY(hipMemcpyAsync(out, res.p, 4, hipMemcpyDeviceToHost, cs.s));
    HIP_TRY(hipStreamSynchronize(cs.s));
    return PIV_OK;
}

piv_status piv_pair_to_pair_change_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                          const float *dx, const float *dy, float *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        return series_median(device, nt, ny, nx, dx, dy, out, true);
    });
}

piv_status piv_spatial_roughness_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                        const float *dx, const float *dy, float *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        return series_median(device, nt, ny, nx, dx, dy, out, false);
    });
}

}  // extern "C"

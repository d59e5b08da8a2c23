// piv_analyze_batch_noalloc: the batched analysis of piv_analyze_batch_device for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  It runs the same kernels on the same descriptors; only where the
// workspace lives differs.
//
// Workspace: one file-scope device array per GPU, `scratch` (kScratchBytes, 64 MiB), that the
// library holds until the process ends.  Host code never names the array: its device address
// on each GPU comes from hipGetSymbolAddress, taken once per device when an analyzer is first
// created there (noalloc_init) and kept per device.  Every call on a GPU shares that GPU's
// array; calls that overlap on one GPU (several host threads, or several streams of one
// thread) race on it and must be serialised by the caller.
//
// Per call: the per-pair descriptors are written into the array by a kernel that takes the
// first pair's descriptor by value (no host staging buffer, no copy), the mask tables (when the
// configuration needs them) are built there from the samples, the per-pass summaries are
// cleared by a kernel, and then each pass is the usual five launches.  Nothing is allocated,
// freed, copied or synchronised.
#include "piv_internal.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

// The pair-k descriptors of one pass from pair 0's: every pointer advances by pair k's share
// of its array (frames frame_stride bytes apart, mask tables sat_stride doubles apart, the
// fields of this pass `cells` vectors apart, the predictor `prev_cells` apart).
__global__ void k_noalloc_desc(PassPair t, ValidatePair v, long frame_stride, long sat_stride,
                               uint32_t cells, uint32_t prev_cells, uint32_t npairs,
                               PassPair *__restrict__ pp, ValidatePair *__restrict__ vp)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const size_t c = (size_t)k * cells;
    PassPair d = t;
    d.img_a = static_cast<const char *>(t.img_a) + (long)k * frame_stride;
    d.img_b = static_cast<const char *>(t.img_b) + (long)k * frame_stride;
    if (t.sat_a) d.sat_a = t.sat_a + (long)k * sat_stride;
    if (t.sat_b) d.sat_b = t.sat_b + (long)k * sat_stride;
    if (t.prev_dx) {
        d.prev_dx = t.prev_dx + (size_t)k * prev_cells;
        d.prev_dy = t.prev_dy + (size_t)k * prev_cells;
    }
    d.origins = t.origins + 4 * c;
    d.sig = t.sig + c;
    d.peaks = t.peaks + c;
    d.out_dx = t.out_dx + c;
    d.out_dy = t.out_dy + c;
    d.out_ppr = t.out_ppr + c;
    d.out_flags = t.out_flags + c;
    pp[k] = d;
    ValidatePair w = v;
    w.raw_dx = d.out_dx;
    w.raw_dy = d.out_dy;
    w.raw_ppr = d.out_ppr;
    w.raw_flags = d.out_flags;
    w.dx = v.dx + c;
    w.dy = v.dy + c;
    w.ppr = v.ppr + c;
    w.flags = v.flags + c;
    w.snapshot = v.snapshot + c;
    w.summary = v.summary + (size_t)k * 5;
    vp[k] = w;
}

// A bump allocator over the scratch array (offsets only; nothing is allocated).
struct Carve {
    char *base;
    size_t off = 0;
    template <class T> T *take(size_t count)
    {
        T *p = reinterpret_cast<T *>(base + off);
        off = pivi::align_up(off + count * sizeof(T), 256);
        return p;
    }
};

}  // namespace

namespace pivi {

// The device address of this GPU's scratch array (device `dev` current), taken once per
// device; later calls return the stored address or the stored failure.
piv_status noalloc_init(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(PIV_ERR_DEVICE, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[dev]) {
        void *p = nullptr;
        g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[dev] = g_err[dev] == hipSuccess ? p : nullptr;
        g_done[dev] = true;
    }
    if (g_err[dev] != hipSuccess)
        return fail(PIV_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                    hipGetErrorString(g_err[dev]));
    return PIV_OK;
}

}  // namespace pivi

using namespace pivi;

extern "C" {

piv_status piv_analyze_batch_noalloc(piv_analyzer *a, const void *pixels_a, const void *pixels_b,
                                     piv_pixel_format fmt, uint32_t npairs, uint32_t height,
                                     uint32_t width, size_t row_stride, size_t frame_stride,
                                     float *dx, float *dy, float *ppr, uint8_t *flags,
                                     uint32_t *summaries, void *stream)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !dx || !dy || !ppr || !flags) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (npairs == 0) return PIV_OK;
        PIV_TRY(check_pixels_args(pixels_a, fmt, height, width, row_stride));
        PIV_TRY(check_pixels_args(pixels_b, fmt, height, width, row_stride));
        const int bpp = bytes_per_px(fmt);
        if (npairs > 1 && frame_stride < (size_t)height * row_stride)
            return fail(PIV_ERR_INVALID_ARGUMENT, "frame_stride %zu below one frame (%zu bytes)",
                        frame_stride, (size_t)height * row_stride);
        std::lock_guard<std::mutex> lk(a->mu);
        const piv_config &c = a->cfg;
        if (c.smooth_sigma > 0.0f || c.normalize_radius > 0)
            return fail(PIV_ERR_UNSUPPORTED,
                        "piv_analyze_batch_noalloc analyses raw samples: conditioning (smooth_sigma, "
                        "normalize_radius) needs frame storage; prepare the frames and call "
                        "piv_analyze_batch_device");
        const int np = (int)a->passes.size();
        constexpr int kMaxPasses = 64;
        if (np > kMaxPasses) return fail(PIV_ERR_CONFIG, "too many passes");
        PIV_TRY(set_device(a->device));
        PIV_TRY(noalloc_init(a->device));
        const int H = (int)height, W = (int)width;
        PassPlan plans[kMaxPasses];
        for (int i = 0; i < np; ++i) {
            PIV_TRY(grid_new(height, width, (uint32_t)a->passes[i], c.overlap, plans[i].grid));
            plans[i].n = a->passes[i];
            plans[i].cells = (size_t)plans[i].grid.nx * plans[i].grid.ny;
        }
        // ---- layout in the scratch array (offsets first, checked against its size) ----
        const bool mask = mask_needed(c, fmt, false);
        const int64_t sat_cells = sat_table_elems(H, W);
        Carve cv{static_cast<char *>(g_ptr[a->device])};
        double *sat_a = nullptr, *sat_b = nullptr, *carry = nullptr;
        if (mask) {
            sat_a = cv.take<double>((size_t)npairs * 2 * sat_cells) + SAT_LEAD;
            sat_b = cv.take<double>((size_t)npairs * 2 * sat_cells) + SAT_LEAD;
            carry = cv.take<double>(sat_carry_bytes(H, W, (int)npairs) / sizeof(double) + 1);
        }
        uint32_t *sums = cv.take<uint32_t>((size_t)np * npairs * 5);
        PassPair *pp[kMaxPasses];
        ValidatePair *vp[kMaxPasses];
        float *raw[kMaxPasses], *val[kMaxPasses], *org[kMaxPasses];
        uint8_t *snap[kMaxPasses], *sig[kMaxPasses];
        PeakRec *rec[kMaxPasses];
        for (int i = 0; i < np; ++i) {
            const size_t n = plans[i].cells * npairs;
            pp[i] = cv.take<PassPair>(npairs);
            vp[i] = cv.take<ValidatePair>(npairs);
            raw[i] = cv.take<float>(3 * n + (n + 3) / 4);            // dx, dy, ppr, then flags
            val[i] = i + 1 < np ? cv.take<float>(3 * n + (n + 3) / 4) : nullptr;
            snap[i] = cv.take<uint8_t>(n);
            org[i] = cv.take<float>(4 * n);
            sig[i] = cv.take<uint8_t>(n);
            rec[i] = cv.take<PeakRec>(n);
        }
        if (cv.off > kScratchBytes)
            return fail(PIV_ERR_INVALID_ARGUMENT,
                        "piv_analyze_batch_noalloc: %u pairs of %ux%u need %zu bytes of workspace, "
                        "more than its %zu; analyse fewer pairs per call",
                        npairs, height, width, cv.off, kScratchBytes);
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : a->stream;
        const int pitch = (int)(row_stride / (size_t)bpp);
        if (mask) {
            HIP_TRY(launch_sat_raw(fmt, pixels_a, pitch, (long)frame_stride, H, W, sat_a, sat_a + sat_cells,
                                   carry, st, (int)npairs, 2 * sat_cells));
            HIP_TRY(launch_sat_raw(fmt, pixels_b, pitch, (long)frame_stride, H, W, sat_b, sat_b + sat_cells,
                                   carry, st, (int)npairs, 2 * sat_cells));
        }
        uint32_t *final_sums = summaries ? summaries : sums + (size_t)(np - 1) * npairs * 5;
        HIP_TRY(launch_zero_u32(sums, (uint32_t)(np * npairs * 5), st));
        if (summaries) HIP_TRY(launch_zero_u32(summaries, npairs * 5, st));
        const float maxdisp = (float)c.windows[0] / 2.0f;
        for (int i = 0; i < np; ++i) {
            const PassPlan &pl = plans[i];
            const size_t n = pl.cells * npairs;
            PassPair d{};
            d.img_a = pixels_a;
            d.img_b = pixels_b;
            d.use_mask = mask ? 1 : 0;
            d.sat_a = sat_a;
            d.sat_b = sat_b;
            d.sat_cells = sat_cells;
            if (i > 0) {
                d.prev_dx = val[i - 1];
                d.prev_dy = val[i - 1] + plans[i - 1].cells * npairs;
                d.prev = plans[i - 1].grid;
                d.has_prev = 1;
            }
            d.out_dx = raw[i];
            d.out_dy = raw[i] + n;
            d.out_ppr = raw[i] + 2 * n;
            d.out_flags = reinterpret_cast<uint8_t *>(raw[i] + 3 * n);
            d.origins = org[i];
            d.sig = sig[i];
            d.peaks = rec[i];
            d.out_peak = nullptr;
            d.H = H;
            d.W = W;
            d.pitch = pitch;
            d.fmt = fmt;
            d.grid = pl.grid;
            d.min_mean = c.min_mean;
            d.min_std = c.min_std;
            d.deform = (int32_t)c.deformation;
            ValidatePair v{};
            const bool last = i + 1 == np;
            v.dx = last ? dx : val[i];
            v.dy = last ? dy : val[i] + n;
            v.ppr = last ? ppr : val[i] + 2 * n;
            v.flags = last ? flags : reinterpret_cast<uint8_t *>(val[i] + 3 * n);
            v.snapshot = snap[i];
            v.summary = last ? final_sums : sums + (size_t)i * npairs * 5;
            v.nx = pl.grid.nx;
            v.ny = pl.grid.ny;
            v.ppr_threshold = c.ppr_threshold;
            v.median_threshold = c.median_threshold;
            v.median_eps = c.median_eps;
            v.max_displacement = maxdisp;
            v.replace = c.replace ? 1 : 0;
            const uint32_t prev_cells = i > 0 ? (uint32_t)plans[i - 1].cells : 0;
            hipLaunchKernelGGL(k_noalloc_desc, dim3((npairs + 63) / 64), dim3(64), 0, st, d, v,
                               (long)frame_stride, (long)(2 * sat_cells), (uint32_t)pl.cells, prev_cells,
                               npairs, pp[i], vp[i]);
            HIP_TRY(hipGetLastError());
            const uint32_t cells = (uint32_t)pl.cells;
            HIP_TRY(launch_place_pass(pp[i], (int)npairs, cells, st));
            const int mp = mf_precision(a, pl.n);
            const int def = i > 0 ? (int)c.deformation : 0;
            if (mp)
                HIP_TRY(launch_correlate_mf(pl.n, fmt, mp, pp[i], (int)npairs, cells, st, def));
            else
                HIP_TRY(launch_correlate(pl.n, fmt, variant_of(a, pl.n, fmt), pp[i], (int)npairs, cells, st, def));
            HIP_TRY(launch_peak_finish(pp[i], (int)npairs, cells, st, mp != 0));
            HIP_TRY(launch_validate(vp[i], (int)npairs, cells, st));
        }
        return PIV_OK;
    });
}

}  // extern "C"

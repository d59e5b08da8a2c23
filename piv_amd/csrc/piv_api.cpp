// C-ABI implementation (include/piv.h): PivAnalyzer / PreparedFrame lifecycle, frame
// conditioning and the multi-pass orchestration over the HIP kernels.
//
// Per pass (multipass.rs:152-192) each kernel is one launch for every window of every pair in
// the call: placement + mask, correlation, sub-pixel fit + assembly, then the two validation
// kernels; the validated field stays in HBM as the next pass's predictor.  All per-pair
// descriptors of a call are uploaded with a single host-to-device copy before the first
// launch, and a call whose descriptors equal the previous call's replays its captured graph.
#include <cstdlib>

#include "piv_internal.hpp"

using namespace piv;

namespace pivi {

thread_local std::string g_err;

piv_status fail(piv_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

// grid.rs:20-44
piv_status grid_new(uint32_t h, uint32_t w, uint32_t window, float overlap, GridDev &g)
{
    if (!(overlap >= 0.0f && overlap < 1.0f))
        return fail(PIV_ERR_CONFIG, "invalid configuration: overlap %g must be in [0, 1)", overlap);
    if (window == 0 || window > w || window > h)
        return fail(PIV_ERR_IMAGE_TOO_SMALL,
                    "image (%ux%u) is smaller than the %u px interrogation window", h, w, window);
    volatile float prod = (float)window * (1.0f - overlap);  // f32, as grid.rs:33
    const float s = std::round(prod);
    int step = s > 0.0f ? (int)s : 0;
    if (step < 1) step = 1;
    g.window = (int)window;
    g.step = step;
    g.nx = ((int)w - (int)window) / step + 1;
    g.ny = ((int)h - (int)window) / step + 1;
    g.x0 = ((int)w - (int)window - (g.nx - 1) * step) / 2;
    g.y0 = ((int)h - (int)window - (g.ny - 1) * step) / 2;
    return PIV_OK;
}

piv_grid to_pub(const GridDev &g)
{
    return piv_grid{(uint32_t)g.window, (uint32_t)g.step, (uint32_t)g.nx,
                    (uint32_t)g.ny,     (uint32_t)g.x0,   (uint32_t)g.y0};
}

// preprocess.rs:161-168 kernel taps (host expf, as the reference's f32 exp)
int gauss_taps(float sigma, std::vector<float> &k)
{
    float rf = std::ceil(3.0f * sigma);
    if (rf < 1.0f) rf = 1.0f;
    const int r = (int)rf;
    k.assign(2 * r + 1, 0.0f);
    for (int i = -r; i <= r; ++i) {
        volatile float num = -(float)(i * i);
        volatile float den = 2.0f * sigma * sigma;
        k[i + r] = std::exp((float)(num / den));
    }
    float total = 0.0f;
    for (int i = 0; i <= 2 * r; ++i) total += k[i];
    for (int i = 0; i <= 2 * r; ++i) k[i] /= total;
    return r;
}

}  // namespace pivi

using namespace pivi;

namespace pivi {

int variant_of(const piv_analyzer *a, int n, int fmt)
{
    auto it = a->variants.find({n, fmt});
    return it == a->variants.end() ? 0 : it->second;
}

int mf_precision(const piv_analyzer *a, int n)
{
    return (a->precision != 0 && correlate_mf_supported(n)) ? a->precision : 0;
}

// PivAnalyzer::new checks (multipass.rs:91-119)
piv_status check_config(const piv_config &c)
{
    if (c.num_windows == 0)
        return fail(PIV_ERR_CONFIG, "invalid configuration: at least one window size is required");
    if (c.num_windows > PIV_MAX_PASSES)
        return fail(PIV_ERR_CONFIG, "invalid configuration: at most %d window sizes", PIV_MAX_PASSES);
    for (uint32_t i = 0; i < c.num_windows; ++i)
        if (c.windows[i] < 8 || c.windows[i] % 2 != 0)
            return fail(PIV_ERR_CONFIG, "invalid configuration: window size %u must be even and >= 8",
                        c.windows[i]);
    if (!(c.overlap >= 0.0f && c.overlap < 1.0f))
        return fail(PIV_ERR_CONFIG, "invalid configuration: overlap %g must be in [0, 1)", c.overlap);
    if (c.final_iterations == 0)
        return fail(PIV_ERR_CONFIG, "invalid configuration: final_iterations must be at least 1");
    if (c.num_windows + c.final_iterations - 1 > 64)
        return fail(PIV_ERR_CONFIG, "invalid configuration: too many passes");
    for (uint32_t i = 0; i < c.num_windows; ++i)
        if (!correlate_supported((int)c.windows[i]))
            return fail(PIV_ERR_UNSUPPORTED,
                        "window size %u is valid in the reference but above this library's largest "
                        "(%d px; every even size from 8 up to it is supported)",
                        c.windows[i], kMaxWindow);
    if (c.deformation > PIV_DEFORM_BICUBIC)
        return fail(PIV_ERR_CONFIG, "invalid configuration: deformation mode %u", c.deformation);
    if (c.smooth_sigma > 0.0f && std::ceil(3.0f * c.smooth_sigma) > 255.0f)
        return fail(PIV_ERR_CONFIG, "smooth_sigma %g too large", c.smooth_sigma);
    return PIV_OK;
}

// The window signal mask (multipass.rs:233-241) can be skipped without changing any result
// only when it provably passes for every window: min_std <= 0 (std >= 0 always) and either
// min_mean = -inf, or the base image is u8 samples without background (all values >= 0 and
// every SAT sum exact in f64: multiples of 2^-31 below 2^22 fit 53 bits), so mean >= 0.
bool mask_needed(const piv_config &c, int in_fmt, bool has_bg)
{
    if (!(c.min_std <= 0.0f)) return true;
    if (c.min_mean == -std::numeric_limits<float>::infinity()) return false;
    if (in_fmt == FMT_U8 && !has_bg && c.min_mean <= 0.0f) return false;
    return true;
}

int screen_tile(const piv_analyzer *a, int in_fmt, bool cond, int H, int W)
{
    const piv_config &c = a->cfg;
    if (in_fmt != FMT_U16 || cond || !(c.min_mean <= 0.0f) || !(c.min_std <= 0.0f)) return 0;
    if (const char *env = std::getenv("PIV_NO_TABLE_SCREEN"))
        if (env[0] == '1') return 0;
    if ((H + 31) / 32 > SCREEN_MAX_BANDS) return 0;
    // the tables' rounding error on any box, (4 (W + H) + 3) 2^-53 H W for samples <= 1,
    // must stay below the smallest non-zero sample 1/65535
    if ((4.0 * ((double)W + H) + 3.0) * (double)H * (double)W * 0x1p-53 >= 0x1p-16) return 0;
    int n = 1 << 30;
    for (int p : a->passes) n = std::min(n, p);
    int t = 1;
    while (t * 2 <= n / 2 && t < 32) t *= 2;
    return t;
}

piv_status screen_for(piv_analyzer *a, const piv_frame *const *fs, size_t count, int n,
                      hipStream_t st)
{
    int need = 1;
    while (need * 2 <= n / 2 && need < 32) need *= 2;
    for (size_t i = 0; i < count; ++i) {
        const piv_frame *f = fs[i];
        if (!f->has_sat || f->screen_t <= need) continue;
        PIV_TRY(a->sat_carry.reserve(sat_carry_bytes(f->H, f->W, 1)));
        const void *img = f->raw_ptr;
        double *sum = const_cast<double *>(f->sat_ptr);
        HIP_TRY(launch_sat_list(f->in_fmt, &img, &sum, 1, f->raw_pitch, f->H, f->W,
                                a->sat_carry.as<double>(), st, need));
        f->screen_t = need;
    }
    return PIV_OK;
}

bool prepare_launches(const piv_config &c, int in_fmt, bool has_bg)
{
    return c.smooth_sigma > 0.0f || c.normalize_radius > 0 || has_bg || mask_needed(c, in_fmt, has_bg);
}

piv_status set_device(int dev)
{
    HIP_TRY(hipSetDevice(dev));
    return PIV_OK;
}

hipError_t prof_event(piv_analyzer *a, hipEvent_t *e)
{
    if (!a->prof_pool.empty()) {
        *e = a->prof_pool.back();
        a->prof_pool.pop_back();
        return hipSuccess;
    }
    // timing only: read by hipEventElapsedTime after the stream is idle, so the record needs
    // no system-scope fence (whose cache writeback and invalidation would otherwise stretch
    // the gap between the kernels it separates)
    return hipEventCreateWithFlags(e, hipEventDisableSystemFence);
}

// Fold completed timing events into the counters (the stream must be idle).
void prof_harvest(piv_analyzer *a)
{
    for (auto &r : a->prof_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
            auto &st = a->prof_stats[{r.name, r.window}];
            std::snprintf(st.name, sizeof st.name, "%s", r.name.c_str());
            st.window = (uint32_t)r.window;
            st.launches += 1;
            st.units += r.units;
            st.total_ms += ms;
            st.bytes += r.bytes;
        }
        a->prof_pool.push_back(r.e0);
        a->prof_pool.push_back(r.e1);
    }
    a->prof_pending.clear();
}

// A frame buffer of at least `bytes`: kept if large enough, else a fitting block from the
// analyzer's pool of released frame memory, else a new allocation.
static piv_status frame_reserve(piv_analyzer *a, DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes) return PIV_OK;
    if (a->pool->take(b, bytes)) return PIV_OK;
    return b.reserve(bytes);
}

// PreparedFrame::new (preprocess.rs:137-153) on device samples.
static piv_status prepare_frame_impl(piv_analyzer *a, piv_frame *f, const float *bg_dev)
{
    const piv_config &c = a->cfg;
    const int H = f->H, W = f->W;
    const size_t n = (size_t)H * W;
    const bool blur = c.smooth_sigma > 0.0f, norm = c.normalize_radius > 0;
    const bool cond = blur || norm || bg_dev != nullptr;
    const bool mask = mask_needed(c, f->in_fmt, bg_dev != nullptr);
    // the blur reads the samples itself (base and blur in one pass)
    const bool fused = blur && blur_fusable(a->taps_r);
    // The base image (background-subtracted samples) is materialised where something reads
    // it: with a background (mask tables, downloads, the image itself when nothing else is
    // configured) or for a conditioning step that reads an image.  Without a background and
    // with the fused blur the tables read the samples, as without conditioning.
    const bool mat_base = cond && (bg_dev != nullptr || !fused);
    f->has_bg = bg_dev != nullptr;
    f->cells = sat_table_elems(H, W);
    hipStream_t st = a->stream;
    f->base_ptr = nullptr;
    if (mat_base) {
        PIV_TRY(frame_reserve(a, f->base, n * sizeof(float)));
        f->base_ptr = f->base.as<float>();
    }
    // ---- base image and blur ----
    const float *cur = f->base_ptr;
    float *image = nullptr;
    if (blur || norm) {
        PIV_TRY(frame_reserve(a, f->image, n * sizeof(float)));
        image = f->image.as<float>();
    }
    if (fused) {
        float *dst = image;
        if (norm) {
            PIV_TRY(a->prep_tmp2.reserve(n * sizeof(float)));
            dst = a->prep_tmp2.as<float>();
        }
        HIP_TRY(launch_blur_fused(f->in_fmt, f->raw_ptr, f->raw_pitch, bg_dev, H, W,
                                  a->d_taps.as<float>(), a->taps_r, f->base_ptr, dst, st));
        cur = dst;
    } else if (mat_base) {
        HIP_TRY(launch_base(f->in_fmt, f->raw_ptr, f->raw_pitch, bg_dev, H, W, f->base_ptr, st));
        if (blur) {
            float *dst = image;
            if (norm) {
                PIV_TRY(a->prep_tmp2.reserve(n * sizeof(float)));
                dst = a->prep_tmp2.as<float>();
            }
            PIV_TRY(a->prep_tmp.reserve(n * sizeof(float)));
            HIP_TRY(launch_blur(cur, H, W, a->d_taps.as<float>(), a->taps_r, a->prep_tmp.as<float>(),
                                dst, st));
            cur = dst;
        }
    }
    // ---- mask tables of the base image ----
    if (mask) {
        PIV_TRY(frame_reserve(a, f->sat, (size_t)f->cells * 2 * sizeof(double)));
        double *sum = f->sat.as<double>() + SAT_LEAD, *sq = sum + f->cells;
        f->sat_ptr = sum;
        // without a materialised base image the base is the normalised raw frame: the table is
        // summed from the samples (piv_frame_download recomputes the base)
        PIV_TRY(a->sat_carry.reserve(sat_carry_bytes(H, W, 1)));
        double *carry = a->sat_carry.as<double>();
        f->screen_t = cond ? 0 : screen_tile(a, f->in_fmt, false, H, W);
        if (mat_base) HIP_TRY(launch_sat(f->base_ptr, H, W, sum, sq, carry, st));
        else HIP_TRY(launch_sat_raw(f->in_fmt, f->raw_ptr, f->raw_pitch, 0, H, W, sum, sq, carry, st,
                                    1, 0, f->screen_t));
        f->has_sat = true;
    }
    if (!cond) {
        f->fmt = f->in_fmt;
        f->img = f->raw_ptr;
        f->pitch = f->raw_pitch;
        return PIV_OK;
    }
    // ---- local normalisation ----
    if (norm) {
        PIV_TRY(a->prep_sat.reserve((size_t)f->cells * 2 * sizeof(double)));
        double *s = a->prep_sat.as<double>() + SAT_LEAD;
        PIV_TRY(a->sat_carry.reserve(sat_carry_bytes(H, W, 1)));
        HIP_TRY(launch_sat(cur, H, W, s, s + f->cells, a->sat_carry.as<double>(), st));
        HIP_TRY(launch_local_norm(cur, H, W, (int)c.normalize_radius, c.normalize_eps, s,
                                  s + f->cells, image, st));
    }
    f->fmt = FMT_F32;
    // with only a background subtracted the base image is the correlation's image
    f->img = (blur || norm) ? static_cast<const void *>(image) : static_cast<const void *>(f->base_ptr);
    f->pitch = W;
    return PIV_OK;
}

// Frames of one size and format at their own addresses (the requests of a serving batch),
// prepared together: without conditioning, their tables come from one launch per kernel for
// up to 64 frames (launch_sat_list); otherwise frame by frame.
static piv_status prepare_frames_list_impl(piv_analyzer *a, piv_frame *const *fs, int count,
                                           const float *bg_dev)
{
    if (count <= 0) return PIV_OK;
    const piv_config &c = a->cfg;
    piv_frame *f0 = fs[0];
    const bool cond = c.smooth_sigma > 0.0f || c.normalize_radius > 0 || bg_dev != nullptr;
    bool same = true;
    for (int i = 1; i < count; ++i)
        same = same && fs[i]->H == f0->H && fs[i]->W == f0->W && fs[i]->in_fmt == f0->in_fmt &&
               fs[i]->raw_pitch == f0->raw_pitch;
    if (cond || !same || !mask_needed(c, f0->in_fmt, false)) {
        for (int i = 0; i < count; ++i) PIV_TRY(prepare_frame_impl(a, fs[i], bg_dev));
        return PIV_OK;
    }
    const int H = f0->H, W = f0->W;
    const int64_t cells = sat_table_elems(H, W);
    const int t = screen_tile(a, f0->in_fmt, false, H, W);
    std::vector<const void *> imgs(count);
    std::vector<double *> sums(count);
    for (int i = 0; i < count; ++i) {
        piv_frame *f = fs[i];
        PIV_TRY(frame_reserve(a, f->sat, (size_t)cells * 2 * sizeof(double)));
        sums[i] = f->sat.as<double>() + SAT_LEAD;
        imgs[i] = f->raw_ptr;
        f->sat_ptr = sums[i];
        f->has_sat = true;
        f->screen_t = t;
        f->has_bg = false;
        f->cells = cells;
        f->base_ptr = nullptr;
        f->fmt = f->in_fmt;
        f->img = f->raw_ptr;
        f->pitch = f->raw_pitch;
    }
    PIV_TRY(a->sat_carry.reserve(sat_carry_bytes(H, W, std::min(count, 64))));
    HIP_TRY(launch_sat_list(f0->in_fmt, imgs.data(), sums.data(), count, f0->raw_pitch, H, W,
                            a->sat_carry.as<double>(), a->stream, t));
    return PIV_OK;
}

// Device storage shared by the frames of one batch preparation, from the analyzer's pool
// when a fitting block is free; it returns there when the last frame is released.
static piv_status batch_storage(piv_analyzer *a, size_t bytes, std::shared_ptr<DevBuf> &store)
{
    std::weak_ptr<BlockPool> wp = a->pool;
    const int dev = a->device;
    store.reset(new DevBuf, [wp, dev](DevBuf *b) {
        // the last frame of the batch is being released: every frame of it has waited for its
        // own last use (piv_frame_release), so the block is free for reuse
        if (b->p) {
            int cur = dev;
            (void)hipGetDevice(&cur);
            if (cur != dev) (void)hipSetDevice(dev);
            if (auto pool = wp.lock()) pool->give(*b);
            if (cur != dev) (void)hipSetDevice(cur);
        }
        b->release();
        delete b;
    });
    return frame_reserve(a, *store, bytes);
}

// PreparedFrame::new (preprocess.rs:137-153) for a batch of frames with conditioning
// (background, blur, local normalisation): the same kernels as prepare_frame, each launched
// once per chunk of frames (frames in blockIdx.y), so that the per-frame kernels are not
// latency-bound launches of their own.  Base images, conditioned images and tables live in
// storage the frames share; the blur and normalisation temporaries are sized for one chunk.
static piv_status condition_frames_batch(piv_analyzer *a, piv_frame *const *fs, int count,
                                         const float *bg_dev, long raw_frame_bytes, bool mask)
{
    constexpr int CHUNK = 32;
    const piv_config &c = a->cfg;
    piv_frame *f0 = fs[0];
    const int H = f0->H, W = f0->W;
    const size_t n = (size_t)H * W;
    const int64_t cells = sat_table_elems(H, W);
    const bool blur = c.smooth_sigma > 0.0f, norm = c.normalize_radius > 0;
    // as prepare_frame_impl: the base image is materialised where something reads it, the
    // blur reads the samples itself
    const bool fused = blur && blur_fusable(a->taps_r);
    const bool mat_base = bg_dev != nullptr || !fused;
    const bool has_image = blur || norm;
    const size_t img_bytes = align_up((size_t)count * n * sizeof(float), 256);
    const size_t base_bytes = mat_base ? img_bytes : 0, image_bytes = has_image ? img_bytes : 0;
    const size_t sat_bytes = mask ? (size_t)count * cells * 2 * sizeof(double) : 0;
    std::shared_ptr<DevBuf> store;
    PIV_TRY(batch_storage(a, base_bytes + image_bytes + sat_bytes, store));
    float *base = mat_base ? store->as<float>() : nullptr;
    float *image = has_image ? reinterpret_cast<float *>(store->as<char>() + base_bytes) : nullptr;
    double *sat = mask ? reinterpret_cast<double *>(store->as<char>() + base_bytes + image_bytes) + SAT_LEAD
                       : nullptr;
    const int k_max = std::min(CHUNK, count);
    if (blur && !fused) PIV_TRY(a->prep_tmp.reserve((size_t)k_max * n * sizeof(float)));
    if (blur && norm) PIV_TRY(a->prep_tmp2.reserve((size_t)k_max * n * sizeof(float)));
    if (norm) PIV_TRY(a->prep_sat.reserve((size_t)k_max * cells * 2 * sizeof(double)));
    if (mask || norm) PIV_TRY(a->sat_carry.reserve(sat_carry_bytes(H, W, k_max)));
    hipStream_t st = a->stream;
    double *carry = a->sat_carry.as<double>();
    for (int i0 = 0; i0 < count; i0 += CHUNK) {
        const int k = std::min(CHUNK, count - i0);
        const void *raw = static_cast<const char *>(f0->raw_ptr) + (long)i0 * raw_frame_bytes;
        float *b0 = mat_base ? base + (size_t)i0 * n : nullptr;
        float *img0 = has_image ? image + (size_t)i0 * n : nullptr;
        const float *cur = b0;
        if (fused) {
            float *dst = norm ? a->prep_tmp2.as<float>() : img0;
            HIP_TRY(launch_blur_fused(f0->in_fmt, raw, f0->raw_pitch, bg_dev, H, W, a->d_taps.as<float>(),
                                      a->taps_r, b0, dst, st, k, raw_frame_bytes));
            cur = dst;
        } else {
            HIP_TRY(launch_base(f0->in_fmt, raw, f0->raw_pitch, bg_dev, H, W, b0, st, k, raw_frame_bytes));
            if (blur) {
                float *dst = norm ? a->prep_tmp2.as<float>() : img0;
                HIP_TRY(launch_blur(cur, H, W, a->d_taps.as<float>(), a->taps_r, a->prep_tmp.as<float>(), dst,
                                    st, k));
                cur = dst;
            }
        }
        if (mask) {
            double *s0 = sat + (size_t)i0 * 2 * cells;
            if (mat_base) HIP_TRY(launch_sat(b0, H, W, s0, s0 + cells, carry, st, k, 2 * cells));
            else HIP_TRY(launch_sat_raw(f0->in_fmt, raw, f0->raw_pitch, raw_frame_bytes, H, W, s0, s0 + cells,
                                        carry, st, k, 2 * cells));
        }
        if (norm) {
            double *ns = a->prep_sat.as<double>() + SAT_LEAD;
            HIP_TRY(launch_sat(cur, H, W, ns, ns + cells, carry, st, k, 2 * cells));
            HIP_TRY(launch_local_norm(cur, H, W, (int)c.normalize_radius, c.normalize_eps, ns, ns + cells,
                                      img0, st, k, 2 * cells));
        }
    }
    for (int i = 0; i < count; ++i) {
        piv_frame *f = fs[i];
        f->has_bg = bg_dev != nullptr;
        f->cells = cells;
        f->base_ptr = mat_base ? base + (size_t)i * n : nullptr;
        f->has_sat = mask;
        f->screen_t = 0;
        f->sat_ptr = mask ? sat + (size_t)i * 2 * cells : nullptr;
        f->fmt = FMT_F32;
        // with only a background subtracted the base image is the correlation's image
        f->img = has_image ? image + (size_t)i * n : base + (size_t)i * n;
        f->pitch = W;
        f->batch = store;
    }
    return PIV_OK;
}

// Preparation of `count` frames of one size and format (frames raw_frame_bytes apart): every
// kernel is launched once for the whole set (per chunk when conditioning), into storage the
// frames share.
static piv_status prepare_frames_batch_impl(piv_analyzer *a, piv_frame *const *fs, int count,
                                            const float *bg_dev, long raw_frame_bytes)
{
    if (count <= 0) return PIV_OK;
    const piv_config &c = a->cfg;
    piv_frame *f0 = fs[0];
    const bool cond = c.smooth_sigma > 0.0f || c.normalize_radius > 0 || bg_dev != nullptr;
    const bool mask = mask_needed(c, f0->in_fmt, bg_dev != nullptr);
    const int H = f0->H, W = f0->W;
    const int64_t cells = sat_table_elems(H, W);
    if (cond) return condition_frames_batch(a, fs, count, bg_dev, raw_frame_bytes, mask);
    std::shared_ptr<DevBuf> store;
    if (mask) {
        // one pass over the raw samples per frame; no base image is kept (piv_frame_download
        // recomputes it).  The tables go into a block from the analyzer's batch pool when one
        // of a fitting size is free.
        const size_t need = (size_t)count * cells * 2 * sizeof(double);
        PIV_TRY(batch_storage(a, need, store));
        double *sat = store->as<double>() + SAT_LEAD;  // entry (0, 0) of frame 0's sums
        const int t = screen_tile(a, f0->in_fmt, false, H, W);
        PIV_TRY(a->sat_carry.reserve(sat_carry_bytes(H, W, count)));
        HIP_TRY(launch_sat_raw(f0->in_fmt, f0->raw_ptr, f0->raw_pitch, raw_frame_bytes, H, W, sat,
                               sat + cells, a->sat_carry.as<double>(), a->stream, count, 2 * cells, t));
        for (int i = 0; i < count; ++i) {
            piv_frame *f = fs[i];
            f->base_ptr = nullptr;
            f->sat_ptr = sat + (size_t)i * 2 * cells;
            f->has_sat = true;
            f->screen_t = t;
            f->batch = store;
        }
    }
    for (int i = 0; i < count; ++i) {
        piv_frame *f = fs[i];
        f->has_bg = false;
        f->cells = cells;
        f->fmt = f->in_fmt;
        f->img = f->raw_ptr;
        f->pitch = f->raw_pitch;
    }
    return PIV_OK;
}

piv_status plan_passes(piv_analyzer *a, int H, int W, std::vector<PassPlan> &plans)
{
    plans.clear();
    for (int n : a->passes) {
        PassPlan p{};
        PIV_TRY(grid_new((uint32_t)H, (uint32_t)W, (uint32_t)n, a->cfg.overlap, p.grid));
        p.n = n;
        p.cells = (size_t)p.grid.nx * p.grid.ny;
        plans.push_back(p);
    }
    return PIV_OK;
}

// Output targets of the last pass (device pointers, [npairs][cells]); null = workspace.
struct BatchOut {
    float *dx = nullptr, *dy = nullptr, *ppr = nullptr;
    uint8_t *flags = nullptr;
};

struct FieldPtrs {
    float *dx, *dy, *ppr;
    uint8_t *flags;
};

// Runs every pass for npairs pairs.  Returns device pointers of each pass's raw and validated
// fields (in the workspace unless `out` redirects the final validated field).
static piv_status run_batch_impl(piv_analyzer *a, const piv_frame *const *fa, const piv_frame *const *fb,
                                 uint32_t npairs, const BatchOut &out, bool traced,
                                 std::vector<PassPlan> &plans, std::vector<FieldPtrs> &raw,
                                 std::vector<FieldPtrs> &val, std::vector<float *> &origins,
                                 std::vector<int32_t *> &peaks, uint32_t **summaries_dev)
{
    const int H = fa[0]->H, W = fa[0]->W;
    for (uint32_t i = 0; i < npairs; ++i) {
        if (!fa[i] || !fb[i]) return fail(PIV_ERR_INVALID_ARGUMENT, "null frame");
        if (fa[i]->H != fb[i]->H || fa[i]->W != fb[i]->W)
            return fail(PIV_ERR_SHAPE_MISMATCH, "frames have different shapes: (%d, %d) vs (%d, %d)",
                        fa[i]->H, fa[i]->W, fb[i]->H, fb[i]->W);
        if (fa[i]->H != H || fa[i]->W != W)
            return fail(PIV_ERR_SHAPE_MISMATCH, "all pairs of a batch must share one frame size");
        if (fa[i]->fmt != fa[0]->fmt || fb[i]->fmt != fa[0]->fmt)
            return fail(PIV_ERR_INVALID_ARGUMENT, "all frames of a batch must share one sample format");
        if (fa[i]->device != a->device || fb[i]->device != a->device)
            return fail(PIV_ERR_INVALID_ARGUMENT, "frame belongs to another device");
    }
    PIV_TRY(plan_passes(a, H, W, plans));
    const size_t np = plans.size();
    {
        int nmin = 1 << 30;
        for (const PassPlan &pl : plans) nmin = std::min(nmin, pl.n);
        PIV_TRY(screen_for(a, fa, npairs, nmin, a->stream));
        PIV_TRY(screen_for(a, fb, npairs, nmin, a->stream));
    }
    // ---- workspace layout ----
    size_t off = 0;
    std::vector<size_t> o_raw(np), o_val(np), o_snap(np), o_org(np), o_sig(np), o_rec(np), o_pk(np);
    std::vector<uint8_t *> snap(np), sigs(np);
    std::vector<PeakRec *> recs(np);
    std::vector<float *> orgs(np);
    for (size_t i = 0; i < np; ++i) {
        const size_t c = plans[i].cells * npairs;
        o_raw[i] = off;
        off = align_up(off + c * (3 * sizeof(float) + 1), 256);
        o_val[i] = off;
        off = align_up(off + c * (3 * sizeof(float) + 1), 256);
        o_snap[i] = off;
        off = align_up(off + c, 256);
        o_org[i] = off;
        off = align_up(off + c * 4 * sizeof(float), 256);
        o_sig[i] = off;
        off = align_up(off + c, 256);
        o_rec[i] = off;
        off = align_up(off + c * sizeof(PeakRec), 256);
        if (traced) {
            o_pk[i] = off;
            off = align_up(off + c * sizeof(int32_t), 256);
        }
    }
    PIV_TRY(a->fields.reserve(off));
    char *base = a->fields.as<char>();
    raw.assign(np, {});
    val.assign(np, {});
    origins.assign(np, nullptr);
    peaks.assign(np, nullptr);
    for (size_t i = 0; i < np; ++i) {
        const size_t c = plans[i].cells * npairs;
        auto carve = [&](size_t o) {
            FieldPtrs f;
            f.dx = reinterpret_cast<float *>(base + o);
            f.dy = f.dx + c;
            f.ppr = f.dy + c;
            f.flags = reinterpret_cast<uint8_t *>(f.ppr + c);
            return f;
        };
        raw[i] = carve(o_raw[i]);
        val[i] = carve(o_val[i]);
        snap[i] = reinterpret_cast<uint8_t *>(base + o_snap[i]);
        orgs[i] = reinterpret_cast<float *>(base + o_org[i]);
        sigs[i] = reinterpret_cast<uint8_t *>(base + o_sig[i]);
        recs[i] = reinterpret_cast<PeakRec *>(base + o_rec[i]);
        if (traced) {
            origins[i] = orgs[i];
            peaks[i] = reinterpret_cast<int32_t *>(base + o_pk[i]);
        }
    }
    if (out.dx) {
        FieldPtrs &l = val[np - 1];
        l.dx = out.dx;
        l.dy = out.dy;
        l.ppr = out.ppr;
        l.flags = out.flags;
    }
    // ---- descriptors ----
    size_t dsz = 0;
    std::vector<size_t> o_pp(np), o_vp(np);
    for (size_t i = 0; i < np; ++i) {
        o_pp[i] = dsz;
        dsz = align_up(dsz + npairs * sizeof(PassPair), 64);
        o_vp[i] = dsz;
        dsz = align_up(dsz + npairs * sizeof(ValidatePair), 64);
    }
    PIV_TRY(a->desc.reserve(dsz));
    PIV_TRY(a->h_desc.reserve(dsz));
    PIV_TRY(a->summaries.reserve(np * npairs * 5 * sizeof(uint32_t)));
    std::vector<char> hbuf(dsz, 0);
    char *hd = hbuf.data();
    char *dd = a->desc.as<char>();
    uint32_t *sum_dev = a->summaries.as<uint32_t>();
    const float maxdisp = (float)a->cfg.windows[0] / 2.0f;
    for (size_t i = 0; i < np; ++i) {
        const PassPlan &pl = plans[i];
        auto *pp = reinterpret_cast<PassPair *>(hd + o_pp[i]);
        auto *vp = reinterpret_cast<ValidatePair *>(hd + o_vp[i]);
        for (uint32_t k = 0; k < npairs; ++k) {
            PassPair d{};
            const piv_frame *A = fa[k], *B = fb[k];
            d.img_a = A->img;
            d.img_b = B->img;
            d.use_mask = (A->has_sat && B->has_sat) ? 1 : 0;
            d.sat_a = A->has_sat ? A->sat_ptr : nullptr;
            d.sat_b = B->has_sat ? B->sat_ptr : nullptr;
            d.sat_cells = A->cells;
            d.screened = screened_bits(A, B);
            if (i > 0) {
                d.prev_dx = val[i - 1].dx + (size_t)k * plans[i - 1].cells;
                d.prev_dy = val[i - 1].dy + (size_t)k * plans[i - 1].cells;
                d.prev = plans[i - 1].grid;
                d.has_prev = 1;
            }
            const size_t co = (size_t)k * pl.cells;
            d.out_dx = raw[i].dx + co;
            d.out_dy = raw[i].dy + co;
            d.out_ppr = raw[i].ppr + co;
            d.out_flags = raw[i].flags + co;
            d.origins = orgs[i] + 4 * co;
            d.sig = sigs[i] + co;
            d.peaks = recs[i] + co;
            d.out_peak = traced ? peaks[i] + co : nullptr;
            d.H = H;
            d.W = W;
            d.pitch = A->pitch;
            d.fmt = A->fmt;
            d.grid = pl.grid;
            d.min_mean = a->cfg.min_mean;
            d.min_std = a->cfg.min_std;
            d.deform = (int32_t)a->cfg.deformation;
            if (B->pitch != A->pitch)
                return fail(PIV_ERR_INVALID_ARGUMENT, "frames of a pair must share a row pitch");
            pp[k] = d;
            ValidatePair v{};
            v.raw_dx = d.out_dx;
            v.raw_dy = d.out_dy;
            v.raw_ppr = d.out_ppr;
            v.raw_flags = d.out_flags;
            v.dx = val[i].dx + co;
            v.dy = val[i].dy + co;
            v.ppr = val[i].ppr + co;
            v.flags = val[i].flags + co;
            v.snapshot = snap[i] + co;
            v.summary = sum_dev + (i * npairs + k) * 5;
            v.nx = pl.grid.nx;
            v.ny = pl.grid.ny;
            v.ppr_threshold = a->cfg.ppr_threshold;
            v.median_threshold = a->cfg.median_threshold;
            v.median_eps = a->cfg.median_eps;
            v.max_displacement = maxdisp;
            v.replace = a->cfg.replace ? 1 : 0;
            vp[k] = v;
        }
    }
    hipStream_t st = a->stream;
    // Replay cache: a call whose descriptors (frames, outputs, workspace, grids) are identical
    // to the previous call's reuses the descriptors already on the device and replays the
    // captured launch sequence as one graph (kernel nodes only).  Profiled calls always launch
    // eagerly so every kernel can be bracketed by events.
    // Lanes: groups of at least 4 pairs each; profiled calls run one chain on `st` so that
    // every kernel is bracketed by events on one stream.
    const uint32_t lanes =
        a->prof_on ? 1u : std::max(1u, std::min((uint32_t)a->lanes, npairs / 4));
    std::vector<char> sig(hbuf);
    {
        const uintptr_t extra[7] = {(uintptr_t)dd, (uintptr_t)dsz, (uintptr_t)np, (uintptr_t)npairs,
                                    (uintptr_t)fa[0]->fmt, (uintptr_t)st, (uintptr_t)lanes};
        sig.insert(sig.end(), reinterpret_cast<const char *>(extra),
                   reinterpret_cast<const char *>(extra) + sizeof extra);
        for (const PassPlan &pl : plans) {
            const uintptr_t pv[4] = {(uintptr_t)pl.n, (uintptr_t)pl.cells,
                                     (uintptr_t)variant_of(a, pl.n, fa[0]->fmt),
                                     (uintptr_t)mf_precision(a, pl.n)};
            sig.insert(sig.end(), reinterpret_cast<const char *>(pv),
                       reinterpret_cast<const char *>(pv) + sizeof pv);
        }
    }
    // The pass chain of pairs [k0, k0 + cnt) on stream s.
    auto chain = [&](uint32_t k0, uint32_t cnt, hipStream_t s) -> piv_status {
        for (size_t i = 0; i < np; ++i) {
            const PassPlan &pl = plans[i];
            const uint32_t total = (uint32_t)(pl.cells * cnt);
            const double cbytes = (double)total * (2.0 * pl.n * pl.n * 4.0 + 49.0);
            const PassPair *ppd = reinterpret_cast<PassPair *>(dd + o_pp[i]) + k0;
            // placement reads the predictor field (8 B/vector of the previous grid, bilinear
            // taps mostly cached) and writes origins + mask (17 B/window)
            HIP_TRY(timed(a, "place", 0, (double)total * 17.0, total, [&] {
                return launch_place_pass(ppd, (int)cnt, (uint32_t)pl.cells, s);
            }));
            const int mp = mf_precision(a, pl.n);
            const int def = i > 0 ? (int)a->cfg.deformation : 0;
            HIP_TRY(timed(a, mp ? "correlate_mf" : "correlate", pl.n, cbytes, total, [&] {
                return mp ? launch_correlate_mf(pl.n, fa[0]->fmt, mp, ppd, (int)cnt,
                                                (uint32_t)pl.cells, s, def)
                          : launch_correlate(pl.n, fa[0]->fmt, variant_of(a, pl.n, fa[0]->fmt),
                                             ppd, (int)cnt, (uint32_t)pl.cells, s, def);
            }));
            // the finishing kernel reads peak records + origins (48 B) and writes 13 B/window
            HIP_TRY(timed(a, "finish", 0, (double)total * 61.0, total, [&] {
                return launch_peak_finish(ppd, (int)cnt, (uint32_t)pl.cells, s, mp != 0);
            }));
            // validation reads the raw field (13 B/vector) and writes the validated one (13 B)
            HIP_TRY(timed(a, "validate", 0, (double)total * 26.0, total, [&] {
                return launch_validate(reinterpret_cast<ValidatePair *>(dd + o_vp[i]) + k0,
                                       (int)cnt, (uint32_t)pl.cells, s);
            }));
        }
        return PIV_OK;
    };
    auto enqueue = [&]() -> piv_status {
        HIP_TRY(launch_zero_u32(sum_dev, (uint32_t)(np * npairs * 5), st));
        if (lanes == 1) return chain(0, npairs, st);
        // fork: every helper stream starts after the work already on `st`
        HIP_TRY(hipEventRecord(a->ev_fork, st));
        for (uint32_t l = 1; l < lanes; ++l) HIP_TRY(hipStreamWaitEvent(a->side[l - 1], a->ev_fork, 0));
        for (uint32_t l = lanes; l-- > 0;) {
            const uint32_t k0 = (uint32_t)((uint64_t)npairs * l / lanes);
            const uint32_t k1 = (uint32_t)((uint64_t)npairs * (l + 1) / lanes);
            PIV_TRY(chain(k0, k1 - k0, l ? a->side[l - 1] : st));
        }
        // join: `st` continues only after every helper stream's chain
        for (uint32_t l = 1; l < lanes; ++l) {
            HIP_TRY(hipEventRecord(a->ev_join[l - 1], a->side[l - 1]));
            HIP_TRY(hipStreamWaitEvent(st, a->ev_join[l - 1], 0));
        }
        return PIV_OK;
    };
    // The call synchronises its stream before it rewrites the staging buffer, which a stream
    // under capture does not allow; repeated identical calls are replayed as a graph instead.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(PIV_ERR_UNSUPPORTED, "the analysis stream is being captured by the caller");
    const bool same = !a->prof_on && sig == a->last_sig;
    if (same && a->graph_exec) {
        HIP_TRY(hipGraphLaunch(a->graph_exec, st));
    } else if (same) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        piv_status cs = enqueue();
        hipError_t ce = hipStreamEndCapture(st, &graph);
        if (cs != PIV_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return cs;
        }
        HIP_TRY(ce);
        hipError_t ie = hipGraphInstantiate(&a->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ie);
        HIP_TRY(hipGraphLaunch(a->graph_exec, st));
    } else {
        if (a->graph_exec) {
            (void)hipGraphExecDestroy(a->graph_exec);
            a->graph_exec = nullptr;
        }
        // the pinned staging buffer may still be read by a previous call's copy: wait for it
        HIP_TRY(hipStreamSynchronize(st));
        prof_harvest(a);
        std::memcpy(a->h_desc.p, hbuf.data(), dsz);
        HIP_TRY(hipMemcpyAsync(dd, a->h_desc.p, dsz, hipMemcpyHostToDevice, st));
        PIV_TRY(enqueue());
        a->last_sig.swap(sig);
    }
    *summaries_dev = sum_dev;
    return PIV_OK;
}

// The entry points below enqueue a frame operation and then record where on the stream its
// frames were last used, also after a failure part-way (work may be in flight).

piv_status prepare_frame(piv_analyzer *a, piv_frame *f, const float *bg_dev)
{
    const piv_status s = prepare_frame_impl(a, f, bg_dev);
    const piv_frame *used = f;
    const piv_status m = mark_use(a, a->stream, &used, 1);
    return s != PIV_OK ? s : m;
}

piv_status prepare_frames_list(piv_analyzer *a, piv_frame *const *fs, int count, const float *bg_dev)
{
    if (count <= 0) return PIV_OK;
    const piv_status s = prepare_frames_list_impl(a, fs, count, bg_dev);
    const piv_status m = mark_use(a, a->stream, fs, (size_t)count);
    return s != PIV_OK ? s : m;
}

piv_status prepare_frames_batch(piv_analyzer *a, piv_frame *const *fs, int count,
                                const float *bg_dev, long raw_frame_bytes)
{
    if (count <= 0) return PIV_OK;
    const piv_status s = prepare_frames_batch_impl(a, fs, count, bg_dev, raw_frame_bytes);
    const piv_status m = mark_use(a, a->stream, fs, (size_t)count);
    return s != PIV_OK ? s : m;
}

piv_status run_batch(piv_analyzer *a, const piv_frame *const *fa, const piv_frame *const *fb,
                     uint32_t npairs, const BatchOut &out, bool traced,
                     std::vector<PassPlan> &plans, std::vector<FieldPtrs> &raw,
                     std::vector<FieldPtrs> &val, std::vector<float *> &origins,
                     std::vector<int32_t *> &peaks, uint32_t **summaries_dev)
{
    if (npairs == 0) return fail(PIV_ERR_INVALID_ARGUMENT, "empty batch");
    const piv_status s = run_batch_impl(a, fa, fb, npairs, out, traced, plans, raw, val, origins,
                                        peaks, summaries_dev);
    piv_status m = mark_use(a, a->stream, fa, npairs);
    if (m == PIV_OK) m = mark_use(a, a->stream, fb, npairs);
    return s != PIV_OK ? s : m;
}

void to_summary(const uint32_t *c, piv_summary *s)
{
    s->no_signal = c[0];
    s->low_ppr = c[1];
    s->median_outliers = c[2];
    s->out_of_range = c[3];
    s->replaced = c[4];
}

piv_status frame_from_device(piv_analyzer *a, const void *pixels, int fmt, uint32_t h,
                             uint32_t w, size_t row_stride, bool owned, const float *bg_dev,
                             piv_frame **out)
{
    auto *f = new (std::nothrow) piv_frame();
    if (!f) return fail(PIV_ERR_OOM, "host allocation failed");
    f->device = a->device;
    f->pool = a->pool;
    f->H = (int)h;
    f->W = (int)w;
    f->in_fmt = fmt;
    f->raw_ptr = pixels;
    f->raw_pitch = (int)(row_stride / bytes_per_px(fmt));
    (void)owned;
    piv_status s = prepare_frame(a, f, bg_dev);
    if (s != PIV_OK) {
        (void)hipStreamSynchronize(a->stream);  // kernels of the failed preparation may still run
        delete f;
        return s;
    }
    *out = f;
    return PIV_OK;
}

// Helper streams (default priority) and fork / join events for `lanes` lanes, if missing.
piv_status ensure_lanes(piv_analyzer *a, int lanes)
{
    if (lanes > 1 && !a->ev_fork) HIP_TRY(hipEventCreateWithFlags(&a->ev_fork, hipEventDisableTiming));
    for (int i = 0; i < lanes - 1; ++i) {
        if (!a->side[i]) HIP_TRY(hipStreamCreateWithFlags(&a->side[i], hipStreamNonBlocking));
        if (!a->ev_join[i]) HIP_TRY(hipEventCreateWithFlags(&a->ev_join[i], hipEventDisableTiming));
    }
    return PIV_OK;
}

piv_status wait_frames(hipStream_t st, const piv_frame *const *fs, size_t n)
{
    const UseEvent *waited[8] = {};
    int nw = 0;
    for (size_t i = 0; i < n; ++i) {
        const UseEvent *ev = fs[i] ? fs[i]->last_use.get() : nullptr;
        if (!ev || !ev->e || ev->st == st) continue;
        if (std::find(waited, waited + nw, ev) != waited + nw) continue;
        HIP_TRY(hipStreamWaitEvent(st, ev->e, 0));
        if (nw < 8) waited[nw++] = ev;
    }
    return PIV_OK;
}

piv_status mark_use(piv_analyzer *a, hipStream_t st, const piv_frame *const *fs, size_t n)
{
    // on a stream the caller is capturing, a recorded event marks no point of execution: the
    // frames keep their previous last use
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return PIV_OK;
    std::shared_ptr<UseEvent> &ev = a->use_events[st];
    if (!ev) {
        auto e = std::make_shared<UseEvent>();
        HIP_TRY(hipEventCreateWithFlags(&e->e, hipEventDisableTiming));
        e->st = st;
        ev = e;
    }
    HIP_TRY(hipEventRecord(ev->e, st));
    for (size_t i = 0; i < n; ++i)
        if (fs[i]) fs[i]->last_use = ev;
    return PIV_OK;
}

piv_status check_pixels_args(const void *pixels, int fmt, uint32_t h, uint32_t w,
                             size_t row_stride)
{
    if (!pixels) return fail(PIV_ERR_INVALID_ARGUMENT, "pixels is null");
    if (fmt != FMT_U8 && fmt != FMT_U16 && fmt != FMT_F32)
        return fail(PIV_ERR_INVALID_ARGUMENT, "unknown pixel format %d", fmt);
    if (h == 0 || w == 0) return fail(PIV_ERR_INVALID_ARGUMENT, "empty frame");
    if (h > (1u << 15) || w > (1u << 15)) return fail(PIV_ERR_INVALID_ARGUMENT, "frame too large");
    const size_t bpp = (size_t)bytes_per_px(fmt);
    if (row_stride < w * bpp || row_stride % bpp != 0)
        return fail(PIV_ERR_INVALID_ARGUMENT, "row_stride %zu invalid for width %u", row_stride, w);
    return PIV_OK;
}

}  // namespace pivi

// =========================================================================================
// C ABI
// =========================================================================================
extern "C" {

const char *piv_last_error(void) { return g_err.c_str(); }

int piv_abi_version(void) { return PIV_ABI_VERSION; }

piv_status piv_device_count(int *count)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!count) return fail(PIV_ERR_INVALID_ARGUMENT, "count is null");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) n = 0;
        *count = n;
        return PIV_OK;
    });
}

void piv_config_default(piv_config *c)
{
    try {
        if (!c) return;
        std::memset(c, 0, sizeof *c);
        c->windows[0] = 96;
        c->windows[1] = 64;
        c->num_windows = 2;
        c->overlap = 0.5f;
        c->final_iterations = 2;
        c->min_mean = 0.0f;
        c->min_std = 0.0f;
        c->ppr_threshold = 1.3f;
        c->median_threshold = 2.0f;
        c->median_eps = 0.1f;
        c->replace = 1;
        c->smooth_sigma = 1.0f;
        c->normalize_radius = 4;
        c->normalize_eps = 0.02f;
    } catch (...) {
    }
}

piv_status piv_grid_new(uint32_t height, uint32_t width, uint32_t window, float overlap,
                        piv_grid *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!out) return fail(PIV_ERR_INVALID_ARGUMENT, "out is null");
        GridDev g{};
        PIV_TRY(grid_new(height, width, window, overlap, g));
        *out = to_pub(g);
        return PIV_OK;
    });
}

}  // extern "C"

namespace pivi {

// An analyzer; `batch_lanes` false (the serving streams' analyzers, whose batches the server
// launches itself) creates no helper streams and pins one lane.
piv_status analyzer_create(const piv_config *cfg, int device, piv_analyzer **out, bool batch_lanes)
{
    {
        if (!cfg || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        PIV_TRY(check_config(*cfg));
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            return fail(PIV_ERR_DEVICE, "no HIP device available");
        if (device < 0 || device >= ndev)
            return fail(PIV_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, ndev);
        PIV_TRY(set_device(device));
        // set-up of the allocation-free entry point's workspace address on this device; a
        // failure is reported by that entry point, not here
        (void)noalloc_init(device);
        auto *a = new (std::nothrow) piv_analyzer();
        if (!a) return fail(PIV_ERR_OOM, "host allocation failed");
        a->cfg = *cfg;
        a->device = device;
        for (uint32_t i = 0; i < cfg->num_windows; ++i) a->passes.push_back((int)cfg->windows[i]);
        for (uint32_t i = 1; i < cfg->final_iterations; ++i)
            a->passes.push_back((int)cfg->windows[cfg->num_windows - 1]);
        hipError_t e = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking);
        // batch lanes (PIV_BATCH_LANES = 1..4 pins the count, else the default of 2 until the
        // autotuner or the caller picks one): helper streams at the default priority for the
        // most lanes, and the events that fork them from and join them into `stream`
        if (const char *env = std::getenv("PIV_BATCH_LANES")) {
            const int v = std::atoi(env);
            if (v >= 1 && v <= piv_analyzer::MAX_LANES) {
                a->lanes = v;
                a->lanes_pinned = true;
            }
        }
        if (e != hipSuccess) {
            delete a;
            return fail(PIV_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
        }
        if (!batch_lanes) {
            a->lanes = 1;
            a->lanes_pinned = true;
        }
        // The autotuner's lane sweep is opt-in (PIV_TUNE_LANES=1): processes that had loaded
        // torch's HIP runtime first ended twice in a host segmentation fault with three or four
        // lanes in use, cause not found.
        if (const char *env = std::getenv("PIV_TUNE_LANES"))
            a->tune_lanes = batch_lanes && !a->lanes_pinned && std::atoi(env) == 1;
        // helper streams for the lane count in use (all counts when the autotuner may switch)
        if (const piv_status ls = ensure_lanes(a, a->tune_lanes ? piv_analyzer::MAX_LANES : a->lanes);
            ls != PIV_OK) {
            delete a;
            return ls;
        }
        if (cfg->smooth_sigma > 0.0f) {
            a->taps_r = gauss_taps(cfg->smooth_sigma, a->taps);
            piv_status s = a->d_taps.reserve(a->taps.size() * sizeof(float));
            if (s == PIV_OK &&
                hipMemcpy(a->d_taps.p, a->taps.data(), a->taps.size() * sizeof(float),
                          hipMemcpyHostToDevice) != hipSuccess)
                s = fail(PIV_ERR_DEVICE, "copying blur taps failed");
            if (s != PIV_OK) {
                delete a;
                return s;
            }
        }
        *out = a;
        return PIV_OK;
    }
}

}  // namespace pivi

extern "C" {

piv_status piv_analyzer_create(const piv_config *cfg, int device, piv_analyzer **out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        return pivi::analyzer_create(cfg, device, out, true);
    });
}

void piv_analyzer_destroy(piv_analyzer *a)
{
    try {
        if (!a) return;
        (void)hipSetDevice(a->device);
        if (a->stream) (void)hipStreamSynchronize(a->stream);
        delete a;
    } catch (...) {
    }
}

uint32_t piv_analyzer_num_passes(const piv_analyzer *a) { return a ? (uint32_t)a->passes.size() : 0; }

piv_status piv_output_grid(const piv_analyzer *a, uint32_t height, uint32_t width, piv_grid *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        return piv_grid_new(height, width, a->cfg.windows[a->cfg.num_windows - 1], a->cfg.overlap,
                            out);
    });
}

piv_status piv_frame_prepare(piv_analyzer *a, const void *pixels, piv_pixel_format fmt,
                             uint32_t height, uint32_t width, size_t row_stride,
                             const float *background, piv_frame **out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        PIV_TRY(check_pixels_args(pixels, fmt, height, width, row_stride));
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        auto *f = new (std::nothrow) piv_frame();
        if (!f) return fail(PIV_ERR_OOM, "host allocation failed");
        f->device = a->device;
        f->pool = a->pool;
        f->H = (int)height;
        f->W = (int)width;
        f->in_fmt = fmt;
        const size_t bpp = (size_t)bytes_per_px(fmt);
        piv_status s = frame_reserve(a, f->raw, (size_t)height * width * bpp);
        DevBuf bg;
        if (s == PIV_OK) {
            hipError_t e = hipMemcpy2DAsync(f->raw.p, width * bpp, pixels, row_stride, width * bpp,
                                            height, hipMemcpyHostToDevice, a->stream);
            if (e != hipSuccess) s = fail(PIV_ERR_DEVICE, "upload: %s", hipGetErrorString(e));
        }
        if (s == PIV_OK && background) {
            s = bg.reserve((size_t)height * width * sizeof(float));
            if (s == PIV_OK &&
                hipMemcpyAsync(bg.p, background, (size_t)height * width * sizeof(float),
                               hipMemcpyHostToDevice, a->stream) != hipSuccess)
                s = fail(PIV_ERR_DEVICE, "background upload failed");
        }
        if (s == PIV_OK) {
            f->raw_ptr = f->raw.p;
            f->raw_pitch = (int)width;
            s = prepare_frame(a, f, background ? bg.as<float>() : nullptr);
        }
        if (s == PIV_OK && hipStreamSynchronize(a->stream) != hipSuccess)
            s = fail(PIV_ERR_DEVICE, "frame preparation failed");
        bg.release();
        if (s != PIV_OK) {
            delete f;
            return s;
        }
        *out = f;
        return PIV_OK;
    });
}

piv_status piv_frame_prepare_device(piv_analyzer *a, const void *pixels, piv_pixel_format fmt,
                                    uint32_t height, uint32_t width, size_t row_stride,
                                    const float *background, piv_frame **out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        PIV_TRY(check_pixels_args(pixels, fmt, height, width, row_stride));
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        PIV_TRY(frame_from_device(a, pixels, fmt, height, width, row_stride, false, background, out));
        return PIV_OK;
    });
}

piv_status piv_frame_prepare_device_batch(piv_analyzer *a, const void *pixels,
                                          piv_pixel_format fmt, uint32_t count, uint32_t height,
                                          uint32_t width, size_t row_stride, size_t frame_stride,
                                          const float *background, piv_frame **out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (count == 0) return PIV_OK;
        PIV_TRY(check_pixels_args(pixels, fmt, height, width, row_stride));
        if (frame_stride < row_stride * height)
            return fail(PIV_ERR_INVALID_ARGUMENT, "frame_stride %zu < %zu", frame_stride, row_stride * height);
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        std::vector<piv_frame *> fs(count, nullptr);
        piv_status s = PIV_OK;
        for (uint32_t i = 0; i < count && s == PIV_OK; ++i) {
            fs[i] = new (std::nothrow) piv_frame();
            if (!fs[i]) {
                s = fail(PIV_ERR_OOM, "host allocation failed");
                break;
            }
            piv_frame *f = fs[i];
            f->device = a->device;
            f->pool = a->pool;
            f->H = (int)height;
            f->W = (int)width;
            f->in_fmt = (int)fmt;
            f->raw_ptr = static_cast<const char *>(pixels) + (size_t)i * frame_stride;
            f->raw_pitch = (int)(row_stride / bytes_per_px((int)fmt));
        }
        if (s == PIV_OK) s = prepare_frames_batch(a, fs.data(), (int)count, background, (long)frame_stride);
        if (s != PIV_OK) {
            // kernels of a failed preparation may still run: the frames' memory is freed after
            (void)hipStreamSynchronize(a->stream);
            for (piv_frame *f : fs) delete f;
            return s;
        }
        for (uint32_t i = 0; i < count; ++i) out[i] = fs[i];
        return PIV_OK;
    });
}

void piv_frame_release(piv_frame *f)
{
    try {
        if (!f) return;
        (void)hipSetDevice(f->device);
        // the frame's last user (its preparation or the last analysis that read it) is done
        // before its memory, or its batch's, is reused
        if (f->last_use && f->last_use->e) (void)hipEventSynchronize(f->last_use->e);
        if (f->raw.p || f->image.p || f->base.p || f->sat.p) {
            if (auto pool = f->pool.lock()) {
                pool->give(f->raw);
                pool->give(f->image);
                pool->give(f->base);
                pool->give(f->sat);
            }
        }
        delete f;
    } catch (...) {
    }
}

piv_status piv_frame_download(const piv_frame *f, float *image, float *base)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!f) return fail(PIV_ERR_INVALID_ARGUMENT, "null frame");
        PIV_TRY(set_device(f->device));
        // the preparation runs asynchronously on the analyzer's stream
        if (f->last_use && f->last_use->e) HIP_TRY(hipEventSynchronize(f->last_use->e));
        const size_t n = (size_t)f->H * f->W;
        DevBuf tmp;
        auto normalized = [&](float **p) -> piv_status {
            PIV_TRY(tmp.reserve(n * sizeof(float)));
            // the original samples, normalised (pitch-aware via base kernel without background)
            HIP_TRY(launch_base(f->in_fmt, f->raw_ptr, f->raw_pitch, nullptr, f->H, f->W,
                                tmp.as<float>(), nullptr));
            *p = tmp.as<float>();
            return PIV_OK;
        };
        if (image) {
            // a conditioned image (own or batch storage) is what the correlation reads; otherwise
            // the image is the normalised samples
            if (f->fmt == FMT_F32 && f->img && f->img != f->raw_ptr) {
                HIP_TRY(hipMemcpy(image, f->img, n * sizeof(float), hipMemcpyDeviceToHost));
            } else {
                float *p = nullptr;
                PIV_TRY(normalized(&p));
                HIP_TRY(hipMemcpy(image, p, n * sizeof(float), hipMemcpyDeviceToHost));
            }
        }
        if (base) {
            if (f->base_ptr) {
                HIP_TRY(hipMemcpy(base, f->base_ptr, n * sizeof(float), hipMemcpyDeviceToHost));
            } else {
                float *p = nullptr;
                PIV_TRY(normalized(&p));
                HIP_TRY(hipMemcpy(base, p, n * sizeof(float), hipMemcpyDeviceToHost));
            }
        }
        tmp.release();
        return PIV_OK;
    });
}

piv_status piv_frame_tables_built(const piv_frame *f, int *built)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!f || !built) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *built = 0;
        if (!f->has_sat) return PIV_OK;
        if (f->screen_t == 0) {
            *built = 1;
            return PIV_OK;
        }
        PIV_TRY(set_device(f->device));
        // the state word is written by the preparation (or a re-screen) on the analyzer's stream
        if (f->last_use && f->last_use->e) HIP_TRY(hipEventSynchronize(f->last_use->e));
        uint32_t state = 0;
        HIP_TRY(hipMemcpy(&state, sat_state_word(f->sat_ptr), sizeof state, hipMemcpyDeviceToHost));
        *built = state != 0 ? 1 : 0;
        return PIV_OK;
    });
}

static piv_status analyze_impl(piv_analyzer *a, const piv_frame *fa, const piv_frame *fb,
                               piv_field_out *out, piv_pass_trace *traces, uint32_t num_traces)
{
    if (!a || !fa || !fb || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
    if (fa->H != fb->H || fa->W != fb->W)
        return fail(PIV_ERR_SHAPE_MISMATCH, "frames have different shapes: (%d, %d) vs (%d, %d)",
                    fa->H, fa->W, fb->H, fb->W);
    if (fa->fmt != fb->fmt)
        return fail(PIV_ERR_INVALID_ARGUMENT, "frames of a pair must share a sample format");
    if (traces && num_traces != a->passes.size())
        return fail(PIV_ERR_INVALID_ARGUMENT, "need %zu traces", a->passes.size());
    std::lock_guard<std::mutex> lk(a->mu);
    PIV_TRY(set_device(a->device));
    std::vector<PassPlan> plans;
    PIV_TRY(plan_passes(a, fa->H, fa->W, plans));
    const size_t cells = plans.back().cells;
    if (out->capacity < cells || !out->dx || !out->dy || !out->peak_ratio || !out->flags)
        return fail(PIV_ERR_INVALID_ARGUMENT, "output arrays need %zu elements", cells);
    if (traces)
        for (size_t i = 0; i < plans.size(); ++i)
            if (traces[i].capacity < plans[i].cells)
                return fail(PIV_ERR_INVALID_ARGUMENT, "trace %zu needs %zu elements", i, plans[i].cells);
    std::vector<FieldPtrs> raw, val;
    std::vector<float *> org;
    std::vector<int32_t *> pk;
    uint32_t *sums = nullptr;
    const piv_frame *A[1] = {fa}, *B[1] = {fb};
    PIV_TRY(wait_frames(a->stream, A, 1));
    PIV_TRY(wait_frames(a->stream, B, 1));
    PIV_TRY(run_batch(a, A, B, 1, BatchOut{}, traces != nullptr, plans, raw, val, org, pk, &sums));
    hipStream_t st = a->stream;
    const size_t np = plans.size();
    PIV_TRY(a->h_summaries.reserve(np * 5 * sizeof(uint32_t)));
    HIP_TRY(hipMemcpyAsync(a->h_summaries.p, sums, np * 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    const FieldPtrs &l = val[np - 1];
    HIP_TRY(hipMemcpyAsync(out->dx, l.dx, cells * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out->dy, l.dy, cells * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out->peak_ratio, l.ppr, cells * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out->flags, l.flags, cells, hipMemcpyDeviceToHost, st));
    if (traces) {
        for (size_t i = 0; i < np; ++i) {
            piv_pass_trace &t = traces[i];
            const size_t c = plans[i].cells;
            auto cp = [&](void *dst, const void *src, size_t bytes) -> piv_status {
                if (dst) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
                return PIV_OK;
            };
            PIV_TRY(cp(t.origins, org[i], c * 16));
            PIV_TRY(cp(t.raw_dx, raw[i].dx, c * 4));
            PIV_TRY(cp(t.raw_dy, raw[i].dy, c * 4));
            PIV_TRY(cp(t.raw_ppr, raw[i].ppr, c * 4));
            PIV_TRY(cp(t.raw_flags, raw[i].flags, c));
            PIV_TRY(cp(t.peak_index, pk[i], c * 4));
            PIV_TRY(cp(t.dx, val[i].dx, c * 4));
            PIV_TRY(cp(t.dy, val[i].dy, c * 4));
            PIV_TRY(cp(t.ppr, val[i].ppr, c * 4));
            PIV_TRY(cp(t.flags, val[i].flags, c));
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    const uint32_t *hs = static_cast<const uint32_t *>(a->h_summaries.p);
    out->grid = to_pub(plans.back().grid);
    to_summary(hs + (np - 1) * 5, &out->summary);
    if (traces)
        for (size_t i = 0; i < np; ++i) {
            traces[i].grid = to_pub(plans[i].grid);
            to_summary(hs + i * 5, &traces[i].summary);
        }
    return PIV_OK;
}

piv_status piv_analyze_pair(piv_analyzer *a, const piv_frame *fa, const piv_frame *fb,
                            piv_field_out *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        return analyze_impl(a, fa, fb, out, nullptr, 0);
    });
}

piv_status piv_analyze_traced(piv_analyzer *a, const piv_frame *fa, const piv_frame *fb,
                              piv_field_out *out, piv_pass_trace *traces, uint32_t num_traces)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!traces) return fail(PIV_ERR_INVALID_ARGUMENT, "traces is null");
        return analyze_impl(a, fa, fb, out, traces, num_traces);
    });
}

piv_status piv_analyze_batch_device(piv_analyzer *a, const piv_frame *const *fa,
                                    const piv_frame *const *fb, uint32_t npairs, float *dx,
                                    float *dy, float *ppr, uint8_t *flags,
                                    piv_summary *summaries, void *stream, int async)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !fa || !fb || !dx || !dy || !ppr || !flags)
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (npairs == 0) return PIV_OK;
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        hipStream_t saved = a->stream;
        if (stream) a->stream = static_cast<hipStream_t>(stream);
        std::vector<PassPlan> plans;
        std::vector<FieldPtrs> raw, val;
        std::vector<float *> org;
        std::vector<int32_t *> pk;
        uint32_t *sums = nullptr;
        BatchOut bo;
        bo.dx = dx;
        bo.dy = dy;
        bo.ppr = ppr;
        bo.flags = flags;
        piv_status s = wait_frames(a->stream, fa, npairs);
        if (s == PIV_OK) s = wait_frames(a->stream, fb, npairs);
        if (s == PIV_OK) s = run_batch(a, fa, fb, npairs, bo, false, plans, raw, val, org, pk, &sums);
        if (s == PIV_OK && summaries) {
            const size_t np = plans.size();
            s = a->h_summaries.reserve(np * npairs * 5 * sizeof(uint32_t));
            if (s == PIV_OK &&
                (hipMemcpyAsync(a->h_summaries.p, sums + (np - 1) * npairs * 5,
                                npairs * 5 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                a->stream) != hipSuccess ||
                 hipStreamSynchronize(a->stream) != hipSuccess))
                s = fail(PIV_ERR_DEVICE, "summary download failed");
            if (s == PIV_OK) {
                const uint32_t *hs = static_cast<const uint32_t *>(a->h_summaries.p);
                for (uint32_t k = 0; k < npairs; ++k) to_summary(hs + k * 5, &summaries[k]);
            }
        } else if (s == PIV_OK && !async) {
            if (hipStreamSynchronize(a->stream) != hipSuccess) s = fail(PIV_ERR_DEVICE, "batch failed");
        }
        a->stream = saved;
        return s;
    });
}

piv_status piv_place_windows(piv_analyzer *a, const piv_grid *grid, uint32_t height,
                             uint32_t width, const piv_grid *prev, const float *prev_dx,
                             const float *prev_dy, float *origins)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !grid || !origins || (prev && (!prev_dx || !prev_dy)))
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        GridDev g{(int)grid->window, (int)grid->step, (int)grid->nx, (int)grid->ny, (int)grid->x0, (int)grid->y0};
        GridDev pg{};
        const size_t n = (size_t)g.nx * g.ny;
        size_t pn = 0;
        if (prev) {
            pg = GridDev{(int)prev->window, (int)prev->step, (int)prev->nx, (int)prev->ny, (int)prev->x0, (int)prev->y0};
            pn = (size_t)pg.nx * pg.ny;
        }
        PIV_TRY(a->stage_in.reserve(2 * pn * sizeof(float) + 16));
        PIV_TRY(a->stage_out.reserve(n * 4 * sizeof(float)));
        float *pdx = a->stage_in.as<float>(), *pdy = pdx + pn;
        hipStream_t st = a->stream;
        if (prev) {
            HIP_TRY(hipMemcpyAsync(pdx, prev_dx, pn * 4, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemcpyAsync(pdy, prev_dy, pn * 4, hipMemcpyHostToDevice, st));
        }
        HIP_TRY(launch_place(g, pg, prev ? 1 : 0, pdx, pdy, (int)height, (int)width,
                             a->stage_out.as<float>(), st));
        HIP_TRY(hipMemcpyAsync(origins, a->stage_out.p, n * 16, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return PIV_OK;
    });
}

piv_status piv_correlate_windows(piv_analyzer *a, const piv_frame *fa, const piv_frame *fb,
                                 uint32_t window, const float *origins, uint32_t count,
                                 float *dx, float *dy, float *ppr, uint8_t *flags,
                                 int32_t *peak_index)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !fa || !fb || !origins || !dx || !dy || !ppr || !flags)
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (fa->H != fb->H || fa->W != fb->W || fa->fmt != fb->fmt || fa->pitch != fb->pitch)
            return fail(PIV_ERR_SHAPE_MISMATCH, "frames differ in shape or format");
        if (!correlate_supported((int)window))
            return fail(PIV_ERR_UNSUPPORTED, "window size %u has no gfx950 kernel", window);
        if ((int)window > fa->W || (int)window > fa->H)
            return fail(PIV_ERR_IMAGE_TOO_SMALL, "window %u larger than the frame", window);
        if (count == 0) return PIV_OK;
        for (uint32_t k = 0; k < count; ++k) {
            // origins must keep the (n+1)^2 bilinear footprint inside the frame
            const float ax = origins[4 * k], ay = origins[4 * k + 1], bx = origins[4 * k + 2], by = origins[4 * k + 3];
            if (!(ax >= 0 && ay >= 0 && bx >= 0 && by >= 0 && ax + window <= fa->W && bx + window <= fa->W &&
                  ay + window <= fa->H && by + window <= fa->H))
                return fail(PIV_ERR_INVALID_ARGUMENT, "origin %u outside the frame", k);
        }
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        {
            const piv_frame *fs[2] = {fa, fb};
            PIV_TRY(wait_frames(a->stream, fs, 2));
        }
        const size_t o_pp = align_up(count * 16, 256);
        const size_t o_sig = o_pp + align_up(sizeof(PassPair), 256);
        const size_t o_rec = o_sig + align_up(count, 256);
        PIV_TRY(a->stage_in.reserve(o_rec + count * sizeof(PeakRec)));
        PIV_TRY(a->stage_out.reserve((size_t)count * 17 + 1024));
        char *din = a->stage_in.as<char>();
        float *d_org = reinterpret_cast<float *>(din);
        auto *d_pp = reinterpret_cast<PassPair *>(din + o_pp);
        float *o_dx = a->stage_out.as<float>(), *o_dy = o_dx + count, *o_ppr = o_dy + count;
        int32_t *o_pk = reinterpret_cast<int32_t *>(o_ppr + count);
        uint8_t *o_fl = reinterpret_cast<uint8_t *>(o_pk + count);
        PIV_TRY(screen_for(a, &fa, 1, (int)window, a->stream));
        PIV_TRY(screen_for(a, &fb, 1, (int)window, a->stream));
        PassPair d{};
        d.img_a = fa->img;
        d.img_b = fb->img;
        d.use_mask = (fa->has_sat && fb->has_sat) ? 1 : 0;
        d.sat_a = fa->has_sat ? fa->sat_ptr : nullptr;
        d.sat_b = fb->has_sat ? fb->sat_ptr : nullptr;
        d.sat_cells = fa->cells;
        d.screened = screened_bits(fa, fb);
        d.origins = d_org;
        d.origins_given = 1;
        d.sig = reinterpret_cast<uint8_t *>(din + o_sig);
        d.peaks = reinterpret_cast<PeakRec *>(din + o_rec);
        d.out_dx = o_dx;
        d.out_dy = o_dy;
        d.out_ppr = o_ppr;
        d.out_flags = o_fl;
        d.out_peak = o_pk;
        d.H = fa->H;
        d.W = fa->W;
        d.pitch = fa->pitch;
        d.fmt = fa->fmt;
        d.grid = GridDev{(int)window, 1, (int)count, 1, 0, 0};
        d.min_mean = a->cfg.min_mean;
        d.min_std = a->cfg.min_std;
        hipStream_t st = a->stream;
        HIP_TRY(hipMemcpyAsync(d_org, origins, count * 16, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_pp, &d, sizeof d, hipMemcpyHostToDevice, st));
        if (const int mp = mf_precision(a, (int)window)) {
            HIP_TRY(launch_place_pass(d_pp, 1, count, st));
            HIP_TRY(launch_correlate_mf((int)window, fa->fmt, mp, d_pp, 1, count, st));
            HIP_TRY(launch_peak_finish(d_pp, 1, count, st, true));
        } else {
            HIP_TRY(launch_pass((int)window, fa->fmt, variant_of(a, (int)window, fa->fmt), d_pp, 1, count, st));
        }
        HIP_TRY(hipMemcpyAsync(dx, o_dx, count * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(dy, o_dy, count * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(ppr, o_ppr, count * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(flags, o_fl, count, hipMemcpyDeviceToHost, st));
        if (peak_index) HIP_TRY(hipMemcpyAsync(peak_index, o_pk, count * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return PIV_OK;
    });
}

piv_status piv_validate_field(piv_analyzer *a, uint32_t ny, uint32_t nx, float *dx, float *dy,
                              const float *ppr, uint8_t *flags, float max_displacement,
                              piv_summary *summary)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !dx || !dy || !ppr || !flags) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (nx == 0 || ny == 0) return fail(PIV_ERR_INVALID_ARGUMENT, "empty field");
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        const size_t n = (size_t)nx * ny;
        const size_t fb = align_up(n * 13, 256);
        PIV_TRY(a->stage_in.reserve(2 * fb + align_up(n, 256) + 1024));
        char *b = a->stage_in.as<char>();
        float *r_dx = reinterpret_cast<float *>(b), *r_dy = r_dx + n, *r_ppr = r_dy + n;
        uint8_t *r_fl = reinterpret_cast<uint8_t *>(r_ppr + n);
        float *v_dx = reinterpret_cast<float *>(b + fb), *v_dy = v_dx + n, *v_ppr = v_dy + n;
        uint8_t *v_fl = reinterpret_cast<uint8_t *>(v_ppr + n);
        uint8_t *snapd = reinterpret_cast<uint8_t *>(b + 2 * fb);
        char *tail = b + 2 * fb + align_up(n, 256);
        auto *d_job = reinterpret_cast<ValidatePair *>(tail);
        auto *d_sum = reinterpret_cast<uint32_t *>(tail + 512);
        ValidatePair v{};
        v.raw_dx = r_dx;
        v.raw_dy = r_dy;
        v.raw_ppr = r_ppr;
        v.raw_flags = r_fl;
        v.dx = v_dx;
        v.dy = v_dy;
        v.ppr = v_ppr;
        v.flags = v_fl;
        v.snapshot = snapd;
        v.summary = d_sum;
        v.nx = (int)nx;
        v.ny = (int)ny;
        v.ppr_threshold = a->cfg.ppr_threshold;
        v.median_threshold = a->cfg.median_threshold;
        v.median_eps = a->cfg.median_eps;
        v.max_displacement = max_displacement;
        v.replace = a->cfg.replace ? 1 : 0;
        uint32_t hs[5];
        hipStream_t st = a->stream;
        HIP_TRY(hipMemcpyAsync(r_dx, dx, n * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(r_dy, dy, n * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(r_ppr, ppr, n * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(r_fl, flags, n, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_job, &v, sizeof v, hipMemcpyHostToDevice, st));
        HIP_TRY(launch_zero_u32(d_sum, 5, st));
        HIP_TRY(launch_validate(d_job, 1, (uint32_t)n, st));
        HIP_TRY(hipMemcpyAsync(dx, v_dx, n * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(dy, v_dy, n * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(flags, v_fl, n, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(hs, d_sum, sizeof hs, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (summary) to_summary(hs, summary);
        return PIV_OK;
    });
}

piv_status piv_background_create(piv_analyzer *a, piv_background_mode mode, uint32_t height,
                                 uint32_t width, piv_background **out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if ((int)mode != 0 && (int)mode != 1) return fail(PIV_ERR_INVALID_ARGUMENT, "unknown background mode");
        if (height == 0 || width == 0 || height > (1u << 15) || width > (1u << 15))
            return fail(PIV_ERR_INVALID_ARGUMENT, "bad frame size %ux%u", height, width);
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        auto *bg = new (std::nothrow) piv_background();
        if (!bg) return fail(PIV_ERR_OOM, "host allocation failed");
        bg->an = a;
        bg->mode = (int)mode;
        bg->H = (int)height;
        bg->W = (int)width;
        piv_status s = bg->acc.reserve((size_t)height * width * sizeof(float));
        if (s != PIV_OK) {
            delete bg;
            return s;
        }
        *out = bg;
        return PIV_OK;
    });
}

piv_status piv_background_add(piv_background *bg, const void *pixels, piv_pixel_format fmt,
                              size_t row_stride, int on_device)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!bg) return fail(PIV_ERR_INVALID_ARGUMENT, "null background");
        PIV_TRY(check_pixels_args(pixels, fmt, (uint32_t)bg->H, (uint32_t)bg->W, row_stride));
        if (bg->finished) return fail(PIV_ERR_INVALID_ARGUMENT, "background already finished");
        piv_analyzer *a = bg->an;
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        const size_t bpp = (size_t)bytes_per_px(fmt);
        const void *src = pixels;
        int pitch = (int)(row_stride / bpp);
        if (!on_device) {
            PIV_TRY(bg->stage.reserve((size_t)bg->H * bg->W * bpp));
            HIP_TRY(hipMemcpy2DAsync(bg->stage.p, bg->W * bpp, pixels, row_stride, bg->W * bpp, bg->H,
                                     hipMemcpyHostToDevice, a->stream));
            src = bg->stage.p;
            pitch = bg->W;
        }
        HIP_TRY(launch_bg_add(fmt, src, pitch, bg->H, bg->W, bg->mode, bg->count == 0 ? 1 : 0,
                              bg->acc.as<float>(), a->stream));
        HIP_TRY(hipStreamSynchronize(a->stream));
        bg->count += 1;
        return PIV_OK;
    });
}

piv_status piv_background_finish(piv_background *bg, float *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!bg) return fail(PIV_ERR_INVALID_ARGUMENT, "null background");
        if (bg->count == 0) return fail(PIV_ERR_INVALID_ARGUMENT, "no frames were added");
        piv_analyzer *a = bg->an;
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        const long n = (long)bg->H * bg->W;
        if (!bg->finished && bg->mode == 1)
            HIP_TRY(launch_bg_finish(bg->acc.as<float>(), n, (float)bg->count, a->stream));
        bg->finished = true;
        if (out) HIP_TRY(hipMemcpyAsync(out, bg->acc.p, n * sizeof(float), hipMemcpyDeviceToHost, a->stream));
        HIP_TRY(hipStreamSynchronize(a->stream));
        return PIV_OK;
    });
}

piv_status piv_background_device(const piv_background *bg, const float **device_ptr)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!bg || !device_ptr) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (!bg->finished) return fail(PIV_ERR_INVALID_ARGUMENT, "background not finished");
        *device_ptr = bg->acc.as<float>();
        return PIV_OK;
    });
}

void piv_background_destroy(piv_background *bg)
{
    try {
        if (!bg) return;
        (void)hipSetDevice(bg->an->device);
        delete bg;
    } catch (...) {
    }
}

piv_status piv_device_alloc(int device, size_t bytes, void **out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!out) return fail(PIV_ERR_INVALID_ARGUMENT, "out is null");
        *out = nullptr;
        PIV_TRY(set_device(device));
        HIP_TRY(hipMalloc(out, bytes ? bytes : 1));
        return PIV_OK;
    });
}

piv_status piv_device_free(int device, void *ptr)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!ptr) return PIV_OK;
        PIV_TRY(set_device(device));
        HIP_TRY(hipFree(ptr));
        return PIV_OK;
    });
}

piv_status piv_copy_to_device(int device, void *dst, const void *src, size_t bytes)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!dst || !src) return fail(PIV_ERR_INVALID_ARGUMENT, "null pointer");
        PIV_TRY(set_device(device));
        HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return PIV_OK;
    });
}

piv_status piv_copy_to_host(int device, void *dst, const void *src, size_t bytes)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!dst || !src) return fail(PIV_ERR_INVALID_ARGUMENT, "null pointer");
        PIV_TRY(set_device(device));
        HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return PIV_OK;
    });
}

piv_status piv_analyzer_synchronize(piv_analyzer *a)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a) return fail(PIV_ERR_INVALID_ARGUMENT, "null analyzer");
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        HIP_TRY(hipStreamSynchronize(a->stream));
        return PIV_OK;
    });
}

piv_status piv_profile_enable(piv_analyzer *a, int enable)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a) return fail(PIV_ERR_INVALID_ARGUMENT, "null analyzer");
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        HIP_TRY(hipStreamSynchronize(a->stream));
        prof_harvest(a);
        a->prof_stats.clear();
        a->prof_on = enable != 0;
        return PIV_OK;
    });
}

piv_status piv_profile_read(piv_analyzer *a, piv_kernel_stat *out, uint32_t capacity,
                            uint32_t *count)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !count) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        HIP_TRY(hipStreamSynchronize(a->stream));
        prof_harvest(a);
        uint32_t i = 0;
        for (auto &kv : a->prof_stats) {
            if (out && i < capacity) out[i] = kv.second;
            ++i;
        }
        *count = i;
        return PIV_OK;
    });
}

piv_status piv_analyzer_set_precision(piv_analyzer *a, piv_precision p)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if ((int)p < 0 || (int)p > 2) return fail(PIV_ERR_INVALID_ARGUMENT, "unknown precision %d", (int)p);
        std::lock_guard<std::mutex> lk(a->mu);
        a->precision = (int)p;
        return PIV_OK;
    });
}

uint32_t piv_kernel_variant_count(uint32_t window)
{
    return correlate_supported((int)window) ? (uint32_t)corr_variants((int)window).count : 0u;
}

piv_status piv_kernel_variant_set(piv_analyzer *a, uint32_t window, piv_pixel_format fmt,
                                  uint32_t variant)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (!correlate_supported((int)window))
            return fail(PIV_ERR_UNSUPPORTED, "window size %u has no gfx950 kernel", window);
        if ((int)fmt != FMT_U8 && (int)fmt != FMT_U16 && (int)fmt != FMT_F32)
            return fail(PIV_ERR_INVALID_ARGUMENT, "unknown pixel format %d", (int)fmt);
        if (variant >= (uint32_t)corr_variants((int)window).count)
            return fail(PIV_ERR_INVALID_ARGUMENT, "window %u has %d variants", window,
                        corr_variants((int)window).count);
        std::lock_guard<std::mutex> lk(a->mu);
        a->variants[{(int)window, (int)fmt}] = (int)variant;
        return PIV_OK;
    });
}

namespace pivi {

// Batch lanes for the caller's batch: the whole batch timed at each lane count by events on
// the analysis stream (which joins every lane) over `repeats` graph replays, after an eager
// and a capturing call; two sweeps, each count keeping its faster one.  Counts that the
// batch's size reduces to a smaller effective count (groups of at least 4 pairs) are skipped.
static piv_status tune_lanes(piv_analyzer *a, const piv_frame *const *fa,
                             const piv_frame *const *fb, uint32_t npairs, uint32_t repeats)
{
    constexpr int ML = piv_analyzer::MAX_LANES;
    const int saved = a->lanes;
    const bool saved_prof = a->prof_on;
    double ms[ML + 1];
    for (double &m : ms) m = 1e30;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        HIP_TRY(he);
    }
    piv_status s = PIV_OK;
    a->prof_on = false;
    for (int sweep = 0; sweep < 2 && s == PIV_OK; ++sweep)
    for (int l = 1; l <= ML && s == PIV_OK; ++l) {
        if (l > 1 && npairs / 4 < (uint32_t)l) continue;  // same chains as a smaller count
        s = ensure_lanes(a, l);
        if (s != PIV_OK) break;
        a->lanes = l;
        std::vector<PassPlan> pl;
        std::vector<FieldPtrs> raw, val;
        std::vector<float *> org;
        std::vector<int32_t *> pk;
        uint32_t *sums = nullptr;
        for (int k = 0; k < 2 && s == PIV_OK; ++k)  // eager, then captured
            s = run_batch(a, fa, fb, npairs, BatchOut{}, false, pl, raw, val, org, pk, &sums);
        if (s != PIV_OK) break;
        if (hipEventRecord(e0, a->stream) != hipSuccess) {
            s = fail(PIV_ERR_DEVICE, "autotune event failed");
            break;
        }
        for (uint32_t r = 0; r < repeats && s == PIV_OK; ++r)
            s = run_batch(a, fa, fb, npairs, BatchOut{}, false, pl, raw, val, org, pk, &sums);
        if (s != PIV_OK) break;
        float t = 0.0f;
        if (hipEventRecord(e1, a->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            s = fail(PIV_ERR_DEVICE, "autotune run failed");
            break;
        }
        ms[l] = std::min(ms[l], (double)t / (double)repeats);
    }
    a->prof_on = saved_prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s != PIV_OK) {
        a->lanes = saved;
        return s;
    }
    int best = 1;
    for (int l = 2; l <= ML; ++l)
        if (ms[l] < ms[best]) best = l;
    a->lanes = best;
    return PIV_OK;
}

}  // namespace pivi

piv_status piv_autotune(piv_analyzer *a, const piv_frame *const *fa, const piv_frame *const *fb,
                        uint32_t npairs, uint32_t repeats, piv_tune_entry *out, uint32_t capacity,
                        uint32_t *count)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !fa || !fb || !count || npairs == 0) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (repeats == 0) repeats = 1;
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        PIV_TRY(wait_frames(a->stream, fa, npairs));
        PIV_TRY(wait_frames(a->stream, fb, npairs));
        const int fmt = fa[0]->fmt;
        std::vector<PassPlan> plans;
        PIV_TRY(plan_passes(a, fa[0]->H, fa[0]->W, plans));
        std::vector<int> sizes;
        for (const PassPlan &pl : plans)
            if (std::find(sizes.begin(), sizes.end(), pl.n) == sizes.end()) sizes.push_back(pl.n);
        int maxv = 1;
        for (int n : sizes) maxv = std::max(maxv, corr_variants(n).count);
        HIP_TRY(hipStreamSynchronize(a->stream));
        prof_harvest(a);
        const bool saved_on = a->prof_on;
        const auto saved_stats = a->prof_stats;
        std::map<int, std::vector<double>> ms;
        for (int n : sizes) ms[n].assign(corr_variants(n).count, 0.0);
        piv_status s = PIV_OK;
        // Two sweeps over the variants, each variant keeping its faster sweep: kernel durations
        // fall over the first launches of a process as the GPU clocks up, which in a single
        // sweep would favour the variants timed last.
        for (int sweep = 0; sweep < 2 && s == PIV_OK; ++sweep)
        for (int v = 0; v < maxv && s == PIV_OK; ++v) {
            for (int n : sizes) a->variants[{n, fmt}] = std::min(v, corr_variants(n).count - 1);
            std::vector<PassPlan> pl;
            std::vector<FieldPtrs> raw, val;
            std::vector<float *> org;
            std::vector<int32_t *> pk;
            uint32_t *sums = nullptr;
            a->prof_on = false;  // one untimed run per variant (first launch of each kernel)
            s = run_batch(a, fa, fb, npairs, BatchOut{}, false, pl, raw, val, org, pk, &sums);
            if (s != PIV_OK) break;
            if (hipStreamSynchronize(a->stream) != hipSuccess) {
                s = fail(PIV_ERR_DEVICE, "autotune run failed");
                break;
            }
            a->prof_stats.clear();
            a->prof_on = true;
            for (uint32_t r = 0; r < repeats && s == PIV_OK; ++r)
                s = run_batch(a, fa, fb, npairs, BatchOut{}, false, pl, raw, val, org, pk, &sums);
            if (s != PIV_OK) break;
            if (hipStreamSynchronize(a->stream) != hipSuccess) {
                s = fail(PIV_ERR_DEVICE, "autotune run failed");
                break;
            }
            prof_harvest(a);
            for (int n : sizes) {
                if (v >= corr_variants(n).count) continue;
                auto it = a->prof_stats.find({"correlate", n});
                if (it != a->prof_stats.end() && it->second.launches) {
                    const double t = it->second.total_ms / (double)it->second.launches;
                    ms[n][v] = sweep == 0 ? t : std::min(ms[n][v], t);
                }
            }
        }
        a->prof_on = saved_on;
        a->prof_stats = saved_stats;
        if (s != PIV_OK) {
            for (int n : sizes) a->variants.erase({n, fmt});
            return s;
        }
        uint32_t i = 0;
        for (int n : sizes) {
            const VariantSet vs = corr_variants(n);
            int best = 0;
            for (int v = 1; v < vs.count; ++v)
                if (ms[n][v] < ms[n][best]) best = v;
            a->variants[{n, fmt}] = best;
            if (out && i < capacity) {
                piv_tune_entry &e = out[i];
                e = piv_tune_entry{};
                e.window = (uint32_t)n;
                e.fmt = fmt;
                e.chosen = (uint32_t)best;
                e.num_variants = (uint32_t)vs.count;
                for (int v = 0; v < vs.count && v < 4; ++v) {
                    e.occupancy[v] = (uint32_t)vs.occ[v];
                    e.ms[v] = (float)ms[n][v];
                }
            }
            ++i;
        }
        *count = i;
        if (a->tune_lanes) PIV_TRY(tune_lanes(a, fa, fb, npairs, repeats));
        return PIV_OK;
    });
}

piv_status piv_batch_lanes_set(piv_analyzer *a, uint32_t lanes)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a) return fail(PIV_ERR_INVALID_ARGUMENT, "null analyzer");
        if (lanes < 1 || lanes > (uint32_t)piv_analyzer::MAX_LANES)
            return fail(PIV_ERR_INVALID_ARGUMENT, "batch lanes %u outside 1..%d", lanes,
                        piv_analyzer::MAX_LANES);
        std::lock_guard<std::mutex> lk(a->mu);
        PIV_TRY(set_device(a->device));
        PIV_TRY(pivi::ensure_lanes(a, (int)lanes));
        a->lanes = (int)lanes;
        return PIV_OK;
    });
}

piv_status piv_batch_lanes_get(const piv_analyzer *a, uint32_t *lanes)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!a || !lanes) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *lanes = (uint32_t)a->lanes;
        return PIV_OK;
    });
}

piv_status piv_extract_plane_device(int device, const void *pixels, piv_pixel_format fmt,
                                    size_t count, uint32_t channels, uint32_t plane, void *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!pixels || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if ((int)fmt != FMT_U8 && (int)fmt != FMT_U16 && (int)fmt != FMT_F32)
            return fail(PIV_ERR_INVALID_ARGUMENT, "unknown pixel format %d", (int)fmt);
        if (channels < 1 || channels > 4 || plane >= channels)
            return fail(PIV_ERR_INVALID_ARGUMENT, "plane %u of %u channels", plane, channels);
        if (count == 0) return PIV_OK;
        PIV_TRY(set_device(device));
        hipStream_t st = nullptr;
        HIP_TRY(hipStreamCreate(&st));
        hipError_t e = launch_extract_plane((int)fmt, pixels, (long)count, (int)channels, (int)plane,
                                            out, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
        HIP_TRY(e);
        return PIV_OK;
    });
}

piv_status piv_normalize_pixels(int device, const void *pixels, piv_pixel_format fmt,
                                size_t count, float *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!pixels || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if ((int)fmt != FMT_U8 && (int)fmt != FMT_U16 && (int)fmt != FMT_F32)
            return fail(PIV_ERR_INVALID_ARGUMENT, "unknown pixel format %d", (int)fmt);
        PIV_TRY(set_device(device));
        DevBuf in, o;
        const size_t bpp = (size_t)bytes_per_px(fmt);
        PIV_TRY(in.reserve(count * bpp + 4));
        PIV_TRY(o.reserve(count * 4 + 4));
        HIP_TRY(hipMemcpy(in.p, pixels, count * bpp, hipMemcpyHostToDevice));
        HIP_TRY(launch_normalize(fmt, in.p, (long)count, o.as<float>(), nullptr));
        HIP_TRY(hipMemcpy(out, o.p, count * 4, hipMemcpyDeviceToHost));
        in.release();
        o.release();
        return PIV_OK;
    });
}

}  // extern "C"

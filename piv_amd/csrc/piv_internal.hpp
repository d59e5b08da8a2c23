// Internal declarations shared by the C-ABI implementation files (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/piv.h"
#include "piv_device.hpp"
#include "piv_kernels.hpp"
#include "variants.hpp"

namespace pivi {
using namespace piv;

// Sets the calling thread's piv_last_error() message and returns `st`.
piv_status fail(piv_status st, const char *fmt, ...);

// Runs the body of a C entry point: a C++ exception never crosses the ABI, it becomes
// PIV_ERR_OOM (std::bad_alloc) or PIV_ERR_DEVICE (anything else) with the message set.
template <class F>
piv_status guard(const char *entry, F &&body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return fail(PIV_ERR_OOM, "%s: host allocation failed", entry);
    } catch (const std::exception &e) {
        return fail(PIV_ERR_DEVICE, "%s: %s", entry, e.what());
    } catch (...) {
        return fail(PIV_ERR_DEVICE, "%s: unexpected exception", entry);
    }
}

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(e_ == hipErrorOutOfMemory ? PIV_ERR_OOM : PIV_ERR_DEVICE, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                    \
    } while (0)

#define PIV_TRY(expr)                                                                         \
    do {                                                                                      \
        piv_status s_ = (expr);                                                               \
        if (s_ != PIV_OK) return s_;                                                          \
    } while (0)

// Device buffer that grows on demand (contents not preserved).
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    piv_status reserve(size_t bytes)
    {
        if (bytes <= cap) return PIV_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = bytes < 256 ? 256 : bytes;
        HIP_TRY(hipMalloc(&p, want));
        cap = want;
        return PIV_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    piv_status reserve(size_t bytes)
    {
        if (bytes <= cap) return PIV_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        size_t want = bytes < 4096 ? 4096 : bytes;
        HIP_TRY(hipHostMalloc(&p, want, hipHostMallocDefault));
        cap = want;
        return PIV_OK;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// An event recorded on one stream after each analyzer call that used a frame there: the
// frame's memory may be reused once the event has completed.
struct UseEvent {
    hipEvent_t e = nullptr;
    hipStream_t st = nullptr;  // the stream the event is recorded on
    ~UseEvent()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

inline int bytes_per_px(int fmt) { return fmt == FMT_U8 ? 1 : (fmt == FMT_U16 ? 2 : 4); }

piv_status grid_new(uint32_t h, uint32_t w, uint32_t window, float overlap, GridDev &g);
piv_grid to_pub(const GridDev &g);
int gauss_taps(float sigma, std::vector<float> &k);

}  // namespace pivi

// The opaque handles of include/piv.h.
using pivi::DevBuf;
using pivi::HostBuf;
using namespace piv;

// Device memory of released frames and frame batches, kept for the next preparation: a
// video loop prepares frames of one size over and over, and hipMalloc / hipFree of their
// images and tables (a synchronising unmap per free) cost more than preparing them.  A
// block is reused for a request of at most twice less than its size; at most kLimit bytes
// are kept.  Shared with the frames and their batch storage, so it outlives the analyzer
// while they do.
struct BlockPool {
    static constexpr size_t kLimit = (size_t)16 << 30;
    std::mutex mu;
    std::vector<std::pair<void *, size_t>> blocks;  // (pointer, capacity)
    size_t held = 0;
    // moves a fitting block into `b` (whose own memory is released first)
    bool take(DevBuf &b, size_t bytes)
    {
        b.release();
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < blocks.size(); ++i)
            if (blocks[i].second >= bytes && blocks[i].second / 2 <= bytes) {
                b.p = blocks[i].first;
                b.cap = blocks[i].second;
                held -= b.cap;
                blocks[i] = blocks.back();
                blocks.pop_back();
                return true;
            }
        return false;
    }
    // takes `b`'s memory (b becomes empty), or leaves it to `b` when full; the caller has
    // made sure no device work still uses it
    void give(DevBuf &b)
    {
        if (!b.p) return;
        std::lock_guard<std::mutex> lk(mu);
        if (held + b.cap > kLimit) return;
        blocks.emplace_back(b.p, b.cap);
        held += b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    ~BlockPool()
    {
        for (auto &bl : blocks) (void)hipFree(bl.first);
    }
};

struct piv_frame {
    int device = 0;
    int H = 0, W = 0;
    int fmt = FMT_F32;         // format the correlation kernel reads `img` in
    const void *img = nullptr; // device
    int pitch = 0;             // pixels
    DevBuf raw;                // owned copy of the input pixels (host input)
    DevBuf image;              // owned conditioned f32 image
    DevBuf base;               // owned background-subtracted base (f32), when materialised
    DevBuf sat;                // owned f64 SAT of base (sum, then sum of squares)
    int64_t cells = 0;
    bool has_sat = false;
    // > 0: the tables were screened with this tile size (see screen_tile); windows of size n
    // may rely on the screen when screen_t <= n / 2 (re-screened with a smaller tile otherwise)
    mutable int screen_t = 0;
    float *base_ptr = nullptr;        // background-subtracted base: own `base` or batch storage
    const double *sat_ptr = nullptr;  // SAT (sum, then sum of squares): own `sat` or batch storage
    std::shared_ptr<DevBuf> batch;    // storage shared by the frames of one batch preparation
    std::weak_ptr<BlockPool> pool;    // where piv_frame_release returns the owned buffers
    int in_fmt = FMT_F32;      // original sample format
    const void *raw_ptr = nullptr;  // device pointer of the original samples
    int raw_pitch = 0;
    bool has_bg = false;
    // the last stream position that used the frame (preparation or analysis): release waits
    // for it instead of for the whole device
    mutable std::shared_ptr<pivi::UseEvent> last_use;
    ~piv_frame()
    {
        raw.release();
        image.release();
        base.release();
        sat.release();
    }
};

struct piv_background {
    piv_analyzer *an = nullptr;
    int mode = 0;
    int H = 0, W = 0;
    uint64_t count = 0;
    bool finished = false;
    DevBuf acc;   // estimate [H*W] f32
    DevBuf stage; // upload of host samples
    ~piv_background()
    {
        acc.release();
        stage.release();
    }
};

struct piv_analyzer {
    piv_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    // A batch's pairs are independent: run_batch splits them into `lanes` groups whose pass
    // chains run side by side, on `stream` and on the helper streams, forked from and joined
    // back into `stream` by events (so one group's short per-vector kernels overlap another
    // group's correlation).  Callers only ever see `stream`.
    static constexpr int MAX_LANES = 4;
    int lanes = 2;
    bool lanes_pinned = false;  // PIV_BATCH_LANES set: the autotuner keeps `lanes`
    bool tune_lanes = false;    // PIV_TUNE_LANES=1: piv_autotune also sweeps the lane count
    hipStream_t side[MAX_LANES - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[MAX_LANES - 1] = {};
    std::mutex mu;
    std::vector<int> passes;
    std::vector<float> taps;
    int taps_r = 0;
    DevBuf d_taps;
    // workspace
    DevBuf fields;       // per-pass raw + validated fields of the current call
    DevBuf desc;         // descriptors
    HostBuf h_desc;
    DevBuf summaries;    // u32 [npass][npairs][5]
    HostBuf h_summaries;
    DevBuf prep_tmp, prep_tmp2, prep_sat;  // conditioning scratch
    DevBuf sat_carry;                      // row carries of the SAT passes
    std::shared_ptr<BlockPool> pool = std::make_shared<BlockPool>();
    DevBuf stage_in, stage_out;            // stage-level entry points
    DevBuf post_part, post_keys, post_out; // flow statistics (piv_flow_stats_device)
    // kernel timing (piv_profile_*)
    struct ProfRec {
        std::string name;
        int window;
        double bytes;
        uint64_t units;
        hipEvent_t e0, e1;
    };
    // correlation kernel variant (variants.hpp) per (window size, sample format), set by
    // piv_autotune; absent means variant 0
    std::map<std::pair<int, int>, int> variants;
    // correlation precision (piv_analyzer_set_precision): 0 FP32 FFT, 1 fp16 / 2 bf16 DFT-GEMM
    int precision = 0;
    // one use event per stream the analyzer has run frames on (UseEvent)
    std::map<hipStream_t, std::shared_ptr<pivi::UseEvent>> use_events;
    // replay cache: the last call's descriptor bytes and the graph of its launch sequence
    std::vector<char> last_sig;
    hipGraphExec_t graph_exec = nullptr;
    bool prof_on = false;
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    std::map<std::pair<std::string, int>, piv_kernel_stat> prof_stats;
    ~piv_analyzer()
    {
        for (auto &r : prof_pending) {
            (void)hipEventDestroy(r.e0);
            (void)hipEventDestroy(r.e1);
        }
        for (auto e : prof_pool) (void)hipEventDestroy(e);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        fields.release();
        desc.release();
        h_desc.release();
        summaries.release();
        h_summaries.release();
        prep_tmp.release();
        prep_tmp2.release();
        prep_sat.release();
        sat_carry.release();
        stage_in.release();
        stage_out.release();
        post_part.release();
        post_keys.release();
        post_out.release();
        d_taps.release();
        for (int i = 0; i < MAX_LANES - 1; ++i) {
            if (side[i]) (void)hipStreamDestroy(side[i]);
            if (ev_join[i]) (void)hipEventDestroy(ev_join[i]);
        }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (stream) (void)hipStreamDestroy(stream);
    }
};


namespace pivi {

struct PassPlan {
    GridDev grid;
    int n;
    size_t cells;  // per pair
};

piv_status set_device(int dev);
int variant_of(const piv_analyzer *a, int n, int fmt);
// DFT-GEMM precision for a pass of window n (0: the FP32 FFT kernel)
int mf_precision(const piv_analyzer *a, int n);
piv_status check_config(const piv_config &c);
// whether the signal mask can change a result (else its tables are not built)
bool mask_needed(const piv_config &c, int in_fmt, bool has_bg);
// Tile size with which an unconditioned frame's tables may be screened for the analyzer's
// passes (the largest power of two <= n / 2 for its smallest window n), or 0 where the tables
// must be built unconditionally.
int screen_tile(const piv_analyzer *a, int in_fmt, bool cond, int H, int W);
// Re-screens frames whose screen tile is too large for windows of size n.
piv_status screen_for(piv_analyzer *a, const piv_frame *const *fs, size_t count, int n,
                      hipStream_t st);
// PassPair::screened bits of a pair of frames
inline int32_t screened_bits(const piv_frame *a, const piv_frame *b)
{
    return (a->has_sat && a->screen_t > 0 ? 1 : 0) | (b->has_sat && b->screen_t > 0 ? 2 : 0);
}
// device address of this GPU's workspace for piv_analyze_batch_noalloc, taken once per device
// (device `dev` current); later calls return the stored address or failure
piv_status noalloc_init(int dev);
piv_status prepare_frame(piv_analyzer *a, piv_frame *f, const float *bg_dev);
piv_status prepare_frames_list(piv_analyzer *a, piv_frame *const *fs, int count, const float *bg_dev);
// whether preparing a frame of this format launches device work (conditioning or tables)
bool prepare_launches(const piv_config &c, int in_fmt, bool has_bg);
piv_status prepare_frames_batch(piv_analyzer *a, piv_frame *const *fs, int count,
                                const float *bg_dev, long raw_frame_bytes);
piv_status plan_passes(piv_analyzer *a, int H, int W, std::vector<PassPlan> &plans);
void to_summary(const uint32_t *c, piv_summary *s);
piv_status frame_from_device(piv_analyzer *a, const void *pixels, int fmt, uint32_t h,
                             uint32_t w, size_t row_stride, bool owned, const float *bg_dev,
                             piv_frame **out);
piv_status check_pixels_args(const void *pixels, int fmt, uint32_t h, uint32_t w,
                             size_t row_stride);
// Records the analyzer's use event of stream `st` after the work enqueued there and makes it
// the last use of the `n` frames.
piv_status mark_use(piv_analyzer *a, hipStream_t st, const piv_frame *const *fs, size_t n);
// Creates the helper streams and events that `lanes` batch lanes use, if missing.
piv_status ensure_lanes(piv_analyzer *a, int lanes);
// piv_analyzer_create; batch_lanes false: no helper streams, one lane (serving analyzers).
piv_status analyzer_create(const piv_config *cfg, int device, piv_analyzer **out, bool batch_lanes);
// Makes `st` wait for the last use of frames last used on another stream: device frames are
// prepared asynchronously on their analyzer's stream, and a caller may analyse them on its own
// stream or through another analyzer.  The public analysis entry points call it; the serving
// scheduler orders its frames itself.
piv_status wait_frames(hipStream_t st, const piv_frame *const *fs, size_t n);

}  // namespace pivi

namespace pivi {
// Timing events from the analyzer's pool (piv_profile_*).
hipError_t prof_event(piv_analyzer *a, hipEvent_t *e);
// Fold completed timing events into the counters (the stream must be idle).
void prof_harvest(piv_analyzer *a);

// Launch `fn` bracketed by timing events when profiling is on.
template <class F>
hipError_t timed(piv_analyzer *a, const char *name, int window, double bytes, uint64_t units, F &&fn)
{
    if (!a->prof_on) return fn();
    piv_analyzer::ProfRec r{name, window, bytes, units, nullptr, nullptr};
    hipError_t e = prof_event(a, &r.e0);
    if (e == hipSuccess) e = prof_event(a, &r.e1);
    if (e == hipSuccess) e = hipEventRecord(r.e0, a->stream);
    if (e != hipSuccess) return e;
    e = fn();
    if (e != hipSuccess) return e;
    e = hipEventRecord(r.e1, a->stream);
    if (e == hipSuccess) a->prof_pending.push_back(r);
    return e;
}
}  // namespace pivi

// Serving: a device-resident request loop with continuous batching (include/piv.h,
// "serving").  The reference analyses one pair per call (piv-cli/src/analyze.rs:428-441);
// this loop keeps that per-pair result and changes only how pairs reach the device.
//
//   submit   -> admission against the HBM budget, the pixels go to a pinned staging slot and
//               are uploaded on the stream's own upload stream;
//   scheduler-> takes every admitted request (up to max_batch_pairs) as soon as a compute lane
//               is free, launches each request's conditioning and tables on its stream's upload
//               stream (all launches come from this one thread); the pass chain of the whole
//               batch runs as one launch per kernel and window size per pass step (requests of
//               different streams, frame sizes and pass schedules share launches), replayed
//               from a captured hipGraph when the batch has a launch structure seen before;
//   completer-> waits for a lane's batch, copies each result to its slot's host buffers and
//               hands it to piv_poll.
// Up to max_inflight batches are on the device at once (one HIP stream per lane), so a batch
// admitted while another runs overlaps with it.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <new>
#include <thread>

#include "piv_internal.hpp"

using namespace piv;
using namespace pivi;

namespace {

struct SrvStream;

// Device and host memory of one request of one stream (reused after release).
struct Slot {
    SrvStream *owner = nullptr;
    DevBuf dev;
    HostBuf host;
    size_t dev_bytes = 0;
    void *raw_a = nullptr, *raw_b = nullptr;        // device samples
    char *h_raw_a = nullptr, *h_raw_b = nullptr;    // pinned staging of the samples
    struct PassWs {
        float *raw;   // dx | dy | ppr (cells each), then flags
        float *val;
        uint8_t *snap, *sig;
        float *org;
        PeakRec *rec;
    };
    std::vector<PassWs> ws;
    float *h_dx = nullptr, *h_dy = nullptr, *h_ppr = nullptr;  // final field (host)
    uint8_t *h_flags = nullptr;
    piv_frame *fa = nullptr, *fb = nullptr;
    hipEvent_t uploaded = nullptr;
};

struct SrvStream {
    uint32_t id = 0;
    piv_analyzer *an = nullptr;  // config, pass plan, conditioning; its stream uploads
    int H = 0, W = 0, fmt = FMT_U8;
    DevBuf bg;
    bool has_bg = false;
    std::vector<PassPlan> plans;
    size_t slot_dev = 0, slot_host = 0;
    std::vector<Slot *> free_slots;
    int in_use = 0;        // slots held by admitted requests not yet released
    int in_flight = 0;     // admitted requests not yet finished (queued or on the device);
                           //   polled results the client has not released are not counted
    std::mutex upload_mu;  // orders uploads + conditioning on the analyzer's stream
};

struct Request {
    uint64_t id = 0;
    SrvStream *st = nullptr;
    Slot *slot = nullptr;
    double t_submit = 0, t_done = 0;
    piv_status status = PIV_OK;
    uint32_t batch_pairs = 0;
    piv_summary summary{};
    bool prepared = false;  // conditioning and tables launched by the scheduler
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf desc, sums, out;
    HostBuf h_desc, h_sums, h_out;
    std::vector<Request *> batch;
    std::vector<size_t> out_off;    // per request offset in `out`
    std::vector<size_t> sum_index;  // per request index of its final summary in `sums`
    size_t out_bytes = 0, sum_count = 0;
    std::map<std::vector<int64_t>, hipGraphExec_t> graphs;
    std::map<std::vector<int64_t>, int> seen;  // launch structures run eagerly so far
    bool busy = false;
    void drop_graphs()
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
        seen.clear();
    }
};

constexpr size_t kMaxGraphsPerLane = 64;
constexpr size_t kMaxQueued = 1 << 16;
// Host pairs of at least this many bytes are staged by the submitting thread together with
// the server's staging helpers (a single thread copies pageable memory at about 16 GB/s, so
// one 2048x2048 u16 pair alone would take a millisecond).
constexpr size_t kSplitStageBytes = (size_t)3 << 20;
constexpr int kStageHelpers = 3;

// One request's staging copy split into row chunks of about 1 MiB, claimed by whichever
// thread (submitter or helper) asks next.  `done[k]` is set once chunk k is in the pinned
// buffer; the submitter queues the uploads in chunk order as chunks complete.
struct StageTask {
    const char *src[2] = {nullptr, nullptr};
    char *stage[2] = {nullptr, nullptr};
    size_t row_stride = 0, rowb = 0;
    int H = 0, chunk_rows = 1, per_frame = 0, total = 0;
    std::atomic<int> next{0};
    std::unique_ptr<std::atomic<int>[]> done;
    // copies the next unclaimed chunk; false when every chunk has been claimed
    bool run_one()
    {
        const int k = next.fetch_add(1, std::memory_order_relaxed);
        if (k >= total) return false;
        const int i = k / per_frame;
        const int h0 = (k % per_frame) * chunk_rows, h1 = std::min(H, h0 + chunk_rows);
        for (int h = h0; h < h1; ++h)
            std::memcpy(stage[i] + (size_t)h * rowb, src[i] + (size_t)h * row_stride, rowb);
        done[k].store(1, std::memory_order_release);
        return true;
    }
};

// Helper threads that join the staging of large host pairs.  A task stays at the front of
// the queue until a helper finds every chunk claimed; the submitter holds its own reference
// and returns only when all chunks are done, so helpers never touch a caller's buffer after
// the call that passed it has returned.
struct StagePool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<StageTask>> tasks;
    std::vector<std::thread> threads;
    bool stop = false;
    // may throw std::system_error (no thread) or std::bad_alloc; the threads started so far
    // stay and are joined by shutdown()
    void start(int n)
    {
        threads.reserve(n);
        for (int i = 0; i < n; ++i) threads.emplace_back([this] { loop(); });
    }
    void loop()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] { return stop || !tasks.empty(); });
            if (stop) return;
            std::shared_ptr<StageTask> t = tasks.front();
            lk.unlock();
            while (t->run_one()) {
            }
            lk.lock();
            if (!tasks.empty() && tasks.front() == t) tasks.pop_front();
        }
    }
    bool post(const std::shared_ptr<StageTask> &t)
    {
        try {
            std::lock_guard<std::mutex> lk(mu);
            tasks.push_back(t);
        } catch (const std::bad_alloc &) {
            return false;
        }
        cv.notify_all();
        return true;
    }
    void shutdown()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (std::thread &th : threads) th.join();
        threads.clear();
    }
};

size_t fields_bytes(size_t c) { return align_up(c * 13, 256); }

}  // namespace

struct piv_server {
    int device = 0;
    piv_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_complete;
    std::vector<SrvStream *> streams;
    std::deque<Request *> queue;  // admitted, waiting for a lane
    size_t unprepared = 0;        // queued requests the scheduler has not prepared yet
    std::deque<Request *> done;   // finished, not yet polled
    std::map<uint64_t, Request *> polled;
    std::vector<Lane> lanes;
    std::deque<int> inflight;     // lanes in launch order
    size_t allocated = 0, peak = 0;
    uint64_t next_id = 1;
    piv_server_stats stats{};
    bool stop = false;
    bool paused = false;  // piv_server_pause: admitted requests wait, nothing is launched
    std::thread scheduler, completer;
    StagePool stagers;    // helpers for the staging copies of large host pairs
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void free_slot(piv_server *s, Slot *sl)
{
    if (sl->uploaded) (void)hipEventDestroy(sl->uploaded);
    delete sl->fa;
    delete sl->fb;
    s->allocated -= sl->dev_bytes;
    sl->dev.release();
    sl->host.release();
    delete sl;
}

// Drop cached free slots (largest streams first) until `need` more bytes fit the budget.
// A stream gives up free slots only while it holds more of them than it has requests in
// flight (queued or on the device; results held by the client count neither way): a stream being served reuses a free slot for its next request as soon as one of
// its results is released, and taking those slots would only have it reallocate them
// (hipFree / hipMalloc, each synchronising) a moment later.  Past that, the submission is
// refused and retried by the client once results are released; slots left over from an
// earlier burst still go to other streams.
bool make_room(piv_server *s, size_t need)
{
    while (s->allocated + need > s->cfg.hbm_budget) {
        SrvStream *best = nullptr;
        for (SrvStream *st : s->streams)
            if (st->free_slots.size() > (size_t)st->in_flight &&
                (!best || st->slot_dev > best->slot_dev))
                best = st;
        if (!best) return false;
        Slot *sl = best->free_slots.back();
        best->free_slots.pop_back();
        free_slot(s, sl);
    }
    return true;
}

piv_status new_slot(piv_server *s, SrvStream *st, Slot **out)
{
    auto *sl = new (std::nothrow) Slot();
    if (!sl) return fail(PIV_ERR_OOM, "host allocation failed");
    sl->owner = st;
    const size_t fb = (size_t)st->H * st->W * bytes_per_px(st->fmt);
    piv_status r = sl->dev.reserve(st->slot_dev);
    if (r == PIV_OK) r = sl->host.reserve(st->slot_host);
    if (r == PIV_OK && hipEventCreateWithFlags(&sl->uploaded, hipEventDisableTiming) != hipSuccess)
        r = fail(PIV_ERR_DEVICE, "event creation failed");
    if (r != PIV_OK) {
        sl->dev.release();
        sl->host.release();
        delete sl;
        return r;
    }
    sl->dev_bytes = st->slot_dev;
    s->allocated += sl->dev_bytes;
    s->peak = std::max(s->peak, s->allocated);
    char *d = sl->dev.as<char>();
    size_t off = 0;
    sl->raw_a = d + off;
    off = align_up(off + fb, 256);
    sl->raw_b = d + off;
    off = align_up(off + fb, 256);
    for (const PassPlan &pl : st->plans) {
        const size_t c = pl.cells;
        Slot::PassWs w{};
        w.raw = reinterpret_cast<float *>(d + off);
        off += fields_bytes(c);
        w.val = reinterpret_cast<float *>(d + off);
        off += fields_bytes(c);
        w.snap = reinterpret_cast<uint8_t *>(d + off);
        off = align_up(off + c, 256);
        w.sig = reinterpret_cast<uint8_t *>(d + off);
        off = align_up(off + c, 256);
        w.org = reinterpret_cast<float *>(d + off);
        off = align_up(off + c * 16, 256);
        w.rec = reinterpret_cast<PeakRec *>(d + off);
        off = align_up(off + c * sizeof(PeakRec), 256);
        sl->ws.push_back(w);
    }
    char *h = static_cast<char *>(sl->host.p);
    const size_t cf = st->plans.back().cells;
    sl->h_raw_a = h;
    sl->h_raw_b = h + align_up(fb, 256);
    char *hr = h + 2 * align_up(fb, 256);
    sl->h_dx = reinterpret_cast<float *>(hr);
    sl->h_dy = sl->h_dx + cf;
    sl->h_ppr = sl->h_dy + cf;
    sl->h_flags = reinterpret_cast<uint8_t *>(sl->h_ppr + cf);
    const size_t pitch_bytes = (size_t)st->W * bytes_per_px(st->fmt);
    std::lock_guard<std::mutex> ul(st->upload_mu);
    r = frame_from_device(st->an, sl->raw_a, st->fmt, (uint32_t)st->H, (uint32_t)st->W, pitch_bytes,
                          false, st->has_bg ? st->bg.as<float>() : nullptr, &sl->fa);
    if (r == PIV_OK)
        r = frame_from_device(st->an, sl->raw_b, st->fmt, (uint32_t)st->H, (uint32_t)st->W,
                              pitch_bytes, false, st->has_bg ? st->bg.as<float>() : nullptr, &sl->fb);
    if (r != PIV_OK) {
        free_slot(s, sl);
        return r;
    }
    *out = sl;
    return PIV_OK;
}

// Descriptors and launch structure of one batch.  Per pass step: the PassPairs of every
// request that has the step, grouped by (window, format) so each group is one contiguous
// correlation launch, then the ValidatePairs of the same requests.
struct BatchPlan {
    struct Group {
        int n, fmt, deform, count, variant;
        uint32_t max_cells;
        size_t pp_off;
    };
    struct Step {
        std::vector<Group> groups;
        int count = 0;
        uint32_t max_cells = 0;
        size_t pp_off = 0, vp_off = 0;
        // block map of the per-window / per-vector kernels (uint32 [count + 1]): requests of
        // different frame sizes get their own block counts instead of the largest one's
        size_t bm_off = 0;
        uint32_t blocks = 0;
    };
    std::vector<Step> steps;
    std::vector<char> host;  // descriptor bytes
    std::vector<int64_t> key;
};

// windows (vectors) per block of the per-window and per-vector kernels
constexpr size_t kStepBlock = 256;

void build_batch(Lane &L, BatchPlan &bp)
{
    const std::vector<Request *> &reqs = L.batch;
    const int B = (int)reqs.size();
    size_t maxsteps = 0;
    for (Request *r : reqs) maxsteps = std::max(maxsteps, r->st->plans.size());
    // final outputs of each request, contiguous in the lane's output region
    L.out_off.assign(B, 0);
    L.sum_index.assign(B, 0);
    size_t off = 0;
    for (int k = 0; k < B; ++k) {
        L.out_off[k] = off;
        off += fields_bytes(reqs[k]->st->plans.back().cells);
        L.sum_index[k] = (reqs[k]->st->plans.size() - 1) * B + k;
    }
    L.out_bytes = off;
    L.sum_count = maxsteps * B * 5;
    // layout
    bp.steps.assign(maxsteps, {});
    size_t doff = 0;
    for (size_t si = 0; si < maxsteps; ++si) {
        BatchPlan::Step &S = bp.steps[si];
        S.pp_off = doff;
        // groups: (window, format, deformation mode of this step), one launch each
        auto group_key = [&](Request *r) {
            const int def = si > 0 ? (int)r->st->an->cfg.deformation : 0;
            return std::array<int, 3>{r->st->plans[si].n, r->slot->fa->fmt, def};
        };
        std::vector<std::array<int, 3>> keys;
        for (Request *r : reqs) {
            if (r->st->plans.size() <= si) continue;
            const std::array<int, 3> kk = group_key(r);
            if (std::find(keys.begin(), keys.end(), kk) == keys.end()) keys.push_back(kk);
        }
        std::sort(keys.begin(), keys.end());
        for (auto &kk : keys) {
            BatchPlan::Group g{kk[0], kk[1], kk[2], 0, -1, 0, doff};
            for (Request *r : reqs) {
                if (r->st->plans.size() <= si) continue;
                if (group_key(r) != kk) continue;
                // the kernel variant of the group's first stream (variants are bit-identical)
                if (g.variant < 0) g.variant = variant_of(r->st->an, kk[0], kk[1]);
                ++g.count;
                g.max_cells = std::max(g.max_cells, (uint32_t)r->st->plans[si].cells);
            }
            doff += g.count * sizeof(PassPair);
            S.groups.push_back(g);
            S.count += g.count;
            S.max_cells = std::max(S.max_cells, g.max_cells);
        }
        doff = align_up(doff, 64);
        S.vp_off = doff;
        doff = align_up(doff + S.count * sizeof(ValidatePair), 64);
        S.bm_off = doff;
        doff = align_up(doff + (S.count + 1) * sizeof(uint32_t), 64);
    }
    bp.host.assign(doff, 0);
    // contents
    char *d_out = L.out.as<char>();
    uint32_t *d_sums = L.sums.as<uint32_t>();
    for (size_t si = 0; si < maxsteps; ++si) {
        BatchPlan::Step &S = bp.steps[si];
        int vi = 0;
        uint32_t *bm = reinterpret_cast<uint32_t *>(bp.host.data() + S.bm_off);
        bm[0] = 0;
        for (BatchPlan::Group &g : S.groups) {
            int gi = 0;
            for (int k = 0; k < B; ++k) {
                Request *r = reqs[k];
                SrvStream *st = r->st;
                if (st->plans.size() <= si) continue;
                if (st->plans[si].n != g.n || r->slot->fa->fmt != g.fmt ||
                    (si > 0 ? (int)st->an->cfg.deformation : 0) != g.deform)
                    continue;
                const PassPlan &pl = st->plans[si];
                const size_t c = pl.cells;
                Slot *sl = r->slot;
                const piv_frame *A = sl->fa, *Bf = sl->fb;
                PassPair d{};
                d.img_a = A->img;
                d.img_b = Bf->img;
                d.use_mask = (A->has_sat && Bf->has_sat) ? 1 : 0;
                d.sat_a = A->has_sat ? A->sat_ptr : nullptr;
                d.sat_b = Bf->has_sat ? Bf->sat_ptr : nullptr;
                d.sat_cells = A->cells;
                // the stream's own analyzer prepared both frames for these passes, so their
                // screen tiles fit every window of the plan
                d.screened = screened_bits(A, Bf);
                if (si > 0) {
                    const size_t pc = st->plans[si - 1].cells;
                    d.prev_dx = sl->ws[si - 1].val;
                    d.prev_dy = sl->ws[si - 1].val + pc;
                    d.prev = st->plans[si - 1].grid;
                    d.has_prev = 1;
                }
                const Slot::PassWs &w = sl->ws[si];
                d.origins = w.org;
                d.sig = w.sig;
                d.peaks = w.rec;
                d.out_dx = w.raw;
                d.out_dy = w.raw + c;
                d.out_ppr = w.raw + 2 * c;
                d.out_flags = reinterpret_cast<uint8_t *>(w.raw + 3 * c);
                d.H = A->H;
                d.W = A->W;
                d.pitch = A->pitch;
                d.fmt = A->fmt;
                d.grid = pl.grid;
                d.min_mean = st->an->cfg.min_mean;
                d.min_std = st->an->cfg.min_std;
                d.deform = (int32_t)st->an->cfg.deformation;
                reinterpret_cast<PassPair *>(bp.host.data() + g.pp_off)[gi++] = d;
                ValidatePair v{};
                v.raw_dx = d.out_dx;
                v.raw_dy = d.out_dy;
                v.raw_ppr = d.out_ppr;
                v.raw_flags = d.out_flags;
                float *vd = w.val;
                if (si + 1 == st->plans.size()) vd = reinterpret_cast<float *>(d_out + L.out_off[k]);
                v.dx = vd;
                v.dy = vd + c;
                v.ppr = vd + 2 * c;
                v.flags = reinterpret_cast<uint8_t *>(vd + 3 * c);
                v.snapshot = w.snap;
                v.summary = d_sums + (si * B + k) * 5;
                v.nx = pl.grid.nx;
                v.ny = pl.grid.ny;
                const piv_config &cf = st->an->cfg;
                v.ppr_threshold = cf.ppr_threshold;
                v.median_threshold = cf.median_threshold;
                v.median_eps = cf.median_eps;
                v.max_displacement = (float)cf.windows[0] / 2.0f;
                v.replace = cf.replace ? 1 : 0;
                reinterpret_cast<ValidatePair *>(bp.host.data() + S.vp_off)[vi] = v;
                bm[vi + 1] = bm[vi] + (uint32_t)((c + kStepBlock - 1) / kStepBlock);
                ++vi;
            }
        }
    }
    for (BatchPlan::Step &S : bp.steps)
        S.blocks = reinterpret_cast<const uint32_t *>(bp.host.data() + S.bm_off)[S.count];
    // launch structure (the graph cache key): descriptor offsets follow from it
    bp.key.clear();
    bp.key.push_back((int64_t)maxsteps);
    bp.key.push_back((int64_t)L.sum_count);
    for (const BatchPlan::Step &S : bp.steps) {
        bp.key.push_back(S.count);
        bp.key.push_back(S.max_cells);
        bp.key.push_back(S.blocks);
        for (const BatchPlan::Group &g : S.groups) {
            bp.key.push_back(g.n);
            bp.key.push_back(g.fmt);
            bp.key.push_back(g.count);
            bp.key.push_back(g.max_cells);
            bp.key.push_back(g.variant);
            bp.key.push_back(g.deform);
        }
    }
}

hipError_t enqueue_batch(Lane &L, const BatchPlan &bp)
{
    hipStream_t st = L.stream;
    const char *dd = L.desc.as<char>();
    hipError_t e = launch_zero_u32(L.sums.as<uint32_t>(), (uint32_t)L.sum_count, st);
    for (const BatchPlan::Step &S : bp.steps) {
        if (e != hipSuccess) break;
        const auto *pp = reinterpret_cast<const PassPair *>(dd + S.pp_off);
        const auto *bm = reinterpret_cast<const uint32_t *>(dd + S.bm_off);
        e = launch_place_pass(pp, S.count, S.max_cells, st, bm, S.blocks);
        for (const BatchPlan::Group &g : S.groups) {
            if (e != hipSuccess) break;
            e = launch_correlate(g.n, g.fmt, g.variant, reinterpret_cast<const PassPair *>(dd + g.pp_off), g.count,
                                 g.max_cells, st, g.deform);
        }
        if (e == hipSuccess) e = launch_peak_finish(pp, S.count, S.max_cells, st, false, bm, S.blocks);
        if (e == hipSuccess)
            e = launch_validate(reinterpret_cast<const ValidatePair *>(dd + S.vp_off), S.count,
                                S.max_cells, st, bm, S.blocks);
    }
    return e;
}

// Launches one batch on lane L (scheduler thread, server lock NOT held).
piv_status launch_batch(piv_server *s, Lane &L)
{
    HIP_TRY(hipSetDevice(s->device));
    // lane buffers: output region and summaries sized for this batch
    size_t need_out = 0, steps = 0;
    for (Request *r : L.batch) {
        need_out += fields_bytes(r->st->plans.back().cells);
        steps = std::max(steps, r->st->plans.size());
    }
    const size_t nsum = steps * L.batch.size() * 5;
    // lane buffers grow at least twofold: a server warming up to its batch sizes reallocates
    // (and drops its graphs) a few times, not at every larger batch
    auto grow = [](size_t need, size_t cap) { return need <= cap ? need : std::max(need, 2 * cap); };
    if (need_out > L.out.cap || nsum * 4 > L.sums.cap) L.drop_graphs();
    PIV_TRY(L.out.reserve(grow(need_out, L.out.cap)));
    PIV_TRY(L.h_out.reserve(grow(need_out, L.h_out.cap)));
    PIV_TRY(L.sums.reserve(grow(nsum * 4, L.sums.cap)));
    PIV_TRY(L.h_sums.reserve(grow(nsum * 4, L.h_sums.cap)));
    BatchPlan bp;
    build_batch(L, bp);
    if (bp.host.size() > L.desc.cap) {
        L.drop_graphs();
        PIV_TRY(L.desc.reserve(grow(bp.host.size(), L.desc.cap)));
    }
    PIV_TRY(L.h_desc.reserve(grow(bp.host.size(), L.h_desc.cap)));
    hipStream_t st = L.stream;
    for (Request *r : L.batch) HIP_TRY(hipStreamWaitEvent(st, r->slot->uploaded, 0));
    std::memcpy(L.h_desc.p, bp.host.data(), bp.host.size());
    HIP_TRY(hipMemcpyAsync(L.desc.p, L.h_desc.p, bp.host.size(), hipMemcpyHostToDevice, st));
    // A structure seen for the first time is launched eagerly (capturing costs more than the
    // launches); one that recurs is captured once and replayed from then on.
    auto it = L.graphs.find(bp.key);
    if (it != L.graphs.end()) {
        HIP_TRY(hipGraphLaunch(it->second, st));
    } else if (++L.seen[bp.key] < 2) {
        if (L.seen.size() > 4096) L.seen.clear();
        HIP_TRY(enqueue_batch(L, bp));
    } else {
        if (L.graphs.size() >= kMaxGraphsPerLane) L.drop_graphs();
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t ce = enqueue_batch(L, bp);
        hipError_t ee = hipStreamEndCapture(st, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            HIP_TRY(ce);
            HIP_TRY(ee);
        }
        hipGraphExec_t ex = nullptr;
        hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ie);
        L.graphs[bp.key] = ex;
        HIP_TRY(hipGraphLaunch(ex, st));
    }
    HIP_TRY(hipMemcpyAsync(L.h_out.p, L.out.p, L.out_bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(L.h_sums.p, L.sums.p, L.sum_count * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(L.done, st));
    return PIV_OK;
}

// Requests whose upload (and conditioning) has finished, oldest first: a batch never waits
// for a slow upload of another stream.
bool upload_done(Request *r) { return hipEventQuery(r->slot->uploaded) == hipSuccess; }

// Conditioning and tables of requests of one stream, launched on its upload stream after
// their uploads (the tables of all their frames in one launch per kernel where the stream's
// configuration allows); `uploaded` is recorded again after them, and the lane that runs a
// request waits for it.
piv_status prepare_requests(SrvStream *rs, const std::vector<Request *> &reqs)
{
    std::lock_guard<std::mutex> ul(rs->upload_mu);
    const float *bg = rs->has_bg ? rs->bg.as<float>() : nullptr;
    std::vector<piv_frame *> fs;
    fs.reserve(2 * reqs.size());
    for (Request *r : reqs) {
        fs.push_back(r->slot->fa);
        fs.push_back(r->slot->fb);
    }
    PIV_TRY(prepare_frames_list(rs->an, fs.data(), (int)fs.size(), bg));
    // raw frames without tables need no device work: the uploads' events stand
    if (prepare_launches(rs->an->cfg, rs->fmt, bg != nullptr))
        for (Request *r : reqs) HIP_TRY(hipEventRecord(r->slot->uploaded, rs->an->stream));
    return PIV_OK;
}

void scheduler_loop(piv_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        int free_lane = -1;
        auto find_lane = [&] {
            for (size_t i = 0; i < s->lanes.size(); ++i)
                if (!s->lanes[i].busy) {
                    free_lane = (int)i;
                    return true;
                }
            return false;
        };
        // woken by a submission to prepare, or by a free lane with requests waiting
        s->cv_sched.wait(lk, [&] {
            if (s->stop) return true;
            if (s->paused || s->queue.empty()) return false;
            return s->unprepared > 0 || find_lane();
        });
        if (s->stop) return;
        // Requests whose upload has landed are prepared at once, while the lanes may still be
        // busy, so a batch launches as soon as a lane frees; all launches come from this
        // thread (launches from several threads contend in the runtime).
        std::vector<Request *> prep;
        for (Request *r : s->queue)
            if (!r->prepared && upload_done(r)) prep.push_back(r);
        if (!prep.empty()) {
            lk.unlock();
            // stream by stream, oldest stream first
            std::vector<SrvStream *> order;
            for (Request *r : prep)
                if (std::find(order.begin(), order.end(), r->st) == order.end()) order.push_back(r->st);
            std::vector<Request *> group;
            for (SrvStream *rs : order) {
                group.clear();
                for (Request *r : prep)
                    if (r->st == rs) group.push_back(r);
                const piv_status ps = prepare_requests(rs, group);
                for (Request *r : group) r->status = ps;
            }
            lk.lock();
            for (Request *r : prep) {
                r->prepared = true;
                s->unprepared -= 1;
            }
        }
        free_lane = -1;
        if (s->paused || !find_lane()) {
            // no lane: wait for one (or, with uploads in flight, look again shortly)
            if (prep.empty()) s->cv_sched.wait_for(lk, std::chrono::microseconds(50));
            continue;
        }
        Lane &L = s->lanes[free_lane];
        L.batch.clear();
        bool failed = false;
        for (auto it = s->queue.begin(); it != s->queue.end() && L.batch.size() < s->cfg.max_batch_pairs;) {
            Request *r = *it;
            if (!r->prepared) {
                ++it;
            } else if (r->status != PIV_OK) {
                // its preparation failed: the request completes with that error
                r->t_done = s->now_ms();
                r->st->in_flight -= 1;
                s->done.push_back(r);
                it = s->queue.erase(it);
                failed = true;
            } else {
                L.batch.push_back(r);
                it = s->queue.erase(it);
            }
        }
        if (failed) s->cv_done.notify_all();
        if (L.batch.empty()) {
            // uploads still in flight: look again shortly (or when a submission arrives)
            if (prep.empty()) s->cv_sched.wait_for(lk, std::chrono::microseconds(50));
            continue;
        }
        L.busy = true;
        lk.unlock();
        piv_status st = launch_batch(s, L);
        lk.lock();
        s->stats.batches += 1;
        s->stats.batched_pairs += L.batch.size();
        if (st != PIV_OK) {
            // nothing of this batch completes on the device: fail its requests
            (void)hipStreamSynchronize(L.stream);
            for (Request *r : L.batch) {
                r->status = st;
                r->t_done = s->now_ms();
                r->st->in_flight -= 1;
                s->done.push_back(r);
            }
            L.batch.clear();
            L.busy = false;
            s->cv_done.notify_all();
            continue;
        }
        s->inflight.push_back(free_lane);
        s->cv_complete.notify_one();
    }
}

void completer_loop(piv_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_complete.wait(lk, [&] { return s->stop || !s->inflight.empty(); });
        if (s->inflight.empty() && s->stop) return;
        const int li = s->inflight.front();
        Lane &L = s->lanes[li];
        lk.unlock();
        const hipError_t e = hipEventSynchronize(L.done);
        const double t = s->now_ms();
        // the lane's requests belong to it until they are queued as done: their results are
        // copied into the slots without the server lock, which submissions and polls take
        const int B = (int)L.batch.size();
        const char *ho = static_cast<const char *>(L.h_out.p);
        const uint32_t *hs = static_cast<const uint32_t *>(L.h_sums.p);
        for (int k = 0; k < B; ++k) {
            Request *r = L.batch[k];
            r->t_done = t;
            r->batch_pairs = (uint32_t)B;
            if (e != hipSuccess) {
                r->status = fail(PIV_ERR_DEVICE, "batch failed: %s", hipGetErrorString(e));
            } else {
                const size_t c = r->st->plans.back().cells;
                const float *src = reinterpret_cast<const float *>(ho + L.out_off[k]);
                Slot *sl = r->slot;
                std::memcpy(sl->h_dx, src, c * 4);
                std::memcpy(sl->h_dy, src + c, c * 4);
                std::memcpy(sl->h_ppr, src + 2 * c, c * 4);
                std::memcpy(sl->h_flags, src + 3 * c, c);
                to_summary(hs + L.sum_index[k] * 5, &r->summary);
                r->status = PIV_OK;
            }
        }
        lk.lock();
        s->inflight.pop_front();
        for (int k = 0; k < B; ++k) {
            L.batch[k]->st->in_flight -= 1;
            s->done.push_back(L.batch[k]);
            s->stats.completed += 1;
        }
        L.batch.clear();
        L.busy = false;
        s->cv_done.notify_all();
        s->cv_sched.notify_one();
    }
}

}  // namespace

extern "C" {

piv_status piv_server_create(int device, const piv_server_config *cfg, piv_server **out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        PIV_TRY(set_device(device));
        auto *s = new (std::nothrow) piv_server();
        if (!s) return fail(PIV_ERR_OOM, "host allocation failed");
        s->device = device;
        if (cfg) s->cfg = *cfg;
        if (s->cfg.hbm_budget == 0) s->cfg.hbm_budget = (size_t)4 << 30;
        if (s->cfg.max_batch_pairs == 0) s->cfg.max_batch_pairs = 64;
        if (s->cfg.max_inflight == 0) s->cfg.max_inflight = 2;
        if (s->cfg.max_inflight > 8) s->cfg.max_inflight = 8;
        s->lanes.resize(s->cfg.max_inflight);
        for (Lane &L : s->lanes) {
            if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&L.done, hipEventDisableTiming | hipEventBlockingSync) != hipSuccess) {
                piv_server_destroy(s);
                return fail(PIV_ERR_DEVICE, "stream creation failed");
            }
        }
        try {
            s->scheduler = std::thread(scheduler_loop, s);
            s->completer = std::thread(completer_loop, s);
        } catch (const std::exception &) {
            piv_server_destroy(s);
            return fail(PIV_ERR_DEVICE, "server threads could not be started");
        }
        try {
            s->stagers.start(kStageHelpers);
        } catch (const std::exception &) {
            // staging helpers are optional: submitters copy their pairs alone without them
        }
        *out = s;
        return PIV_OK;
    });
}

void piv_server_destroy(piv_server *s)
{
    try {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_sched.notify_all();
        s->cv_complete.notify_all();
        s->stagers.shutdown();
        if (s->scheduler.joinable()) s->scheduler.join();
        // the completer drains every launched batch before it exits
        if (s->completer.joinable()) s->completer.join();
        (void)hipSetDevice(s->device);
        for (Lane &L : s->lanes) {
            if (L.stream) (void)hipStreamSynchronize(L.stream);
            L.drop_graphs();
            L.desc.release();
            L.sums.release();
            L.out.release();
            L.h_desc.release();
            L.h_sums.release();
            L.h_out.release();
            if (L.done) (void)hipEventDestroy(L.done);
            if (L.stream) (void)hipStreamDestroy(L.stream);
        }
        std::vector<Request *> all(s->queue.begin(), s->queue.end());
        all.insert(all.end(), s->done.begin(), s->done.end());
        for (auto &kv : s->polled) all.push_back(kv.second);
        for (Request *r : all) {
            if (r->slot) free_slot(s, r->slot);
            delete r;
        }
        for (SrvStream *st : s->streams) {
            if (st->an) (void)hipStreamSynchronize(st->an->stream);
            for (Slot *sl : st->free_slots) free_slot(s, sl);
            st->bg.release();
            piv_analyzer_destroy(st->an);
            delete st;
        }
        delete s;
    } catch (...) {
    }
}

piv_status piv_server_add_stream(piv_server *s, const piv_config *cfg, uint32_t height,
                                 uint32_t width, piv_pixel_format fmt, const float *background,
                                 uint32_t *stream_id)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!s || !cfg || !stream_id) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if ((int)fmt != FMT_U8 && (int)fmt != FMT_U16 && (int)fmt != FMT_F32)
            return fail(PIV_ERR_INVALID_ARGUMENT, "unknown pixel format %d", (int)fmt);
        auto *st = new (std::nothrow) SrvStream();
        if (!st) return fail(PIV_ERR_OOM, "host allocation failed");
        // the server launches the stream's batches on its own lanes: no helper streams
        piv_status r = analyzer_create(cfg, s->device, &st->an, false);
        if (r != PIV_OK) {
            delete st;
            return r;
        }
        st->H = (int)height;
        st->W = (int)width;
        st->fmt = (int)fmt;
        r = plan_passes(st->an, st->H, st->W, st->plans);
        // the conditioning scratch at its final size, so that no submission allocates (its users,
        // serialised by upload_mu, are ordered on the analyzer's stream)
        if (r == PIV_OK) {
            const size_t n = (size_t)height * width, cells = (size_t)sat_table_elems(st->H, st->W);
            const piv_config &c = st->an->cfg;
            r = st->an->sat_carry.reserve(sat_carry_bytes(st->H, st->W, 64));
            if (r == PIV_OK && c.smooth_sigma > 0.0f) r = st->an->prep_tmp.reserve(n * sizeof(float));
            if (r == PIV_OK && c.smooth_sigma > 0.0f && c.normalize_radius > 0)
                r = st->an->prep_tmp2.reserve(n * sizeof(float));
            if (r == PIV_OK && c.normalize_radius > 0) r = st->an->prep_sat.reserve(cells * 2 * sizeof(double));
        }
        if (r == PIV_OK && background) {
            const size_t n = (size_t)height * width;
            r = st->bg.reserve(n * 4);
            if (r == PIV_OK && hipMemcpy(st->bg.p, background, n * 4, hipMemcpyHostToDevice) != hipSuccess)
                r = fail(PIV_ERR_DEVICE, "background upload failed");
            st->has_bg = r == PIV_OK;
        }
        if (r != PIV_OK) {
            st->bg.release();
            piv_analyzer_destroy(st->an);
            delete st;
            return r;
        }
        // slot sizes (the layout new_slot carves)
        const size_t fb = (size_t)height * width * bytes_per_px((int)fmt);
        size_t dev = 2 * align_up(fb, 256);
        for (const PassPlan &pl : st->plans) {
            const size_t c = pl.cells;
            dev += 2 * fields_bytes(c) + 2 * align_up(c, 256) + align_up(c * 16, 256) +
                   align_up(c * sizeof(PeakRec), 256);
        }
        // conditioned frames hold their own image, base and SAT buffers
        const size_t n = (size_t)height * width;
        const size_t cells = (size_t)sat_table_elems(st->H, st->W);
        const piv_config &c = st->an->cfg;
        if (c.smooth_sigma > 0.0f || c.normalize_radius > 0 || background) dev += 2 * 2 * n * 4;
        dev += 2 * cells * 16;  // SAT (upper bound; only when the mask is needed)
        st->slot_dev = dev;
        st->slot_host = 2 * align_up(fb, 256) + fields_bytes(st->plans.back().cells);
        std::lock_guard<std::mutex> lk(s->mu);
        st->id = (uint32_t)s->streams.size();
        s->streams.push_back(st);
        *stream_id = st->id;
        return PIV_OK;
    });
}

namespace {

// Admission + slot, then (outside the server lock) staging/upload or device binding and
// conditioning on the stream's upload stream, then the queue.
piv_status submit_impl(piv_server *s, uint32_t stream_id, const void *pa, const void *pb,
                       size_t row_stride, bool on_device, uint64_t *request_id)
{
    if (!s || !pa || !pb || !request_id) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
    const double t_arrival = s->now_ms();
    std::unique_lock<std::mutex> lk(s->mu);
    if (stream_id >= s->streams.size()) return fail(PIV_ERR_INVALID_ARGUMENT, "unknown stream %u", stream_id);
    SrvStream *st = s->streams[stream_id];
    const size_t bpp = (size_t)bytes_per_px(st->fmt);
    const size_t rowb = (size_t)st->W * bpp;
    if (row_stride == 0) row_stride = rowb;
    if (row_stride < rowb || row_stride % bpp != 0)
        return fail(PIV_ERR_INVALID_ARGUMENT, "row_stride %zu invalid for width %d", row_stride, st->W);
    if (s->queue.size() >= kMaxQueued) {
        s->stats.rejected += 1;
        return fail(PIV_ERR_ADMISSION, "request queue full");
    }
    PIV_TRY(set_device(s->device));
    Slot *sl = nullptr;
    if (!st->free_slots.empty()) {
        sl = st->free_slots.back();
        st->free_slots.pop_back();
    } else {
        if (!make_room(s, st->slot_dev)) {
            s->stats.rejected += 1;
            return fail(PIV_ERR_ADMISSION, "HBM budget of %zu bytes exhausted (%zu held by requests)",
                        s->cfg.hbm_budget, s->allocated);
        }
        PIV_TRY(new_slot(s, st, &sl));
    }
    st->in_use += 1;
    st->in_flight += 1;
    lk.unlock();
    piv_status r = PIV_OK;
    {
        hipStream_t us = st->an->stream;
        if (on_device) {
            // the caller's samples are read in place until the result is released
            std::lock_guard<std::mutex> ul(st->upload_mu);
            sl->fa->raw_ptr = pa;
            sl->fb->raw_ptr = pb;
            sl->fa->raw_pitch = sl->fb->raw_pitch = (int)(row_stride / bpp);
            if (hipEventRecord(sl->uploaded, us) != hipSuccess) r = fail(PIV_ERR_DEVICE, "event record failed");
        } else {
            // staged: the caller's buffers are free when this returns.  Copied into the slot's
            // own pinned staging in chunks of about 1 MiB without the stream's lock (requests of
            // one stream stage in parallel from several threads); each chunk's upload is queued
            // as soon as it is staged, so the DMA of one chunk overlaps the copy of the next.
            // Large pairs are copied by the server's staging helpers and this thread together.
            const int chunk_rows = (int)std::max<size_t>(1, ((size_t)1 << 20) / rowb);
            const size_t pair_bytes = 2 * (size_t)st->H * rowb;
            // no C++ exception leaves the C ABI: a failed host allocation is PIV_ERR_OOM
            std::shared_ptr<StageTask> t;
            try {
                t = std::make_shared<StageTask>();
                t->per_frame = (st->H + chunk_rows - 1) / chunk_rows;
                t->total = 2 * t->per_frame;
                t->done.reset(new std::atomic<int>[t->total]());
            } catch (const std::bad_alloc &) {
                t.reset();
                r = fail(PIV_ERR_OOM, "host allocation failed");
            }
            if (t) {
                t->src[0] = static_cast<const char *>(pa);
                t->src[1] = static_cast<const char *>(pb);
                t->stage[0] = sl->h_raw_a;
                t->stage[1] = sl->h_raw_b;
                t->row_stride = row_stride;
                t->rowb = rowb;
                t->H = st->H;
                t->chunk_rows = chunk_rows;
                char *dst[2] = {static_cast<char *>(sl->raw_a), static_cast<char *>(sl->raw_b)};
                // when the task cannot be queued, this thread copies every chunk itself
                if (pair_bytes >= kSplitStageBytes && !s->stagers.threads.empty())
                    (void)s->stagers.post(t);
                int queued = 0;
                // queues the uploads of the staged chunks that follow the last one queued (after
                // an upload error the remaining chunks are still waited for, not uploaded)
                auto queue_staged = [&]() {
                    while (queued < t->total && t->done[queued].load(std::memory_order_acquire)) {
                        if (r == PIV_OK) {
                            const int i = queued / t->per_frame;
                            const int h0 = (queued % t->per_frame) * chunk_rows;
                            const int h1 = std::min(st->H, h0 + chunk_rows);
                            std::lock_guard<std::mutex> ul(st->upload_mu);
                            if (hipMemcpyAsync(dst[i] + (size_t)h0 * rowb, t->stage[i] + (size_t)h0 * rowb,
                                               (size_t)(h1 - h0) * rowb, hipMemcpyHostToDevice,
                                               us) != hipSuccess)
                                r = fail(PIV_ERR_DEVICE, "upload failed");
                        }
                        ++queued;
                    }
                };
                while (t->run_one()) queue_staged();
                while (queued < t->total) {
                    queue_staged();
                    if (queued < t->total) std::this_thread::yield();
                }
            }
            std::lock_guard<std::mutex> ul(st->upload_mu);
            sl->fa->raw_ptr = sl->raw_a;
            sl->fb->raw_ptr = sl->raw_b;
            sl->fa->raw_pitch = sl->fb->raw_pitch = st->W;
            // conditioning and tables are launched by the scheduler when the request is taken
            // up (one thread issues those launches)
            if (r == PIV_OK && hipEventRecord(sl->uploaded, us) != hipSuccess)
                r = fail(PIV_ERR_DEVICE, "event record failed");
        }
    }
    lk.lock();
    if (r != PIV_OK) {
        st->free_slots.push_back(sl);
        st->in_use -= 1;
        st->in_flight -= 1;
        return r;
    }
    auto *req = new Request();
    req->id = s->next_id++;
    req->st = st;
    req->slot = sl;
    req->t_submit = t_arrival;
    s->queue.push_back(req);
    s->unprepared += 1;
    s->stats.submitted += 1;
    *request_id = req->id;
    s->cv_sched.notify_one();
    return PIV_OK;
}

}  // namespace

piv_status piv_submit(piv_server *s, uint32_t stream_id, const void *pixels_a,
                      const void *pixels_b, size_t row_stride, uint64_t *request_id)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        return submit_impl(s, stream_id, pixels_a, pixels_b, row_stride, false, request_id);
    });
}

piv_status piv_submit_device(piv_server *s, uint32_t stream_id, const void *pixels_a,
                             const void *pixels_b, size_t row_stride, uint64_t *request_id)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        return submit_impl(s, stream_id, pixels_a, pixels_b, row_stride, true, request_id);
    });
}

piv_status piv_submit_device_many(piv_server *s, uint32_t count, const uint32_t *stream_ids,
                                  const void *const *pixels_a, const void *const *pixels_b,
                                  size_t row_stride, uint64_t *request_ids, uint32_t *submitted)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!s || !stream_ids || !pixels_a || !pixels_b || !request_ids || !submitted)
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        *submitted = 0;
        for (uint32_t i = 0; i < count; ++i) {
            const piv_status r = submit_impl(s, stream_ids[i], pixels_a[i], pixels_b[i], row_stride, true,
                                             request_ids + i);
            if (r != PIV_OK) return r;
            *submitted = i + 1;
        }
        return PIV_OK;
    });
}

piv_status piv_poll(piv_server *s, piv_result *out, uint32_t capacity, uint32_t *count, int wait_ms)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!s || !count || (!out && capacity)) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->done.empty() && wait_ms != 0) {
            auto pred = [&] { return !s->done.empty(); };
            if (wait_ms < 0) s->cv_done.wait(lk, pred);
            else s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
        }
        uint32_t n = 0;
        while (n < capacity && !s->done.empty()) {
            Request *r = s->done.front();
            s->done.pop_front();
            piv_result &o = out[n++];
            o = piv_result{};
            o.request_id = r->id;
            o.stream_id = r->st->id;
            o.status = r->status;
            o.grid = to_pub(r->st->plans.back().grid);
            if (r->status == PIV_OK) {
                o.dx = r->slot->h_dx;
                o.dy = r->slot->h_dy;
                o.peak_ratio = r->slot->h_ppr;
                o.flags = r->slot->h_flags;
            }
            o.summary = r->summary;
            o.submit_ms = r->t_submit;
            o.done_ms = r->t_done;
            o.batch_pairs = r->batch_pairs;
            s->polled[r->id] = r;
        }
        *count = n;
        return PIV_OK;
    });
}

piv_status piv_result_release(piv_server *s, uint64_t request_id)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!s) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->polled.find(request_id);
        if (it == s->polled.end()) return fail(PIV_ERR_INVALID_ARGUMENT, "unknown request %llu",
                                               (unsigned long long)request_id);
        Request *r = it->second;
        s->polled.erase(it);
        if (r->slot) {
            r->st->free_slots.push_back(r->slot);
            r->st->in_use -= 1;
        }
        delete r;
        s->cv_sched.notify_one();
        return PIV_OK;
    });
}

piv_status piv_server_pause(piv_server *s, int paused)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!s) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->paused = paused != 0;
        }
        s->cv_sched.notify_one();
        return PIV_OK;
    });
}

piv_status piv_server_stats_read(piv_server *s, piv_server_stats *out)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!s || !out) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->stats;
        out->hbm_allocated = s->allocated;
        out->hbm_peak = s->peak;
        out->queued = s->queue.size();
        return PIV_OK;
    });
}

}  // extern "C"

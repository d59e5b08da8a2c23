// piv-analyze: the `piv analyze` host loop (piv-cli/src/analyze.rs:243-523) over the C ABI of
// libpiv_amd, for frame stacks already decoded to a .npy file [frames, height, width] of u8,
// u16 or f32 samples (the ffmpeg decode of the reference is outside this library).
//
// Same frame range, pairing, background sampling and output as the reference: pairs
// (a, a + stride) for a in start, start + pair_step, ... < end - stride; background = min or
// mean of N frames sampled at k (total - 1) / (N - 1) across the whole clip; dt = stride / fps;
// pair time = capture time of the pair midpoint; one pair_AAAAAA_BBBBBB.npz per pair plus
// run.json in the output directory.  The pairs are analysed in batches on the GPU (one launch
// per kernel and pass for a whole batch).
#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/piv.h"

namespace {

// Work that must stop before the process exits (background upload threads, writer threads),
// run by die() in reverse order of registration.  die() is called from the main thread only.
std::vector<std::function<void()>> g_cleanup;

[[noreturn]] void die(const std::string &msg)
{
    std::fprintf(stderr, "error: %s\n", msg.c_str());
    static bool dying = false;
    if (!dying) {  // a hook that fails does not run the hooks again
        dying = true;
        for (auto it = g_cleanup.rbegin(); it != g_cleanup.rend(); ++it) (*it)();
    }
    std::exit(1);
}

// Runs `f` when the scope ends, also when an exception unwinds it.
struct OnExit {
    std::function<void()> f;
    ~OnExit()
    {
        if (f) f();
    }
};

void check(piv_status s, const char *what)
{
    if (s != PIV_OK) die(std::string(what) + ": " + piv_last_error());
}

struct Args {
    std::string input, output;
    bool force = false;
    long start = -1, end = -1;
    double start_s = -1, end_s = -1;
    int stride = 1, pair_step = 1;
    double fps = 0;  // capture rate; required for physical time (default 1 frame per second)
    bool roi = false;
    uint32_t rx = 0, ry = 0, rw = 0, rh = 0;
    std::string background = "min:40";
    float smooth_sigma = 1.0f;
    uint32_t normalize_radius = 4;
    std::vector<uint32_t> windows{96, 64};
    float overlap = 0.5f;
    uint32_t final_iterations = 2;
    float min_mean = 0.04f, min_std = 0.015f;
    float ppr = -1, median_threshold = -1, median_eps = -1;
    bool no_replace = false;
    std::string format = "npz";
    double px_per_mm = 0;
    uint32_t batch = 64;
    std::string precision = "fp32";
    std::string deformation = "none";
    std::string channel = "luma";
    int device = 0;
};

void usage()
{
    std::fprintf(stderr,
                 "usage: piv-analyze INPUT.npy -o OUTPUT_DIR [--force]\n"
                 "  frames:        --start N | --start-s S, --end N | --end-s S, --stride 1, --pair-step 1,\n"
                 "                 --capture-fps FPS (default 1)\n"
                 "  image:         --channel luma|green|yplane, --roi x,y,w,h,\n"
                 "                 --background none|min:N|mean:N (min:40),\n"
                 "                 --smooth-sigma 1, --normalize-radius 4\n"
                 "  interrogation: --windows 96,64, --overlap 0.5, --final-iterations 2,\n"
                 "                 --min-mean 0.04, --min-std 0.015,\n"
                 "                 --deformation none|bilinear|bicubic (between passes; none)\n"
                 "  validation:    --ppr-threshold, --median-threshold, --median-eps, --no-replace\n"
                 "  output:        --format npz (hdf5 needs libhdf5), --px-per-mm X\n"
                 "  device:        --batch 64, --precision fp32|fp16|bf16, --device 0\n");
}

Args parse(int argc, char **argv)
{
    Args a;
    auto need = [&](int &i) -> std::string {
        if (i + 1 >= argc) die(std::string("missing value for ") + argv[i]);
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        if (s == "-h" || s == "--help") {
            usage();
            std::exit(0);
        } else if (s == "-o" || s == "--output") a.output = need(i);
        else if (s == "--force") a.force = true;
        else if (s == "--start") a.start = std::stol(need(i));
        else if (s == "--end") a.end = std::stol(need(i));
        else if (s == "--start-s") a.start_s = std::stod(need(i));
        else if (s == "--end-s") a.end_s = std::stod(need(i));
        else if (s == "--stride") a.stride = std::stoi(need(i));
        else if (s == "--pair-step") a.pair_step = std::stoi(need(i));
        else if (s == "--capture-fps") {
            const std::string v = need(i);
            const size_t sl = v.find('/');
            a.fps = sl == std::string::npos ? std::stod(v)
                                            : std::stod(v.substr(0, sl)) / std::stod(v.substr(sl + 1));
        } else if (s == "--roi") {
            const std::string v = need(i);
            if (std::sscanf(v.c_str(), "%u,%u,%u,%u", &a.rx, &a.ry, &a.rw, &a.rh) != 4)
                die("--roi expects x,y,w,h");
            a.roi = true;
        } else if (s == "--background") a.background = need(i);
        else if (s == "--channel") a.channel = need(i);
        else if (s == "--smooth-sigma") a.smooth_sigma = std::stof(need(i));
        else if (s == "--normalize-radius") a.normalize_radius = (uint32_t)std::stoul(need(i));
        else if (s == "--windows") {
            a.windows.clear();
            std::string v = need(i);
            size_t p = 0;
            while (p <= v.size()) {
                const size_t c = v.find(',', p);
                a.windows.push_back((uint32_t)std::stoul(v.substr(p, c - p)));
                if (c == std::string::npos) break;
                p = c + 1;
            }
        } else if (s == "--overlap") a.overlap = std::stof(need(i));
        else if (s == "--final-iterations") a.final_iterations = (uint32_t)std::stoul(need(i));
        else if (s == "--min-mean") a.min_mean = std::stof(need(i));
        else if (s == "--min-std") a.min_std = std::stof(need(i));
        else if (s == "--ppr-threshold") a.ppr = std::stof(need(i));
        else if (s == "--median-threshold") a.median_threshold = std::stof(need(i));
        else if (s == "--median-eps") a.median_eps = std::stof(need(i));
        else if (s == "--no-replace") a.no_replace = true;
        else if (s == "--format") a.format = need(i);
        else if (s == "--px-per-mm") a.px_per_mm = std::stod(need(i));
        else if (s == "--batch") a.batch = (uint32_t)std::stoul(need(i));
        else if (s == "--precision") a.precision = need(i);
        else if (s == "--deformation") a.deformation = need(i);
        else if (s == "--device") a.device = std::stoi(need(i));
        else if (!s.empty() && s[0] == '-') die("unknown option " + s);
        else if (a.input.empty()) a.input = s;
        else die("unexpected argument " + s);
    }
    if (a.input.empty() || a.output.empty()) {
        usage();
        std::exit(2);
    }
    return a;
}

// A C-order .npy array [frames][height][width] (one plane) or [frames][height][width][3|4]
// (interleaved RGB / RGBA samples), memory-mapped.
struct Npy {
    const unsigned char *base = nullptr;
    size_t size = 0;
    const unsigned char *data = nullptr;
    piv_pixel_format fmt = PIV_PIXEL_U8;
    size_t bpp = 1;
    size_t frames = 0, height = 0, width = 0, channels = 1;
};

Npy open_npy(const std::string &path)
{
    Npy n;
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) die("cannot open " + path);
    struct stat st {};
    if (fstat(fd, &st) != 0) die("cannot stat " + path);
    n.size = (size_t)st.st_size;
    void *m = mmap(nullptr, n.size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (m == MAP_FAILED) die("cannot map " + path);
    n.base = static_cast<const unsigned char *>(m);
    if (n.size < 10 || std::memcmp(n.base, "\x93NUMPY", 6) != 0) die(path + " is not a .npy file");
    const int major = n.base[6];
    size_t hl, off;
    if (major == 1) {
        hl = n.base[8] | (n.base[9] << 8);
        off = 10;
    } else {
        hl = n.base[8] | (n.base[9] << 8) | ((size_t)n.base[10] << 16) | ((size_t)n.base[11] << 24);
        off = 12;
    }
    const std::string h(reinterpret_cast<const char *>(n.base + off), hl);
    n.data = n.base + off + hl;
    auto field = [&](const char *key) {
        const size_t p = h.find(key);
        if (p == std::string::npos) die(std::string(".npy header without ") + key);
        return p + std::strlen(key);
    };
    const size_t d = h.find('\'', field("'descr':")) + 1;
    const std::string descr = h.substr(d, h.find('\'', d) - d);
    if (descr == "|u1" || descr == "u1") n.fmt = PIV_PIXEL_U8, n.bpp = 1;
    else if (descr == "<u2") n.fmt = PIV_PIXEL_U16, n.bpp = 2;
    else if (descr == "<f4") n.fmt = PIV_PIXEL_F32, n.bpp = 4;
    else die("unsupported sample type " + descr + " (u8, u16 or f32)");
    if (h.find("True", field("'fortran_order':")) < h.find(',', field("'fortran_order':")))
        die("Fortran-order arrays are not supported");
    const size_t sp = h.find('(', field("'shape':")) + 1;
    const std::string shp = h.substr(sp, h.find(')', sp) - sp);
    std::vector<size_t> dims;
    size_t p = 0;
    while (p < shp.size()) {
        const size_t c = shp.find(',', p);
        const std::string t = shp.substr(p, c == std::string::npos ? std::string::npos : c - p);
        if (t.find_first_not_of(' ') != std::string::npos) dims.push_back(std::stoul(t));
        if (c == std::string::npos) break;
        p = c + 1;
    }
    if (dims.size() == 4 && (dims[3] == 3 || dims[3] == 4)) n.channels = dims[3];
    else if (dims.size() != 3)
        die("expected a [frames, height, width] or [frames, height, width, 3|4] array");
    n.frames = dims[0];
    n.height = dims[1];
    n.width = dims[2];
    if (n.data + n.frames * n.height * n.width * n.channels * n.bpp > n.base + n.size)
        die("truncated .npy file");
    return n;
}

std::string json_str(const std::string &s)
{
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}

}  // namespace

int run(int argc, char **argv);

int main(int argc, char **argv)
{
    // an exception unwinds run() (its upload and writer threads are stopped on the way) and
    // ends the process with the tool's own error status
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        die(std::string("unexpected error: ") + e.what());
    } catch (...) {
        die("unexpected error");
    }
}

int run(int argc, char **argv)
{
    const Args A = parse(argc, argv);
    const auto t0 = std::chrono::steady_clock::now();
    if (A.format == "hdf5")
        die("this build has no HDF5 support (libhdf5 is not available); use --format npz");
    if (A.format != "npz") die("unknown format " + A.format);
    const Npy in = open_npy(A.input);
    const double fps = A.fps > 0 ? A.fps : 1.0;
    // decode.rs:12-26,114-133 channel selection.  A one-plane clip already is the plane ffmpeg's
    // filters would have produced (its gray conversion, or the stored Y plane); an RGB clip
    // gives the green plane (gbrp + extractplanes=g is a plain copy for RGB sources).  ffmpeg's
    // gray conversion of RGB, and a Y plane of RGB, are not reproduced.
    if (A.channel != "luma" && A.channel != "green" && A.channel != "yplane")
        die("unknown channel " + A.channel + " (luma, green or yplane)");
    if (in.channels == 1 && A.channel == "green")
        die("--channel green needs a colour clip [frames, height, width, 3|4]; this clip has one "
            "plane");
    if (in.channels > 1 && A.channel != "green")
        die("--channel " + A.channel + " of an RGB clip is an ffmpeg conversion this tool does not "
            "reproduce; decode the clip to that plane, or use --channel green");

    // frame range (analyze.rs:251-269)
    auto to_frame = [&](double s) { return (long)std::max(0.0, std::round(s * fps)); };
    const long total = (long)in.frames;
    const long start = A.start >= 0 ? A.start : (A.start_s >= 0 ? to_frame(A.start_s) : 0);
    long end = A.end >= 0 ? A.end : (A.end_s >= 0 ? to_frame(A.end_s) : total);
    end = std::min(end, total);
    if (A.stride < 1 || A.pair_step < 1) die("--stride and --pair-step must be >= 1");
    if (!(start + A.stride < end))
        die("frame range [" + std::to_string(start) + ", " + std::to_string(end) +
            ") is too short for stride " + std::to_string(A.stride) + " (clip has " +
            std::to_string(total) + " frames)");
    std::vector<long> pair_starts;
    for (long a = start; a < end - A.stride; a += A.pair_step) pair_starts.push_back(a);
    const double dt = A.stride / fps;

    // output directory
    namespace fs = std::filesystem;
    if (fs::exists(A.output)) {
        if (!A.force) die(A.output + " exists; pass --force to overwrite");
        fs::remove_all(A.output);
    }
    fs::create_directories(A.output);

    // region of interest
    uint32_t H = (uint32_t)in.height, W = (uint32_t)in.width, x0 = 0, y0 = 0;
    if (A.roi) {
        if (A.rw == 0 || A.rh == 0 || A.rx + A.rw > W || A.ry + A.rh > H) die("--roi outside the frame");
        x0 = A.rx;
        y0 = A.ry;
        W = A.rw;
        H = A.rh;
    }
    const size_t row_stride = in.width * in.bpp;
    const size_t frame_bytes = in.height * row_stride;
    const size_t roi_off = (size_t)y0 * row_stride + (size_t)x0 * in.bpp;

    // analyzer (PivConfig / PreprocessConfig from the options)
    piv_config cfg;
    piv_config_default(&cfg);
    if (A.windows.empty() || A.windows.size() > PIV_MAX_PASSES) die("--windows: 1 to 16 sizes");
    for (size_t i = 0; i < A.windows.size(); ++i) cfg.windows[i] = A.windows[i];
    cfg.num_windows = (uint32_t)A.windows.size();
    cfg.overlap = A.overlap;
    cfg.final_iterations = A.final_iterations;
    cfg.min_mean = A.min_mean;
    cfg.min_std = A.min_std;
    if (A.ppr >= 0) cfg.ppr_threshold = A.ppr;
    if (A.median_threshold >= 0) cfg.median_threshold = A.median_threshold;
    if (A.median_eps >= 0) cfg.median_eps = A.median_eps;
    cfg.replace = A.no_replace ? 0 : 1;
    cfg.smooth_sigma = A.smooth_sigma;
    cfg.normalize_radius = A.normalize_radius;
    cfg.deformation = A.deformation == "none"       ? PIV_DEFORM_NONE
                      : A.deformation == "bilinear" ? PIV_DEFORM_BILINEAR
                      : A.deformation == "bicubic"  ? PIV_DEFORM_BICUBIC
                                                    : (die("unknown deformation " + A.deformation), PIV_DEFORM_NONE);
    piv_analyzer *an = nullptr;
    check(piv_analyzer_create(&cfg, A.device, &an), "analyzer");
    if (A.precision != "fp32") {
        const piv_precision p = A.precision == "fp16" ? PIV_PRECISION_FP16
                                : A.precision == "bf16" ? PIV_PRECISION_BF16
                                                        : (die("unknown precision " + A.precision), PIV_PRECISION_FP32);
        check(piv_analyzer_set_precision(an, p), "precision");
    }
    piv_grid grid;
    check(piv_output_grid(an, H, W, &grid), "grid");

    // the clip in HBM (one plane).  Uploaded in 32 MiB pieces by several threads at once (the
    // mapped file's pages are faulted in and staged by all of them), in the background: the
    // loop waits only for the frames each batch needs, so the upload of the rest of the clip
    // overlaps the analysis of its start.
    struct ClipUpload {
        const size_t kPiece = (size_t)32 << 20;
        int device = 0;
        char *dst = nullptr;
        const unsigned char *src = nullptr;
        size_t bytes = 0, pieces = 0, prefix = 0;  // pieces [0, prefix) are in HBM
        std::atomic<size_t> next{0};
        std::vector<char> done;
        std::mutex mu;
        std::condition_variable cv;
        std::string err;
        std::vector<std::thread> ts;
        void start(int dev, void *d, const unsigned char *s, size_t b)
        {
            device = dev;
            dst = static_cast<char *>(d);
            src = s;
            bytes = b;
            pieces = (b + kPiece - 1) / kPiece;
            done.assign(pieces, 0);
            const unsigned nt = (unsigned)std::min<size_t>(std::max<size_t>(pieces, 1), 8);
            for (unsigned i = 0; i < nt; ++i) ts.emplace_back([this] { work(); });
        }
        void work()
        {
            for (size_t k; (k = next.fetch_add(1)) < pieces;) {
                const size_t off = k * kPiece, len = std::min(kPiece, bytes - off);
                const bool ok = piv_copy_to_device(device, dst + off, src + off, len) == PIV_OK;
                std::lock_guard<std::mutex> lk(mu);
                if (!ok && err.empty()) err = piv_last_error();
                done[k] = 1;
                while (prefix < pieces && done[prefix]) ++prefix;
                cv.notify_all();
            }
        }
        // blocks until the first `b` bytes are in HBM
        void wait_bytes(size_t b)
        {
            std::string e;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !err.empty() || prefix == pieces || prefix * kPiece >= b; });
                e = err;
            }
            if (!e.empty()) die("upload: " + e);  // lock released: die() joins the workers
        }
        // workers claim no further piece; the pieces in progress finish, then the threads end
        void abort()
        {
            next.store(pieces);
            for (auto &t : ts)
                if (t.joinable()) t.join();
            ts.clear();
        }
        ~ClipUpload() { abort(); }
        void finish()
        {
            wait_bytes(bytes);
            for (auto &t : ts) t.join();
            ts.clear();
        }
    };
    ClipUpload up;
    g_cleanup.push_back([&up] { up.abort(); });
    void *clip = nullptr;
    check(piv_device_alloc(A.device, in.frames * frame_bytes, &clip), "device memory for the clip");
    if (in.channels == 1) {
        up.start(A.device, clip, in.data, in.frames * frame_bytes);
    } else {
        void *rgb = nullptr;
        const size_t bytes = in.frames * frame_bytes * in.channels;
        check(piv_device_alloc(A.device, bytes, &rgb), "device memory for the colour clip");
        up.start(A.device, rgb, in.data, bytes);
        up.finish();
        check(piv_extract_plane_device(A.device, rgb, in.fmt, in.frames * in.height * in.width,
                                       (uint32_t)in.channels, 1, clip),
              "green plane");
        piv_device_free(A.device, rgb);
    }
    auto frame_ptr = [&](long f) {
        return static_cast<const char *>(clip) + (size_t)f * frame_bytes + roi_off;
    };

    // background (analyze.rs:370-385): N frames sampled across the whole clip
    const float *bg_dev = nullptr;
    piv_background *bg = nullptr;
    std::string bg_spec = "none";
    if (A.background != "none") {
        const size_t c = A.background.find(':');
        const std::string mode = A.background.substr(0, c);
        long n = c == std::string::npos ? 40 : std::stol(A.background.substr(c + 1));
        if ((mode != "min" && mode != "mean") || n < 1) die("--background: none, min:N or mean:N");
        n = std::min(n, total);
        up.finish();  // frames sampled across the whole clip
        check(piv_background_create(an, mode == "min" ? PIV_BACKGROUND_MIN : PIV_BACKGROUND_MEAN, H, W, &bg),
              "background");
        for (long k = 0; k < n; ++k) {
            const long idx = n > 1 ? k * (total - 1) / (n - 1) : 0;
            check(piv_background_add(bg, frame_ptr(idx), in.fmt, row_stride, 1), "background frame");
        }
        check(piv_background_finish(bg, nullptr), "background");
        check(piv_background_device(bg, &bg_dev), "background");
        bg_spec = A.background;
        std::fprintf(stderr, "background: %s of %ld frames sampled across the clip\n", mode.c_str(), n);
    }

    // pairs, in batches
    const size_t cells = (size_t)grid.nx * grid.ny, B = std::max<uint32_t>(1, A.batch);
    void *ddx, *ddy, *dpr, *dfl;
    check(piv_device_alloc(A.device, B * cells * 4, &ddx), "outputs");
    check(piv_device_alloc(A.device, B * cells * 4, &ddy), "outputs");
    check(piv_device_alloc(A.device, B * cells * 4, &dpr), "outputs");
    check(piv_device_alloc(A.device, B * cells, &dfl), "outputs");
    std::vector<piv_summary> sums(B);
    std::map<long, piv_frame *> held;  // prepared frames still needed
    // The frames of a batch that no earlier batch prepared.  Evenly spaced runs of them are
    // prepared together (one launch per kernel for the run, piv_frame_prepare_device_batch);
    // the rest one by one.
    auto prepare_frames = [&](std::vector<long> need) {
        std::sort(need.begin(), need.end());
        need.erase(std::unique(need.begin(), need.end()), need.end());
        std::vector<long> miss;
        for (long f : need)
            if (!held.count(f)) miss.push_back(f);
        for (size_t i = 0; i < miss.size();) {
            size_t j = i + 1;  // [i, j): one evenly spaced run
            if (j < miss.size()) {
                const long d = miss[j] - miss[i];
                while (j + 1 < miss.size() && miss[j + 1] - miss[j] == d) ++j;
                ++j;
            }
            if (j - i >= 2) {
                const size_t step = (size_t)(miss[i + 1] - miss[i]) * frame_bytes;
                std::vector<piv_frame *> out(j - i);
                check(piv_frame_prepare_device_batch(an, frame_ptr(miss[i]), in.fmt, (uint32_t)(j - i), H,
                                                     W, row_stride, step, bg_dev, out.data()),
                      "frames");
                for (size_t k = i; k < j; ++k) held[miss[k]] = out[k - i];
            } else {
                piv_frame *pf = nullptr;
                check(piv_frame_prepare_device(an, frame_ptr(miss[i]), in.fmt, H, W, row_stride, bg_dev, &pf),
                      "frame");
                held[miss[i]] = pf;
            }
            i = j;
        }
    };

    // Pair archives are compressed and written by a pool of writer threads while the GPU
    // analyses the next batch; at most two batches of results wait for the writers.
    struct HostBatch {
        std::vector<float> dx, dy, ppr;
        std::vector<uint8_t> flags;
    };
    struct WriteJob {
        std::shared_ptr<HostBatch> hb;
        size_t k;
        long a;
        double time_s;
    };
    std::deque<WriteJob> jobs;
    std::mutex jmu;
    std::condition_variable jcv, jdone;
    bool closing = false;
    std::string write_error;
    const double m_per_px = A.px_per_mm > 0 ? 1e-3 / A.px_per_mm : 0.0;
    auto writer = [&]() {
        for (;;) {
            WriteJob j;
            {
                std::unique_lock<std::mutex> lk(jmu);
                jcv.wait(lk, [&] { return closing || !jobs.empty(); });
                if (jobs.empty()) return;
                j = std::move(jobs.front());
                jobs.pop_front();
            }
            const size_t o = j.k * cells;
            const piv_status st = piv_write_pair_npz(A.output.c_str(), (uint64_t)j.a, (uint64_t)(j.a + A.stride),
                                                     j.time_s, dt, m_per_px, x0, y0, &grid, j.hb->dx.data() + o,
                                                     j.hb->dy.data() + o, j.hb->ppr.data() + o,
                                                     j.hb->flags.data() + o);
            std::lock_guard<std::mutex> lk(jmu);
            if (st != PIV_OK && write_error.empty()) write_error = piv_last_error();
            jdone.notify_all();
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nwriters = std::min(16u, std::max(1u, hw - 1));
    std::vector<std::thread> writers;
    for (unsigned i = 0; i < nwriters; ++i) writers.emplace_back(writer);
    auto stop_writers = [&]() {
        {
            std::lock_guard<std::mutex> lk(jmu);
            closing = true;
        }
        jcv.notify_all();
        for (auto &t : writers) t.join();
        writers.clear();
    };
    OnExit writers_guard{stop_writers};
    g_cleanup.push_back(stop_writers);

    uint64_t valid = 0, vectors = 0;
    // seconds of the loop per phase: waiting for the clip's upload, preparation, analysis, flow
    // statistics, download, and waiting for the writers
    double ph_upload = 0, ph_prep = 0, ph_an = 0, ph_stats = 0, ph_down = 0, ph_wait = 0;
    auto lap = [](std::chrono::steady_clock::time_point &t) {
        const auto now = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(now - t).count();
        t = now;
        return s;
    };
    std::vector<double> series_t;
    std::vector<piv_flow_stats> series_stats;
    const auto t_loop = std::chrono::steady_clock::now();
    for (size_t p0 = 0; p0 < pair_starts.size(); p0 += B) {
        const size_t n = std::min(B, pair_starts.size() - p0);
        std::vector<long> need;
        for (size_t k = 0; k < n; ++k) {
            need.push_back(pair_starts[p0 + k]);
            need.push_back(pair_starts[p0 + k] + A.stride);
        }
        auto tp = std::chrono::steady_clock::now();
        up.wait_bytes((size_t)(*std::max_element(need.begin(), need.end()) + 1) * frame_bytes);
        ph_upload += lap(tp);
        prepare_frames(need);
        ph_prep += lap(tp);
        std::vector<const piv_frame *> fa(n), fb(n);
        for (size_t k = 0; k < n; ++k) {
            fa[k] = held.at(pair_starts[p0 + k]);
            fb[k] = held.at(pair_starts[p0 + k] + A.stride);
        }
        check(piv_analyze_batch_device(an, fa.data(), fb.data(), (uint32_t)n, (float *)ddx, (float *)ddy,
                                       (float *)dpr, (uint8_t *)dfl, sums.data(), nullptr, 0),
              "analysis");
        ph_an += lap(tp);
        // analyze.rs:483-489: per-pair flow statistics from the device-resident fields
        series_stats.resize(p0 + n);
        check(piv_flow_stats_device(an, &grid, (uint32_t)n, (const float *)ddx, (const float *)ddy,
                                    (const uint8_t *)dfl, dt, series_stats.data() + p0),
              "flow statistics");
        ph_stats += lap(tp);
        auto hb = std::make_shared<HostBatch>();
        hb->dx.resize(n * cells);
        hb->dy.resize(n * cells);
        hb->ppr.resize(n * cells);
        hb->flags.resize(n * cells);
        check(piv_copy_to_host(A.device, hb->dx.data(), ddx, n * cells * 4), "download");
        check(piv_copy_to_host(A.device, hb->dy.data(), ddy, n * cells * 4), "download");
        check(piv_copy_to_host(A.device, hb->ppr.data(), dpr, n * cells * 4), "download");
        check(piv_copy_to_host(A.device, hb->flags.data(), dfl, n * cells), "download");
        for (size_t i = 0; i < n * cells; ++i) valid += (hb->flags[i] & PIV_FLAG_INVALID) == 0;
        vectors += n * cells;
        ph_down += lap(tp);
        {
            std::unique_lock<std::mutex> lk(jmu);
            jdone.wait(lk, [&] { return jobs.size() <= B || !write_error.empty(); });
            if (!write_error.empty()) {
                lk.unlock();
                stop_writers();
                die("output: " + write_error);
            }
            for (size_t k = 0; k < n; ++k) {
                const long a = pair_starts[p0 + k];
                const double time_s = (a + A.stride / 2.0) / fps;
                series_t.push_back(time_s);
                jobs.push_back(WriteJob{hb, k, a, time_s});
            }
        }
        jcv.notify_all();
        ph_wait += lap(tp);
        // drop frames no later pair needs
        const long keep_from = p0 + n < pair_starts.size() ? pair_starts[p0 + n] : end;
        for (auto it = held.begin(); it != held.end();) {
            if (it->first < keep_from) {
                piv_frame_release(it->second);
                it = held.erase(it);
            } else {
                ++it;
            }
        }
    }
    {
        auto tw = std::chrono::steady_clock::now();
        stop_writers();
        ph_wait += lap(tw);
    }
    if (!write_error.empty()) die("output: " + write_error);
    up.finish();
    std::fprintf(stderr, "loop: waiting for the upload %.3f s, preparation %.3f s, analysis %.3f s, "
                         "flow statistics %.3f s, download %.3f s, waiting for %u writers %.3f s\n",
                 ph_upload, ph_prep, ph_an, ph_stats, ph_down, nwriters, ph_wait);
    // before the loop: opening the clip, the output directory, the analyzer (the device's
    // first use), the start of the clip's upload and the background estimate (which waits
    // for the whole upload)
    std::fprintf(stderr, "set-up: %.3f s before the loop\n",
                 std::chrono::duration<double>(t_loop - t0).count());
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
    for (auto &kv : held) piv_frame_release(kv.second);

    // analyze.rs:549: the run's flow statistics next to the pair archives
    check(piv_write_stats_npz(A.output.c_str(), (uint32_t)series_t.size(), series_t.data(),
                              series_stats.data()),
          "stats.npz");
    std::vector<float> noise;
    for (const auto &s : series_stats)
        if (std::isfinite(s.noise_px)) noise.push_back(s.noise_px);
    if (!noise.empty()) {
        std::nth_element(noise.begin(), noise.begin() + noise.size() / 2, noise.end());
        std::fprintf(stderr, "flow: neighbour noise median %.4f px over %zu pairs\n",
                     noise[noise.size() / 2], noise.size());
    }

    // run.json (write.rs:20-49 RunMetadata, the fields this host loop knows)
    std::string ws;
    for (size_t i = 0; i < A.windows.size(); ++i) ws += (i ? ", " : "") + std::to_string(A.windows[i]);
    char buf[4096];
    std::snprintf(
        buf, sizeof buf,
        "{\n  \"tool_version\": \"piv-amd %d\",\n  \"source_path\": %s,\n  \"source_size\": %zu,\n"
        "  \"camera\": null,\n  \"frame_width\": %zu,\n  \"frame_height\": %zu,\n  \"roi\": %s,\n"
        "  \"channel\": %s,\n"
        "  \"capture_fps\": %.17g,\n  \"start_frame\": %ld,\n  \"end_frame\": %ld,\n  \"stride\": %d,\n"
        "  \"pair_step\": %d,\n  \"dt_s\": %.17g,\n  \"background\": %s,\n"
        "  \"preprocess\": {\"smooth_sigma\": %.9g, \"normalize_radius\": %u, \"normalize_eps\": %.9g},\n"
        "  \"piv\": {\"windows\": [%s], \"overlap\": %.9g, \"final_iterations\": %u,\n"
        "          \"mask\": {\"min_mean\": %.9g, \"min_std\": %.9g},\n"
        "          \"validation\": {\"ppr_threshold\": %.9g, \"median_threshold\": %.9g, "
        "\"median_eps\": %.9g, \"replace\": %s}},\n"
        "  \"scale\": %s,\n  \"precision\": %s\n}\n",
        piv_abi_version(), json_str(A.input).c_str(), in.size, in.width, in.height,
        A.roi ? ("{\"x\": " + std::to_string(A.rx) + ", \"y\": " + std::to_string(A.ry) + ", \"w\": " +
                 std::to_string(A.rw) + ", \"h\": " + std::to_string(A.rh) + "}").c_str()
              : "null",
        json_str(A.channel).c_str(), fps, start, end, A.stride, A.pair_step, dt,
        bg_spec == "none" ? "null" : json_str(bg_spec).c_str(), cfg.smooth_sigma, cfg.normalize_radius,
        cfg.normalize_eps, ws.c_str(), cfg.overlap, cfg.final_iterations, cfg.min_mean, cfg.min_std,
        cfg.ppr_threshold, cfg.median_threshold, cfg.median_eps, cfg.replace ? "true" : "false",
        m_per_px > 0 ? ("{\"m_per_px\": " + std::to_string(m_per_px) + "}").c_str() : "null",
        json_str(A.precision).c_str());
    FILE *f = std::fopen((A.output + "/run.json").c_str(), "w");
    if (!f || std::fputs(buf, f) < 0 || std::fclose(f) != 0) die("writing run.json failed");

    const double all = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "analysed %zu pairs in %.3f s (%.1f pairs/s on the GPU loop, %.3f s total); "
                         "%.1f %% of %llu vectors valid -> %s\n",
                 pair_starts.size(), secs, pair_starts.size() / std::max(secs, 1e-9), all,
                 vectors ? 100.0 * valid / vectors : 0.0, (unsigned long long)vectors, A.output.c_str());
    piv_device_free(A.device, ddx);
    piv_device_free(A.device, ddy);
    piv_device_free(A.device, dpr);
    piv_device_free(A.device, dfl);
    piv_device_free(A.device, clip);
    if (bg) piv_background_destroy(bg);
    piv_analyzer_destroy(an);
    munmap(const_cast<unsigned char *>(in.base), in.size);
    return 0;
}

/*
 * piv.h -- C ABI of the MI355X PIV analysis engine (libpiv_amd.so).
 *
 * A drop-in for the per-pair hot path of the Rust `piv analyze` pipeline (agoodman/piv,
 * crate piv-core).  Plain C types only: pointers, sizes, POD structs.  Every call returns a
 * piv_status; the message of the last failure on the calling thread is piv_last_error().
 * No C++ exception crosses this boundary.  Handles are safe to use from several host threads
 * (calls on one analyzer are serialised internally).
 *
 * Each entry point names the reference interface it replaces (paths relative to
 * rust/crates/piv-core/src unless stated).
 */
#ifndef PIV_AMD_H
#define PIV_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIV_ABI_VERSION 2
#define PIV_MAX_PASSES 16

/* PivError (error.rs:3-20) plus device-side conditions. */
typedef enum {
    PIV_OK = 0,
    PIV_ERR_CONFIG = 1,          /* PivError::Config */
    PIV_ERR_SHAPE_MISMATCH = 2,  /* PivError::ShapeMismatch */
    PIV_ERR_IMAGE_TOO_SMALL = 3, /* PivError::ImageTooSmall */
    PIV_ERR_DEVICE = 4,          /* HIP runtime / kernel failure */
    PIV_ERR_OOM = 5,             /* device allocation failed */
    PIV_ERR_ADMISSION = 6,       /* serving: request refused by the HBM budget */
    PIV_ERR_CANCELLED = 7,       /* serving: request cancelled */
    PIV_ERR_INVALID_ARGUMENT = 8,/* null pointer, bad enum, buffer too small */
    PIV_ERR_UNSUPPORTED = 9      /* valid in the reference, not implemented on this device */
} piv_status;

/* Sample formats; u8 and u16 are normalised as decode.rs:227-234 does (v/255, v/65535). */
typedef enum { PIV_PIXEL_U8 = 0, PIV_PIXEL_U16 = 1, PIV_PIXEL_F32 = 2 } piv_pixel_format;

/* flags (field.rs:6-20) */
#define PIV_FLAG_NO_SIGNAL 1u
#define PIV_FLAG_LOW_PPR 2u
#define PIV_FLAG_MEDIAN_OUTLIER 4u
#define PIV_FLAG_REPLACED 8u
#define PIV_FLAG_OUT_OF_RANGE 16u
#define PIV_FLAG_INVALID 23u

/* Window deformation between passes (no counterpart in the reference, which only offsets the
 * windows rigidly by the predictor at the window centre, multipass.rs:207-231).  With
 * BILINEAR or BICUBIC, every pass after the first keeps that rigid placement and in addition
 * moves each pixel of window A by -delta and of window B by +delta, delta = (D(p) - D(c)) / 2,
 * where D is the bilinear interpolant of the previous pass's validated field, p the pixel's
 * position and c the window centre; the windows are resampled bilinearly or with 4x4 cubic
 * (Lagrange) weights.  A uniform predictor gives delta = 0 and, with BILINEAR, exactly the
 * rigid result.  NONE (the default) is the reference's method. */
typedef enum {
    PIV_DEFORM_NONE = 0,
    PIV_DEFORM_BILINEAR = 1,
    PIV_DEFORM_BICUBIC = 2
} piv_deformation;

/* PivConfig + MaskConfig (multipass.rs:22-63), ValidationConfig (validate.rs:7-37) and
 * PreprocessConfig (preprocess.rs:8-29) in one struct, plus the deformation mode. */
typedef struct {
    uint32_t windows[PIV_MAX_PASSES]; /* window size per pass, largest first */
    uint32_t num_windows;
    float overlap;                    /* [0, 1) */
    uint32_t final_iterations;        /* >= 1 */
    float min_mean, min_std;          /* MaskConfig */
    float ppr_threshold, median_threshold, median_eps; /* ValidationConfig */
    int32_t replace;
    float smooth_sigma;               /* PreprocessConfig */
    uint32_t normalize_radius;
    float normalize_eps;
    uint32_t deformation;             /* piv_deformation (no counterpart in the reference) */
} piv_config;

/* Grid (grid.rs:9-17) */
typedef struct {
    uint32_t window, step, nx, ny, x0, y0;
} piv_grid;

/* ValidationSummary (validate.rs:39-46) */
typedef struct {
    uint64_t no_signal, low_ppr, median_outliers, out_of_range, replaced;
} piv_summary;

typedef struct piv_analyzer piv_analyzer; /* PivAnalyzer (multipass.rs:85-88) */
typedef struct piv_frame piv_frame;       /* PreparedFrame (preprocess.rs:130-135), on device */

/* Caller-owned output of one pair: VectorField (field.rs:25-44) + summary. Arrays [ny*nx]. */
typedef struct {
    float *dx, *dy, *peak_ratio;
    uint8_t *flags;
    size_t capacity;     /* elements available in each array */
    piv_grid grid;       /* filled */
    piv_summary summary; /* filled (final pass, as PairResult::summary) */
} piv_field_out;

/* PassTrace (multipass.rs:72-81) plus the integer peak index that peak parity checks need.
 * All arrays caller-owned, [nx*ny] (origins [nx*ny*4]); any may be NULL. */
typedef struct {
    size_t capacity;
    piv_grid grid;             /* filled */
    float *origins;            /* [ax, ay, bx, by] per cell */
    float *raw_dx, *raw_dy, *raw_ppr;
    uint8_t *raw_flags;
    int32_t *peak_index;       /* iy*n + ix of the correlation maximum, -1 if no peak */
    float *dx, *dy, *ppr;      /* validated */
    uint8_t *flags;
    piv_summary summary;
} piv_pass_trace;

const char *piv_last_error(void);
int piv_abi_version(void);
piv_status piv_device_count(int *count);

/* PivConfig::default() etc.: windows [96, 64], overlap 0.5, final_iterations 2, mask 0/0,
 * validation 1.3 / 2 / 0.1 / replace, preprocessing sigma 1, radius 4, eps 0.02; no
 * deformation. */
void piv_config_default(piv_config *cfg);

/* Grid::new (grid.rs:20-44) */
piv_status piv_grid_new(uint32_t height, uint32_t width, uint32_t window, float overlap,
                        piv_grid *out);

/* PivAnalyzer::new (multipass.rs:91-119): validates the config, binds a device. */
piv_status piv_analyzer_create(const piv_config *cfg, int device, piv_analyzer **out);
void piv_analyzer_destroy(piv_analyzer *a);
/* Number of passes: windows ++ [last] * (final_iterations - 1) (multipass.rs:165-171) */
uint32_t piv_analyzer_num_passes(const piv_analyzer *a);

/* PivAnalyzer::output_grid (multipass.rs:126-133) */
piv_status piv_output_grid(const piv_analyzer *a, uint32_t height, uint32_t width,
                           piv_grid *out);

/* PreparedFrame::new (preprocess.rs:137-153) from HOST pixels (row_stride in bytes,
 * background: optional host float [h*w]).  The frame lives on the analyzer's device. */
piv_status piv_frame_prepare(piv_analyzer *a, const void *pixels, piv_pixel_format fmt,
                             uint32_t height, uint32_t width, size_t row_stride,
                             const float *background, piv_frame **out);
/* Same from DEVICE pixels already resident in HBM (no host copy).  When no conditioning is
 * configured (sigma 0, radius 0, no background) the frame references `pixels` without a copy
 * and the caller keeps them alive until piv_frame_release; otherwise it is conditioned into
 * frame-owned memory.  background: optional device float [h*w].  The preparation runs
 * asynchronously on the analyzer's stream (as PreparedFrame::new's result, the frame is
 * ready for every call of this library: analyses on the analyzer's stream follow it, one on
 * a stream the caller passes waits for it, and download, release and the table query
 * synchronise with it); `pixels` and `background` must stay unchanged until it has run. */
piv_status piv_frame_prepare_device(piv_analyzer *a, const void *pixels, piv_pixel_format fmt,
                                    uint32_t height, uint32_t width, size_t row_stride,
                                    const float *background, piv_frame **out);
/* `count` frames at pixels + i * frame_stride bytes (device memory, kept unchanged while the
 * frames live), prepared together: one launch per conditioning kernel for the whole set where
 * the configuration allows, asynchronously as piv_frame_prepare_device.  out[i] receives
 * frame i (release each with piv_frame_release). */
piv_status piv_frame_prepare_device_batch(piv_analyzer *a, const void *pixels,
                                          piv_pixel_format fmt, uint32_t count, uint32_t height,
                                          uint32_t width, size_t row_stride, size_t frame_stride,
                                          const float *background, piv_frame **out);
void piv_frame_release(piv_frame *f);
/* PreparedFrame's public `image` (and the background-subtracted base behind `stats`), as
 * host float [h*w]; either pointer may be NULL. */
piv_status piv_frame_download(const piv_frame *f, float *image, float *base);
/* Whether the frame's signal-mask tables (IntegralImage, preprocess.rs:84-124) exist on the
 * device: *built = 0 when the frame has no tables (the mask needs none) or when they were
 * screened out.  No counterpart in the reference (a diagnostic): with the default mask a u16
 * frame without conditioning gets tables only if it holds an aligned all-zero tile of half
 * the analyzer's smallest window, since only all-zero windows can depend on them. */
piv_status piv_frame_tables_built(const piv_frame *f, int *built);

/* BackgroundEstimator (preprocess.rs:31-72): per-pixel min or mean over sample frames, on
 * the analyzer's device.  add() takes HOST samples (on_device = 0) or DEVICE samples
 * (on_device = 1), normalised as decode.rs:227-234; finish() downloads the estimate (host
 * float [h*w]) and keeps it on the device, where piv_background_device() returns it for
 * piv_frame_prepare_device. */
typedef enum { PIV_BACKGROUND_MIN = 0, PIV_BACKGROUND_MEAN = 1 } piv_background_mode;
typedef struct piv_background piv_background;
piv_status piv_background_create(piv_analyzer *a, piv_background_mode mode, uint32_t height,
                                 uint32_t width, piv_background **out);
piv_status piv_background_add(piv_background *bg, const void *pixels, piv_pixel_format fmt,
                              size_t row_stride, int on_device);
piv_status piv_background_finish(piv_background *bg, float *out);
piv_status piv_background_device(const piv_background *bg, const float **device_ptr);
void piv_background_destroy(piv_background *bg);

/* PivAnalyzer::analyze (multipass.rs:136-138) */
piv_status piv_analyze_pair(piv_analyzer *a, const piv_frame *fa, const piv_frame *fb,
                            piv_field_out *out);
/* PivAnalyzer::analyze_traced (multipass.rs:142-150); traces[num_traces] with
 * num_traces == piv_analyzer_num_passes */
piv_status piv_analyze_traced(piv_analyzer *a, const piv_frame *fa, const piv_frame *fb,
                              piv_field_out *out, piv_pass_trace *traces, uint32_t num_traces);

/* Batched analysis of `npairs` independent pairs (the CLI's per-pair loop, analyze.rs:422-523,
 * run as one launch per pass): fa[i], fb[i] -> DEVICE outputs dx/dy/ppr/flags of
 * [npairs][ny][nx] and summaries [npairs] (host, may be NULL).  `stream` is a hipStream_t
 * (NULL: the analyzer's stream); the call returns once the work is enqueued when
 * `async` != 0, else after it completes. */
piv_status piv_analyze_batch_device(piv_analyzer *a, const piv_frame *const *fa,
                                    const piv_frame *const *fb, uint32_t npairs, float *dx,
                                    float *dy, float *ppr, uint8_t *flags,
                                    piv_summary *summaries, void *stream, int async);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the analysis
 * of piv_analyze_batch_device, with the same kernels and bitwise the same results, on `npairs`
 * pairs of raw DEVICE samples -- pair i is pixels_a + i * frame_stride and pixels_b +
 * i * frame_stride (a clip of consecutive pairs: pixels_b = pixels_a + frame_stride) -- into
 * DEVICE outputs dx/dy/ppr/flags [npairs][ny][nx] and, when not NULL, device summaries
 * uint32 [npairs][5] (no_signal, low_ppr, median_outliers, out_of_range, replaced).
 * All workspace (per-pass fields, descriptors, mask tables) lives in one 64 MiB device array
 * per GPU that the library holds for the life of the process: the call allocates, frees,
 * copies and synchronises nothing; it only enqueues kernels on `stream` (NULL: the analyzer's
 * stream) and returns.  A call whose workspace exceeds 64 MiB fails with
 * PIV_ERR_INVALID_ARGUMENT (analyse fewer pairs per call); a configuration with conditioning
 * (smooth_sigma > 0 or normalize_radius > 0) fails with PIV_ERR_UNSUPPORTED.  Every call on a
 * GPU shares that GPU's workspace: calls that overlap on one GPU race, so the caller
 * serialises them (one stream, or waits between streams). */
piv_status piv_analyze_batch_noalloc(piv_analyzer *a, const void *pixels_a, const void *pixels_b,
                                     piv_pixel_format fmt, uint32_t npairs, uint32_t height,
                                     uint32_t width, size_t row_stride, size_t frame_stride,
                                     float *dx, float *dy, float *ppr, uint8_t *flags,
                                     uint32_t *summaries, void *stream);

/* ---- stage-level entry points (parity tests replay each stage from the Rust fixtures) --- */

/* multipass.rs:207-231 window placement for `grid`; prev == NULL on the first pass.
 * prev_dx/prev_dy host float [prev->ny*prev->nx]; origins host float [grid->nx*ny*4]. */
piv_status piv_place_windows(piv_analyzer *a, const piv_grid *grid, uint32_t height,
                             uint32_t width, const piv_grid *prev, const float *prev_dx,
                             const float *prev_dy, float *origins);

/* PivAnalyzer::pass given origins (multipass.rs:233-259 with correlate.rs:99-139): mask,
 * correlation and assembly of the raw field.  origins host [count*4]; outputs host [count]
 * (peak_index may be NULL). */
piv_status piv_correlate_windows(piv_analyzer *a, const piv_frame *fa, const piv_frame *fb,
                                 uint32_t window, const float *origins, uint32_t count,
                                 float *dx, float *dy, float *ppr, uint8_t *flags,
                                 int32_t *peak_index);

/* validate (validate.rs:51-130) in place on a host field [ny*nx] using the analyzer's
 * ValidationConfig and max_displacement. */
piv_status piv_validate_field(piv_analyzer *a, uint32_t ny, uint32_t nx, float *dx, float *dy,
                              const float *ppr, uint8_t *flags, float max_displacement,
                              piv_summary *summary);

/* ---- device memory plumbing (no counterpart in the reference) --------------------------
 * For callers without their own HIP allocator (tests, the Python binding): plain device
 * buffers and copies on the given device, and a wait for an analyzer's queued work. */
piv_status piv_device_alloc(int device, size_t bytes, void **out);
piv_status piv_device_free(int device, void *ptr);
piv_status piv_copy_to_device(int device, void *dst, const void *src, size_t bytes);
piv_status piv_copy_to_host(int device, void *dst, const void *src, size_t bytes);
piv_status piv_analyzer_synchronize(piv_analyzer *a);

/* ---- output -------------------------------------------------------------------------------
 * NpzWriter::write (piv-io/src/write.rs:129-165): <dir>/pair_AAAAAA_BBBBBB.npz, deflated, with
 * x, y (grid centres + ROI offset), dx, dy, u, v (= d / dt), u_mps, v_mps (only when
 * m_per_px > 0), peak_ratio, flags, frames [a, b], t (pair midpoint, seconds).  Host arrays of
 * grid->ny * grid->nx values. */
piv_status piv_write_pair_npz(const char *dir, uint64_t frame_a, uint64_t frame_b, double time_s,
                              double dt_s, double m_per_px, uint32_t roi_x, uint32_t roi_y,
                              const piv_grid *grid, const float *dx, const float *dy,
                              const float *peak_ratio, const uint8_t *flags);

/* ---- measurement (no counterpart in the reference) ------------------------------------ */

/* Per-kernel timing: when enabled, every kernel the analyzer launches is bracketed by HIP
 * events recorded on the stream it is launched on.  `bytes` is the algorithmic traffic of
 * the launches (window-gather model: 2 n^2 4 + 49 bytes per window for a correlation pass). */
typedef struct {
    char name[32];       /* "correlate", "validate", "prepare" */
    uint32_t window;     /* window size for correlate, else 0 */
    uint64_t launches;
    uint64_t units;      /* windows (correlate) or vectors (validate) processed */
    double total_ms;
    double bytes;
} piv_kernel_stat;

/* enable != 0 turns timing on and clears the counters; 0 turns it off. */
piv_status piv_profile_enable(piv_analyzer *a, int enable);
/* Waits for the analyzer's work, then copies up to `capacity` stats; *count = total. */
piv_status piv_profile_read(piv_analyzer *a, piv_kernel_stat *out, uint32_t capacity,
                            uint32_t *count);

/* ---- serving (no counterpart in the reference, which analyses one pair per call:
 *      piv-cli/src/analyze.rs:428-441) ------------------------------------------------------
 * A server runs a device-resident request loop with continuous batching.  Each stream is one
 * frame source with its own frame size, sample format, pass schedule, conditioning and
 * background (kept in HBM).  Pairs submitted while others are on the device join the next
 * batch as soon as a compute lane is free; a batch runs one launch per kernel and window size
 * per pass step for all its requests, whatever their streams.  Every result is the one
 * piv_analyze_pair gives for the same pair and configuration (pairs are independent). */
typedef struct piv_server piv_server;

typedef struct {
    size_t hbm_budget;         /* device bytes request slots may hold at once (0: 4 GiB).  A
                                * released slot stays cached for its stream; another stream
                                * takes it over the budget only while the owner caches more
                                * slots than it has requests in flight. */
    uint32_t max_batch_pairs;  /* pairs per batch (0: 64) */
    uint32_t max_inflight;     /* batches on the device at once, one HIP stream each (0: 2) */
} piv_server_config;

typedef struct {
    uint64_t request_id;
    uint32_t stream_id;
    piv_status status;
    piv_grid grid;                   /* final pass grid: ny * nx values below */
    const float *dx, *dy, *peak_ratio;
    const uint8_t *flags;            /* host memory owned by the server until release */
    piv_summary summary;
    double submit_ms, done_ms;       /* server clock: submission and result on the host */
    uint32_t batch_pairs;            /* pairs of the batch it ran in */
} piv_result;

typedef struct {
    uint64_t submitted, completed, rejected, batches, batched_pairs;
    size_t hbm_allocated, hbm_peak;  /* request slots (cached free slots included) */
    size_t queued;
} piv_server_stats;

piv_status piv_server_create(int device, const piv_server_config *cfg, piv_server **out);
/* Finishes the batches on the device, then frees everything (unpolled results included). */
void piv_server_destroy(piv_server *s);
/* background: host f32 [height * width] or null; it is uploaded once and stays in HBM. */
piv_status piv_server_add_stream(piv_server *s, const piv_config *cfg, uint32_t height,
                                 uint32_t width, piv_pixel_format fmt, const float *background,
                                 uint32_t *stream_id);
/* Copies the pair (row_stride bytes per row, 0 = packed) and queues it.  PIV_ERR_ADMISSION:
 * the HBM budget or the queue is full and nothing was queued (retry after releasing results).
 * Concurrent calls, for one stream or several, stage their copies in parallel; pairs of 3 MiB
 * or more are staged by the calling thread and the server's staging threads together. */
piv_status piv_submit(piv_server *s, uint32_t stream_id, const void *pixels_a,
                      const void *pixels_b, size_t row_stride, uint64_t *request_id);
/* As piv_submit for frames already in device memory on the server's GPU (complete when
 * called): they are read in place, so keep them unchanged until the result is released. */
piv_status piv_submit_device(piv_server *s, uint32_t stream_id, const void *pixels_a,
                             const void *pixels_b, size_t row_stride, uint64_t *request_id);
/* `count` requests of frames in device memory in one call (each as piv_submit_device, in
 * order), for a caller with many pairs ready.  Stops at the first request the server cannot
 * admit (HBM budget or full queue): returns PIV_ERR_ADMISSION with *submitted the number
 * admitted (their ids in request_ids), and the caller retries the rest after polling.  No
 * counterpart in the reference, whose CLI analyses one pair per call. */
piv_status piv_submit_device_many(piv_server *s, uint32_t count, const uint32_t *stream_ids,
                                  const void *const *pixels_a, const void *const *pixels_b,
                                  size_t row_stride, uint64_t *request_ids, uint32_t *submitted);
/* Up to `capacity` finished results, oldest first; waits up to wait_ms (-1: forever) for one. */
piv_status piv_poll(piv_server *s, piv_result *out, uint32_t capacity, uint32_t *count,
                    int wait_ms);
/* Returns a polled result's memory (and its request slot) to the server. */
piv_status piv_result_release(piv_server *s, uint64_t request_id);
/* paused != 0: admitted requests wait and no batch is launched (batches on the device
 * finish); 0 resumes.  Used to drain or to hand a burst to the device as full batches. */
piv_status piv_server_pause(piv_server *s, int paused);
piv_status piv_server_stats_read(piv_server *s, piv_server_stats *out);

/* ---- correlation precision (no counterpart in the reference, which is FP32) ----------- */

typedef enum {
    PIV_PRECISION_FP32 = 0, /* FFT correlation in FP32: the reference's arithmetic */
    PIV_PRECISION_FP16 = 1, /* DFT-matrix GEMMs on the matrix cores, fp16 operands */
    PIV_PRECISION_BF16 = 2  /* the same with bf16 operands */
} piv_precision;

/* FP16 / BF16: passes with 16, 32 or 64 px windows compute the correlation as dense DFT-matrix
 * GEMMs (16-bit operands, FP32 accumulation, windows scaled to unit RMS), fit the peak and
 * accumulate the displacement in FP64; other window sizes keep the FP32 FFT.  Results are no
 * longer bit-comparable with the reference: integer peaks can move on near-ties and the
 * sub-pixel displacement carries the 16-bit rounding (see the DFT-GEMM accuracy tests). */
piv_status piv_analyzer_set_precision(piv_analyzer *a, piv_precision p);

/* ---- autotuning (no counterpart in the reference) -------------------------------------- */

/* The correlation kernel of each window size has launch variants (occupancy targets of the
 * register allocator).  Variants differ only in scheduling: every variant gives bit-identical
 * results.  piv_autotune sweeps the variants twice; per sweep it runs the analyzer's pass plan
 * on the caller's batch once untimed and `repeats` times timed per variant.  It keeps the
 * fastest variant per window size (each timed by its faster sweep) for the frames' sample
 * format on this analyzer, and reports one entry per distinct window size. */
typedef struct {
    uint32_t window;
    int32_t fmt;             /* piv_pixel_format of the frames */
    uint32_t chosen;         /* variant index now used by the analyzer */
    uint32_t num_variants;
    uint32_t occupancy[4];   /* waves per SIMD targeted by each variant */
    float ms[4];             /* mean kernel time per launch of each variant (faster sweep) */
} piv_tune_entry;

/* Number of variants of the correlation kernel for a window size (0 if unsupported). */
uint32_t piv_kernel_variant_count(uint32_t window);
/* Pin a variant for (window, fmt) on this analyzer (as piv_autotune would choose it). */
piv_status piv_kernel_variant_set(piv_analyzer *a, uint32_t window, piv_pixel_format fmt,
                                  uint32_t variant);

piv_status piv_autotune(piv_analyzer *a, const piv_frame *const *frames_a,
                        const piv_frame *const *frames_b, uint32_t npairs, uint32_t repeats,
                        piv_tune_entry *out, uint32_t capacity, uint32_t *count);

/* Batch lanes: piv_analyze_batch* split a batch's pairs into this many groups (1..4, at least
 * 4 pairs each) whose pass chains run side by side on helper streams, so that one group's
 * short per-vector kernels and launch tails overlap another group's correlation.  Results do
 * not depend on it.  The default is 2, or PIV_BATCH_LANES in the environment at analyzer
 * creation, which also pins it.  With PIV_TUNE_LANES=1 at creation (and no pin),
 * piv_autotune ends by timing the caller's batch (graph replays, two sweeps) at each count
 * and keeps the fastest. */
piv_status piv_batch_lanes_set(piv_analyzer *a, uint32_t lanes);
piv_status piv_batch_lanes_get(const piv_analyzer *a, uint32_t *lanes);

/* Frame source channel selection on decoded clips (piv-io/src/decode.rs:12-26,124-133): one
 * plane of `count` pixels of interleaved samples, device to device: out[i] = pixels[i *
 * channels + plane].  The green plane of an RGB clip is ffmpeg's `format=gbrp,extractplanes=g`
 * of an RGB source (channels 3 or 4, plane 1). */
piv_status piv_extract_plane_device(int device, const void *pixels, piv_pixel_format fmt,
                                    size_t count, uint32_t channels, uint32_t plane, void *out);

/* decode.rs:227-234 sample normalisation on the device (host in/out), for parity checks. */
piv_status piv_normalize_pixels(int device, const void *pixels, piv_pixel_format fmt,
                                size_t count, float *out);

/* ---- flow statistics of validated fields ----------------------------------------------- */

/* FlowStats (flowstats.rs:9-25) plus the neighbour noise estimate (flowstats.rs:103-137). */
typedef struct {
    float vorticity_mean;     /* 1/s for px/s input; + = counter-clockwise on screen; NaN if none */
    float vorticity_abs_mean; /* mean |vorticity|, NaN if none */
    float ke_mean;            /* 1/2 <u^2 + v^2> over usable vectors (px^2/s^2), NaN if none */
    float ke_sum;             /* 1/2 sum (u^2 + v^2) step^2 (px^4/s^2), 0 if none */
    float coverage;           /* usable vectors / grid cells */
    float noise_px;           /* neighbour_noise of dx, dy (px); NaN with < 20 residuals */
    uint32_t n_usable, n_vorticity;
} piv_flow_stats;

/* Replaces the per-pair statistics of `piv analyze` (piv-cli/src/analyze.rs:483-489): usable
 * vectors (piv-io/src/flow.rs:13-17: finite and accepted or REPLACED), pair_stats (:20-34: velocity
 * = dx * f32(1 / dt_s), grid step as the spacing) and neighbour_noise of the raw displacements
 * (flowstats.rs:106-137), for `npairs` validated fields of one grid, device-resident
 * [npairs][ny][nx] (dx, dy in px, flags).  `out` is host memory [npairs].  Runs on the
 * analyzer's stream and device; profiled as "flowstats". */
piv_status piv_flow_stats_device(piv_analyzer *a, const piv_grid *grid, uint32_t npairs,
                                 const float *dx, const float *dy, const uint8_t *flags,
                                 double dt_s, piv_flow_stats *out);

/* piv-io/src/flow.rs:133-145 write_stats_npz: `dir`/stats.npz with t (f64 [npairs], pair midpoint time)
 * and vorticity_mean, vorticity_abs_mean, ke_mean, ke_sum, coverage (f32 [npairs]), deflated. */
piv_status piv_write_stats_npz(const char *dir, uint32_t npairs, const double *t,
                               const piv_flow_stats *stats);

/* ---- time-series filters of a run (smooth.rs) ------------------------------------------ */
/* Series are device-resident [nt][ny][nx] (pair, row, column); masks are u8 0/1.  Each call
 * runs on its own stream of `device` and returns when the results are written. */

typedef enum { PIV_TEMPORAL_MEAN = 0, PIV_TEMPORAL_MEDIAN = 1 } piv_temporal_kernel;

/* smooth.rs:44-130 temporal(): moving mean or median over `window` (odd, <= 63) consecutive
 * pairs of usable vectors, outputs only where `gate`, NaN below max(min_support, 1) samples;
 * leave_out: offsets skipped (each non-zero one replaced by the next further out). */
piv_status piv_smooth_temporal_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                      const float *dx, const float *dy, const uint8_t *usable,
                                      const uint8_t *gate, piv_temporal_kernel kernel,
                                      uint32_t window, uint32_t min_support,
                                      const int32_t *leave_out, uint32_t n_leave, float *out_dx,
                                      float *out_dy, uint16_t *support);

/* smooth.rs:137-208 spatial(): per-pair Gaussian normalized convolution, sigma in cells
 * (> 0, radius ceil(2.5 sigma) <= 32); leave_out >= 0 skips that Chebyshev radius, < 0 none. */
piv_status piv_smooth_spatial_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                     const float *dx, const float *dy, const uint8_t *gate,
                                     float sigma, float min_weight, int32_t leave_out,
                                     float *out_dx, float *out_dy, uint16_t *support);

/* smooth.rs:303-310 mask(): NaN where not usable. */
piv_status piv_mask_device(int device, size_t n, const float *values, const uint8_t *usable,
                           float *out);

/* smooth.rs:213-232 cross_validation_errors(): |raw - prediction| of the usable vectors in
 * index order into out[0 .. *count) (device); *count on the host. */
piv_status piv_cross_validation_errors_device(int device, size_t n, const float *raw_dx,
                                              const float *raw_dy, const uint8_t *usable,
                                              const float *pred_dx, const float *pred_dy,
                                              float *out, uint64_t *count);

/* smooth.rs:236-249 median_error(): median of the finite errors where keep (NULL: all);
 * NaN and 0 when there are none.  Results on the host. */
piv_status piv_median_error_device(int device, const float *errors, size_t n, const uint8_t *keep,
                                   float *median, uint64_t *count);

/* smooth.rs:252-274 pair_to_pair_change() and 277-300 spatial_roughness() (host result). */
piv_status piv_pair_to_pair_change_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                          const float *dx, const float *dy, float *out);
piv_status piv_spatial_roughness_device(int device, uint32_t nt, uint32_t ny, uint32_t nx,
                                        const float *dx, const float *dy, float *out);

#ifdef __cplusplus
}
#endif
#endif /* PIV_AMD_H */

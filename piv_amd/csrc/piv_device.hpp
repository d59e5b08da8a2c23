// Structures shared by the host orchestration (piv_api.cpp) and the HIP kernels.
#pragma once

#include <cstdint>

namespace piv {

enum PixelFormat : int32_t { FMT_U8 = 0, FMT_U16 = 1, FMT_F32 = 2 };

// Grid::new result (piv-core/src/grid.rs:9-44)
struct GridDev {
    int32_t window, step, nx, ny, x0, y0;
};

// flags (piv-core/src/field.rs:6-20)
constexpr uint8_t NO_SIGNAL = 1, LOW_PPR = 2, MEDIAN_OUTLIER = 4, REPLACED = 8, OUT_OF_RANGE = 16;
constexpr uint8_t INVALID = NO_SIGNAL | LOW_PPR | MEDIAN_OUTLIER | OUT_OF_RANGE;

// Correlation peak of one window (correlate.rs:215-303 before the sub-pixel fit): the
// correlation kernel writes it, the finishing kernel turns it into the raw vector.
struct PeakRec {
    float best, second;       // highest value, highest local maximum outside the footprint
    float mx, qx, my, qy;     // neighbours of the peak at -1 / +1 along x and along y
    int32_t index;            // integer peak iy*n + ix
    int32_t has_peak;         // 0: masked or flat window (NO_SIGNAL)
};

// Summed-area table layout: a table of (H+1) x (W+1) f64 entries with a row pitch of
// sat_pitch(W) entries (a multiple of 16, 128 B) and entry (0, 0) SAT_LEAD entries past a
// 128 B boundary, so that table column 1 -- the first column of every 32-column strip the
// table kernels write -- starts a 128 B line in every row.  A frame's sum and sum-of-squares
// tables are sat_table_elems(H, W) entries apart.
constexpr int SAT_ALIGN = 16, SAT_LEAD = SAT_ALIGN - 1;
__host__ __device__ constexpr int64_t sat_pitch(int W)
{
    return ((int64_t)W + 1 + SAT_ALIGN - 1) / SAT_ALIGN * SAT_ALIGN;
}
__host__ __device__ constexpr int64_t sat_table_elems(int H, int W)
{
    return SAT_ALIGN + ((int64_t)H + 1) * sat_pitch(W);
}

// One image pair's share of one correlation pass.  A launch covers many pairs (the whole
// batch, or windows of one size from several serving requests): block b of a launch with
// bpp blocks per pair works on pair b / bpp.
struct PassPair {
    const void *img_a;        // conditioned image (FMT_F32) or raw pixels (fused u8/u16/f32)
    const void *img_b;
    const double *sat_a;      // f64 summed-area tables of the background-subtracted frame:
    const double *sat_b;      //   entry (0, 0) of the sums; the squares sat_cells further on
    const float *prev_dx;     // previous pass's validated field (predictor), or null
    const float *prev_dy;
    float *origins;           // [nx*ny*4] window origins (ax, ay, bx, by): written by the
                              //   placement kernel, or given by the caller (origins_given)
    uint8_t *sig;             // [nx*ny] 1 if the window passes the signal mask
    PeakRec *peaks;           // [nx*ny] correlation peaks
    float *out_dx, *out_dy, *out_ppr;
    uint8_t *out_flags;
    int32_t *out_peak;        // optional [nx*ny] integer peak iy*n+ix, -1 if no peak
    int64_t sat_cells;        // sat_table_elems(H, W)
    int32_t H, W, pitch;      // pitch in pixels
    int32_t fmt;
    GridDev grid, prev;
    int32_t has_prev;
    int32_t use_mask;
    int32_t origins_given;
    float min_mean, min_std;
    int32_t deform;           // window deformation from the predictor (passes with has_prev):
                              //   0 none (rigid offset), 1 bilinear, 2 bicubic resampling
    int32_t screened;         // bit 0 (frame A), bit 1 (frame B): the frame's tables were
                              //   screened; its state word (sat_state) says whether they
                              //   were built, and unbuilt tables pass the (default) mask
};

// Screened mask tables (u16 frames without conditioning, masks with min_mean <= 0 and
// min_std <= 0): the SAT_LEAD entries before entry (0, 0) of a frame's sums hold a u32 state
// word (byte 0: 1 if the tables were built) and one flag byte per 32-row band (from byte
// SCREEN_FLAGS: 1 if the band holds an aligned all-zero tile).
constexpr int SCREEN_FLAGS = 8;
constexpr int SCREEN_MAX_BANDS = SAT_LEAD * 8 - SCREEN_FLAGS;
__host__ __device__ inline uint32_t *sat_state_word(double *sum)
{
    return reinterpret_cast<uint32_t *>(sum - SAT_LEAD);
}
__host__ __device__ inline const uint32_t *sat_state_word(const double *sum)
{
    return reinterpret_cast<const uint32_t *>(sum - SAT_LEAD);
}
__host__ __device__ inline uint8_t *screen_flags(double *sum)
{
    return reinterpret_cast<uint8_t *>(sum - SAT_LEAD) + SCREEN_FLAGS;
}
__host__ __device__ inline const uint8_t *screen_flags(const double *sum)
{
    return reinterpret_cast<const uint8_t *>(sum - SAT_LEAD) + SCREEN_FLAGS;
}

// One image pair's validation job (piv-core/src/validate.rs:51-130).
struct ValidatePair {
    const float *raw_dx, *raw_dy, *raw_ppr;
    const uint8_t *raw_flags;
    float *dx, *dy, *ppr;
    uint8_t *flags;
    uint8_t *snapshot;        // flags after the median test ([nx*ny] scratch)
    uint32_t *summary;        // 5 counters: no_signal, low_ppr, median_outliers, out_of_range, replaced
    int32_t nx, ny;
    float ppr_threshold, median_threshold, median_eps, max_displacement;
    int32_t replace;
    int32_t pad_;
};

}  // namespace piv

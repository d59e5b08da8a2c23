// HIP kernels for the PIV per-pair path on gfx950 (MI355X).
//
//   k_correlate<N>  (correlate.hpp, instantiated per size in corr_n*.hip) one launch per pass for every window of every pair in the launch:
//                   predictor sample + symmetric window placement, signal mask from the
//                   f64 summed-area tables, fused pixel ingest (u8/u16 normalisation in
//                   registers) + bilinear window load + mean removal, packed 2-D FFT
//                   cross-correlation held in VGPRs with LDS transposes, argmax, peak
//                   footprint, secondary local maximum, 3-point Gaussian sub-pixel fit.
//   k_validate      normalised-median validation + replacement on 16x16 tiles of the
//                   vector grid with a 3-cell halo in LDS (one launch per pass).
//   k_*prep*        frame conditioning: normalise / background subtract, f64 SAT,
//                   separable Gaussian blur, local normalisation.
//
// Scalar arithmetic that the reference specifies (placement, sampling, bilinear weights,
// validation, preprocessing) is evaluated in the reference's operation order and compiled
// with -ffp-contract=off, so it is bit-exact with piv-core; the FFT and peak fit use their
// own (fma) arithmetic and agree to FP32 rounding.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "piv_common.hpp"
#include "piv_kernels.hpp"
#include "corr_sizes.hpp"

namespace piv {

#define PIV_DECL_FFT(N)                                                                       \
    hipError_t launch_correlate_##N(int fmt, int variant, int deform, const PassPair *pairs,       \
                                    int npairs, uint32_t max_cells, hipStream_t st);
#define PIV_DECL_MF(N)                                                                        \
    hipError_t launch_correlate_mf_##N(int fmt, int precision, int deform, const PassPair *pairs, \
                                       int npairs, uint32_t max_cells, hipStream_t st);
PIV_FFT_SIZES(PIV_DECL_FFT)
PIV_MF_SIZES(PIV_DECL_MF)
hipError_t launch_correlate_any_32(int fmt, int deform, const PassPair *pairs, int npairs,
                                   uint32_t max_cells, hipStream_t st);
hipError_t launch_correlate_any_64(int fmt, int deform, const PassPair *pairs, int npairs,
                                   uint32_t max_cells, hipStream_t st);
hipError_t launch_correlate_any_128(int fmt, int deform, const PassPair *pairs, int npairs,
                                    uint32_t max_cells, hipStream_t st);

// ---------------------------------------------------------------------------------------
// validation (validate.rs:51-171): two launches per pass, one thread per vector
//   k_val_median : per-vector tests of the vector and its 3x3 neighbours (validate.rs:59-78),
//                  then the normalised median test (validate.rs:80-100) -> flags after test
//   k_val_replace: replacement of rejected vectors by the median of usable neighbours at
//                  radius 1, else 2 (validate.rs:102-128) -> validated field
// Medians sort total_cmp keys (f32::total_cmp, validate.rs:164) with a sorting network.
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ int32_t total_key(float v)
{
    const int32_t b = __float_as_int(v);
    return b ^ (int32_t)((uint32_t)(b >> 31) >> 1);
}

__device__ __forceinline__ float from_key(int32_t k)
{
    return __int_as_float(k ^ (int32_t)((uint32_t)(k >> 31) >> 1));
}

#define CSWAP(a, b)                   \
    do {                              \
        const int32_t lo_ = min(a, b); \
        b = max(a, b);                \
        a = lo_;                      \
    } while (0)

// Batcher odd-even merge network, 19 comparators; empty slots hold INT_MAX (sorted last)
__device__ __forceinline__ void sort8(int32_t (&k)[8])
{
    CSWAP(k[0], k[1]); CSWAP(k[2], k[3]); CSWAP(k[4], k[5]); CSWAP(k[6], k[7]);
    CSWAP(k[0], k[2]); CSWAP(k[1], k[3]); CSWAP(k[4], k[6]); CSWAP(k[5], k[7]);
    CSWAP(k[1], k[2]); CSWAP(k[5], k[6]);
    CSWAP(k[0], k[4]); CSWAP(k[1], k[5]); CSWAP(k[2], k[6]); CSWAP(k[3], k[7]);
    CSWAP(k[2], k[4]); CSWAP(k[3], k[5]);
    CSWAP(k[1], k[2]); CSWAP(k[3], k[4]); CSWAP(k[5], k[6]);
}

// validate.rs:163-171 median of the cnt smallest (valid) sorted keys
template <int M>
__device__ __forceinline__ float median_sorted(const int32_t (&k)[M], int cnt)
{
    const int ilo = (cnt - 1) >> 1, ihi = cnt >> 1;
    int32_t lo = k[0], hi = k[0];
#pragma unroll
    for (int i = 1; i < M; ++i) {
        lo = i == ilo ? k[i] : lo;
        hi = i == ihi ? k[i] : hi;
    }
    if (cnt & 1) return from_key(hi);
    return 0.5f * (from_key(lo) + from_key(hi));
}

__device__ __forceinline__ float median8(const float (&v)[8], int cnt)
{
    int32_t k[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) k[i] = i < cnt ? total_key(v[i]) : INT32_MAX;
    sort8(k);
    return median_sorted<8>(k, cnt);
}

// validate.rs:156-161 residual
__device__ __forceinline__ float residual8(float value, const float (&nb)[8], int cnt, float eps)
{
    const float med = median8(nb, cnt);
    float dev[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) dev[i] = fabsf(nb[i] - med);
    return fabsf(value - med) / (median8(dev, cnt) + eps);
}

// median of up to 24 values (radius-2 replacement, only when no radius-1 neighbour is
// usable): insertion into a sorted key list
__device__ float median24(const float *v, int cnt)
{
    int32_t k[24];
    for (int i = 0; i < cnt; ++i) {
        int32_t x = total_key(v[i]);
        int j = i;
        while (j > 0 && k[j - 1] > x) {
            k[j] = k[j - 1];
            --j;
        }
        k[j] = x;
    }
    const int ilo = (cnt - 1) >> 1, ihi = cnt >> 1;
    if (cnt & 1) return from_key(k[ihi]);
    return 0.5f * (from_key(k[ilo]) + from_key(k[ihi]));
}

// glibc hypotf: sqrt of the exact double sum of squares, rounded to float
__device__ __forceinline__ float hypot_rs(float u, float v)
{
    if (isinf(u) || isinf(v)) return INFINITY;
    const double du = (double)u, dv = (double)v;
    return (float)sqrt(du * du + dv * dv);
}

// validate.rs:59-78 per-vector tests
__device__ __forceinline__ uint8_t vector_flags(const ValidatePair &J, long i)
{
    uint8_t f = J.raw_flags[i];
    if (f & NO_SIGNAL) return f;
    if (hypot_rs(J.raw_dx[i], J.raw_dy[i]) > J.max_displacement) f |= OUT_OF_RANGE;
    if (J.raw_ppr[i] < J.ppr_threshold) f |= LOW_PPR;
    return f;
}

// add per-wave counts to a pair's counters (one atomic per wave and counter)
__device__ __forceinline__ void count_add(uint32_t *ctr, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m && (threadIdx.x & 63) == (uint32_t)__ffsll((long long)m) - 1)
        atomicAdd(ctr, (uint32_t)__popcll(m));
}

constexpr int VBLOCK = 256;
// Vectors whose usability and components a k_val_median block stages in LDS: its own 256 and
// the grid rows above and below them (256 + 2 nx + 2), for grids up to 895 vectors wide.
constexpr int VSTAGE = 2048;

// Pair and block-in-pair of block b: bpp blocks per pair, or, with a block map (serving
// batches of frames of different sizes), pair p owns blocks begins[p] .. begins[p + 1] - 1.
__device__ __forceinline__ void block_pair(uint32_t b, uint32_t bpp, const uint32_t *begins,
                                           int npairs, uint32_t &p, uint32_t &blk)
{
    if (begins) {
        p = find_pair_u(begins, npairs, b);
        blk = b - begins[p];
    } else {
        p = b / bpp;
        blk = b - p * bpp;
    }
}

__global__ __launch_bounds__(VBLOCK) void k_val_median(const ValidatePair *__restrict__ jobs,
                                                       uint32_t bpp, const uint32_t *begins,
                                                       int njobs)
{
    uint32_t p, blk;
    block_pair(blockIdx.x, bpp, begins, njobs, p, blk);
    const ValidatePair &J = jobs[p];
    const int nx = J.nx, ny = J.ny;
    const int c = (int)(blk * VBLOCK + threadIdx.x);
    const bool in = c < nx * ny;
    // Each vector's per-vector tests run once per block instead of once per neighbourhood
    // that holds it (nine times): the block first stages the usability and components of
    // every vector its threads' 3x3 neighbourhoods read.  Same arithmetic, same results.
    __shared__ uint8_t s_ok[VSTAGE];
    __shared__ float s_u[VSTAGE], s_v[VSTAGE];
    const int lo = (int)(blk * VBLOCK) - nx - 1;
    const int span = VBLOCK + 2 * nx + 2;
    const bool staged = span <= VSTAGE;  // uniform over the block
    if (staged) {
        for (int t = threadIdx.x; t < span; t += VBLOCK) {
            const int j = lo + t;
            const bool inside = j >= 0 && j < nx * ny;
            s_ok[t] = inside && (vector_flags(J, j) & INVALID) == 0;
            s_u[t] = inside ? J.raw_dx[j] : 0.0f;
            s_v[t] = inside ? J.raw_dy[j] : 0.0f;
        }
        __syncthreads();
    }
    uint8_t f = 0;
    bool med_out = false;
    if (in) {
        const int ix = c % nx, iy = c / nx;
        f = vector_flags(J, c);
        if ((f & INVALID) == 0) {
            float nu[8], nv[8];
            int cnt = 0;
#pragma unroll
            for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
                for (int ox = -1; ox <= 1; ++ox) {
                    if (oy == 0 && ox == 0) continue;
                    const int x = ix + ox, y = iy + oy;
                    bool use = false;
                    float u = 0.0f, v = 0.0f;
                    if (x >= 0 && y >= 0 && x < nx && y < ny) {
                        const int j = y * nx + x;
                        if (staged) {
                            use = s_ok[j - lo] != 0;
                            u = s_u[j - lo];
                            v = s_v[j - lo];
                        } else {
                            use = (vector_flags(J, j) & INVALID) == 0;
                            u = J.raw_dx[j];
                            v = J.raw_dy[j];
                        }
                    }
                    // compact: slot cnt takes this neighbour
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (use && q == cnt) {
                            nu[q] = u;
                            nv[q] = v;
                        }
                    cnt += use ? 1 : 0;
                }
            if (cnt >= 2) {
                const float ru = residual8(J.raw_dx[c], nu, cnt, J.median_eps);
                const float rv = residual8(J.raw_dy[c], nv, cnt, J.median_eps);
                if (sqrtf(ru * ru + rv * rv) > J.median_threshold) {
                    f |= MEDIAN_OUTLIER;
                    med_out = true;
                }
            }
        }
        J.snapshot[c] = f;  // flags after the median test (snapshot for replacement)
    }
    const bool ns = in && (f & NO_SIGNAL);
    count_add(J.summary + 0, ns);
    count_add(J.summary + 1, in && !ns && (f & LOW_PPR));
    count_add(J.summary + 2, med_out);
    count_add(J.summary + 3, in && !ns && (f & OUT_OF_RANGE));
}

__global__ __launch_bounds__(VBLOCK) void k_val_replace(const ValidatePair *__restrict__ jobs,
                                                        uint32_t bpp, const uint32_t *begins,
                                                        int njobs)
{
    uint32_t p, blk;
    block_pair(blockIdx.x, bpp, begins, njobs, p, blk);
    const ValidatePair &J = jobs[p];
    const int nx = J.nx, ny = J.ny;
    const int c = (int)(blk * VBLOCK + threadIdx.x);
    const bool in = c < nx * ny;
    bool rep = false;
    if (in) {
        const int ix = c % nx, iy = c / nx;
        const uint8_t f = J.snapshot[c];
        float u = J.raw_dx[c], v = J.raw_dy[c];
        uint8_t fo = f;
        if (J.replace && (f & NO_SIGNAL) == 0 && (f & INVALID) != 0) {
            float nu[8], nv[8];
            int cnt = 0;
#pragma unroll
            for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
                for (int ox = -1; ox <= 1; ++ox) {
                    if (oy == 0 && ox == 0) continue;
                    const int x = ix + ox, y = iy + oy;
                    bool use = false;
                    float a = 0.0f, b = 0.0f;
                    if (x >= 0 && y >= 0 && x < nx && y < ny) {
                        const long j = (long)y * nx + x;
                        use = (J.snapshot[j] & INVALID) == 0;
                        a = J.raw_dx[j];
                        b = J.raw_dy[j];
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (use && q == cnt) {
                            nu[q] = a;
                            nv[q] = b;
                        }
                    cnt += use ? 1 : 0;
                }
            if (cnt > 0) {
                u = median8(nu, cnt);
                v = median8(nv, cnt);
                rep = true;
            } else {
                float wu[24], wv[24];
                int n2 = 0;
                for (int y = max(iy - 2, 0); y <= min(iy + 2, ny - 1); ++y)
                    for (int x = max(ix - 2, 0); x <= min(ix + 2, nx - 1); ++x) {
                        if (x == ix && y == iy) continue;
                        const long j = (long)y * nx + x;
                        if ((J.snapshot[j] & INVALID) == 0) {
                            wu[n2] = J.raw_dx[j];
                            wv[n2] = J.raw_dy[j];
                            ++n2;
                        }
                    }
                if (n2 > 0) {
                    u = median24(wu, n2);
                    v = median24(wv, n2);
                    rep = true;
                } else {
                    u = NAN;
                    v = NAN;
                }
            }
            if (rep) fo = f | REPLACED;
        }
        J.dx[c] = u;
        J.dy[c] = v;
        J.ppr[c] = J.raw_ppr[c];
        J.flags[c] = fo;
    }
    count_add(J.summary + 4, rep);
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------------------
// frame conditioning (preprocess.rs:137-216)
// ---------------------------------------------------------------------------------------

// base = normalise(raw) [- background, clamped at 0]
template <int FMT>
__global__ void k_base(const void *__restrict__ raw, int pitch, long raw_frame_bytes,
                       const float *__restrict__ bg, int H, int W, float *__restrict__ out)
{
    // blockIdx.y: frame of a batch (frames raw_frame_bytes apart, outputs H*W apart)
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    raw = static_cast<const char *>(raw) + (long)blockIdx.y * raw_frame_bytes;
    out += (long)blockIdx.y * H * W;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    float v = load_pixel<FMT>(raw, (long)y * pitch + x);
    if (bg) {
        v = v - bg[i];
        v = v > 0.0f ? v : 0.0f;  // f32::max(v, 0.0)
    }
    out[i] = v;
}

// SAT (preprocess.rs:84-110): per row, running sums rs, rq of v and v^2 in f64 in column
// order; t[y+1][x+1] = t[y][x+1] + rs (the same for rq), down each column in row order.
// Two passes, the same additions in the same order:
//   k_sat_carry:  each row's (rs, rq) after every SAT_SW-th column, the carries into the
//                 column strips;
//   k_sat_strips: one wave per strip of SAT_SW columns walks down the frame in bands of
//                 SAT_BAND rows, continues each row's sums from its carry across the strip,
//                 then carries each column down the band.
// The samples are read twice and the table written once; the f64 chains are serial per row
// and per column whatever the split, so the parallelism comes from rows x strips x frames.
// In both kernels lanes 0-31 run the sums and lanes 32-63 the squares of 32 rows (or of the
// strip's columns), so every lane carries one chain.  (16-column strips, twice the strip
// waves per frame, were measured too: a serving batch's few frames gained within the noise,
// a 2000-frame video batch lost 15 %.)
constexpr int SAT_SW = 32, SAT_BAND = 32, SATK_TW = 64;
// lanes that load one strip column of a band (each loads every SAT_LPC-th row)
constexpr int SAT_LPC = 64 / SAT_SW;

// Frames of a launch at their own addresses (a serving batch's frames in their request
// slots): frame blockIdx.y reads img[y] and writes the tables at sum[y] (entry (0, 0) of the
// sums; the squares sat_table_elems(H, W) entries later).  Passed by value.
constexpr int SAT_LIST_MAX = 64;
struct SatList {
    const void *img[SAT_LIST_MAX];
    double *sum[SAT_LIST_MAX];
};

// one frame's carries; carry: [nb boundaries][2: sum, square][H rows], boundary b holding
// the running sums through column SAT_SW * (b + 1) - 1
template <int FMT>
__device__ __forceinline__ void sat_carry_frame(const void *__restrict__ img, int pitch, int H,
                                                int W, double *__restrict__ carry, int nb)
{
    __shared__ float tin[SAT_BAND][SATK_TW + 1];
    const int lane = threadIdx.x, r = lane & 31, sq = lane >> 5, y0 = blockIdx.x * SAT_BAND;
    const int rows = min(SAT_BAND, H - y0);
    const int ntile = (nb * SAT_SW + SATK_TW - 1) / SATK_TW;
    // raw samples from clamped (in-frame) addresses: all of a tile's loads are in flight at
    // once; samples outside the frame are zeroed when the tile is normalised into LDS
    raw_t<FMT> pre[SAT_BAND];
    int xt = 0;  // first column of the fetched tile
    auto fetch = [&](int t) {
        xt = t * SATK_TW;
        const int xc = min(xt + lane, W - 1);
#pragma unroll
        for (int i = 0; i < SAT_BAND; ++i)
            pre[i] = load_raw<FMT>(img, (long)min(y0 + i, H - 1) * pitch + xc);
    };
    if (ntile > 0) fetch(0);
    double acc = 0.0;
    for (int t = 0; t < ntile; ++t) {
        const bool okx = xt + lane < W;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SAT_BAND; ++i)
            tin[i][lane] = (i < rows && okx) ? normalise_raw<FMT>(pre[i]) : 0.0f;
        __syncthreads();
        if (t + 1 < ntile) fetch(t + 1);
#pragma unroll
        for (int h = 0; h < SATK_TW / SAT_SW; ++h) {
            float v[SAT_SW];
#pragma unroll
            for (int k = 0; k < SAT_SW; ++k) v[k] = tin[r][h * SAT_SW + k];
#pragma unroll
            for (int k = 0; k < SAT_SW; ++k) {
                const double d = (double)v[k], dd = d * d;
                acc += sq ? dd : d;
            }
            const int b = t * (SATK_TW / SAT_SW) + h;
            if (b < nb && r < rows) carry[((long)b * 2 + sq) * H + y0 + r] = acc;
        }
    }
}

// ---- screened tables --------------------------------------------------------------------
// With the default mask (min_mean <= 0, min_std <= 0) a u16 frame's tables only decide windows
// whose n x n box is all zero: any other box sums to at least 1/65535, above the tables'
// rounding error on frames up to about 2580^2 (the host checks the bound), so its mean is
// positive and its standard deviation never negative.  Every n x n box contains an aligned
// t x t tile for t <= n / 2, so a frame without an all-zero aligned tile needs no tables.
// k_sat_screen flags each 32-row band holding such a tile (full tiles only); the table
// kernels then build the tables only for flagged frames, and record whether they did in the
// frame's state word for the placement kernel.

// u16 frames whose rows are whole 16 B lines (W and the pitch multiples of 8 samples, the base
// 16 B aligned): each lane ORs 8 samples per load over a tile row's T rows; a tile of T = 4
// is half a lane's samples, of T >= 8 the samples of T / 8 neighbouring lanes.
template <int T>
__device__ __forceinline__ bool screen_band_wide(const uint16_t *__restrict__ img, int pitch, int W,
                                                 int y0, int ntr, int lane, int wave)
{
    const int ng = W / 8;  // 8-sample groups per row
    bool zero = false;
    for (int gb = wave * 64; gb < ng; gb += 256) {  // 64 groups per wave (tiles never straddle)
        const int gi = gb + lane;
        const bool ing = gi < ng;
        const long gc = (long)min(gi, ng - 1) * 8;
        for (int j = 0; j < ntr; ++j) {
            uint32_t a0 = 0, a1 = 0;  // samples 0-3 and 4-7 of the lane's group
#pragma unroll
            for (int r0 = 0; r0 < T; r0 += 8) {
                uint4 v[T < 8 ? T : 8];
#pragma unroll
                for (int r = 0; r < (T < 8 ? T : 8); ++r)
                    v[r] = *reinterpret_cast<const uint4 *>(img + (long)(y0 + j * T + r0 + r) * pitch + gc);
#pragma unroll
                for (int r = 0; r < (T < 8 ? T : 8); ++r) {
                    a0 |= v[r].x | v[r].y;
                    a1 |= v[r].z | v[r].w;
                }
            }
            if constexpr (T == 4) {
                zero = zero || __ballot(ing && (a0 == 0 || a1 == 0)) != 0;
            } else {
                // lanes whose 8 samples are all zero; groups past the frame count as non-zero
                const unsigned long long z = __ballot(ing && (a0 | a1) == 0);
                unsigned long long m = z;
                if constexpr (T >= 16) m &= (z >> 1) & 0x5555555555555555ull;
                if constexpr (T >= 32) m &= (z >> 2) & (z >> 3) & 0x1111111111111111ull;
                zero = zero || m != 0;
            }
        }
    }
    return zero;
}

// one 32-row band of a frame: flag = 1 if it holds an aligned all-zero t x t tile (t | 32)
template <int FMT>
__device__ __forceinline__ void screen_band(const void *__restrict__ img, int pitch, int H, int W,
                                            int t, uint8_t *__restrict__ flag)
{
    __shared__ int s_zero;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_zero = 0;
    __syncthreads();
    const int y0 = blockIdx.x * SAT_BAND;
    const int ntr = max(0, min(SAT_BAND, H - y0)) / t;  // full tile rows of the band
    const unsigned long long gmask = t >= 64 ? ~0ull : ((1ull << t) - 1);
    bool zero = false;
    if (FMT == FMT_U16 && W % 8 == 0 && pitch % 8 == 0 &&
        (reinterpret_cast<uintptr_t>(img) & 15) == 0) {
        const uint16_t *p = static_cast<const uint16_t *>(img);
        switch (t) {
        case 4: zero = screen_band_wide<4>(p, pitch, W, y0, ntr, lane, wave); break;
        case 8: zero = screen_band_wide<8>(p, pitch, W, y0, ntr, lane, wave); break;
        case 16: zero = screen_band_wide<16>(p, pitch, W, y0, ntr, lane, wave); break;
        default: zero = screen_band_wide<32>(p, pitch, W, y0, ntr, lane, wave); break;
        }
        if (zero) s_zero = 1;
        __syncthreads();
        if (tid == 0) *flag = (uint8_t)s_zero;
        return;
    }
    for (int cb = wave * 64; cb < W; cb += 256) {  // 64-column chunks (tiles never straddle)
        const int x = cb + lane;
        const bool inx = x < W;
        const long xc = min(x, W - 1);
        for (int j = 0; j < ntr; ++j) {
            uint32_t acc = 0;
            for (int r = 0; r < t; ++r) acc |= (uint32_t)load_raw<FMT>(img, (long)(y0 + j * t + r) * pitch + xc);
            // columns past the frame count as non-zero: tiles crossing the edge are not full
            const unsigned long long nz = __ballot(acc != 0 || !inx);
            for (int g = 0; g < 64; g += t) zero = zero || ((nz >> g) & gmask) == 0;
        }
    }
    if (zero) s_zero = 1;
    __syncthreads();
    if (tid == 0) *flag = (uint8_t)s_zero;
}

template <int FMT>
__global__ __launch_bounds__(256) void k_sat_screen(const void *__restrict__ img, int pitch,
                                                    long frame_bytes, int H, int W, int t,
                                                    double *__restrict__ sum0, long sat_stride)
{
    screen_band<FMT>(static_cast<const char *>(img) + (long)blockIdx.y * frame_bytes, pitch, H, W, t,
                     screen_flags(sum0 + (long)blockIdx.y * sat_stride) + blockIdx.x);
}

template <int FMT>
__global__ __launch_bounds__(256) void k_sat_screen_list(SatList list, int pitch, int H, int W, int t)
{
    screen_band<FMT>(list.img[blockIdx.y], pitch, H, W, t, screen_flags(list.sum[blockIdx.y]) + blockIdx.x);
}

// 64-lane table blocks: does the screen require this frame's tables?
__device__ __forceinline__ bool screen_needs(const double *sum, int H)
{
    const int nbands = (H + SAT_BAND - 1) / SAT_BAND, l = threadIdx.x;
    const uint8_t *f = screen_flags(sum);
    const bool any = (l < nbands && f[l] != 0) || (l + 64 < nbands && f[l + 64] != 0);
    return __ballot(any) != 0;
}

template <int FMT>
__global__ __launch_bounds__(64) void k_sat_carry(const void *__restrict__ img, int pitch,
                                                  long frame_bytes, int H, int W,
                                                  double *__restrict__ carry, int nb,
                                                  const double *__restrict__ sum0, long sat_stride,
                                                  int screened)
{
    if (screened && !screen_needs(sum0 + (long)blockIdx.y * sat_stride, H)) return;
    sat_carry_frame<FMT>(static_cast<const char *>(img) + (long)blockIdx.y * frame_bytes, pitch, H,
                         W, carry + (long)blockIdx.y * nb * 2 * H, nb);
}

template <int FMT>
__global__ __launch_bounds__(64) void k_sat_carry_list(SatList list, int pitch, int H, int W,
                                                       double *__restrict__ carry, int nb, int screened)
{
    if (screened && !screen_needs(list.sum[blockIdx.y], H)) return;
    sat_carry_frame<FMT>(list.img[blockIdx.y], pitch, H, W, carry + (long)blockIdx.y * nb * 2 * H, nb);
}

// one frame's tables, strip blockIdx.x
template <int FMT>
__device__ __forceinline__ void sat_strips_frame(const void *__restrict__ img, int pitch, int H,
                                                 int W, const double *__restrict__ carry, int nb,
                                                 double *__restrict__ sum,
                                                 double *__restrict__ sumsq)
{
    __shared__ float tin[SAT_BAND][SAT_SW + 1];
    __shared__ double tv[2][SAT_BAND][SAT_SW + 1];
    const int lane = threadIdx.x, c = lane & 31, sq = lane >> 5, strip = blockIdx.x;
    // this lane's table: sums (lanes 0-31) or squares (32-63)
    double *tab = sq ? sumsq : sum;
    const long s = (long)sat_pitch(W);
    // the column walk: lane (c, sq) carries strip column c < SAT_SW down the frame
    const bool colw = c < SAT_SW;
    const int x = strip * SAT_SW + c;  // sample column of table column x + 1
    if (colw && x < W) tab[x + 1] = 0.0;  // row 0
    if (strip == 0 && c == 0) tab[0] = 0.0;
    const double *crow = strip > 0 ? carry + ((long)(strip - 1) * 2 + sq) * H : nullptr;
    // band loads: lane (lc, lg) fetches strip column lc of rows lg, lg + SAT_LPC, ..., as raw
    // samples from clamped addresses, zeroed outside the frame when normalised into LDS
    const int lc = lane % SAT_SW, lg = lane / SAT_SW;
    const int xl = strip * SAT_SW + lc;
    raw_t<FMT> pre[SAT_BAND / SAT_LPC];
    const int xc = min(xl, W - 1);
    auto fetch = [&](int y0) {
#pragma unroll
        for (int i = 0; i < SAT_BAND / SAT_LPC; ++i)
            pre[i] = load_raw<FMT>(img, (long)min(y0 + SAT_LPC * i + lg, H - 1) * pitch + xc);
    };
    auto fetch_carry = [&](int y) {
        if (!crow) return 0.0;
        const double v = crow[min(y, H - 1)];
        return y < H ? v : 0.0;
    };
    fetch(0);
    double ci = fetch_carry(c);  // carry of row y0 + c into this strip
    double cv = 0.0;                               // column chain, carried across bands
    for (int y0 = 0; y0 < H; y0 += SAT_BAND) {
        const int rows = min(SAT_BAND, H - y0);
        if (strip == 0)
            for (int r = c; r < rows; r += 32) tab[(long)(y0 + r + 1) * s] = 0.0;  // column 0
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SAT_BAND / SAT_LPC; ++i)
            tin[SAT_LPC * i + lg][lc] =
                (y0 + SAT_LPC * i + lg < H && xl < W) ? normalise_raw<FMT>(pre[i]) : 0.0f;
        __syncthreads();
        double acc = ci;
        if (y0 + SAT_BAND < H) {
            fetch(y0 + SAT_BAND);
            ci = fetch_carry(y0 + SAT_BAND + c);
        }
        // lane (c, sq): row y0 + c across the strip, from its carry
        {
            float v[SAT_SW];
#pragma unroll
            for (int k = 0; k < SAT_SW; ++k) v[k] = tin[c][k];
#pragma unroll
            for (int k = 0; k < SAT_SW; ++k) {
                const double d = (double)v[k], dd = d * d;
                acc += sq ? dd : d;
                tv[sq][c][k] = acc;
            }
        }
        __syncthreads();
        // lane (c, sq): column x down the band
        if (colw && x < W) {
#pragma unroll
            for (int r0 = 0; r0 < SAT_BAND; r0 += 8) {
                double col[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) col[r] = tv[sq][r0 + r][c];
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    if (r0 + r < rows) {
                        cv = cv + col[r];
                        tab[(long)(y0 + r0 + r + 1) * s + x + 1] = cv;
                    }
                }
            }
        }
    }
}

// screened frames: every strip block records in the state word whether the tables are built
// (the same value from every block)
__device__ __forceinline__ bool screen_strips(double *sum, int H)
{
    const bool need = screen_needs(sum, H);
    if (threadIdx.x == 0) *sat_state_word(sum) = need ? 1u : 0u;
    return need;
}

template <int FMT>
__global__ __launch_bounds__(64) void k_sat_strips(const void *__restrict__ img, int pitch,
                                                   long frame_bytes, int H, int W,
                                                   const double *__restrict__ carry, int nb,
                                                   double *__restrict__ sum,
                                                   double *__restrict__ sumsq, long sat_stride,
                                                   int screened)
{
    const long y = blockIdx.y;
    if (screened && !screen_strips(sum + y * sat_stride, H)) return;
    sat_strips_frame<FMT>(static_cast<const char *>(img) + y * frame_bytes, pitch, H, W,
                          carry + y * nb * 2 * H, nb, sum + y * sat_stride, sumsq + y * sat_stride);
}

template <int FMT>
__global__ __launch_bounds__(64) void k_sat_strips_list(SatList list, int pitch, int H, int W,
                                                        const double *__restrict__ carry, int nb,
                                                        int screened)
{
    const long y = blockIdx.y;
    double *sum = list.sum[y];
    if (screened && !screen_strips(sum, H)) return;
    sat_strips_frame<FMT>(list.img[y], pitch, H, W, carry + y * nb * 2 * H, nb, sum,
                          sum + sat_table_elems(H, W));
}

// gaussian_blur (preprocess.rs:161-195) fused with the base image (PreparedFrame::new,
// preprocess.rs:137-153: the normalised samples, minus the background, clamped at 0): a
// workgroup holds the base samples of a BT_H x BT_W tile of outputs and its halo in LDS, with
// the halo's coordinates clamped into the frame as the reference clamps its taps, blurs the
// rows of the tile and halo rows, then the columns.  Taps are summed in kernel order with a
// rounded product then a rounded sum, as the two separate passes do, so the image is the same
// bit for bit; the intermediate row-blurred image never leaves LDS.  base_out (optional)
// receives the tile's base samples.
constexpr int BT_W = 64, BT_H = 16, BT_RMAX = 16;

template <int FMT>
__global__ __launch_bounds__(256) void k_blur_tiled(const void *__restrict__ raw, int pitch,
                                                    long raw_frame_bytes, const float *__restrict__ bg,
                                                    int H, int W, const float *__restrict__ taps,
                                                    int r, float *__restrict__ base_out,
                                                    float *__restrict__ out)
{
    __shared__ float sb[(BT_H + 2 * BT_RMAX) * (BT_W + 2 * BT_RMAX)];  // base, tile + halo
    __shared__ float sr[(BT_H + 2 * BT_RMAX) * BT_W];                  // row-blurred
    __shared__ float stp[2 * BT_RMAX + 1];
    const long fy = blockIdx.z;  // frame of a batch
    raw = static_cast<const char *>(raw) + fy * raw_frame_bytes;
    out += fy * (long)H * W;
    if (base_out) base_out += fy * (long)H * W;
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * BT_W, y0 = blockIdx.y * BT_H;
    const int tw = BT_W + 2 * r, th = BT_H + 2 * r, nt = 2 * r + 1;
    for (int i = tid; i < nt; i += 256) stp[i] = taps[i];
    for (int i = tid; i < th * tw; i += 256) {
        const int ly = i / tw, lx = i - ly * tw;
        const int gy = min(max(y0 - r + ly, 0), H - 1), gx = min(max(x0 - r + lx, 0), W - 1);
        float v = load_pixel<FMT>(raw, (long)gy * pitch + gx);
        if (bg) {
            v = v - bg[(long)gy * W + gx];
            v = v > 0.0f ? v : 0.0f;  // f32::max(v, 0.0)
        }
        sb[i] = v;
    }
    __syncthreads();
    if (base_out) {
        for (int i = tid; i < BT_H * BT_W; i += 256) {
            const int ly = i / BT_W, lx = i - ly * BT_W;
            if (y0 + ly < H && x0 + lx < W)
                base_out[(long)(y0 + ly) * W + x0 + lx] = sb[(ly + r) * tw + lx + r];
        }
    }
    // rows: output column x reads base columns clamp(x + j - r), halo column x - x0 + j
    for (int i = tid; i < th * BT_W; i += 256) {
        const int ly = i / BT_W, lx = i - ly * BT_W;
        const float *row = sb + ly * tw + lx;
        float acc = 0.0f;
        for (int j = 0; j < nt; ++j) acc += stp[j] * row[j];
        sr[i] = acc;
    }
    __syncthreads();
    // columns: output row y reads row-blurred rows clamp(y + j - r), halo row y - y0 + j
    for (int i = tid; i < BT_H * BT_W; i += 256) {
        const int ly = i / BT_W, lx = i - ly * BT_W;
        if (y0 + ly >= H || x0 + lx >= W) continue;
        const float *col = sr + ly * BT_W + lx;
        float acc = 0.0f;
        for (int j = 0; j < nt; ++j) acc += stp[j] * col[j * BT_W];
        out[(long)(y0 + ly) * W + x0 + lx] = acc;
    }
}

// gaussian_blur rows: sum over taps in kernel order, clamped edges
__global__ void k_blur_rows(const float *__restrict__ in, int H, int W,
                            const float *__restrict__ taps, int r, float *__restrict__ out)
{
    // blockIdx.y: frame of a batch (images H*W apart)
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    in += (long)blockIdx.y * H * W;
    out += (long)blockIdx.y * H * W;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const float *row = in + (long)y * W;
    float acc = 0.0f;
    for (int j = 0; j <= 2 * r; ++j) {
        int sx = x + j - r;
        sx = sx < 0 ? 0 : (sx > W - 1 ? W - 1 : sx);
        acc += taps[j] * row[sx];
    }
    out[i] = acc;
}

__global__ void k_blur_cols(const float *__restrict__ in, int H, int W,
                            const float *__restrict__ taps, int r, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    in += (long)blockIdx.y * H * W;
    out += (long)blockIdx.y * H * W;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    float acc = 0.0f;
    for (int j = 0; j <= 2 * r; ++j) {
        int sy = y + j - r;
        sy = sy < 0 ? 0 : (sy > H - 1 ? H - 1 : sy);
        acc += taps[j] * in[(long)sy * W + x];
    }
    out[i] = acc;
}

// local_normalize: (v - mean) / (std + eps) over the clipped (2r+1)^2 box
__global__ void k_local_norm(const float *__restrict__ in, int H, int W, int radius, float eps,
                             const double *__restrict__ sum, const double *__restrict__ sq,
                             float *__restrict__ out, long sat_stride)
{
    // blockIdx.y: frame of a batch (images H*W apart, tables sat_stride apart)
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    in += (long)blockIdx.y * H * W;
    out += (long)blockIdx.y * H * W;
    sum += (long)blockIdx.y * sat_stride;
    sq += (long)blockIdx.y * sat_stride;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const int y0 = y > radius ? y - radius : 0, y1 = min(y + radius + 1, H);
    const int x0 = x > radius ? x - radius : 0, x1 = min(x + radius + 1, W);
    float m, sd;
    sat_stats(sum, sq, W, x0, y0, x1 - x0, y1 - y0, m, sd);
    out[i] = (in[i] - m) / (sd + eps);
}

// BackgroundEstimator::add (preprocess.rs:56-63): the first frame initialises the estimate,
// later frames are folded in with f32::min (NaN-ignoring) or added in f32.
template <int FMT>
__global__ void k_bg_add(const void *__restrict__ raw, int pitch, int H, int W, int mode,
                         int first, float *__restrict__ acc)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const float v = load_pixel<FMT>(raw, (long)y * pitch + x);
    if (first) acc[i] = v;
    else if (mode == 0) acc[i] = fminf(acc[i], v);
    else acc[i] = acc[i] + v;
}

// BackgroundEstimator::finish (preprocess.rs:65-71): mean = sum / count in f32
__global__ void k_bg_finish(float *__restrict__ acc, long n, float count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = acc[i] / count;
}

// pixel normalisation only (for the exhaustive u8/u16 check and frame downloads)
template <int FMT>
__global__ void k_normalize(const void *__restrict__ raw, long count, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = load_pixel<FMT>(raw, i);
}

// window placement only (the stage-level parity entry point)
__global__ void k_place(GridDev g, GridDev prev, int has_prev, const float *__restrict__ prev_dx,
                        const float *__restrict__ prev_dy, int H, int W,
                        float *__restrict__ origins)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.nx * g.ny) return;
    const int ix = k % g.nx, iy = k / g.nx, n = g.window;
    float px = 0.0f, py = 0.0f;
    if (has_prev)
        sample_field(prev, prev_dx, prev_dy, center(g.x0, ix, g.step, n),
                     center(g.y0, iy, g.step, n), px, py);
    const float max_x = (float)(W - n > 1 ? W - n - 1 : 0);
    const float max_y = (float)(H - n > 1 ? H - n - 1 : 0);
    float ax, bx, ay, by;
    place(g.x0 + ix * g.step, px, max_x, ax, bx);
    place(g.y0 + iy * g.step, py, max_y, ay, by);
    origins[4 * k + 0] = ax;
    origins[4 * k + 1] = ay;
    origins[4 * k + 2] = bx;
    origins[4 * k + 3] = by;
}

// ---------------------------------------------------------------------------------------
// Per-window kernels around the correlation kernel, one thread per window; block b of a
// launch with bpp blocks per pair works on pair b / bpp (pair descriptors stay scalar).
//   k_place_pass  : window placement (multipass.rs:207-231) and signal mask
//                   (multipass.rs:233-241) -> origins, sig
//   k_peak_finish : sub-pixel fit and peak ratio (correlate.rs:262-303) and assembly of
//                   the raw vector (multipass.rs:238-259)
// ---------------------------------------------------------------------------------------
constexpr int WBLOCK = 256;

__global__ __launch_bounds__(WBLOCK) void k_place_pass(const PassPair *__restrict__ pairs,
                                                       uint32_t bpp, const uint32_t *begins,
                                                       int npairs)
{
    // window rows k and k + 2 read the same table rows (at 50 % overlap the bottom corners of
    // one row of windows are the top corners of the row two below): keep them in one L2
    const uint32_t b = xcd_block();
    uint32_t p, blk;
    block_pair(b, bpp, begins, npairs, p, blk);
    const PassPair &pp = pairs[p];
    const GridDev g = pp.grid;
    const uint32_t k = blk * WBLOCK + threadIdx.x;
    if (k >= (uint32_t)(g.nx * g.ny)) return;
    const int n = g.window;
    float4 o;
    if (pp.origins_given) {
        o = reinterpret_cast<const float4 *>(pp.origins)[k];
    } else {
        const int ix = (int)(k % (uint32_t)g.nx), iy = (int)(k / (uint32_t)g.nx);
        float px = 0.0f, py = 0.0f;
        if (pp.has_prev)
            sample_field(pp.prev, pp.prev_dx, pp.prev_dy, center(g.x0, ix, g.step, n),
                         center(g.y0, iy, g.step, n), px, py);
        const float max_x = (float)(pp.W - n > 1 ? pp.W - n - 1 : 0);
        const float max_y = (float)(pp.H - n > 1 ? pp.H - n - 1 : 0);
        place(g.x0 + ix * g.step, px, max_x, o.x, o.z);
        place(g.y0 + iy * g.step, py, max_y, o.y, o.w);
        reinterpret_cast<float4 *>(pp.origins)[k] = o;
    }
    bool sig = true;
    if (pp.use_mask) {
        float m, sd;
        // a screened frame whose tables were not built has no all-zero box: it passes
        if (!(pp.screened & 1) || *sat_state_word(pp.sat_a) != 0) {
            sat_stats(pp.sat_a, pp.sat_a + pp.sat_cells, pp.W, (int)roundf(o.x), (int)roundf(o.y),
                      n, n, m, sd);
            if (!(m >= pp.min_mean && sd >= pp.min_std)) sig = false;
        }
        if (!(pp.screened & 2) || *sat_state_word(pp.sat_b) != 0) {
            sat_stats(pp.sat_b, pp.sat_b + pp.sat_cells, pp.W, (int)roundf(o.z), (int)roundf(o.w),
                      n, n, m, sd);
            if (!(m >= pp.min_mean && sd >= pp.min_std)) sig = false;
        }
    }
    pp.sig[k] = sig ? 1 : 0;
}

// correlate.rs:262-276: three-point Gaussian fit, parabolic when a value is not positive,
// clamped to +-0.5, non-finite -> 0
__device__ __forceinline__ float subpixel(float m, float c, float q)
{
    float d;
    if (m > 0.0f && c > 0.0f && q > 0.0f) {
        const float lm = logf(m), lc = logf(c), lq = logf(q);
        d = (lm - lq) / (2.0f * (lm - 2.0f * lc + lq));
    } else {
        d = (m - q) / (2.0f * (m - 2.0f * c + q));
    }
    return isfinite(d) ? clamp_rs(d, -0.5f, 0.5f) : 0.0f;
}

// the same fit in FP64 (reduced-precision passes)
__device__ __forceinline__ double subpixel64(float m32, float c32, float q32)
{
    const double m = m32, c = c32, q = q32;
    double d;
    if (m > 0.0 && c > 0.0 && q > 0.0) {
        const double lm = log(m), lc = log(c), lq = log(q);
        d = (lm - lq) / (2.0 * (lm - 2.0 * lc + lq));
    } else {
        d = (m - q) / (2.0 * (m - 2.0 * c + q));
    }
    return isfinite(d) ? (d < -0.5 ? -0.5 : (d > 0.5 ? 0.5 : d)) : 0.0;
}

// F64 = true: the reduced-precision passes fit the peak and accumulate the displacement in
// FP64 (only the final value is rounded to f32).
template <bool F64>
__global__ __launch_bounds__(WBLOCK) void k_peak_finish(const PassPair *__restrict__ pairs,
                                                        uint32_t bpp, const uint32_t *begins,
                                                        int npairs)
{
    uint32_t p, blk;
    block_pair(blockIdx.x, bpp, begins, npairs, p, blk);
    const PassPair &pp = pairs[p];
    const GridDev g = pp.grid;
    const uint32_t k = blk * WBLOCK + threadIdx.x;
    if (k >= (uint32_t)(g.nx * g.ny)) return;
    const PeakRec r = pp.peaks[k];
    if (r.has_peak) {
        const int n = g.window, h2 = n / 2;
        const float4 o = reinterpret_cast<const float4 *>(pp.origins)[k];
        const int pby = r.index / n, pbx = r.index - pby * n;
        // correlate.rs:278-284: indices at or past n/2 wrap to negative displacements
        const float wx = pbx >= h2 ? (float)pbx - (float)n : (float)pbx;
        const float wy = pby >= h2 ? (float)pby - (float)n : (float)pby;
        float ppr = 1000.0f;  // MAX_PPR (correlate.rs:12)
        if (r.second > 0.0f) {
            ppr = r.best / r.second;
            ppr = (ppr > 1000.0f || ppr != ppr) ? 1000.0f : ppr;
        }
        if constexpr (F64) {
            const double sx = subpixel64(r.mx, r.best, r.qx), sy = subpixel64(r.my, r.best, r.qy);
            pp.out_dx[k] = (float)(((double)o.z - (double)o.x) + ((double)wx + sx));
            pp.out_dy[k] = (float)(((double)o.w - (double)o.y) + ((double)wy + sy));
        } else {
            const float sx = subpixel(r.mx, r.best, r.qx), sy = subpixel(r.my, r.best, r.qy);
            pp.out_dx[k] = o.z - o.x + (wx + sx);
            pp.out_dy[k] = o.w - o.y + (wy + sy);
        }
        pp.out_ppr[k] = ppr;
        pp.out_flags[k] = 0;
        if (pp.out_peak) pp.out_peak[k] = r.index;
    } else {
        pp.out_dx[k] = NAN;
        pp.out_dy[k] = NAN;
        pp.out_ppr[k] = 0.0f;
        pp.out_flags[k] = NO_SIGNAL;
        if (pp.out_peak) pp.out_peak[k] = -1;
    }
}

// ---------------------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------------------

// sizes with an instantiated register-resident FFT kernel (corr_sizes.hpp)
static bool fft_kernel(int n)
{
#define PIV_CASE_SUP(N) case N:
    switch (n) {
    PIV_FFT_SIZES(PIV_CASE_SUP) return true;
    default: return false;
    }
#undef PIV_CASE_SUP
}

bool correlate_supported(int n) { return n >= 8 && n <= kMaxWindow && n % 2 == 0; }

bool correlate_generic(int n, int deform) { return !fft_kernel(n) || (deform != 0 && n > 64); }

hipError_t launch_correlate(int n, int fmt, int variant, const PassPair *pairs, int npairs,
                            uint32_t max_cells, hipStream_t st, int deform)
{
    if (!correlate_supported(n)) return hipErrorInvalidValue;
    if (correlate_generic(n, deform)) {
        if (n <= 32) return launch_correlate_any_32(fmt, deform, pairs, npairs, max_cells, st);
        if (n <= 64) return launch_correlate_any_64(fmt, deform, pairs, npairs, max_cells, st);
        return launch_correlate_any_128(fmt, deform, pairs, npairs, max_cells, st);
    }
#define PIV_CASE_FFT(N)                                                                       \
    case N: return launch_correlate_##N(fmt, variant, deform, pairs, npairs, max_cells, st);
    switch (n) {
    PIV_FFT_SIZES(PIV_CASE_FFT)
    default: return hipErrorInvalidValue;
    }
#undef PIV_CASE_FFT
}

hipError_t launch_validate(const ValidatePair *jobs, int njobs, uint32_t max_cells,
                           hipStream_t st, const uint32_t *begins, uint32_t total_blocks)
{
    const uint32_t bpp = (max_cells + VBLOCK - 1) / VBLOCK;
    const uint32_t blocks = begins ? total_blocks : bpp * (uint32_t)njobs;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_val_median, dim3(blocks), dim3(VBLOCK), 0, st, jobs, bpp, begins, njobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_val_replace, dim3(blocks), dim3(VBLOCK), 0, st, jobs, bpp, begins, njobs);
    return hipGetLastError();
}

hipError_t launch_place_pass(const PassPair *pairs, int npairs, uint32_t max_cells, hipStream_t st,
                             const uint32_t *begins, uint32_t total_blocks)
{
    const uint32_t bpp = (max_cells + WBLOCK - 1) / WBLOCK;
    const uint32_t blocks = begins ? total_blocks : bpp * (uint32_t)npairs;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_place_pass, dim3(blocks), dim3(WBLOCK), 0, st, pairs, bpp, begins, npairs);
    return hipGetLastError();
}

hipError_t launch_peak_finish(const PassPair *pairs, int npairs, uint32_t max_cells, hipStream_t st,
                              bool f64, const uint32_t *begins, uint32_t total_blocks)
{
    const uint32_t bpp = (max_cells + WBLOCK - 1) / WBLOCK;
    const uint32_t blocks = begins ? total_blocks : bpp * (uint32_t)npairs;
    if (blocks == 0) return hipSuccess;
    if (f64)
        hipLaunchKernelGGL(k_peak_finish<true>, dim3(blocks), dim3(WBLOCK), 0, st, pairs, bpp, begins, npairs);
    else
        hipLaunchKernelGGL(k_peak_finish<false>, dim3(blocks), dim3(WBLOCK), 0, st, pairs, bpp, begins, npairs);
    return hipGetLastError();
}

bool correlate_mf_supported(int n) { return n == 16 || n == 32 || n == 64; }

hipError_t launch_correlate_mf(int n, int fmt, int precision, const PassPair *pairs, int npairs,
                               uint32_t max_cells, hipStream_t st, int deform)
{
#define PIV_CASE_MF(N)                                                                        \
    case N: return launch_correlate_mf_##N(fmt, precision, deform, pairs, npairs, max_cells, st);
    switch (n) {
    PIV_MF_SIZES(PIV_CASE_MF)
    default: return hipErrorInvalidValue;
    }
#undef PIV_CASE_MF
}

hipError_t launch_pass(int n, int fmt, int variant, const PassPair *pairs, int npairs,
                       uint32_t max_cells, hipStream_t st)
{
    hipError_t e = launch_place_pass(pairs, npairs, max_cells, st);
    if (e == hipSuccess) e = launch_correlate(n, fmt, variant, pairs, npairs, max_cells, st);
    if (e == hipSuccess) e = launch_peak_finish(pairs, npairs, max_cells, st);
    return e;
}

hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }
constexpr int MAX_GRID_Y = 65535;  // frames per launch (gridDim.y)

hipError_t launch_base(int fmt, const void *raw, int pitch, const float *bg, int H, int W,
                       float *out, hipStream_t st, int nframes, long raw_frame_bytes)
{
    const long n = (long)H * W;
    for (int f0 = 0; f0 < nframes; f0 += MAX_GRID_Y) {
        const dim3 grid(nblk(n, 256), (unsigned)min(MAX_GRID_Y, nframes - f0));
        const void *r = static_cast<const char *>(raw) + (long)f0 * raw_frame_bytes;
        float *o = out + (long)f0 * n;
        switch (fmt) {
        case FMT_U8: hipLaunchKernelGGL(k_base<FMT_U8>, grid, dim3(256), 0, st, r, pitch, raw_frame_bytes, bg, H, W, o); break;
        case FMT_U16: hipLaunchKernelGGL(k_base<FMT_U16>, grid, dim3(256), 0, st, r, pitch, raw_frame_bytes, bg, H, W, o); break;
        default: hipLaunchKernelGGL(k_base<FMT_F32>, grid, dim3(256), 0, st, r, pitch, raw_frame_bytes, bg, H, W, o); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

bool blur_fusable(int r) { return r >= 1 && r <= BT_RMAX; }

hipError_t launch_blur_fused(int fmt, const void *raw, int pitch, const float *bg, int H, int W,
                             const float *taps, int r, float *base_out, float *out, hipStream_t st,
                             int nframes, long raw_frame_bytes)
{
    if (!blur_fusable(r)) return hipErrorInvalidValue;
    for (int f0 = 0; f0 < nframes; f0 += MAX_GRID_Y) {
        const dim3 grid(nblk(W, BT_W), nblk(H, BT_H), (unsigned)min(MAX_GRID_Y, nframes - f0));
        const void *rw = static_cast<const char *>(raw) + (long)f0 * raw_frame_bytes;
        const long o = (long)f0 * H * W;
        float *bo = base_out ? base_out + o : nullptr;
        switch (fmt) {
        case FMT_U8: hipLaunchKernelGGL(k_blur_tiled<FMT_U8>, grid, dim3(256), 0, st, rw, pitch, raw_frame_bytes, bg, H, W, taps, r, bo, out + o); break;
        case FMT_U16: hipLaunchKernelGGL(k_blur_tiled<FMT_U16>, grid, dim3(256), 0, st, rw, pitch, raw_frame_bytes, bg, H, W, taps, r, bo, out + o); break;
        default: hipLaunchKernelGGL(k_blur_tiled<FMT_F32>, grid, dim3(256), 0, st, rw, pitch, raw_frame_bytes, bg, H, W, taps, r, bo, out + o); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int sat_boundaries(int W) { return (W + SAT_SW - 1) / SAT_SW - 1; }

size_t sat_carry_bytes(int H, int W, int nframes)
{
    return (size_t)nframes * (size_t)sat_boundaries(W) * 2 * (size_t)H * sizeof(double);
}

hipError_t launch_sat_raw(int fmt, const void *raw, int pitch, long raw_frame_bytes, int H, int W,
                          double *sum, double *sq, double *carry, hipStream_t st, int nframes,
                          long sat_stride, int screen_t)
{
    const int nb = sat_boundaries(W);
    const int scr = screen_t > 0 ? 1 : 0;
    if (scr && (fmt != FMT_U16 || screen_t > SAT_BAND || SAT_BAND % screen_t != 0 ||
                nblk(H, SAT_BAND) > (unsigned)SCREEN_MAX_BANDS))
        return hipErrorInvalidValue;
    for (int f0 = 0; f0 < nframes; f0 += MAX_GRID_Y) {
        const unsigned nf = (unsigned)min(MAX_GRID_Y, nframes - f0);
        const void *r = static_cast<const char *>(raw) + (long)f0 * raw_frame_bytes;
        double *su = sum + (long)f0 * sat_stride, *sqq = sq + (long)f0 * sat_stride;
        if (scr) {
            hipLaunchKernelGGL(k_sat_screen<FMT_U16>, dim3(nblk(H, SAT_BAND), nf), dim3(256), 0, st, r,
                               pitch, raw_frame_bytes, H, W, screen_t, su, sat_stride);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        if (nb > 0) {
            const dim3 grid(nblk(H, SAT_BAND), nf);
            switch (fmt) {
            case FMT_U8: hipLaunchKernelGGL(k_sat_carry<FMT_U8>, grid, dim3(64), 0, st, r, pitch, raw_frame_bytes, H, W, carry, nb, su, sat_stride, scr); break;
            case FMT_U16: hipLaunchKernelGGL(k_sat_carry<FMT_U16>, grid, dim3(64), 0, st, r, pitch, raw_frame_bytes, H, W, carry, nb, su, sat_stride, scr); break;
            default: hipLaunchKernelGGL(k_sat_carry<FMT_F32>, grid, dim3(64), 0, st, r, pitch, raw_frame_bytes, H, W, carry, nb, su, sat_stride, scr); break;
            }
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        const dim3 grid(nblk(W, SAT_SW), nf);
        switch (fmt) {
        case FMT_U8: hipLaunchKernelGGL(k_sat_strips<FMT_U8>, grid, dim3(64), 0, st, r, pitch, raw_frame_bytes, H, W, carry, nb, su, sqq, sat_stride, scr); break;
        case FMT_U16: hipLaunchKernelGGL(k_sat_strips<FMT_U16>, grid, dim3(64), 0, st, r, pitch, raw_frame_bytes, H, W, carry, nb, su, sqq, sat_stride, scr); break;
        default: hipLaunchKernelGGL(k_sat_strips<FMT_F32>, grid, dim3(64), 0, st, r, pitch, raw_frame_bytes, H, W, carry, nb, su, sqq, sat_stride, scr); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_sat_list(int fmt, const void *const *imgs, double *const *sums, int nframes,
                           int pitch, int H, int W, double *carry, hipStream_t st, int screen_t)
{
    const int nb = sat_boundaries(W);
    const int scr = screen_t > 0 ? 1 : 0;
    if (scr && (fmt != FMT_U16 || screen_t > SAT_BAND || SAT_BAND % screen_t != 0 ||
                nblk(H, SAT_BAND) > (unsigned)SCREEN_MAX_BANDS))
        return hipErrorInvalidValue;
    for (int f0 = 0; f0 < nframes; f0 += SAT_LIST_MAX) {
        const int nf = min(SAT_LIST_MAX, nframes - f0);
        SatList list{};
        for (int i = 0; i < nf; ++i) {
            list.img[i] = imgs[f0 + i];
            list.sum[i] = sums[f0 + i];
        }
        if (scr) {
            hipLaunchKernelGGL(k_sat_screen_list<FMT_U16>, dim3(nblk(H, SAT_BAND), (unsigned)nf), dim3(256), 0,
                               st, list, pitch, H, W, screen_t);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        if (nb > 0) {
            const dim3 grid(nblk(H, SAT_BAND), (unsigned)nf);
            switch (fmt) {
            case FMT_U8: hipLaunchKernelGGL(k_sat_carry_list<FMT_U8>, grid, dim3(64), 0, st, list, pitch, H, W, carry, nb, scr); break;
            case FMT_U16: hipLaunchKernelGGL(k_sat_carry_list<FMT_U16>, grid, dim3(64), 0, st, list, pitch, H, W, carry, nb, scr); break;
            default: hipLaunchKernelGGL(k_sat_carry_list<FMT_F32>, grid, dim3(64), 0, st, list, pitch, H, W, carry, nb, scr); break;
            }
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        const dim3 grid(nblk(W, SAT_SW), (unsigned)nf);
        switch (fmt) {
        case FMT_U8: hipLaunchKernelGGL(k_sat_strips_list<FMT_U8>, grid, dim3(64), 0, st, list, pitch, H, W, carry, nb, scr); break;
        case FMT_U16: hipLaunchKernelGGL(k_sat_strips_list<FMT_U16>, grid, dim3(64), 0, st, list, pitch, H, W, carry, nb, scr); break;
        default: hipLaunchKernelGGL(k_sat_strips_list<FMT_F32>, grid, dim3(64), 0, st, list, pitch, H, W, carry, nb, scr); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_sat(const float *img, int H, int W, double *sum, double *sq, double *carry,
                      hipStream_t st, int nframes, long sat_stride)
{
    return launch_sat_raw(FMT_F32, img, W, (long)H * W * (long)sizeof(float), H, W, sum, sq, carry,
                          st, nframes, sat_stride);
}

hipError_t launch_blur(const float *in, int H, int W, const float *taps, int r, float *tmp,
                       float *out, hipStream_t st, int nframes)
{
    const long n = (long)H * W;
    for (int f0 = 0; f0 < nframes; f0 += MAX_GRID_Y) {
        const dim3 grid(nblk(n, 256), (unsigned)min(MAX_GRID_Y, nframes - f0));
        const long o = (long)f0 * n;
        hipLaunchKernelGGL(k_blur_rows, grid, dim3(256), 0, st, in + o, H, W, taps, r, tmp + o);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_blur_cols, grid, dim3(256), 0, st, tmp + o, H, W, taps, r, out + o);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_local_norm(const float *in, int H, int W, int radius, float eps,
                             const double *sum, const double *sq, float *out, hipStream_t st,
                             int nframes, long sat_stride)
{
    const long n = (long)H * W;
    for (int f0 = 0; f0 < nframes; f0 += MAX_GRID_Y) {
        const dim3 grid(nblk(n, 256), (unsigned)min(MAX_GRID_Y, nframes - f0));
        const long o = (long)f0 * n, so = (long)f0 * sat_stride;
        hipLaunchKernelGGL(k_local_norm, grid, dim3(256), 0, st, in + o, H, W, radius, eps,
                           sum + so, sq + so, out + o, sat_stride);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_bg_add(int fmt, const void *raw, int pitch, int H, int W, int mode, int first,
                         float *acc, hipStream_t st)
{
    const long n = (long)H * W;
    if (fmt == FMT_U8)
        hipLaunchKernelGGL(k_bg_add<FMT_U8>, dim3(nblk(n, 256)), dim3(256), 0, st, raw, pitch, H, W, mode, first, acc);
    else if (fmt == FMT_U16)
        hipLaunchKernelGGL(k_bg_add<FMT_U16>, dim3(nblk(n, 256)), dim3(256), 0, st, raw, pitch, H, W, mode, first, acc);
    else
        hipLaunchKernelGGL(k_bg_add<FMT_F32>, dim3(nblk(n, 256)), dim3(256), 0, st, raw, pitch, H, W, mode, first, acc);
    return hipGetLastError();
}

hipError_t launch_bg_finish(float *acc, long n, float count, hipStream_t st)
{
    hipLaunchKernelGGL(k_bg_finish, dim3(nblk(n, 256)), dim3(256), 0, st, acc, n, count);
    return hipGetLastError();
}

// One plane of interleaved samples (decode.rs:124-133: the green plane of an RGB source is
// ffmpeg's `format=gbrp,extractplanes=g`, a plain copy of every pixel's second sample).
template <typename T>
__global__ void k_extract_plane(const T *__restrict__ src, long count, int channels, int plane,
                                T *__restrict__ dst)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i * channels + plane];
}

hipError_t launch_extract_plane(int fmt, const void *src, long count, int channels, int plane,
                                void *dst, hipStream_t st)
{
    const dim3 g(nblk(count, 256)), b(256);
    if (fmt == FMT_U8)
        hipLaunchKernelGGL(k_extract_plane<uint8_t>, g, b, 0, st, static_cast<const uint8_t *>(src),
                           count, channels, plane, static_cast<uint8_t *>(dst));
    else if (fmt == FMT_U16)
        hipLaunchKernelGGL(k_extract_plane<uint16_t>, g, b, 0, st, static_cast<const uint16_t *>(src),
                           count, channels, plane, static_cast<uint16_t *>(dst));
    else
        hipLaunchKernelGGL(k_extract_plane<float>, g, b, 0, st, static_cast<const float *>(src),
                           count, channels, plane, static_cast<float *>(dst));
    return hipGetLastError();
}

hipError_t launch_normalize(int fmt, const void *raw, long count, float *out, hipStream_t st)
{
    if (fmt == FMT_U8)
        hipLaunchKernelGGL(k_normalize<FMT_U8>, dim3(nblk(count, 256)), dim3(256), 0, st, raw, count, out);
    else if (fmt == FMT_U16)
        hipLaunchKernelGGL(k_normalize<FMT_U16>, dim3(nblk(count, 256)), dim3(256), 0, st, raw, count, out);
    else
        hipLaunchKernelGGL(k_normalize<FMT_F32>, dim3(nblk(count, 256)), dim3(256), 0, st, raw, count, out);
    return hipGetLastError();
}

hipError_t launch_place(const GridDev &g, const GridDev &prev, int has_prev, const float *pdx,
                        const float *pdy, int H, int W, float *origins, hipStream_t st)
{
    const int n = g.nx * g.ny;
    hipLaunchKernelGGL(k_place, dim3(nblk(n, 64)), dim3(64), 0, st, g, prev, has_prev, pdx, pdy, H,
                       W, origins);
    return hipGetLastError();
}

}  // namespace piv

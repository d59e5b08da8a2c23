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

namespace piv {

hipError_t launch_correlate_8(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);
hipError_t launch_correlate_12(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);
hipError_t launch_correlate_16(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);
hipError_t launch_correlate_24(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);
hipError_t launch_correlate_32(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);
hipError_t launch_correlate_48(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);
hipError_t launch_correlate_64(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);
hipError_t launch_correlate_96(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st);

// ---------------------------------------------------------------------------------------
// validation (validate.rs:51-171)
// ---------------------------------------------------------------------------------------

constexpr int VT = 16;            // output tile
constexpr int VH = 3;             // halo: median test +-1 feeding replacement +-2
constexpr int VR = VT + 2 * VH;   // 22
constexpr uint8_t ABSENT = 0x81;  // outside the grid: NO_SIGNAL-like, never usable

// f32::total_cmp order as a signed key
__device__ __forceinline__ int32_t total_key(float v)
{
    const int32_t b = __float_as_int(v);
    return b ^ (int32_t)((uint32_t)(b >> 31) >> 1);
}

// k-th smallest (0-based) of v[0..cnt) in total_cmp order
template <int M>
__device__ __forceinline__ float select_k(const float (&v)[M], int cnt, int kth)
{
    float out = 0.0f;
#pragma unroll
    for (int i = 0; i < M; ++i) {
        if (i < cnt) {
            const int32_t ki = total_key(v[i]);
            int lt = 0, le = 0;
#pragma unroll
            for (int j = 0; j < M; ++j) {
                if (j < cnt) {
                    const int32_t kj = total_key(v[j]);
                    lt += kj < ki;
                    le += kj <= ki;
                }
            }
            if (lt <= kth && kth < le) out = v[i];
        }
    }
    return out;
}

// validate.rs:163-171 median (sort by total_cmp; even count -> 0.5 (a + b))
template <int M>
__device__ __forceinline__ float median_of(const float (&v)[M], int cnt)
{
    if (cnt & 1) return select_k<M>(v, cnt, cnt / 2);
    const float a = select_k<M>(v, cnt, cnt / 2 - 1), b = select_k<M>(v, cnt, cnt / 2);
    return 0.5f * (a + b);
}

// validate.rs:156-161 residual
template <int M>
__device__ __forceinline__ float residual(float value, const float (&nb)[M], int cnt, float eps)
{
    const float med = median_of<M>(nb, cnt);
    float dev[M];
#pragma unroll
    for (int i = 0; i < M; ++i) dev[i] = fabsf(nb[i] - med);
    return fabsf(value - med) / (median_of<M>(dev, cnt) + eps);
}

// glibc hypotf: sqrt of the exact double sum of squares, rounded to float
__device__ __forceinline__ float hypot_rs(float u, float v)
{
    if (isinf(u) || isinf(v)) return INFINITY;
    const double du = (double)u, dv = (double)v;
    return (float)sqrt(du * du + dv * dv);
}

__global__ __launch_bounds__(256) void k_validate(const ValidatePair *__restrict__ jobs,
                                                  const uint32_t *__restrict__ tile_begins,
                                                  int njobs, uint32_t *__restrict__ tile_counts)
{
    __shared__ float s_dx[VR * VR], s_dy[VR * VR];
    __shared__ uint8_t s_f1[VR * VR], s_f2[VR * VR];
    __shared__ uint32_t s_cnt[5];

    const uint32_t tile = blockIdx.x;
    const uint32_t jp = find_pair_u(tile_begins, njobs, tile);
    const ValidatePair &J = jobs[jp];
    const uint32_t lt = tile - J.tile_begin;
    const int tx0 = (int)(lt % (uint32_t)J.tiles_x) * VT, ty0 = (int)(lt / (uint32_t)J.tiles_x) * VT;
    const int nx = J.nx, ny = J.ny;
    if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0;
    __syncthreads();

    // per-vector tests (validate.rs:59-78) over tile + halo
    uint32_t c_ns = 0, c_oor = 0, c_low = 0, c_med = 0, c_rep = 0;
    for (int i = threadIdx.x; i < VR * VR; i += blockDim.x) {
        const int gx = tx0 - VH + i % VR, gy = ty0 - VH + i / VR;
        uint8_t f = ABSENT;
        float u = NAN, v = NAN;
        if (gx >= 0 && gy >= 0 && gx < nx && gy < ny) {
            const long gi = (long)gy * nx + gx;
            f = J.raw_flags[gi];
            u = J.raw_dx[gi];
            v = J.raw_dy[gi];
            const bool own = gx >= tx0 && gy >= ty0 && gx < tx0 + VT && gy < ty0 + VT;
            if (f & NO_SIGNAL) {
                c_ns += own;
            } else {
                if (hypot_rs(u, v) > J.max_displacement) {
                    f |= OUT_OF_RANGE;
                    c_oor += own;
                }
                if (J.raw_ppr[gi] < J.ppr_threshold) {
                    f |= LOW_PPR;
                    c_low += own;
                }
            }
        }
        s_dx[i] = u;
        s_dy[i] = v;
        s_f1[i] = f;
    }
    __syncthreads();

    // normalised median test (validate.rs:80-100) on tile + 2-cell halo
    for (int i = threadIdx.x; i < VR * VR; i += blockDim.x) {
        const int lx = i % VR, ly = i / VR;
        uint8_t f = s_f1[i];
        if (lx >= 1 && ly >= 1 && lx < VR - 1 && ly < VR - 1 && (f & INVALID) == 0) {
            float nu[8], nv[8];
            int cnt = 0;
#pragma unroll
            for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
                for (int ox = -1; ox <= 1; ++ox) {
                    if (oy == 0 && ox == 0) continue;
                    const int j = i + oy * VR + ox;
                    if ((s_f1[j] & INVALID) == 0) {
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                            if (q == cnt) {
                                nu[q] = s_dx[j];
                                nv[q] = s_dy[j];
                            }
                        ++cnt;
                    }
                }
            if (cnt >= 2) {
                const float ru = residual<8>(s_dx[i], nu, cnt, J.median_eps);
                const float rv = residual<8>(s_dy[i], nv, cnt, J.median_eps);
                if (sqrtf(ru * ru + rv * rv) > J.median_threshold) {
                    f |= MEDIAN_OUTLIER;
                    const int gx = tx0 - VH + lx, gy = ty0 - VH + ly;
                    c_med += (gx >= tx0 && gy >= ty0 && gx < tx0 + VT && gy < ty0 + VT);
                }
            }
        }
        s_f2[i] = f;
    }
    __syncthreads();

    // replacement (validate.rs:102-128) and output for the tile
    {
        const int lx = VH + threadIdx.x % VT, ly = VH + threadIdx.x / VT;
        const int gx = tx0 + threadIdx.x % VT, gy = ty0 + threadIdx.x / VT;
        if (gx < nx && gy < ny) {
            const int i = ly * VR + lx;
            uint8_t f = s_f2[i];
            float u = s_dx[i], v = s_dy[i];
            if (J.replace && (f & NO_SIGNAL) == 0 && (f & INVALID) != 0) {
                bool done = false;
#pragma unroll
                for (int radius = 1; radius <= 2 && !done; ++radius) {
                    float nu[24], nv[24];
                    int cnt = 0;
                    for (int oy = -radius; oy <= radius; ++oy)
                        for (int ox = -radius; ox <= radius; ++ox) {
                            if (oy == 0 && ox == 0) continue;
                            const int j = i + oy * VR + ox;
                            if ((s_f2[j] & INVALID) == 0) {
#pragma unroll
                                for (int q = 0; q < 24; ++q)
                                    if (q == cnt) {
                                        nu[q] = s_dx[j];
                                        nv[q] = s_dy[j];
                                    }
                                ++cnt;
                            }
                        }
                    if (cnt > 0) {
                        u = median_of<24>(nu, cnt);
                        v = median_of<24>(nv, cnt);
                        f |= REPLACED;
                        ++c_rep;
                        done = true;
                    }
                }
                if (!done) {
                    u = NAN;
                    v = NAN;
                }
            }
            const long gi = (long)gy * nx + gx;
            J.dx[gi] = u;
            J.dy[gi] = v;
            J.ppr[gi] = J.raw_ppr[gi];
            J.flags[gi] = f;
        }
    }
    atomicAdd(&s_cnt[0], c_ns);
    atomicAdd(&s_cnt[1], c_low);
    atomicAdd(&s_cnt[2], c_med);
    atomicAdd(&s_cnt[3], c_oor);
    atomicAdd(&s_cnt[4], c_rep);
    __syncthreads();
    if (threadIdx.x < 5) tile_counts[5 * (size_t)tile + threadIdx.x] = s_cnt[threadIdx.x];
}

// per-pair sum of tile counters (deterministic, no global atomics)
__global__ void k_summarize(const ValidatePair *__restrict__ jobs, int njobs,
                            const uint32_t *__restrict__ tile_counts, uint32_t ntiles_total)
{
    const int jp = blockIdx.x * blockDim.x + threadIdx.x;
    if (jp >= njobs) return;
    const ValidatePair &J = jobs[jp];
    const uint32_t end = jp + 1 < njobs ? jobs[jp + 1].tile_begin : ntiles_total;
    uint32_t acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t t = J.tile_begin; t < end; ++t)
        for (int c = 0; c < 5; ++c) acc[c] += tile_counts[5 * (size_t)t + c];
    for (int c = 0; c < 5; ++c) J.summary[c] = acc[c];
}

// ---------------------------------------------------------------------------------------
// frame conditioning (preprocess.rs:137-216)
// ---------------------------------------------------------------------------------------

// base = normalise(raw) [- background, clamped at 0]
template <int FMT>
__global__ void k_base(const void *__restrict__ raw, int pitch, const float *__restrict__ bg,
                       int H, int W, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    float v = load_pixel<FMT>(raw, (long)y * pitch + x);
    if (bg) {
        v = v - bg[i];
        v = v > 0.0f ? v : 0.0f;  // f32::max(v, 0.0)
    }
    out[i] = v;
}

// SAT, step 1: running row sums in f64 (one row per thread), written at (y+1, x+1)
__global__ void k_sat_rows(const float *__restrict__ img, int H, int W, double *__restrict__ sum,
                           double *__restrict__ sq)
{
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const long s = (long)W + 1;
    if (y == 0) {
        for (int x = 0; x <= W; ++x) {
            sum[x] = 0.0;
            sq[x] = 0.0;
        }
    }
    if (y >= H) return;
    double rs = 0.0, rq = 0.0;
    const float *row = img + (long)y * W;
    double *ps = sum + (long)(y + 1) * s, *pq = sq + (long)(y + 1) * s;
    ps[0] = 0.0;
    pq[0] = 0.0;
    for (int x = 0; x < W; ++x) {
        const double v = (double)row[x];
        rs += v;
        rq += v * v;
        ps[x + 1] = rs;
        pq[x + 1] = rq;
    }
}

// SAT, step 2: t[y+1][x] = t[y][x] + rowsum[y][x], down each column (one column per thread)
__global__ void k_sat_cols(int H, int W, double *__restrict__ sum, double *__restrict__ sq)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (x > W) return;
    const long s = (long)W + 1;
    double cs = 0.0, cq = 0.0;
    for (int y = 1; y <= H; ++y) {
        cs = cs + sum[(long)y * s + x];
        cq = cq + sq[(long)y * s + x];
        sum[(long)y * s + x] = cs;
        sq[(long)y * s + x] = cq;
    }
}

// gaussian_blur rows: sum over taps in kernel order, clamped edges
__global__ void k_blur_rows(const float *__restrict__ in, int H, int W,
                            const float *__restrict__ taps, int r, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
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
                             float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const int y0 = y > radius ? y - radius : 0, y1 = min(y + radius + 1, H);
    const int x0 = x > radius ? x - radius : 0, x1 = min(x + radius + 1, W);
    float m, sd;
    sat_stats(sum, sq, W, x0, y0, x1 - x0, y1 - y0, m, sd);
    out[i] = (in[i] - m) / (sd + eps);
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
// host-side launchers
// ---------------------------------------------------------------------------------------

bool correlate_supported(int n)
{
    switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: return true;
    default: return false;
    }
}

hipError_t launch_correlate(int n, int fmt, const PassPair *pairs, int npairs,
                            uint32_t max_cells, hipStream_t st)
{
    switch (n) {
    case 8: return launch_correlate_8(fmt, pairs, npairs, max_cells, st);
    case 12: return launch_correlate_12(fmt, pairs, npairs, max_cells, st);
    case 16: return launch_correlate_16(fmt, pairs, npairs, max_cells, st);
    case 24: return launch_correlate_24(fmt, pairs, npairs, max_cells, st);
    case 32: return launch_correlate_32(fmt, pairs, npairs, max_cells, st);
    case 48: return launch_correlate_48(fmt, pairs, npairs, max_cells, st);
    case 64: return launch_correlate_64(fmt, pairs, npairs, max_cells, st);
    case 96: return launch_correlate_96(fmt, pairs, npairs, max_cells, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_validate(const ValidatePair *jobs, const uint32_t *tile_begins, int njobs,
                           uint32_t ntiles, uint32_t *tile_counts, hipStream_t st)
{
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_validate, dim3(ntiles), dim3(256), 0, st, jobs, tile_begins, njobs,
                       tile_counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_summarize, dim3((njobs + 63) / 64), dim3(64), 0, st, jobs, njobs,
                       tile_counts, ntiles);
    return hipGetLastError();
}

static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_base(int fmt, const void *raw, int pitch, const float *bg, int H, int W,
                       float *out, hipStream_t st)
{
    const long n = (long)H * W;
    switch (fmt) {
    case FMT_U8: hipLaunchKernelGGL(k_base<FMT_U8>, dim3(nblk(n, 256)), dim3(256), 0, st, raw, pitch, bg, H, W, out); break;
    case FMT_U16: hipLaunchKernelGGL(k_base<FMT_U16>, dim3(nblk(n, 256)), dim3(256), 0, st, raw, pitch, bg, H, W, out); break;
    default: hipLaunchKernelGGL(k_base<FMT_F32>, dim3(nblk(n, 256)), dim3(256), 0, st, raw, pitch, bg, H, W, out); break;
    }
    return hipGetLastError();
}

hipError_t launch_sat(const float *img, int H, int W, double *sum, double *sq, hipStream_t st)
{
    hipLaunchKernelGGL(k_sat_rows, dim3(nblk(H, 64)), dim3(64), 0, st, img, H, W, sum, sq);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sat_cols, dim3(nblk(W, 64)), dim3(64), 0, st, H, W, sum, sq);
    return hipGetLastError();
}

hipError_t launch_blur(const float *in, int H, int W, const float *taps, int r, float *tmp,
                       float *out, hipStream_t st)
{
    const long n = (long)H * W;
    hipLaunchKernelGGL(k_blur_rows, dim3(nblk(n, 256)), dim3(256), 0, st, in, H, W, taps, r, tmp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_blur_cols, dim3(nblk(n, 256)), dim3(256), 0, st, tmp, H, W, taps, r, out);
    return hipGetLastError();
}

hipError_t launch_local_norm(const float *in, int H, int W, int radius, float eps,
                             const double *sum, const double *sq, float *out, hipStream_t st)
{
    const long n = (long)H * W;
    hipLaunchKernelGGL(k_local_norm, dim3(nblk(n, 256)), dim3(256), 0, st, in, H, W, radius, eps,
                       sum, sq, out);
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

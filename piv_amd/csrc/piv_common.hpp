// Device helpers shared by the PIV kernels: scalar restatements of piv-core that must be
// bit-exact (placement, predictor sampling, SAT statistics, pixel normalisation, bilinear
// window load).  Built with -ffp-contract=off: every expression rounds as the Rust source.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "piv_device.hpp"

namespace piv {


// ---------------------------------------------------------------------------------------
// scalar helpers that restate piv-core exactly
// ---------------------------------------------------------------------------------------

// Rust f32::clamp: if x < min { min } ; if x > max { max }
__device__ __forceinline__ float clamp_rs(float x, float lo, float hi)
{
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

// grid.rs:55-66
__device__ __forceinline__ float center(int o0, int i, int step, int window)
{
    return (float)(o0 + i * step) + ((float)window - 1.0f) / 2.0f;
}

// field.rs:56-81 VectorField::sample (bilinear, NaN-skipping, clamped)
__device__ __forceinline__ void sample_field(const GridDev &g, const float *__restrict__ dx,
                             const float *__restrict__ dy, float x, float y, float &u, float &v)
{
    const float fx = clamp_rs((x - center(g.x0, 0, g.step, g.window)) / (float)g.step, 0.0f,
                              (float)(g.nx - 1));
    const float fy = clamp_rs((y - center(g.y0, 0, g.step, g.window)) / (float)g.step, 0.0f,
                              (float)(g.ny - 1));
    const int ix = (int)floorf(fx), iy = (int)floorf(fy);
    const float tx = fx - (float)ix, ty = fy - (float)iy;
    const int ix1 = min(ix + 1, g.nx - 1), iy1 = min(iy + 1, g.ny - 1);
    const int rr[4] = {iy, iy, iy1, iy1};
    const int cc[4] = {ix, ix1, ix, ix1};
    const float ww[4] = {(1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty};
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float su = dx[rr[k] * g.nx + cc[k]], sv = dy[rr[k] * g.nx + cc[k]];
        if (ww[k] > 0.0f && isfinite(su) && isfinite(sv)) {
            a0 = a0 + ww[k] * su;
            a1 = a1 + ww[k] * sv;
            a2 = a2 + ww[k];
        }
    }
    if (a2 > 0.0f) {
        u = a0 / a2;
        v = a1 / a2;
    } else {
        u = 0.0f;
        v = 0.0f;
    }
}

// Window deformation (no counterpart in the reference, which offsets windows rigidly):
// the predictor's displacement at pixel (x, y) as the bilinear interpolant of the grid field
// in lerp form, a + t (b - a) along x then y, so that equal neighbours give exactly their
// value (a uniform field deforms nothing).  Positions clamp to the grid as
// VectorField::sample does; where a neighbour is not finite the weighted NaN-skipping form
// of VectorField::sample (field.rs:56-81) is used instead.
__device__ __forceinline__ void deform_sample(const GridDev &g, const float *__restrict__ dx,
                                              const float *__restrict__ dy, float x, float y,
                                              float &u, float &v)
{
    const float fx = clamp_rs((x - center(g.x0, 0, g.step, g.window)) / (float)g.step, 0.0f,
                              (float)(g.nx - 1));
    const float fy = clamp_rs((y - center(g.y0, 0, g.step, g.window)) / (float)g.step, 0.0f,
                              (float)(g.ny - 1));
    const int ix = (int)floorf(fx), iy = (int)floorf(fy);
    const float tx = fx - (float)ix, ty = fy - (float)iy;
    const int ix1 = min(ix + 1, g.nx - 1), iy1 = min(iy + 1, g.ny - 1);
    const float u00 = dx[iy * g.nx + ix], u01 = dx[iy * g.nx + ix1];
    const float u10 = dx[iy1 * g.nx + ix], u11 = dx[iy1 * g.nx + ix1];
    const float v00 = dy[iy * g.nx + ix], v01 = dy[iy * g.nx + ix1];
    const float v10 = dy[iy1 * g.nx + ix], v11 = dy[iy1 * g.nx + ix1];
    // both forms are evaluated and one selected (straight-line code: the windows' loaders
    // call this once per pixel row, unrolled)
    const float ut = u00 + tx * (u01 - u00), ub = u10 + tx * (u11 - u10);
    const float vt = v00 + tx * (v01 - v00), vb = v10 + tx * (v11 - v10);
    const float ul = ut + ty * (ub - ut), vl = vt + ty * (vb - vt);
    const float w[4] = {(1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty};
    const float su[4] = {u00, u01, u10, u11}, sv[4] = {v00, v01, v10, v11};
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    bool all = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool fin = isfinite(su[k]) && isfinite(sv[k]);
        all = all && fin;
        const bool use = w[k] > 0.0f && fin;
        a0 = use ? a0 + w[k] * su[k] : a0;
        a1 = use ? a1 + w[k] * sv[k] : a1;
        a2 = use ? a2 + w[k] : a2;
    }
    const bool any = a2 > 0.0f;
    const float uw = any ? a0 / a2 : 0.0f, vw = any ? a1 / a2 : 0.0f;
    u = all ? ul : uw;
    v = all ? vl : vw;
}

// deform_sample split into its separable parts, for loaders that evaluate it on a window's
// pixel grid: the x part of a column and the y part of a row are computed once, and the
// predictor taps of one grid row pair are loaded once while successive pixel rows fall in it.
// Every value equals deform_sample's.
struct DeformAxis {
    int i0, i1;  // grid nodes below and above
    float t;     // fraction between them
};

__device__ __forceinline__ DeformAxis deform_axis(int o0, int step, int window, int count, float x)
{
    const float f = clamp_rs((x - center(o0, 0, step, window)) / (float)step, 0.0f,
                             (float)(count - 1));
    DeformAxis a;
    a.i0 = (int)floorf(f);
    a.t = f - (float)a.i0;
    a.i1 = min(a.i0 + 1, count - 1);
    return a;
}

// The predictor's taps of one grid row pair (iy, iy1) at a column's nodes (ix, ix1), and
// their lerps along x; `fin` when all eight are finite (the lerp form applies).
struct DeformRow {
    int iy = -1;
    float u00, u01, u10, u11, v00, v01, v10, v11;
    float ut, ub, vt, vb;
    bool fin;
};

__device__ __forceinline__ void deform_row(const GridDev &g, const float *__restrict__ dx,
                                           const float *__restrict__ dy, const DeformAxis &ax,
                                           const DeformAxis &ay, DeformRow &R)
{
    R.iy = ay.i0;
    const int r0 = ay.i0 * g.nx, r1 = ay.i1 * g.nx;
    R.u00 = dx[r0 + ax.i0];
    R.u01 = dx[r0 + ax.i1];
    R.u10 = dx[r1 + ax.i0];
    R.u11 = dx[r1 + ax.i1];
    R.v00 = dy[r0 + ax.i0];
    R.v01 = dy[r0 + ax.i1];
    R.v10 = dy[r1 + ax.i0];
    R.v11 = dy[r1 + ax.i1];
    R.fin = isfinite(R.u00) && isfinite(R.u01) && isfinite(R.u10) && isfinite(R.u11) &&
            isfinite(R.v00) && isfinite(R.v01) && isfinite(R.v10) && isfinite(R.v11);
    R.ut = R.u00 + ax.t * (R.u01 - R.u00);
    R.ub = R.u10 + ax.t * (R.u11 - R.u10);
    R.vt = R.v00 + ax.t * (R.v01 - R.v00);
    R.vb = R.v10 + ax.t * (R.v11 - R.v10);
}

// deform_sample at the point whose axis parts are ax, ay, from the row pair's taps R
__device__ __forceinline__ void deform_eval(const DeformAxis &ax, const DeformAxis &ay,
                                            const DeformRow &R, float &u, float &v)
{
    if (R.fin) {
        u = R.ut + ay.t * (R.ub - R.ut);
        v = R.vt + ay.t * (R.vb - R.vt);
        return;
    }
    const float tx = ax.t, ty = ay.t;
    const float w[4] = {(1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty};
    const float su[4] = {R.u00, R.u01, R.u10, R.u11}, sv[4] = {R.v00, R.v01, R.v10, R.v11};
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (w[k] > 0.0f && isfinite(su[k]) && isfinite(sv[k])) {
            a0 = a0 + w[k] * su[k];
            a1 = a1 + w[k] * sv[k];
            a2 = a2 + w[k];
        }
    u = a2 > 0.0f ? a0 / a2 : 0.0f;
    v = a2 > 0.0f ? a1 / a2 : 0.0f;
}

// 4-point Lagrange weights at nodes -1, 0, 1, 2 for a fraction t in [0, 1): the cubic
// resampling kernel of the deformation (reproduces polynomials of degree <= 3 exactly).
__device__ __forceinline__ void lagrange4(float t, float (&l)[4])
{
    const float tp1 = t + 1.0f, tm1 = t - 1.0f, tm2 = t - 2.0f;
    l[0] = -(t * tm1 * tm2) / 6.0f;
    l[1] = (tp1 * tm1 * tm2) / 2.0f;
    l[2] = -(tp1 * t * tm2) / 2.0f;
    l[3] = (tp1 * t * tm1) / 6.0f;
}

// multipass.rs:224-229 place closure
__device__ __forceinline__ void place(int origin, float shift, float maxv, float &a, float &b)
{
    shift = clamp_rs(shift, -maxv, maxv);
    const float lo = fmaxf(-shift, 0.0f);
    const float hi = fminf(maxv - shift, maxv);
    const float a0 = clamp_rs((float)origin - shift / 2.0f, lo, hi);
    a = a0;
    b = a0 + shift;
}

// Workgroup id remapped so that each of the 8 XCDs (which take workgroups round-robin, each
// with its own L2) works on one contiguous range of the grid: neighbouring workgroups that
// read the same lines then share an L2.  A bijection of [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t L = blockIdx.x, nb = gridDim.x;
    const uint32_t q = nb >> 3, rem = nb & 7, xcd = L & 7;
    return xcd * q + min(xcd, rem) + (L >> 3);
}

// preprocess.rs:112-124 IntegralImage::stats
__device__ __forceinline__ void sat_stats(const double *__restrict__ sum,
                                          const double *__restrict__ sq, int W, int x, int y,
                                          int bw, int bh, float &mean, float &sd)
{
    const long s = (long)sat_pitch(W);
    const double n = (double)((long)bw * bh);
    const long i11 = (long)(y + bh) * s + (x + bw), i01 = (long)y * s + (x + bw);
    const long i10 = (long)(y + bh) * s + x, i00 = (long)y * s + x;
    const double m = (sum[i11] - sum[i01] - sum[i10] + sum[i00]) / n;
    const double var = (sq[i11] - sq[i01] - sq[i10] + sq[i00]) / n - m * m;
    mean = (float)m;
    sd = (float)sqrt(var > 0.0 ? var : 0.0);
}

// A sample as stored (u8 / u16 widened to 32 bits, f32 as is).  A kernel that loads many
// samples ahead keeps them raw, so the loads stay independent of the arithmetic until the
// values are used.
template <int FMT>
using raw_t = std::conditional_t<FMT == FMT_F32, float, uint32_t>;

template <int FMT>
__device__ __forceinline__ raw_t<FMT> load_raw(const void *__restrict__ base, long idx)
{
    if constexpr (FMT == FMT_U8) return static_cast<const uint8_t *>(base)[idx];
    else if constexpr (FMT == FMT_U16) return static_cast<const uint16_t *>(base)[idx];
    else return static_cast<const float *>(base)[idx];
}

// decode.rs:227-234: b / 255 and v / 65535 as IEEE quotients.  q = v * (1/D) is off by one
// ulp for some v; one fma residual step r = v - q*D, q += r * (1/D) gives the correctly
// rounded quotient for every u8 and u16 value (checked exhaustively in the tests).
template <int FMT>
__device__ __forceinline__ float normalise_raw(raw_t<FMT> raw)
{
    if constexpr (FMT == FMT_F32) {
        return raw;
    } else {
        constexpr float D = FMT == FMT_U8 ? 255.0f : 65535.0f;
        constexpr float inv = 1.0f / D;
        const float v = (float)raw;
        const float q = v * inv;
        const float r = __builtin_fmaf(-q, D, v);
        return __builtin_fmaf(r, inv, q);
    }
}

template <int FMT>
__device__ __forceinline__ float load_pixel(const void *__restrict__ base, long idx)
{
    return normalise_raw<FMT>(load_raw<FMT>(base, idx));
}

// correlate.rs:182-202 load_window, one column per lane: col[r] = window(r, c).
// Lanes of a window read consecutive pixels of each image row (coalesced).  Reads are
// clamped into the image: a window exactly as wide (or tall) as the image has a zero weight
// on the extra column (row) the reference slices past the edge.
template <int N, int FMT>
__device__ __forceinline__ void load_column(const void *__restrict__ img, int pitch, int H, int W,
                                            float x, float y, int c, float (&col)[N])
{
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const float tx = x - (float)x0, ty = y - (float)y0;
    const int xa = min(x0 + c, W - 1), xb = min(x0 + c + 1, W - 1);
    if (tx < 1e-4f && ty < 1e-4f) {
#pragma unroll
        for (int r = 0; r < N; ++r) col[r] = load_pixel<FMT>(img, (long)min(y0 + r, H - 1) * pitch + xa);
    } else {
        const float w00 = (1.0f - tx) * (1.0f - ty), w01 = tx * (1.0f - ty);
        const float w10 = (1.0f - tx) * ty, w11 = tx * ty;
        const long r0 = (long)min(y0, H - 1) * pitch;
        float t0 = load_pixel<FMT>(img, r0 + xa), t1 = load_pixel<FMT>(img, r0 + xb);
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const long rb = (long)min(y0 + r + 1, H - 1) * pitch;
            const float b0 = load_pixel<FMT>(img, rb + xa), b1 = load_pixel<FMT>(img, rb + xb);
            col[r] = w00 * t0 + w01 * t1 + w10 * b0 + w11 * b1;
            t0 = b0;
            t1 = b1;
        }
    }
}

__device__ __forceinline__ uint32_t find_pair_u(const uint32_t *begins, int npairs, uint32_t k)
{
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begins[mid] <= k) lo = mid; else hi = mid - 1;
    }
    return (uint32_t)lo;
}

}  // namespace piv

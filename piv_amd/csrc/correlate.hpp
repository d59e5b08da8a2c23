// The correlation-pass kernel template (one instantiation per window size, see corr_n*.hip).
//
// One launch covers every window of every pair of a pass.  Work decomposition:
//   - a workgroup holds WPB windows of ONE pair, so the pair descriptor and the image base
//     pointers are wave-uniform (scalar registers, SGPR-based global loads): one wave of
//     64 / N windows when N divides 64, three waves of 192 / N windows at N = 24 and 48
//     (packed: some windows cross a wave boundary, and no lane idles), else as many whole
//     windows as one wave holds;
//   - one lane per window line: lane c first owns column c of the window, later row c;
//   - workgroup ids are remapped so that each of the 8 XCDs works on a contiguous range of
//     windows (consecutive windows share most of their pixels: one XCD's L2 serves them).
//
// Per window (correlate.rs:99-139 restated for the GPU):
//   1. origins and the signal mask come from the placement kernel (one thread per window);
//   2. fused ingest: raw u8/u16 samples are read straight from HBM and the window is sampled
//      bilinearly in raw units (the correlation coefficient is invariant to the decode scale
//      1/255 or 1/65535, which only rescales the energy threshold);
//   3. z = a + i b, 2-D FFT: column FFT in registers, LDS transpose, row FFT;
//   4. cross spectrum conj(FA) FB from Z[k] and Z[-k], with the DC bin cleared (the mean
//      removal of load_window); the window energies from the samples (N dividing 64) or
//      from Parseval;
//   5. Hermitian inverse (windows of at most 64 lanes, N >= 16): the half spectrum's columns,
//      then row pairs, each transform split over two lanes; else row IFFTs of rows 0..N/2
//      and column-pair IFFTs (two real columns per complex transform); scaled into an LDS
//      correlation plane;
//   6. find_peak (correlate.rs:215-303): first row-major argmax, footprint half width,
//      highest local maximum outside it; the peak record goes to the finishing kernel
//      (one thread per window), which does the 3-point fit, the peak ratio and assembly.
#pragma once

#include <type_traits>

#include "fft.hpp"
#include "piv_common.hpp"
#include "variants.hpp"

namespace piv {

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

template <int N>
struct KCfg {
    // N divides 64: windows tile a wave exactly, and butterfly reductions over a window's
    // lanes (power-of-two widths) apply
    static constexpr bool IN_WAVE = (64 % N == 0);
    // the least common multiple of N and the wave width
    static constexpr int LCM = N / gcd_c(N, 64) * 64;
    // N below 64 not dividing it, whose one-wave workgroups leave a quarter of the wave idle
    // and whose windows tile at most four waves (24 and 48: three waves): the workgroup packs
    // LCM / N windows back to back, some across a wave boundary.  Elsewhere the workgroup
    // barriers cost more than the idle lanes (measured: 12 packed in three waves, 10, 20 and
    // 40 in five, 14 and 28 in seven are all slower): those sizes keep one-wave workgroups.
    static constexpr bool PACKED = N < 64 && !IN_WAVE && LCM <= 256 && 64 % N >= 16;
    // windows per workgroup: packed, or as many whole windows as one wave holds (N <= 64;
    // one at N = 40 and 56, the rest of the wave idle), else one window of N lanes
    static constexpr int WPB = PACKED ? LCM / N : (N <= 64 ? 64 / N : 1);
    static constexpr int BLOCK = WPB * N;
    static constexpr int P = N + 1;   // odd LDS row stride: conflict-free rows and columns
    static constexpr int PL = N * P;
    // every window inside one one-wave workgroup: cross-lane exchanges may use shuffles and
    // ballots, and workgroup barriers cost nothing
    static constexpr bool ONE_WAVE = N <= 64 && !PACKED;
    // at most one wave's width per window (one-wave and packed workgroups): one LDS plane per
    // window, transposes in two passes (real, then imaginary parts), the half-spectrum inverse
    static constexpr bool NARROW = N <= 64;
};

// Sum over the N lanes of one window (all lanes receive the total).
template <int N>
__device__ __forceinline__ float window_sum(float v, float *red, int lane)
{
    if constexpr (KCfg<N>::IN_WAVE) {
#pragma unroll
        for (int m = N / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, N);
        return v;
    } else {
        __syncthreads();
        red[lane] = v;
        __syncthreads();
        float s = 0.0f;
        for (int i = 0; i < N; ++i) s += red[i];
        return s;
    }
}

template <int FMT>
__device__ __forceinline__ float raw_sample(const void *__restrict__ base, uint32_t idx)
{
    if constexpr (FMT == FMT_U8) {
        return (float)((const __attribute__((address_space(1))) uint8_t *)base)[idx];
    } else if constexpr (FMT == FMT_U16) {
        return (float)((const __attribute__((address_space(1))) uint16_t *)base)[idx];
    } else {
        return ((const __attribute__((address_space(1))) float *)base)[idx];
    }
}

// A frame as a raw buffer resource for the interior loaders: a row's offset is wave-uniform
// (the scalar offset of the load), the lane's column is the vector offset, and the sample to
// its right is the instruction's immediate offset, so a row costs no vector address
// arithmetic.  Offsets are 32-bit: frames whose extent is 2^31 bytes or more take the
// clamped loader (frame_buf_ok).
template <int FMT>
constexpr uint32_t sample_bytes()
{
    return FMT == FMT_U8 ? 1u : (FMT == FMT_U16 ? 2u : 4u);
}

template <int FMT>
__device__ __forceinline__ uint64_t frame_extent_bytes(int H, int W, int pitch)
{
    return ((uint64_t)(H - 1) * (uint64_t)pitch + (uint64_t)W) * sample_bytes<FMT>();
}

template <int FMT>
__device__ __forceinline__ bool frame_buf_ok(int H, int W, int pitch)
{
    return frame_extent_bytes<FMT>(H, W, pitch) < (1ull << 31);
}

template <int FMT>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_buf(const void *img, int H, int W, int pitch)
{
    // word 3: the untyped-access format of gfx9 buffer resources
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(img), (short)0,
                                             (int)frame_extent_bytes<FMT>(H, W, pitch), 0x00020000);
}

template <int FMT>
__device__ __forceinline__ float buf_sample(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff)
{
    if constexpr (FMT == FMT_U8) {
        return (float)__builtin_amdgcn_raw_buffer_load_b8(rs, (int)voff, (int)soff, 0);
    } else if constexpr (FMT == FMT_U16) {
        return (float)__builtin_amdgcn_raw_buffer_load_b16(rs, (int)voff, (int)soff, 0);
    } else {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, 0));
    }
}

// correlate.rs:182-202: window column c, sampled at fractional origin (x, y) in raw units.
// The integer path (tx, ty < 1e-4) copies samples without interpolation, as the reference.
// Rows and columns are clamped into the frame: a window exactly as wide (tall) as the frame
// has zero weight on the sample the reference would slice past the edge.
template <int N, int FMT>
__device__ __forceinline__ void load_column_raw(const void *__restrict__ img, int pitch, int H,
                                                int W, float x, float y, int c, float (&col)[N])
{
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const float tx = x - (float)x0, ty = y - (float)y0;
    const bool ints = tx < 1e-4f && ty < 1e-4f;
    const uint32_t xa = (uint32_t)min(x0 + c, W - 1), xb = (uint32_t)min(x0 + c + 1, W - 1);
    const uint32_t off0 = (uint32_t)y0 * (uint32_t)pitch, last = (uint32_t)(H - 1) * (uint32_t)pitch;
    // Branch-free: both samples are always loaded so the row loads issue back to back; the
    // integer path selects the unweighted sample.
    float hprev = 0.0f;
#pragma unroll
    for (int r = 0; r <= N; ++r) {
        const uint32_t row = min(off0 + (uint32_t)r * (uint32_t)pitch, last);
        const float p0 = raw_sample<FMT>(img, row + xa);
        const float p1 = raw_sample<FMT>(img, row + xb);
        const float h = ints ? p0 : __builtin_fmaf(tx, p1 - p0, p0);
        if (r > 0) col[r - 1] = ints ? hprev : __builtin_fmaf(ty, h - hprev, hprev);
        hprev = h;
        // bound the loads in flight (and their registers) to 16 rows at a time (the clamped
        // path, with two addresses per row)
        if ((r & 15) == 15) __builtin_amdgcn_sched_barrier(0);
    }
}

// Does the (N+1)^2 bilinear footprint of the window at (x, y) lie inside the frame?
template <int N>
__device__ __forceinline__ bool interior(float x, float y, int W, int H)
{
    return x >= 0.0f && y >= 0.0f && (int)x + N + 1 <= W && (int)y + N + 1 <= H;
}

// load_column_raw for a window whose footprint lies inside the frame (of an extent below
// 2^31 bytes): no clamping, and for integer samples the integer path is taken by zeroing the
// weights (0 * (p1 - p0) + p0 is exactly p0 for finite samples), so each row is two buffer
// loads and two lerps.
template <int N, int FMT>
__device__ __forceinline__ void load_column_fast(__amdgpu_buffer_rsrc_t rs, int pitch, float x,
                                                 float y, int c, float (&col)[N])
{
    constexpr uint32_t BPS = sample_bytes<FMT>();
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    float tx = x - (float)x0, ty = y - (float)y0;
    const bool ints = tx < 1e-4f && ty < 1e-4f;
    const uint32_t vo = ((uint32_t)y0 * (uint32_t)pitch + (uint32_t)(x0 + c)) * BPS;
    const uint32_t rb = (uint32_t)pitch * BPS;  // row bytes (wave-uniform)
    if constexpr (FMT != FMT_F32) {
        tx = ints ? 0.0f : tx;
        ty = ints ? 0.0f : ty;
    }
    float hprev = 0.0f;
#pragma unroll
    for (int r = 0; r <= N; ++r) {
        const float p0 = buf_sample<FMT>(rs, vo, (uint32_t)r * rb);
        const float p1 = buf_sample<FMT>(rs, vo + BPS, (uint32_t)r * rb);
        float h;
        if constexpr (FMT == FMT_F32) h = ints ? p0 : __builtin_fmaf(tx, p1 - p0, p0);
        else h = __builtin_fmaf(tx, p1 - p0, p0);
        if (r > 0) {
            if constexpr (FMT == FMT_F32) col[r - 1] = ints ? hprev : __builtin_fmaf(ty, h - hprev, hprev);
            else col[r - 1] = __builtin_fmaf(ty, h - hprev, hprev);
        }
        hprev = h;
        if ((r & 31) == 31) __builtin_amdgcn_sched_barrier(0);
    }
}

// Both fractional parts below 1e-4: the reference copies samples without interpolation
// (correlate.rs:186-188).
__device__ __forceinline__ bool integer_origin(float x, float y)
{
    return x - floorf(x) < 1e-4f && y - floorf(y) < 1e-4f;
}

// load_column_fast for integer origins: one sample per row, no interpolation.
template <int N, int FMT>
__device__ __forceinline__ void load_column_int(__amdgpu_buffer_rsrc_t rs, int pitch, float x,
                                                float y, int c, float (&col)[N])
{
    constexpr uint32_t BPS = sample_bytes<FMT>();
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const uint32_t vo = ((uint32_t)y0 * (uint32_t)pitch + (uint32_t)(x0 + c)) * BPS;
    const uint32_t rb = (uint32_t)pitch * BPS;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        col[r] = buf_sample<FMT>(rs, vo, (uint32_t)r * rb);
        if ((r & 31) == 31) __builtin_amdgcn_sched_barrier(0);
    }
}

// The sample at (x, y) with both indices clamped into the frame, in raw units.
template <int FMT>
__device__ __forceinline__ float clamped_sample(const void *__restrict__ img, int pitch, int H,
                                                int W, int x, int y)
{
    x = min(max(x, 0), W - 1);
    y = min(max(y, 0), H - 1);
    return raw_sample<FMT>(img, (uint32_t)y * (uint32_t)pitch + (uint32_t)x);
}

// One sample of a deformed window in raw units: the pixel whose rigid position has integer
// part (xi, yi) and fractional part (fx, fy), where the fractions already carry the pixel's
// deformation shift and may leave [0, 1).  MODE 1: bilinear as the rigid loaders compute it
// (a row lerp along x, then along y), so a zero shift gives their exact value; MODE 2: 4x4
// Lagrange.  Both fractions below 1e-4: the sample itself, as correlate.rs:186-188 does for
// a window.  Sample indices clamp into the frame.
template <int FMT, int MODE>
__device__ __forceinline__ float deformed_sample(const void *__restrict__ img, int pitch, int H,
                                                 int W, int xi, int yi, float fx, float fy)
{
    const float ixf = floorf(fx), iyf = floorf(fy);
    const float wx = fx - ixf, wy = fy - iyf;
    const int x = xi + (int)ixf, y = yi + (int)iyf;
    const bool ints = wx < 1e-4f && wy < 1e-4f;
#define at(xx, yy) clamped_sample<FMT>(img, pitch, H, W, (xx), (yy))
    if constexpr (MODE == 1) {
        const float p00 = at(x, y), p01 = at(x + 1, y), p10 = at(x, y + 1), p11 = at(x + 1, y + 1);
        const float ha = __builtin_fmaf(wx, p01 - p00, p00);
        const float hb = __builtin_fmaf(wx, p11 - p10, p10);
        const float v = __builtin_fmaf(wy, hb - ha, ha);
        return ints ? p00 : v;
    } else {
        float lx[4], ly[4];
        lagrange4(wx, lx);
        lagrange4(wy, ly);
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float h = 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) h = __builtin_fmaf(lx[i], at(x - 1 + i, y - 1 + j), h);
            v = __builtin_fmaf(ly[j], h, v);
        }
        return ints ? at(x, y) : v;
    }
#undef at
}

// Column c of both windows of a deformed pair.  The windows keep the rigid placement of the
// predictor at the window centre (origins (ax, ay), (bx, by)); pixel (c, r) then moves by
// -h in A and +h in B, h = (D(P) - D(C)) / 2, with D the predictor's deformation interpolant
// (deform_sample), P = (mx + c, my + r) the pixel's position midway between its A and B
// samples and C the window centre.  A uniform predictor gives h = 0 exactly and so the rigid
// samples.  D's x part is the lane's (one column), its y part is the row's, and the taps of
// a grid row pair are loaded once while the rows stay in it.  Both windows' columns are
// sampled by a plain loop into two LDS planes (row stride N + 1) and read back with constant
// indices, so the per-pixel work stays one loop body and the column registers never spill.
// Every lane of the workgroup calls this (it holds barriers); windows without signal get
// zeros.  With pb == nullptr (a kernel without room for a second plane) window A and then
// window B go through pa, and the interpolant is evaluated for each.
template <int N, int FMT, int MODE>
__device__ __forceinline__ void load_columns_deformed(const PassPair &pp, bool sig, float ax,
                                                      float ay, float bx, float by, int c,
                                                      float *pa, float *pb, float (&za)[N],
                                                      float (&zb)[N])
{
    constexpr int P = N + 1;
    const GridDev &g = pp.prev;
    const float mx = (ax + bx) * 0.5f, my = (ay + by) * 0.5f;
    const int xa = (int)floorf(ax), ya = (int)floorf(ay), xb = (int)floorf(bx), yb = (int)floorf(by);
    const float txa = ax - (float)xa, tya = ay - (float)ya, txb = bx - (float)xb, tyb = by - (float)yb;
    float uc = 0.0f, vc = 0.0f;
    if (sig) deform_sample(g, pp.prev_dx, pp.prev_dy, mx + ((float)N - 1.0f) / 2.0f,
                           my + ((float)N - 1.0f) / 2.0f, uc, vc);
    const DeformAxis dxa = deform_axis(g.x0, g.step, g.window, g.nx, mx + (float)c);
    // which: 0 both windows (pb), 1 window A only, 2 window B only (into pa)
    auto sample_rows = [&](int which) {
        DeformRow R;
        for (int r = 0; r < N; ++r) {
            const DeformAxis dya = deform_axis(g.y0, g.step, g.window, g.ny, my + (float)r);
            if (dya.i0 != R.iy) deform_row(g, pp.prev_dx, pp.prev_dy, dxa, dya, R);
            float u, v;
            deform_eval(dxa, dya, R, u, v);
            const float hx = (u - uc) * 0.5f, hy = (v - vc) * 0.5f;
            if (which != 2)
                pa[r * P + c] = deformed_sample<FMT, MODE>(pp.img_a, pp.pitch, pp.H, pp.W, xa + c, ya + r,
                                                           txa - hx, tya - hy);
            const float vb = which != 1 ? deformed_sample<FMT, MODE>(pp.img_b, pp.pitch, pp.H, pp.W, xb + c,
                                                                     yb + r, txb + hx, tyb + hy)
                                        : 0.0f;
            if (which == 0) pb[r * P + c] = vb;
            if (which == 2) pa[r * P + c] = vb;
        }
    };
    if (pb) {
        if (sig) sample_rows(0);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < N; ++r) {
            za[r] = sig ? pa[r * P + c] : 0.0f;
            zb[r] = sig ? pb[r * P + c] : 0.0f;
        }
        __syncthreads();
    } else {
        if (sig) sample_rows(1);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < N; ++r) za[r] = sig ? pa[r * P + c] : 0.0f;
        __syncthreads();
        if (sig) sample_rows(2);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < N; ++r) zb[r] = sig ? pa[r * P + c] : 0.0f;
        __syncthreads();
    }
}

// energy threshold of correlate.rs:110 in raw sample units
template <int FMT>
__device__ __forceinline__ double energy_floor()
{
    if constexpr (FMT == FMT_U8) return 1e-12 * 255.0 * 255.0;
    else if constexpr (FMT == FMT_U16) return 1e-12 * 65535.0 * 65535.0;
    else return 1e-12;
}

// Window pair of one lane's column: raw samples bilinearly sampled at the window origins
// (zeros for a masked window), with the approximate window means removed (the DC bin is
// cleared exactly later, which is what the reference's mean removal amounts to).
template <int N, int FMT, int DEF>
__device__ __forceinline__ void ingest(const PassPair &pp, bool sig, float ax, float ay, float bx,
                                       float by, int lane, float *red, float *pl, float *pl2,
                                       float (&zr)[N], float (&zi)[N])
{
    // every window of the wave keeps its (N+1)^2 footprint inside the frame: no clamping
    // (buffer loads: frames of an extent below 2^31 bytes)
    const bool fast = frame_buf_ok<FMT>(pp.H, pp.W, pp.pitch) &&
                      __all(!sig || (interior<N>(ax, ay, pp.W, pp.H) && interior<N>(bx, by, pp.W, pp.H)));
    // every window of the wave at integer origins (a first pass without predictor)
    const bool ints = __all(!sig || (integer_origin(ax, ay) && integer_origin(bx, by)));
    // DEF: the launch's deformation mode (a pass with a predictor); it is a template
    // parameter so that the rigid kernels carry none of the deformed loader's registers
    if (DEF != 0 && pp.has_prev) {
        load_columns_deformed<N, FMT, DEF>(pp, sig, ax, ay, bx, by, lane, pl, pl2, zr, zi);
    } else if (sig) {
        if (fast) {
            const __amdgpu_buffer_rsrc_t ra = frame_buf<FMT>(pp.img_a, pp.H, pp.W, pp.pitch);
            const __amdgpu_buffer_rsrc_t rb = frame_buf<FMT>(pp.img_b, pp.H, pp.W, pp.pitch);
            if (ints) {
                load_column_int<N, FMT>(ra, pp.pitch, ax, ay, lane, zr);
                load_column_int<N, FMT>(rb, pp.pitch, bx, by, lane, zi);
            } else {
                load_column_fast<N, FMT>(ra, pp.pitch, ax, ay, lane, zr);
                load_column_fast<N, FMT>(rb, pp.pitch, bx, by, lane, zi);
            }
        } else {
            load_column_raw<N, FMT>(pp.img_a, pp.pitch, pp.H, pp.W, ax, ay, lane, zr);
            load_column_raw<N, FMT>(pp.img_b, pp.pitch, pp.H, pp.W, bx, by, lane, zi);
        }
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r) {
            zr[r] = 0.0f;
            zi[r] = 0.0f;
        }
    }
    {
        // Approximate means only reduce the dynamic range of the transform: the DC bin is
        // cleared exactly in the cross spectrum, which is what the reference's mean removal
        // amounts to.
        float sa = 0.0f, sb = 0.0f;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            sa += zr[r];
            sb += zi[r];
        }
        constexpr float inv = 1.0f / (float)(N * N);
        const float ma = window_sum<N>(sa, red, lane) * inv;
        const float mb = window_sum<N>(sb, red, lane) * inv;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            zr[r] -= ma;
            zi[r] -= mb;
        }
    }
}

// DIF half of an N-point inverse: lane half hp transforms v[n] + (-1)^hp v[n + N/2], the odd
// half twiddled by e^{+2 pi i n / N} (tw[hp * N/2 + n]), as an N/2-point inverse; the outputs are
// the transform's entries 2k + hp.
template <int N>
__device__ __forceinline__ void dif_half_inverse(const float (&vr)[N], const float (&vi)[N], int hp,
                                                 const float2 *tw, float (&ur)[N / 2],
                                                 float (&ui)[N / 2])
{
    constexpr int H2 = N / 2;
    const float sg = hp ? -1.0f : 1.0f;
#pragma unroll
    for (int n = 0; n < H2; ++n) {
        const float ar = __builtin_fmaf(sg, vr[n + H2], vr[n]);
        const float ai = __builtin_fmaf(sg, vi[n + H2], vi[n]);
        if (n == 0) {
            ur[n] = ar;
            ui[n] = ai;
        } else {
            const float2 w = tw[hp * H2 + n];
            ur[n] = __builtin_fmaf(ar, w.x, -(ai * w.y));
            ui[n] = __builtin_fmaf(ar, w.y, ai * w.x);
        }
    }
    pivfft::fft<H2, 1>(ur, ui);
}

// The inverse 2-D transform of a window's Hermitian cross spectrum (the correlation is real)
// for windows inside one wave, every lane running two N/2-point transforms.  Lane ky holds
// spectrum row ky in (zr, zi); lane (t, hp) = t + hp N/2 of the window takes part in:
//   1. the columns of the half spectrum, inverted along ky: columns 1 .. N/2-1, and columns 0
//      and N/2 packed as one complex transform (both are real after it, the spectrum being
//      Hermitian along ky in them); column t by lanes (t, 0) and (t, 1), one decimation-in-
//      frequency half each, giving the column's rows 2k + hp;
//   2. rows t and t + N/2 packed as one complex transform (both are real after it), their
//      missing half spectrum the conjugate mirror G[y][N - kx] = conj(G[y][kx]), inverted along
//      kx by the same two lanes, giving the rows' columns 2k + hp.
// The two LDS exchanges move real and imaginary parts through the one plane in two rounds;
// the plane then receives the correlation scaled by `scale`.
template <int N>
__device__ __forceinline__ void inverse_half_spectrum(float (&zr)[N], float (&zi)[N], float *pr,
                                                      int lane, const float2 *tw, float scale)
{
    // odd stride of the half-width plane: conflict-free rows and columns
    constexpr int H2 = N / 2, HS = H2 | 1, P = KCfg<N>::P;
    static_assert(N % 2 == 0, "even window");
    const int hp = lane / H2, t = lane - hp * H2;
    // ---- 1. columns of the half spectrum: slot 0 = X[ky][0] + i X[ky][N/2] ----
    const float s0r = zr[0] - zi[H2], s0i = zi[0] + zr[H2];
    float vr[N], vi[N];
    __syncthreads();
    pr[lane * HS] = s0r;
#pragma unroll
    for (int s = 1; s < H2; ++s) pr[lane * HS + s] = zr[s];
    __syncthreads();
#pragma unroll
    for (int ky = 0; ky < N; ++ky) vr[ky] = pr[ky * HS + t];
    __syncthreads();
    pr[lane * HS] = s0i;
#pragma unroll
    for (int s = 1; s < H2; ++s) pr[lane * HS + s] = zi[s];
    __syncthreads();
#pragma unroll
    for (int ky = 0; ky < N; ++ky) vi[ky] = pr[ky * HS + t];
    float ur[H2], ui[H2];
    dif_half_inverse<N>(vr, vi, hp, tw, ur, ui);  // G[2k + hp][t] (t = 0: G[.][0] + i G[.][N/2])
    // ---- 2. row pairs (t, t + N/2): plane row y, slot s = G[y][s] ----
    float g0r[H2], g1r[H2], g0i[H2], g1i[H2];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < H2; ++k) pr[(2 * k + hp) * HS + t] = ur[k];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < H2; ++s) {
        g0r[s] = pr[t * HS + s];
        g1r[s] = pr[(t + H2) * HS + s];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < H2; ++k) pr[(2 * k + hp) * HS + t] = ui[k];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < H2; ++s) {
        g0i[s] = pr[t * HS + s];
        g1i[s] = pr[(t + H2) * HS + s];
    }
    // R[kx] = G[t][kx] + i G[t + N/2][kx]; G[y][0] and G[y][N/2] are the real and imaginary
    // parts of slot 0
#pragma unroll
    for (int n = 0; n < H2; ++n) {
        if (n == 0) {
            vr[0] = g0r[0];
            vi[0] = g1r[0];
            vr[H2] = g0i[0];
            vi[H2] = g1i[0];
        } else {
            const int k = H2 - n;
            vr[n] = g0r[n] - g1i[n];
            vi[n] = g0i[n] + g1r[n];
            vr[n + H2] = g0r[k] + g1i[k];
            vi[n + H2] = g1r[k] - g0i[k];
        }
    }
    dif_half_inverse<N>(vr, vi, hp, tw, ur, ui);  // c[t][2k + hp] + i c[t + N/2][2k + hp]
    __syncthreads();
#pragma unroll
    for (int k = 0; k < H2; ++k) {
        pr[t * P + 2 * k + hp] = ur[k] * scale;
        pr[(t + H2) * P + 2 * k + hp] = ui[k] * scale;
    }
}

// IEEE max of finite values without the operand canonicalisation the compiler adds to fmaxf
// when it cannot prove an operand canonical (values loaded from LDS); used only where every
// value is finite.
__device__ __forceinline__ float max3_finite(float a, float b, float c)
{
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ float max_finite(float a, float b)
{
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// find_peak (correlate.rs:215-303) on a window's correlation plane in LDS (row stride N+1),
// lane = row: first row-major argmax, footprint half width, highest local maximum outside
// it; lane 0 of the window writes the peak record.
template <int N, int FMT>
__device__ __forceinline__ void find_peak_plane(const float *pr, int lane, int tid, float *pkv,
                                                int *pki, bool active, bool has_peak,
                                                PeakRec *rec)
{
    using C = KCfg<N>;
    constexpr int P = C::P, H2 = N / 2;
    float zr[N];
    // ---- find_peak (correlate.rs:215-303), lane = row -----------------------------------
    // NaN can only reach the plane from non-finite float input; the integer formats take
    // the comparison shortcuts that assume finite values.
    constexpr bool MAYBE_NAN = (FMT == FMT_F32);
    // finite planes take the raw-max form of the 3x3 scan (no operand canonicalisation)
    constexpr bool RAW_MAX = !MAYBE_NAN;
    // (1) first maximum in row-major order (strict >): row scan, then across rows
#pragma unroll
    for (int x = 0; x < N; ++x) zr[x] = pr[lane * P + x];
    float best = -INFINITY;
    int bidx = 0;
    if constexpr (C::IN_WAVE && RAW_MAX) {
        // finite plane: the row maxima by a max3 tree, the window maximum across the rows,
        // then the first row holding it and the first column of that row holding it (ballots
        // of exact equality), which is the first maximum in row-major order
        float rm = zr[0];
#pragma unroll
        for (int x = 1; x < N; x += 2) rm = x + 1 < N ? max3_finite(rm, zr[x], zr[x + 1]) : max_finite(rm, zr[x]);
        float b = rm;
#pragma unroll
        for (int m = 1; m < N; m <<= 1) b = max_finite(b, __shfl_xor(b, m, N));
        best = b;
        const int base = tid - lane;
        const unsigned long long full = N == 64 ? ~0ull : ((1ull << (N & 63)) - 1);
        const unsigned long long rows = (__ballot(rm == b) >> base) & full;
        const int py = rows ? __builtin_ctzll(rows) : 0;
        const float v = pr[py * P + lane];  // lane = column of row py
        const unsigned long long cols = (__ballot(v == b) >> base) & full;
        bidx = py * N + (cols ? __builtin_ctzll(cols) : 0);
    } else {
        float rb = -INFINITY;
        int ri = 0;
#pragma unroll
        for (int x = 0; x < N; ++x) {
            const bool gt = zr[x] > rb;
            rb = gt ? zr[x] : rb;
            ri = gt ? x : ri;
        }
        ri += lane * N;
        if constexpr (C::IN_WAVE) {
            // lexicographic (value desc, index asc) maximum: same as the first maximum
#pragma unroll
            for (int m = 1; m < N; m <<= 1) {
                const float ov = __shfl_xor(rb, m, N);
                const int oi = __shfl_xor(ri, m, N);
                const bool take = (ov > rb) || (ov == rb && oi < ri);
                rb = take ? ov : rb;
                ri = take ? oi : ri;
            }
            best = rb;
            bidx = ri;
        } else {
            pkv[lane] = rb;
            pki[lane] = ri;
            __syncthreads();
            for (int r = 0; r < N; ++r) {
                const float v = pkv[r];
                if (v > best) {
                    best = v;
                    bidx = pki[r];
                }
            }
        }
        if (!(best > -INFINITY)) bidx = 0;  // no maximum at all: (0, 0) as the reference
    }
    const int pby = bidx / N, pbx = bidx - pby * N;
    // (2) footprint: half width at half maximum along +-x, +-y (k < N/2 while the value at
    //     distance k is > best / 2).  Lane l of the window reads the row of the peak at
    //     distance l to the right and the column at distance l downwards; distances to the
    //     left/up are lanes N - l.  The walk lengths are runs of set bits in ballots.
    int hw = 0;
    {
        const float half_best = 0.5f * best;
        const float hv = pr[pby * P + (pbx + lane) % N];
        const float vv = pr[((pby + lane) % N) * P + pbx];
        if constexpr (C::ONE_WAVE) {
            const int base = tid - lane;
            const unsigned long long full = N == 64 ? ~0ull : ((1ull << (N & 63)) - 1);
            const unsigned long long hm = (__ballot(hv > half_best) >> base) & full;
            const unsigned long long vm = (__ballot(vv > half_best) >> base) & full;
            // forward: lanes 0, 1, 2, ...; backward: lanes 0, N-1, N-2, ...
            auto run_fwd = [&](unsigned long long m) {
                const unsigned long long inv = ~m & full;
                return inv ? __builtin_ctzll(inv) : N;
            };
            auto run_bwd = [&](unsigned long long m) {
                if (!(m & 1ull)) return 0;
                // bits N-1, N-2, ... : count leading ones of the N-bit field below bit 0
                const unsigned long long rest = (m >> 1) << (64 - (N - 1)) ;
                const unsigned long long inv = ~rest;
                const int lead = N - 1 > 0 ? (inv ? __builtin_clzll(inv) : 64) : 0;
                return 1 + (lead < N - 1 ? lead : N - 1);
            };
            const int k0 = min(run_fwd(vm), H2), k1 = min(run_bwd(vm), H2);
            const int k2 = min(run_fwd(hm), H2), k3 = min(run_bwd(hm), H2);
            hw = max(max(k0, k1), max(k2, k3));
        } else {
            const int sy[4] = {1, N - 1, 0, 0};
            const int sx[4] = {0, 0, 1, N - 1};
            (void)hv;
            (void)vv;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                int kk = 0;
                while (kk < H2 && pr[((pby + kk * sy[d]) % N) * P + (pbx + kk * sx[d]) % N] > half_best)
                    ++kk;
                hw = kk > hw ? kk : hw;
            }
        }
    }
    int exclude = (int)ceilf(1.5f * (float)hw);
    exclude = exclude < 2 ? 2 : exclude;
    // (3) highest local maximum outside the footprint (all 8 neighbours <= v)
    float second = -INFINITY;
    {
        float cand = -INFINITY;
        const int ru = (lane + N - 1) % N, rd = (lane + 1) % N;
        int dyc = abs(lane - pby);
        dyc = dyc < N - dyc ? dyc : N - dyc;
        const bool row_far = dyc > exclude;
        // Columns of this row outside the footprint as a bit mask (N <= 64): every column
        // when the row is far, else all but the circular run pbx - exclude .. pbx + exclude.
        using Mask = std::conditional_t<(N > 32), unsigned long long, unsigned int>;
        constexpr int MB = 8 * (int)sizeof(Mask);
        constexpr Mask full = N == MB ? ~Mask(0) : ((Mask(1) << (N % MB)) - 1);
        Mask allowed = full;
        if constexpr (N <= 64) {
            if (!row_far) {
                const int len = 2 * exclude + 1;
                if (len >= N) {
                    allowed = 0;
                } else {
                    const Mask run = (Mask(1) << len) - 1;
                    const int s = (pbx - exclude + N) % N;
                    const Mask ex = s == 0 ? run : (Mask)(((run << s) | (run >> (N - s))) & full);
                    allowed = full & ~ex;
                }
            }
        }
        // vertical max of (row above, row, row below) per column, sliding along x
        float u1 = pr[ru * P + 0], d1 = pr[rd * P + 0];
        float vm_prev, vm_cur;
        if constexpr (!RAW_MAX) {
            vm_prev = fmaxf(fmaxf(pr[ru * P + N - 1], zr[N - 1]), pr[rd * P + N - 1]);
            vm_cur = fmaxf(fmaxf(u1, zr[0]), d1);
        } else {
            vm_prev = max3_finite(pr[ru * P + N - 1], zr[N - 1], pr[rd * P + N - 1]);
            vm_cur = max3_finite(u1, zr[0], d1);
        }
#pragma unroll
        for (int x = 0; x < N; ++x) {
            const int xn = x + 1 < N ? x + 1 : 0;
            const float u2 = pr[ru * P + xn], d2 = pr[rd * P + xn];
            const float v = zr[x];
            float vm_next;
            if constexpr (!RAW_MAX) vm_next = fmaxf(fmaxf(u2, zr[xn]), d2);
            else vm_next = max3_finite(u2, zr[xn], d2);
            bool far;
            if constexpr (N <= 64) {
                far = (allowed >> x) & 1;
            } else {
                int dxc = abs(x - pbx);
                dxc = dxc < N - dxc ? dxc : N - dxc;
                far = row_far || dxc > exclude;
            }
            bool local;
            if constexpr (MAYBE_NAN) {
                const float m0 = zr[(x + N - 1) % N], m2 = zr[xn];
                const float u0 = pr[ru * P + (x + N - 1) % N], d0 = pr[rd * P + (x + N - 1) % N];
                local = u0 <= v && u1 <= v && u2 <= v && m0 <= v && m2 <= v && d0 <= v &&
                        d1 <= v && d2 <= v;
                if (local && far && v > cand) cand = v;
            } else if constexpr (!RAW_MAX) {
                // finite values: max(cand, v) is the strict-> update of the reference
                local = fmaxf(fmaxf(vm_prev, vm_next), fmaxf(u1, d1)) <= v;
                cand = (local && far) ? fmaxf(cand, v) : cand;
            } else {
                // the 3x3 maximum (v included) is <= v exactly when every neighbour is; the max
                // is taken unconditionally so the update is a select, not a branch
                local = max3_finite(vm_prev, vm_cur, vm_next) <= v;
                const float up = max_finite(cand, v);
                cand = (local && far) ? up : cand;
            }
            vm_prev = vm_cur;
            vm_cur = vm_next;
            u1 = u2;
            d1 = d2;
        }
        if constexpr (C::IN_WAVE) {
#pragma unroll
            for (int m = 1; m < N; m <<= 1) {
                const float o = __shfl_xor(cand, m, N);
                if constexpr (!RAW_MAX) cand = fmaxf(cand, o);
                else cand = max_finite(cand, o);
            }
            second = cand;
        } else {
            __syncthreads();
            pkv[lane] = cand;
            __syncthreads();
            for (int r = 0; r < N; ++r) second = pkv[r] > second ? pkv[r] : second;
        }
    }
    if (lane == 0 && active) {
        PeakRec r;
        r.best = best;
        r.second = second;
        r.mx = pr[pby * P + (pbx + N - 1) % N];
        r.qx = pr[pby * P + (pbx + 1) % N];
        r.my = pr[((pby + N - 1) % N) * P + pbx];
        r.qy = pr[((pby + 1) % N) * P + pbx];
        r.index = bidx;
        r.has_peak = has_peak ? 1 : 0;
        *rec = r;
    }
}

template <int N, int FMT, int OCC, int DEF>
__global__ __launch_bounds__(KCfg<N>::BLOCK) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_correlate(const PassPair *__restrict__ pairs, uint32_t bpp, uint32_t nblocks)
{
    using C = KCfg<N>;
    constexpr int WPB = C::WPB, P = C::P, PL = C::PL, H2 = N / 2;
    __shared__ float s_re[WPB * PL];
    // a second plane for the deformed loader (window B)
    __shared__ float s_im[DEF == 0 ? 1 : WPB * PL];
    __shared__ float s_red[WPB * N];
    __shared__ float s_pkv[WPB * N];
    __shared__ int s_pki[WPB * N];
    // inverse column pass split over two lanes per column pair (N >= 24; at N = 16 the
    // shorter transform saves less than the twiddles cost): entry (p, n) = 1 for the even half
    // (p = 0), e^{+2 pi i n / N} for the odd half (p = 1), n < N / 2
    constexpr bool SPLIT_INV = N >= 24;
    // windows of at most one wave's width: the inverse by half-spectrum columns and row pairs,
    // every lane running two N/2-point transforms (inverse_half_spectrum)
    constexpr bool HALF_INV = C::NARROW && N >= 16;
    __shared__ float2 s_tw[(SPLIT_INV || HALF_INV) ? N : 1];

    // ---- XCD-aware block order: XCD x (blocks L = x, x+8, ...) gets a contiguous range ----
    const uint32_t L = blockIdx.x;
    const uint32_t q = nblocks >> 3, rem = nblocks & 7, xcd = L & 7;
    const uint32_t Lr = xcd * q + min(xcd, rem) + (L >> 3);
    const uint32_t p = Lr / bpp, blk = Lr - p * bpp;

    const int tid = threadIdx.x;
    const int wl = tid / N;
    const int lane = tid - wl * N;
    if constexpr (SPLIT_INV || HALF_INV) {
        if (tid < N) {
            const int hp = tid / H2, n = tid - hp * H2;
            s_tw[tid] = hp ? make_float2(pivfft::Tw<N>::t.c[n], pivfft::Tw<N>::t.s[n])
                           : make_float2(1.0f, 0.0f);
        }
    }
    const PassPair &pp = pairs[p];
    const GridDev g = pp.grid;
    const uint32_t cells = (uint32_t)(g.nx * g.ny);
    const uint32_t k = blk * WPB + wl;
    const bool active = k < cells;
    float *pr = s_re + wl * PL;
    float *pim = DEF == 0 ? s_im : s_im + wl * PL;
    float *red = s_red + wl * N;
    float *pkv = s_pkv + wl * N;
    int *pki = s_pki + wl * N;

    // ---- origins and signal mask from the placement kernel ------------------------------
    float ax = 0.f, ay = 0.f, bx = 0.f, by = 0.f;
    bool sig = false;
    if (active) {
        const float4 o = reinterpret_cast<const float4 *>(pp.origins)[k];
        ax = o.x;
        ay = o.y;
        bx = o.z;
        by = o.w;
        sig = pp.sig[k] != 0;
    }

    // ---- fused ingest + mean removal --------------------------------------------------
    float zr[N], zi[N];
    ingest<N, FMT, DEF>(pp, sig, ax, ay, bx, by, lane, red, pr, pim, zr, zi);
    // Window energies.  Windows tiling the wave (N | 64) sum (a - m)^2 over the lane's column
    // here (m the approximate mean), with the exact-mean correction from the column sums the
    // column transform leaves in bin 0: two fma per sample instead of four per spectral bin.
    // Other sizes sum |2A|^2 and |2B|^2 over the spectrum (Parseval): their window sums go
    // through LDS, and two more of them next to the transforms cost registers.
    constexpr bool SPATIAL_E = C::IN_WAVE;
    float qa = 0.0f, qb = 0.0f;
    if constexpr (SPATIAL_E) {
#pragma unroll
        for (int r = 0; r < N; ++r) {
            qa = __builtin_fmaf(zr[r], zr[r], qa);
            qb = __builtin_fmaf(zi[r], zi[r], qb);
        }
        // summed here, before the transform: the opaque use keeps the compiler from sinking
        // the sums past it, which would keep a copy of every sample live through the transform
        asm volatile("" : "+v"(qa), "+v"(qb));
    }

    // ---- forward transform of z = a + i b ----------------------------------------------
    // Transposes go through LDS.  A window of at most one wave's width uses a single float
    // plane and moves real and imaginary parts in two passes (the barriers of a one-wave
    // workgroup cost nothing), halving the LDS footprint, which bounds occupancy at N = 32,
    // 48 and 64.
    pivfft::fft<N, -1>(zr, zi);  // along rows r; lane = column
    // sum (a - mean)^2 = sum (a - m)^2 - (sum (a - m))^2 / N^2, in raw units squared
    float ea32 = 0.0f, eb32 = 0.0f;
    if constexpr (SPATIAL_E) {
        constexpr float inv_n2 = 1.0f / (float)(N * N);
        const float wsa = window_sum<N>(zr[0], red, lane), wsb = window_sum<N>(zi[0], red, lane);
        ea32 = window_sum<N>(qa, red, lane) - wsa * wsa * inv_n2;
        eb32 = window_sum<N>(qb, red, lane) - wsb * wsb * inv_n2;
    }
    {
#pragma unroll
        for (int r = 0; r < N; ++r) pr[r * P + lane] = zr[r];
        __syncthreads();
#pragma unroll
        for (int x = 0; x < N; ++x) zr[x] = pr[lane * P + x];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < N; ++r) pr[r * P + lane] = zi[r];
        __syncthreads();
#pragma unroll
        for (int x = 0; x < N; ++x) zi[x] = pr[lane * P + x];
    }
    pivfft::fft<N, -1>(zr, zi);  // lane ky holds Z[ky][kx]

    // ---- cross spectrum conj(FA) FB, DC cleared -------------------------------------------
    // 2A = Z[k] + conj(Z[-k]),  2B = (Z[k] - conj(Z[-k])) / i,  X' = conj(2A) 2B = 4 X
    float ea2 = 0.0f, eb2 = 0.0f;  // Parseval energies (sizes not tiling the wave)
    {
        const int mrow = (N - lane) % N;
        if constexpr (C::ONE_WAVE) {
            // The partner lane (row -ky) supplies Z[-ky][-kx] by shuffle.  Columns x and N-x
            // are done together so both originals are read before either is overwritten.  The
            // half-spectrum inverse reads columns 0 .. N/2 only: with it the columns past N/2
            // are not formed, and their Parseval energies are those of columns 1 .. N/2-1
            // counted twice (|2A(-k)| = |2A(k)|).
            const int src = (tid - lane) + mrow;  // partner lane in this wave
            float ea2d = 0.0f, eb2d = 0.0f;      // Parseval terms counted twice
#pragma unroll
            for (int x = 0; x <= H2; ++x) {
                const int mx = (N - x) % N;
                const bool mirror = !HALF_INV && mx != x;
                const float qr = __shfl(zr[mx], src, 64), qi = __shfl(zi[mx], src, 64);
                float qr2 = 0.0f, qi2 = 0.0f;
                if (mirror) {
                    qr2 = __shfl(zr[x], src, 64);
                    qi2 = __shfl(zi[x], src, 64);
                }
                {
                    const float a2r = zr[x] + qr, a2i = zi[x] - qi;
                    const float b2r = zi[x] + qi, b2i = qr - zr[x];
                    const bool dc = (x == 0) && (lane == 0);
                    if constexpr (!SPATIAL_E) {
                        if (HALF_INV && x != 0 && x != H2) {
                            ea2d = __builtin_fmaf(a2r, a2r, __builtin_fmaf(a2i, a2i, ea2d));
                            eb2d = __builtin_fmaf(b2r, b2r, __builtin_fmaf(b2i, b2i, eb2d));
                        } else {
                            ea2 = dc ? ea2 : __builtin_fmaf(a2r, a2r, __builtin_fmaf(a2i, a2i, ea2));
                            eb2 = dc ? eb2 : __builtin_fmaf(b2r, b2r, __builtin_fmaf(b2i, b2i, eb2));
                        }
                    }
                    const float xr = __builtin_fmaf(a2r, b2r, a2i * b2i);
                    const float xi = __builtin_fmaf(a2r, b2i, -(a2i * b2r));
                    zr[x] = dc ? 0.0f : xr;
                    zi[x] = dc ? 0.0f : xi;
                }
                if (mirror) {
                    const float a2r = zr[mx] + qr2, a2i = zi[mx] - qi2;
                    const float b2r = zi[mx] + qi2, b2i = qr2 - zr[mx];
                    if constexpr (!SPATIAL_E) {
                        ea2 = __builtin_fmaf(a2r, a2r, __builtin_fmaf(a2i, a2i, ea2));
                        eb2 = __builtin_fmaf(b2r, b2r, __builtin_fmaf(b2i, b2i, eb2));
                    }
                    zr[mx] = __builtin_fmaf(a2r, b2r, a2i * b2i);
                    zi[mx] = __builtin_fmaf(a2r, b2i, -(a2i * b2r));
                }
                if ((x & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
            if constexpr (!SPATIAL_E && HALF_INV) {
                ea2 = __builtin_fmaf(2.0f, ea2d, ea2);
                eb2 = __builtin_fmaf(2.0f, eb2d, eb2);
            }
        } else if constexpr (C::PACKED) {
            // A window may cross a wave boundary: Z[-ky][-kx] comes through LDS.  Only the
            // columns the half-spectrum inverse reads (0 .. N/2) are formed, the energies as
            // above, and only their partner columns N - x are exchanged, in one round: row ky
            // of the plane holds Re Z[ky][m] in slot m - N/2 and Im Z[ky][m] in slot m for
            // m = N/2 .. N-1, Re Z[ky][0] in slot N, and the window's reduction row holds
            // Im Z[ky][0].
            static_assert(HALF_INV && !SPATIAL_E, "packed sizes (24, 48) invert half spectra");
            float ea2d = 0.0f, eb2d = 0.0f;  // Parseval terms counted twice
            auto cross = [&](int x, float qr, float qi) {
                const float a2r = zr[x] + qr, a2i = zi[x] - qi;
                const float b2r = zi[x] + qi, b2i = qr - zr[x];
                const bool dc = (x == 0) && (lane == 0);
                if (x != 0 && x != H2) {
                    ea2d = __builtin_fmaf(a2r, a2r, __builtin_fmaf(a2i, a2i, ea2d));
                    eb2d = __builtin_fmaf(b2r, b2r, __builtin_fmaf(b2i, b2i, eb2d));
                } else {
                    ea2 = dc ? ea2 : __builtin_fmaf(a2r, a2r, __builtin_fmaf(a2i, a2i, ea2));
                    eb2 = dc ? eb2 : __builtin_fmaf(b2r, b2r, __builtin_fmaf(b2i, b2i, eb2));
                }
                const float xr = __builtin_fmaf(a2r, b2r, a2i * b2i);
                const float xi = __builtin_fmaf(a2r, b2i, -(a2i * b2r));
                zr[x] = dc ? 0.0f : xr;
                zi[x] = dc ? 0.0f : xi;
            };
            __syncthreads();
#pragma unroll
            for (int m = H2; m < N; ++m) {
                pr[lane * P + m - H2] = zr[m];
                pr[lane * P + m] = zi[m];
            }
            pr[lane * P + N] = zr[0];
            red[lane] = zi[0];
            __syncthreads();
            cross(0, pr[mrow * P + N], red[mrow]);
#pragma unroll
            for (int x = 1; x <= H2; ++x) cross(x, pr[mrow * P + H2 - x], pr[mrow * P + N - x]);
            ea2 = __builtin_fmaf(2.0f, ea2d, ea2);
            eb2 = __builtin_fmaf(2.0f, eb2d, eb2);
        } else {
            // Windows of two waves (N > 64): Z[-ky][-kx] comes through the window's plane in
            // rounds of S column pairs (x, N - x), real and imaginary parts of both columns in
            // 4 S <= N + 1 slots per row, so that no column is rewritten before its partner
            // has read it.
            constexpr int S = (N + 1) / 4;
            auto cross = [&](int x, float qr, float qi) {
                const float a2r = zr[x] + qr, a2i = zi[x] - qi;
                const float b2r = zi[x] + qi, b2i = qr - zr[x];
                const bool dc = (x == 0) && (lane == 0);
                if constexpr (!SPATIAL_E) {
                    ea2 = dc ? ea2 : __builtin_fmaf(a2r, a2r, __builtin_fmaf(a2i, a2i, ea2));
                    eb2 = dc ? eb2 : __builtin_fmaf(b2r, b2r, __builtin_fmaf(b2i, b2i, eb2));
                }
                const float xr = __builtin_fmaf(a2r, b2r, a2i * b2i);
                const float xi = __builtin_fmaf(a2r, b2i, -(a2i * b2r));
                zr[x] = dc ? 0.0f : xr;
                zi[x] = dc ? 0.0f : xi;
            };
#pragma unroll
            for (int x0 = 0; x0 <= H2; x0 += S) {
                __syncthreads();
#pragma unroll
                for (int j = 0; j < S; ++j) {
                    const int x = x0 + j;
                    if (x <= H2) {
                        const int mx = (N - x) % N;
                        pr[lane * P + 4 * j] = zr[x];
                        pr[lane * P + 4 * j + 1] = zi[x];
                        pr[lane * P + 4 * j + 2] = zr[mx];
                        pr[lane * P + 4 * j + 3] = zi[mx];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < S; ++j) {
                    const int x = x0 + j;
                    if (x <= H2) {
                        const int mx = (N - x) % N;
                        const float *q = pr + mrow * P + 4 * j;
                        cross(x, q[2], q[3]);
                        if (mx != x) cross(mx, q[0], q[1]);
                    }
                }
            }
        }
    }
    if constexpr (!SPATIAL_E) {
        // sum_k |2A_k|^2 = 4 N^2 sum (a - mean)^2  (Parseval)
        constexpr float inv4n2 = 1.0f / (4.0f * (float)(N * N));
        ea32 = window_sum<N>(ea2, red, lane) * inv4n2;
        eb32 = window_sum<N>(eb2, red, lane) * inv4n2;
    }
    const double ea = (double)ea32, eb = (double)eb32;
    const double floor_e = energy_floor<FMT>();
    const bool has_peak = sig && !(ea <= floor_e || eb <= floor_e);

    // correlate.rs:134-137: corr = C / (n^2 sqrt(ea eb)), the 1/4 undoes X' = 4 X
    const float denom = (float)((double)(N * N) * sqrt(ea * eb));
    const float scale = 0.25f / denom;
    if constexpr (HALF_INV) {
        inverse_half_spectrum<N>(zr, zi, pr, lane, s_tw, scale);
    } else {
    // ---- Hermitian inverse ------------------------------------------------------------
    // Rows ky = 0..N/2 are inverted along kx (Y[ky][x]); the other rows are conjugates.
    // Then column pairs (2j, 2j+1) are inverted along ky as one complex transform
    //   V[ky] = Y[ky][2j] + i Y[ky][2j+1]  ->  C[.][2j] + i C[.][2j+1].
    pivfft::fft<N, 1>(zr, zi);  // lane ky (<= N/2) holds Y[ky][x]
    const int jj = lane < H2 ? lane : lane - H2;
    {
        // Row owners write V's real parts (virtual rows ky and N-ky of the half plane), keep
        // V's imaginary parts in zi, then the column-pair lanes gather them in two passes.
        constexpr int HW = H2 + 1;  // odd row stride of the half-width plane
        const int s = lane;
        // rows a lane does not own are written to a dummy row N (branch-free stores)
        const int row_own = s <= H2 ? s : N;
        const int row_mir = (s >= 1 && s < H2) ? N - s : N;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < H2; ++j) {
            const float yr0 = zr[2 * j], yi0 = zi[2 * j], yr1 = zr[2 * j + 1], yi1 = zi[2 * j + 1];
            pr[row_own * HW + j] = yr0 - yi1;
            pr[row_mir * HW + j] = yr0 + yi1;
            zi[2 * j] = yi0 + yr1;
            zi[2 * j + 1] = yr1 - yi0;
        }
        __syncthreads();
#pragma unroll
        for (int ky = 0; ky < N; ++ky) zr[ky] = pr[ky * HW + jj];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < H2; ++j) {
            pr[row_own * HW + j] = zi[2 * j];
            pr[row_mir * HW + j] = zi[2 * j + 1];
        }
        __syncthreads();
#pragma unroll
        for (int ky = 0; ky < N; ++ky) zi[ky] = pr[ky * HW + jj];
    }
    const int c0 = 2 * jj, c1 = 2 * jj + 1;
    if constexpr (SPLIT_INV) {
        // Column pair jj by two lanes (decimation in frequency): lane half hp transforms
        // V[n] + (-1)^hp V[n + N/2], the odd half twiddled by e^{+2 pi i n / N}, with an N/2-point
        // inverse, and owns rows y = 2k + hp of the two columns.
        const int hp = lane / H2;
        const float sg = hp ? -1.0f : 1.0f;
        float ur[H2], ui[H2];
#pragma unroll
        for (int n = 0; n < H2; ++n) {
            const float ar = __builtin_fmaf(sg, zr[n + H2], zr[n]);
            const float ai = __builtin_fmaf(sg, zi[n + H2], zi[n]);
            if (n == 0) {
                ur[n] = ar;
                ui[n] = ai;
            } else {
                const float2 w = s_tw[hp * H2 + n];
                ur[n] = __builtin_fmaf(ar, w.x, -(ai * w.y));
                ui[n] = __builtin_fmaf(ar, w.y, ai * w.x);
            }
        }
        pivfft::fft<H2, 1>(ur, ui);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < H2; ++k) {
            pr[(2 * k + hp) * P + c0] = ur[k] * scale;
            pr[(2 * k + hp) * P + c1] = ui[k] * scale;
        }
    } else {
        pivfft::fft<N, 1>(zr, zi);
        __syncthreads();
        if (lane < H2) {
#pragma unroll
            for (int y = 0; y < N; ++y) {
                pr[y * P + c0] = zr[y] * scale;
                pr[y * P + c1] = zi[y] * scale;
            }
        }
    }
    }
    __syncthreads();

    find_peak_plane<N, FMT>(pr, lane, tid, pkv, pki, active, has_peak, pp.peaks + k);
}

template <int N, int FMT, int OCC, int DEF = 0>
hipError_t launch_corr_t(const PassPair *pairs, int npairs, uint32_t max_cells, hipStream_t st)
{
    using C = KCfg<N>;
    const uint32_t bpp = (max_cells + C::WPB - 1) / C::WPB;
    const uint32_t blocks = bpp * (uint32_t)npairs;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_correlate<N, FMT, OCC, DEF>), dim3(blocks), dim3(C::BLOCK), 0, st, pairs, bpp,
                       blocks);
    return hipGetLastError();
}

// deform != 0 (a deformed pass) runs the default variant's occupancy with the deformed loader;
// only sizes N <= 64 have deformed instances here (larger windows deform in the
// generic-size kernel)
template <int N, int FMT>
hipError_t launch_corr_v(int variant, int deform, const PassPair *pairs, int npairs,
                         uint32_t max_cells, hipStream_t st)
{
    constexpr VariantSet vs = corr_variants(N);
    if constexpr (KCfg<N>::NARROW) {
        // the deformed loader's second LDS plane bounds the workgroups per CU (160 KiB of LDS,
        // 32 waves): the kernel's shared arrays, two planes, the three per-lane rows and the
        // twiddles
        using C = KCfg<N>;
        constexpr int lds = 2 * C::WPB * C::PL * 4 + 3 * C::WPB * N * 4 + N * 8;
        constexpr int wpg = (C::BLOCK + 63) / 64;
        constexpr int wgs = 163840 / lds < 32 / wpg ? 163840 / lds : 32 / wpg;
        constexpr int occ_lds = (wgs * wpg + 3) / 4;
        constexpr int occ = vs.occ[0] < occ_lds ? vs.occ[0] : (occ_lds < 1 ? 1 : occ_lds);
        if (deform == 1) return launch_corr_t<N, FMT, occ, 1>(pairs, npairs, max_cells, st);
        if (deform == 2) return launch_corr_t<N, FMT, occ, 2>(pairs, npairs, max_cells, st);
    } else {
        if (deform != 0) return hipErrorInvalidValue;
    }
    if (variant == 0) return launch_corr_t<N, FMT, vs.occ[0]>(pairs, npairs, max_cells, st);
    if constexpr (vs.count > 1)
        if (variant == 1) return launch_corr_t<N, FMT, vs.occ[1]>(pairs, npairs, max_cells, st);
    if constexpr (vs.count > 2)
        if (variant == 2) return launch_corr_t<N, FMT, vs.occ[2]>(pairs, npairs, max_cells, st);
    return hipErrorInvalidValue;
}

template <int N>
hipError_t launch_corr_n(int fmt, int variant, int deform, const PassPair *pairs, int npairs,
                         uint32_t max_cells, hipStream_t st)
{
    switch (fmt) {
    case FMT_U8: return launch_corr_v<N, FMT_U8>(variant, deform, pairs, npairs, max_cells, st);
    case FMT_U16: return launch_corr_v<N, FMT_U16>(variant, deform, pairs, npairs, max_cells, st);
    default: return launch_corr_v<N, FMT_F32>(variant, deform, pairs, npairs, max_cells, st);
    }
}

}  // namespace piv

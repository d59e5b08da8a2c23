// Correlation pass for window sizes without a register-resident FFT kernel: any even n up to
// NMAX (128), every window of every pair of a pass in one launch, one 256-lane workgroup per
// window.  The reference accepts every even n >= 8 (multipass.rs:96-100; realfft plans any
// length, correlate.rs:53-65); the sizes with an instantiated k_correlate<N> (8, 10, 12, 14,
// 16, 20, 24, 28, 32, 40, 48, 56, 64, 96) use it, every other size comes here, and so do
// deformed passes of windows above 64.
//
// Per window (correlate.rs:99-139):
//   1. the window pair as one complex plane z = a + i b in LDS (row stride n + 1), sampled as
//      k_correlate samples it (bilinear in raw units, or deformed), approximate means removed;
//   2. 2-D DFT as two passes of direct length-n DFTs (along x, then y): each output is a sum
//      over one line of the plane times twiddles from a table computed in double; a thread
//      keeps several outputs of one line in registers, so a chunk of lines is transformed in
//      place between two barriers;
//   3. cross spectrum conj(FA) FB from Z[k] and Z[-k] (one thread per bin pair), the DC bin
//      cleared, the window energies by Parseval;
//   4. inverse along y, then along x keeping the real part, scaled into a float plane that
//      overlays the first half of the complex one;
//   5. find_peak (correlate.rs:215-303) by block reductions, the peak record for k_peak_finish.
// Direct DFTs cost n^3 per pass instead of n^2 log n: a fallback for uncommon sizes, where the
// reference's own rustfft runs Bluestein or Rader for large prime factors.
#pragma once

#include "correlate.hpp"

namespace piv {

constexpr int kAnyBlock = 256;
constexpr int kAnyOuts = 8;  // outputs per thread per chunk of lines

// Lines per chunk: the largest power of two L <= n with ceil(n L / 256) <= 8, so that the
// 256 / L threads of each line cover its n outputs with at most 8 each.
__device__ __forceinline__ int any_lines(int n)
{
    int L = 1;
    while (2 * L <= n && 2 * L * n <= kAnyBlock * kAnyOuts) L *= 2;
    return L;
}

// One pass of length-n DFTs over the lines of the plane z (row stride S), in place.  ROWS:
// along x for every row, else along y for every column.  SIGN -1 forward, +1 inverse.
// REAL_OUT: the real part of the result goes to the float plane zf (row stride S) instead.
template <bool ROWS, int SIGN, bool REAL_OUT>
__device__ __forceinline__ void any_dft_pass(float2 *z, float *zf, const float2 *tw, int n, int S)
{
    const int t = threadIdx.x;
    const int L = any_lines(n);
    const int tpl = kAnyBlock / L;  // threads per line
    const int li = t % L, tk = t / L;
    for (int l0 = 0; l0 < n; l0 += L) {
        const int line = l0 + li;
        const bool on = line < n;
        float ar[kAnyOuts], ai[kAnyOuts];
        int idx[kAnyOuts];
#pragma unroll
        for (int j = 0; j < kAnyOuts; ++j) {
            ar[j] = 0.0f;
            ai[j] = 0.0f;
            idx[j] = 0;
        }
        if (on) {
            for (int m = 0; m < n; ++m) {
                const float2 v = ROWS ? z[line * S + m] : z[m * S + line];
#pragma unroll
                for (int j = 0; j < kAnyOuts; ++j) {
                    const int k = tk + tpl * j;
                    if (k < n) {
                        const float2 w = tw[idx[j]];
                        const float ws = SIGN < 0 ? -w.y : w.y;
                        ar[j] = __builtin_fmaf(v.x, w.x, __builtin_fmaf(-v.y, ws, ar[j]));
                        ai[j] = __builtin_fmaf(v.x, ws, __builtin_fmaf(v.y, w.x, ai[j]));
                        idx[j] += k;
                        idx[j] -= idx[j] >= n ? n : 0;
                    }
                }
            }
        }
        __syncthreads();
        if (on) {
#pragma unroll
            for (int j = 0; j < kAnyOuts; ++j) {
                const int k = tk + tpl * j;
                if (k < n) {
                    if constexpr (REAL_OUT) zf[line * S + k] = ar[j];
                    else if constexpr (ROWS) z[line * S + k] = make_float2(ar[j], ai[j]);
                    else z[k * S + line] = make_float2(ar[j], ai[j]);
                }
            }
        }
        __syncthreads();
    }
}

// Block-wide sum of one float per thread (all threads receive it).
__device__ __forceinline__ float any_sum(float v, float *red)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

template <int FMT, int DEF, int NMAX>
__global__ __launch_bounds__(kAnyBlock) void k_correlate_any(const PassPair *__restrict__ pairs,
                                                            uint32_t bpp)
{
    __shared__ float2 s_z[NMAX * (NMAX + 1)];
    __shared__ float2 s_tw[NMAX];
    __shared__ float s_red[8];
    __shared__ float s_bv[4];
    __shared__ int s_bi[4];
    __shared__ int s_hw;

    const uint32_t b = xcd_block();
    const uint32_t p = b / bpp, k = b - p * bpp;
    const PassPair &pp = pairs[p];
    const GridDev g = pp.grid;
    const int n = g.window, S = n + 1, nn = n * n;
    const int t = threadIdx.x;
    if (k >= (uint32_t)(g.nx * g.ny)) return;  // uniform over the workgroup

    for (int j = t; j < n; j += kAnyBlock) {
        double sv, cv;
        sincospi(2.0 * (double)j / (double)n, &sv, &cv);
        s_tw[j] = make_float2((float)cv, (float)sv);
    }
    const float4 o = reinterpret_cast<const float4 *>(pp.origins)[k];
    const bool sig = pp.sig[k] != 0;

    // ---- window pair as z = a + i b (raw units), sampled per pixel ----------------------
    const bool deformed = DEF != 0 && pp.has_prev;
    float uc = 0.0f, vc = 0.0f;
    const float mx = (o.x + o.z) * 0.5f, my = (o.y + o.w) * 0.5f;
    if (deformed && sig)
        deform_sample(pp.prev, pp.prev_dx, pp.prev_dy, mx + ((float)n - 1.0f) / 2.0f,
                      my + ((float)n - 1.0f) / 2.0f, uc, vc);
    const int xa = (int)floorf(o.x), ya = (int)floorf(o.y), xb = (int)floorf(o.z), yb = (int)floorf(o.w);
    const float txa = o.x - (float)xa, tya = o.y - (float)ya, txb = o.z - (float)xb, tyb = o.w - (float)yb;
    float sa = 0.0f, sb = 0.0f;
    for (int q = t; q < nn; q += kAnyBlock) {
        const int r = q / n, c = q - r * n;
        float va = 0.0f, vb = 0.0f;
        if (sig) {
            float hx = 0.0f, hy = 0.0f;
            if (deformed) {
                float u, v;
                deform_sample(pp.prev, pp.prev_dx, pp.prev_dy, mx + (float)c, my + (float)r, u, v);
                hx = (u - uc) * 0.5f;
                hy = (v - vc) * 0.5f;
            }
            constexpr int MODE = DEF == 2 ? 2 : 1;
            va = deformed_sample<FMT, MODE>(pp.img_a, pp.pitch, pp.H, pp.W, xa + c, ya + r, txa - hx, tya - hy);
            vb = deformed_sample<FMT, MODE>(pp.img_b, pp.pitch, pp.H, pp.W, xb + c, yb + r, txb + hx, tyb + hy);
        }
        s_z[r * S + c] = make_float2(va, vb);
        sa += va;
        sb += vb;
    }
    {
        const float ma = any_sum(sa, s_red) / (float)nn;
        const float mb = any_sum(sb, s_red + 4) / (float)nn;
        for (int q = t; q < nn; q += kAnyBlock) {
            const int r = q / n, c = q - r * n;
            float2 &zz = s_z[r * S + c];
            zz.x -= ma;
            zz.y -= mb;
        }
    }
    __syncthreads();

    // ---- forward 2-D DFT -------------------------------------------------------------
    any_dft_pass<true, -1, false>(s_z, nullptr, s_tw, n, S);
    any_dft_pass<false, -1, false>(s_z, nullptr, s_tw, n, S);

    // ---- cross spectrum conj(FA) FB (times 4), DC cleared; energies by Parseval ---------
    // 2A = Z[k] + conj(Z[-k]),  2B = (Z[k] - conj(Z[-k])) / i;  X[-k] = conj(X[k])
    float ea2 = 0.0f, eb2 = 0.0f;
    for (int q = t; q < nn; q += kAnyBlock) {
        const int ky = q / n, kx = q - ky * n;
        const int my_ = ky ? n - ky : 0, mx_ = kx ? n - kx : 0;
        const int qm = my_ * n + mx_;
        if (qm < q) continue;
        const float2 zk = s_z[ky * S + kx], zm = s_z[my_ * S + mx_];
        const float a2r = zk.x + zm.x, a2i = zk.y - zm.y;
        const float b2r = zk.y + zm.y, b2i = zm.x - zk.x;
        float2 x = make_float2(__builtin_fmaf(a2r, b2r, a2i * b2i), __builtin_fmaf(a2r, b2i, -(a2i * b2r)));
        if (q == 0) {
            x = make_float2(0.0f, 0.0f);
        } else {
            const float w = qm == q ? 1.0f : 2.0f;
            ea2 = __builtin_fmaf(w, __builtin_fmaf(a2r, a2r, a2i * a2i), ea2);
            eb2 = __builtin_fmaf(w, __builtin_fmaf(b2r, b2r, b2i * b2i), eb2);
        }
        s_z[ky * S + kx] = x;
        if (qm != q) s_z[my_ * S + mx_] = make_float2(x.x, -x.y);
    }
    const double norm = 1.0 / (4.0 * (double)nn);
    const double ea = (double)any_sum(ea2, s_red) * norm;
    const double eb = (double)any_sum(eb2, s_red + 4) * norm;
    const double floor_e = energy_floor<FMT>();
    const bool has_peak = sig && !(ea <= floor_e || eb <= floor_e);

    // ---- inverse: along y, then along x keeping the real part ---------------------------
    any_dft_pass<false, 1, false>(s_z, nullptr, s_tw, n, S);
    float *pr = reinterpret_cast<float *>(s_z);
    any_dft_pass<true, 1, true>(s_z, pr, s_tw, n, S);
    // correlate.rs:134-137: corr = C / (n^2 sqrt(ea eb)); the 1/4 undoes X' = 4 X
    const float scale = 0.25f / (float)((double)nn * sqrt(ea * eb));
    for (int q = t; q < nn; q += kAnyBlock) {
        const int r = q / n, c = q - r * n;
        pr[r * S + c] *= scale;
    }
    __syncthreads();

    // ---- find_peak (correlate.rs:215-303) -------------------------------------------------
    // (1) first maximum in row-major order (strict >)
    float bv = -INFINITY;
    int bi = 0x7fffffff;
    for (int q = t; q < nn; q += kAnyBlock) {
        const float v = pr[(q / n) * S + q % n];
        if (v > bv) {
            bv = v;
            bi = q;
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const float ov = __shfl_xor(bv, m, 64);
        const int oi = __shfl_xor(bi, m, 64);
        if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    if ((t & 63) == 0) {
        s_bv[t >> 6] = bv;
        s_bi[t >> 6] = bi;
    }
    __syncthreads();
    float best = s_bv[0];
    int bidx = s_bi[0];
    for (int w = 1; w < 4; ++w)
        if (s_bv[w] > best || (s_bv[w] == best && s_bi[w] < bidx)) {
            best = s_bv[w];
            bidx = s_bi[w];
        }
    if (!(best > -INFINITY)) bidx = 0;
    const int pby = bidx / n, pbx = bidx - pby * n;
    // (2) footprint: run lengths along +-x, +-y while the value exceeds best / 2 (k < n/2)
    if (t == 0) {
        const float half_best = 0.5f * best;
        int hw = 0;
        const int sy[4] = {1, n - 1, 0, 0}, sx[4] = {0, 0, 1, n - 1};
        for (int d = 0; d < 4; ++d) {
            int kk = 0;
            while (kk < n / 2 && pr[((pby + kk * sy[d]) % n) * S + (pbx + kk * sx[d]) % n] > half_best) ++kk;
            hw = kk > hw ? kk : hw;
        }
        s_hw = hw;
    }
    __syncthreads();
    int exclude = (int)ceilf(1.5f * (float)s_hw);
    exclude = exclude < 2 ? 2 : exclude;
    // (3) highest local maximum (all 8 neighbours <= v) farther than `exclude` (circular
    //     Chebyshev distance) from the peak
    float cand = -INFINITY;
    for (int q = t; q < nn; q += kAnyBlock) {
        const int y = q / n, x = q - y * n;
        int dy = abs(y - pby), dx = abs(x - pbx);
        dy = dy < n - dy ? dy : n - dy;
        dx = dx < n - dx ? dx : n - dx;
        if ((dy > dx ? dy : dx) <= exclude) continue;
        const float v = pr[y * S + x];
        if (!(v > cand)) continue;
        bool local = true;
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox)
                if (oy | ox) local = local && pr[((y + n + oy) % n) * S + (x + n + ox) % n] <= v;
        if (local) cand = v;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) cand = fmaxf(cand, __shfl_xor(cand, m, 64));
    __syncthreads();
    if ((t & 63) == 0) s_bv[t >> 6] = cand;
    __syncthreads();
    if (t == 0) {
        const float second = fmaxf(fmaxf(s_bv[0], s_bv[1]), fmaxf(s_bv[2], s_bv[3]));
        PeakRec rec;
        rec.best = best;
        rec.second = second;
        rec.mx = pr[pby * S + (pbx + n - 1) % n];
        rec.qx = pr[pby * S + (pbx + 1) % n];
        rec.my = pr[((pby + n - 1) % n) * S + pbx];
        rec.qy = pr[((pby + 1) % n) * S + pbx];
        rec.index = bidx;
        rec.has_peak = has_peak ? 1 : 0;
        pp.peaks[k] = rec;
    }
}

template <int FMT, int DEF, int NMAX>
hipError_t launch_corr_any_t(const PassPair *pairs, int npairs, uint32_t max_cells, hipStream_t st)
{
    const uint32_t blocks = max_cells * (uint32_t)npairs;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_correlate_any<FMT, DEF, NMAX>), dim3(blocks), dim3(kAnyBlock), 0, st,
                       pairs, max_cells);
    return hipGetLastError();
}

template <int NMAX>
hipError_t launch_corr_any_n(int fmt, int deform, const PassPair *pairs, int npairs,
                             uint32_t max_cells, hipStream_t st)
{
#define PIV_ANY_F(F)                                                                          \
    (deform == 1   ? launch_corr_any_t<F, 1, NMAX>(pairs, npairs, max_cells, st)             \
     : deform == 2 ? launch_corr_any_t<F, 2, NMAX>(pairs, npairs, max_cells, st)             \
                   : launch_corr_any_t<F, 0, NMAX>(pairs, npairs, max_cells, st))
    switch (fmt) {
    case FMT_U8: return PIV_ANY_F(FMT_U8);
    case FMT_U16: return PIV_ANY_F(FMT_U16);
    default: return PIV_ANY_F(FMT_F32);
    }
#undef PIV_ANY_F
}

}  // namespace piv

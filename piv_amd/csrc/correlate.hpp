// The correlation-pass kernel template (one instantiation per window size, see corr_n*.hip).
#pragma once

#include "fft.hpp"
#include "piv_common.hpp"

namespace piv {
// ---------------------------------------------------------------------------------------
// correlation pass kernel
// ---------------------------------------------------------------------------------------

// Windows per workgroup: one wave when N divides 64, three waves (192 lanes) when N divides
// 192, otherwise one window.  The complex transpose plane per window is N x (N+1) floats
// for each of re/im (odd row stride: conflict-free column and row access).
template <int N>
struct KCfg {
    static constexpr int WPB = (64 % N == 0) ? 64 / N : ((192 % N == 0 && N <= 48 && N != 48) ? 192 / N : 1);
    static constexpr int BLOCK = WPB * N;
    static constexpr int P = N + 1;
    static constexpr int PL = N * P;
};

template <int N, int FMT>
__global__ __launch_bounds__(KCfg<N>::BLOCK) void k_correlate(const PassPair *__restrict__ pairs,
                                                              const uint32_t *__restrict__ begins,
                                                              int npairs, uint32_t total,
                                                              uint32_t wpp)
{
    using C = KCfg<N>;
    constexpr int WPB = C::WPB, P = C::P, PL = C::PL, H2 = N / 2;
    __shared__ float s_re[WPB * PL];
    __shared__ float s_im[WPB * PL];
    __shared__ double s_red[WPB * 2 * N];
    __shared__ float s_pkv[WPB * N];
    __shared__ int s_pki[WPB * N];

    const int tid = threadIdx.x;
    const int wl = tid / N;
    const int lane = tid - wl * N;
    const uint32_t gw = blockIdx.x * WPB + wl;
    const bool active = gw < total;
    float *pr = s_re + wl * PL;
    float *pi = s_im + wl * PL;
    double *red = s_red + wl * 2 * N;
    float *pkv = s_pkv + wl * N;
    int *pki = s_pki + wl * N;

    // ---- which pair / cell ------------------------------------------------------------
    uint32_t p = 0, k = 0;
    if (active) {
        if (wpp) {
            p = gw / wpp;
            k = gw - p * wpp;
        } else {
            p = find_pair_u(begins, npairs, gw);
            k = gw - begins[p];
        }
    }
    const PassPair &pp = pairs[p];
    const GridDev g = pp.grid;
    const int ix = (int)(k % (uint32_t)g.nx), iy = (int)(k / (uint32_t)g.nx);

    // ---- placement (multipass.rs:207-231) -------------------------------------------
    float ax = 0.f, ay = 0.f, bx = 0.f, by = 0.f;
    if (active) {
        if (pp.in_origins) {
            ax = pp.in_origins[4 * k + 0];
            ay = pp.in_origins[4 * k + 1];
            bx = pp.in_origins[4 * k + 2];
            by = pp.in_origins[4 * k + 3];
        } else {
            float px = 0.0f, py = 0.0f;
            if (pp.has_prev)
                sample_field(pp.prev, pp.prev_dx, pp.prev_dy, center(g.x0, ix, g.step, N),
                             center(g.y0, iy, g.step, N), px, py);
            const float max_x = (float)(pp.W - N > 1 ? pp.W - N - 1 : 0);
            const float max_y = (float)(pp.H - N > 1 ? pp.H - N - 1 : 0);
            place(g.x0 + ix * g.step, px, max_x, ax, bx);
            place(g.y0 + iy * g.step, py, max_y, ay, by);
        }
    }

    // ---- signal mask (multipass.rs:233-241) -------------------------------------------
    bool sig = active;
    if (active && pp.use_mask) {
        float m, sd;
        sat_stats(pp.sat_a, pp.sat_a + pp.sat_cells, pp.W, (int)roundf(ax), (int)roundf(ay), N,
                  N, m, sd);
        if (!(m >= pp.min_mean && sd >= pp.min_std)) sig = false;
        sat_stats(pp.sat_b, pp.sat_b + pp.sat_cells, pp.W, (int)roundf(bx), (int)roundf(by), N,
                  N, m, sd);
        if (!(m >= pp.min_mean && sd >= pp.min_std)) sig = false;
    }

    // ---- fused ingest: window columns (correlate.rs:182-202) -----------------------------
    float zr[N], zi[N];
    if (sig) {
        load_column<N, FMT>(pp.img_a, pp.pitch, pp.H, pp.W, ax, ay, lane, zr);
        load_column<N, FMT>(pp.img_b, pp.pitch, pp.H, pp.W, bx, by, lane, zi);
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r) {
            zr[r] = 0.0f;
            zi[r] = 0.0f;
        }
    }

    // ---- mean removal and energy (correlate.rs:203-209) -------------------------------
    {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            sa += (double)zr[r];
            sb += (double)zi[r];
        }
        red[lane] = sa;
        red[N + lane] = sb;
    }
    __syncthreads();
    double ea = 0.0, eb = 0.0;
    {
        double ta = 0.0, tb = 0.0;
        for (int i = 0; i < N; ++i) {
            ta += red[i];
            tb += red[N + i];
        }
        const float ma = (float)(ta / (double)(N * N));
        const float mb = (float)(tb / (double)(N * N));
#pragma unroll
        for (int r = 0; r < N; ++r) {
            zr[r] -= ma;
            zi[r] -= mb;
            ea += (double)zr[r] * (double)zr[r];
            eb += (double)zi[r] * (double)zi[r];
        }
    }
    __syncthreads();
    red[lane] = ea;
    red[N + lane] = eb;
    __syncthreads();
    ea = 0.0;
    eb = 0.0;
    for (int i = 0; i < N; ++i) {
        ea += red[i];
        eb += red[N + i];
    }
    const bool has_peak = sig && !(ea <= 1e-12 || eb <= 1e-12);

    // ---- forward transform of z = a + i b ----------------------------------------------
    pivfft::fft<N, -1>(zr, zi);  // along rows r, lane = column c
    // transpose: lane c holds Z1[ky][c]; lane ky takes row ky
#pragma unroll
    for (int r = 0; r < N; ++r) {
        pr[r * P + lane] = zr[r];
        pi[r * P + lane] = zi[r];
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < N; ++x) {
        zr[x] = pr[lane * P + x];
        zi[x] = pi[lane * P + x];
    }
    pivfft::fft<N, -1>(zr, zi);  // lane ky holds Z[ky][kx]
    __syncthreads();
#pragma unroll
    for (int x = 0; x < N; ++x) {
        pr[lane * P + x] = zr[x];
        pi[lane * P + x] = zi[x];
    }
    __syncthreads();

    // ---- cross spectrum conj(FA) * FB (correlate.rs:128-131) --------------------------
    // 2A = Z[k] + conj(Z[-k]),  2B = (Z[k] - conj(Z[-k])) / i ;  X' = conj(2A) (2B) = 4 X
    {
        const int mrow = (N - lane) % N;
#pragma unroll
        for (int x = 0; x < N; ++x) {
            const int mx = (N - x) % N;
            const float qr = pr[mrow * P + mx], qi = pi[mrow * P + mx];
            const float a2r = zr[x] + qr, a2i = zi[x] - qi;
            const float b2r = zi[x] + qi, b2i = qr - zr[x];
            zr[x] = __builtin_fmaf(a2r, b2r, a2i * b2i);
            zi[x] = __builtin_fmaf(a2r, b2i, -(a2i * b2r));
        }
    }
    // rows ky = 0..N/2 of the Hermitian cross spectrum: inverse along kx
    pivfft::fft<N, 1>(zr, zi);  // lane ky (<= N/2) holds Y[ky][x]
    __syncthreads();
    if (lane <= H2) {
#pragma unroll
        for (int x = 0; x < N; ++x) {
            pr[lane * P + x] = zr[x];
            pi[lane * P + x] = zi[x];
        }
    }
    __syncthreads();
    // column pair (2j, 2j+1), inverse along ky: C[.][2j] + i C[.][2j+1]
    {
        const int j = lane < H2 ? lane : lane - H2;
        const int c0 = 2 * j, c1 = 2 * j + 1;
#pragma unroll
        for (int ky = 0; ky < N; ++ky) {
            if (ky <= H2) {
                const float y0r = pr[ky * P + c0], y0i = pi[ky * P + c0];
                const float y1r = pr[ky * P + c1], y1i = pi[ky * P + c1];
                zr[ky] = y0r - y1i;
                zi[ky] = y0i + y1r;
            } else {
                const int s = N - ky;
                const float y0r = pr[s * P + c0], y0i = pi[s * P + c0];
                const float y1r = pr[s * P + c1], y1i = pi[s * P + c1];
                zr[ky] = y0r + y1i;
                zi[ky] = y1r - y0i;
            }
        }
        pivfft::fft<N, 1>(zr, zi);
        // normalisation (correlate.rs:134-137): `1.0 / (n^2 sqrt(ea eb)) as f32` divides in
        // f32 after the cast; the 1/4 removes the factor of the packed cross spectrum.
        const float denom = (float)((double)(N * N) * sqrt(ea * eb));
        const float scale = 1.0f / denom;
        __syncthreads();
        if (lane < H2) {
#pragma unroll
            for (int y = 0; y < N; ++y) {
                pr[y * P + c0] = (zr[y] * 0.25f) * scale;
                pr[y * P + c1] = (zi[y] * 0.25f) * scale;
            }
        }
    }
    __syncthreads();

    // ---- find_peak (correlate.rs:215-303), lane = row -----------------------------------
    // (1) first maximum in row-major order (strict >)
    {
        float best = -INFINITY;
        int bi = 0;
        for (int x = 0; x < N; ++x) {
            const float v = pr[lane * P + x];
            if (v > best) {
                best = v;
                bi = lane * N + x;
            }
        }
        pkv[lane] = best;
        pki[lane] = bi;
    }
    __syncthreads();
    float best = -INFINITY;
    int bidx = 0;
    for (int r = 0; r < N; ++r) {
        const float v = pkv[r];
        if (v > best) {
            best = v;
            bidx = pki[r];
        }
    }
    const int pby = bidx / N, pbx = bidx - pby * N;
    // (2) footprint: half width at half maximum along +-x, +-y
    int hw = 0;
    {
        const float half_best = 0.5f * best;
        const int sy[4] = {1, N - 1, 0, 0};
        const int sx[4] = {0, 0, 1, N - 1};
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            int kk = 0;
            while (kk < H2 && pr[((pby + kk * sy[d]) % N) * P + (pbx + kk * sx[d]) % N] > half_best)
                ++kk;
            hw = kk > hw ? kk : hw;
        }
    }
    int exclude = (int)ceilf(1.5f * (float)hw);
    exclude = exclude < 2 ? 2 : exclude;
    // (3) highest local maximum outside the footprint
    {
        float cand = -INFINITY;
        const int ru = (lane + N - 1) % N, rd = (lane + 1) % N;
        int dyc = abs(lane - pby);
        dyc = dyc < N - dyc ? dyc : N - dyc;
        float u0 = pr[ru * P + N - 1], u1 = pr[ru * P + 0];
        float m0 = pr[lane * P + N - 1], m1 = pr[lane * P + 0];
        float d0 = pr[rd * P + N - 1], d1 = pr[rd * P + 0];
        for (int x = 0; x < N; ++x) {
            const int xn = x + 1 < N ? x + 1 : 0;
            const float u2 = pr[ru * P + xn], m2 = pr[lane * P + xn], d2 = pr[rd * P + xn];
            int dxc = abs(x - pbx);
            dxc = dxc < N - dxc ? dxc : N - dxc;
            const int cd = dyc > dxc ? dyc : dxc;
            const float v = m1;
            if (v > cand && cd > exclude && u0 <= v && u1 <= v && u2 <= v && m0 <= v && m2 <= v &&
                d0 <= v && d1 <= v && d2 <= v)
                cand = v;
            u0 = u1; u1 = u2;
            m0 = m1; m1 = m2;
            d0 = d1; d1 = d2;
        }
        __syncthreads();
        pkv[lane] = cand;
    }
    __syncthreads();
    if (lane == 0 && active) {
        float second = -INFINITY;
        for (int r = 0; r < N; ++r) second = pkv[r] > second ? pkv[r] : second;
        // (4) three-point Gaussian (parabolic fallback) sub-pixel fit
        float sub[2];
#pragma unroll
        for (int axis = 0; axis < 2; ++axis) {
            const float m = axis == 0 ? pr[pby * P + (pbx + N - 1) % N] : pr[((pby + N - 1) % N) * P + pbx];
            const float c = best;
            const float q = axis == 0 ? pr[pby * P + (pbx + 1) % N] : pr[((pby + 1) % N) * P + pbx];
            float d;
            if (m > 0.0f && c > 0.0f && q > 0.0f) {
                const float lm = (float)log((double)m), lc = (float)log((double)c),
                            lq = (float)log((double)q);
                d = (lm - lq) / (2.0f * (lm - 2.0f * lc + lq));
            } else {
                d = (m - q) / (2.0f * (m - 2.0f * c + q));
            }
            sub[axis] = isfinite(d) ? clamp_rs(d, -0.5f, 0.5f) : 0.0f;
        }
        // (5) wrap to signed displacement, peak ratio
        const float wx = pbx >= H2 ? (float)pbx - (float)N : (float)pbx;
        const float wy = pby >= H2 ? (float)pby - (float)N : (float)pby;
        float ppr = 1000.0f;
        if (second > 0.0f) {
            ppr = best / second;
            ppr = (ppr > 1000.0f || ppr != ppr) ? 1000.0f : ppr;
        }
        // assembly (multipass.rs:238-259)
        if (has_peak) {
            pp.out_dx[k] = bx - ax + (wx + sub[0]);
            pp.out_dy[k] = by - ay + (wy + sub[1]);
            pp.out_ppr[k] = ppr;
            pp.out_flags[k] = 0;
            if (pp.out_peak) pp.out_peak[k] = bidx;
        } else {
            pp.out_dx[k] = NAN;
            pp.out_dy[k] = NAN;
            pp.out_ppr[k] = 0.0f;
            pp.out_flags[k] = NO_SIGNAL;
            if (pp.out_peak) pp.out_peak[k] = -1;
        }
        if (pp.out_origins) {
            pp.out_origins[4 * k + 0] = ax;
            pp.out_origins[4 * k + 1] = ay;
            pp.out_origins[4 * k + 2] = bx;
            pp.out_origins[4 * k + 3] = by;
        }
    }
}


template <int N, int FMT>
hipError_t launch_corr_t(const PassPair *pairs, const uint32_t *begins, int npairs,
                                uint32_t total, uint32_t wpp, hipStream_t st)
{
    using C = KCfg<N>;
    const uint32_t blocks = (total + C::WPB - 1) / C::WPB;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_correlate<N, FMT>), dim3(blocks), dim3(C::BLOCK), 0, st, pairs, begins,
                       npairs, total, wpp);
    return hipGetLastError();
}

template <int N>
hipError_t launch_corr_n(int fmt, const PassPair *pairs, const uint32_t *begins,
                                int npairs, uint32_t total, uint32_t wpp, hipStream_t st)
{
    switch (fmt) {
    case FMT_U8: return launch_corr_t<N, FMT_U8>(pairs, begins, npairs, total, wpp, st);
    case FMT_U16: return launch_corr_t<N, FMT_U16>(pairs, begins, npairs, total, wpp, st);
    default: return launch_corr_t<N, FMT_F32>(pairs, begins, npairs, total, wpp, st);
    }
}

}  // namespace piv

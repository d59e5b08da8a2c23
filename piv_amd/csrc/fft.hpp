// Register-resident mixed-radix FFT for one line of length N held by one lane.
//
// Design (gfx950, wave64): the correlation kernels give every lane a whole line of an
// interrogation window (one row or one column), so a length-N transform is straight-line
// VALU code on N complex values in VGPRs -- no shuffles, no LDS inside a line.  Lines are
// exchanged between lanes only at the 2-D transpose points, through LDS.
//
// The transform is a Stockham autosort FFT (natural order in and out) with radix-4 stages
// first, then radix 2, 3, 5 and 7, then any other odd prime up to 31 as one symmetric-pair
// DFT stage, so every N whose prime factors are at most 31 is covered.  All loops
// are fully unrolled, indices and twiddles are compile-time constants: twiddles come from a
// constexpr table evaluated in double and rounded to float, and multiplications by 1, -1, +-i are
// elided.  Complex products use explicit fma so the rest of the library can be built with
// -ffp-contract=off (bit-exact scalar paths) without slowing the butterflies down.
#pragma once

#include <hip/hip_runtime.h>

namespace pivfft {

// ---- constexpr cos/sin of 2*pi*k/n (compile-time twiddles) ---------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x)
{
    double term = x, sum = x;
    for (int i = 1; i < 14; ++i) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x)
{
    double term = 1.0, sum = 1.0;
    for (int i = 1; i < 14; ++i) {
        term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
// cos and sin of 2*pi*k/n via octant reduction on the exact rational k/n
constexpr void cos_sin_2pi(long k, long n, double &c, double &s)
{
    k %= n;
    if (k < 0) k += n;
    // angle = 2 pi k / n ; reduce to first octant using 8k/n
    const long q8 = (8 * k) / n;  // octant 0..7
    // residual fraction within octant, expressed as angle in [0, pi/4)
    const double frac = (double)(8 * k - q8 * n) / (double)n;  // in [0,1)
    const double a = frac * (kPi / 4.0);
    const double ca = taylor_cos(a), sa = taylor_sin(a);
    // angle = q8*pi/4 + a
    const double h = 0.70710678118654752440084436210484903;
    switch (q8) {
    case 0: c = ca; s = sa; break;
    case 1: c = h * (ca - sa); s = h * (ca + sa); break;  // pi/4 + a
    case 2: c = -sa; s = ca; break;                       // pi/2 + a
    case 3: c = -h * (ca + sa); s = h * (ca - sa); break; // 3pi/4 + a
    case 4: c = -ca; s = -sa; break;
    case 5: c = -h * (ca - sa); s = -h * (ca + sa); break;
    case 6: c = sa; s = -ca; break;
    default: c = h * (ca + sa); s = -h * (ca - sa); break;
    }
    // octant 1,3,5,7 via the h-combination lose a little accuracy; recompute directly from
    // the mirrored octant for those instead
    if (q8 & 1) {
        const double frac2 = (double)((q8 + 1) * n - 8 * k) / (double)n;  // (0,1]
        const double b = frac2 * (kPi / 4.0);
        const double cb = taylor_cos(b), sb = taylor_sin(b);
        // angle = (q8+1)*pi/4 - b
        switch (q8) {
        case 1: c = sb; s = cb; break;     // pi/2 - b
        case 3: c = -cb; s = sb; break;    // pi - b
        case 5: c = -sb; s = -cb; break;   // 3pi/2 - b
        default: c = cb; s = -sb; break;   // 2pi - b
        }
    }
}

template <int N>
struct Twiddles {
    float c[N];
    float s[N];
    constexpr Twiddles() : c(), s()
    {
        for (int k = 0; k < N; ++k) {
            double cc = 0, ss = 0;
            cos_sin_2pi(k, N, cc, ss);
            c[k] = (float)cc;
            s[k] = (float)ss;
        }
    }
};

template <int N>
struct Tw {
    static constexpr Twiddles<N> t{};
};

constexpr int pick_radix(int rem)
{
    return rem % 4 == 0 ? 4
         : rem % 2 == 0 ? 2
         : rem % 3 == 0 ? 3
         : rem % 5 == 0 ? 5
         : rem % 7 == 0 ? 7
                        : rem;
}

constexpr bool supported(int n)
{
    if (n < 2) return false;
    for (int p : {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31})
        while (n % p == 0) n /= p;
    return n == 1;
}

// multiply (re, im) by W_N^t = exp(S * 2 pi i t / N)
template <int N, int S>
__device__ __forceinline__ void twiddle(float &re, float &im, int t)
{
    t %= N;
    if (t == 0) return;
    if (4 * t == N) {  // S*i
        const float r = re;
        if (S < 0) { re = im; im = -r; } else { re = -im; im = r; }
        return;
    }
    if (2 * t == N) { re = -re; im = -im; return; }
    if (4 * t == 3 * N) {  // -S*i
        const float r = re;
        if (S < 0) { re = -im; im = r; } else { re = im; im = -r; }
        return;
    }
    const float c = Tw<N>::t.c[t];
    const float s = S < 0 ? -Tw<N>::t.s[t] : Tw<N>::t.s[t];
    const float r = re;
    re = __builtin_fmaf(r, c, -im * s);
    im = __builtin_fmaf(r, s, im * c);
}

// Twiddles with their scale folded out (the FMA form of a twiddled butterfly):
//   W = m (1 + i q)  with m = cos, q = tan   when |cos| >= |sin|  (form 1),
//   W = m (q + i)    with m = sin, q = cot   otherwise            (form 2),
// so v W = m y with y = v (1 + i q) or v (q + i), two fma per complex value, |q| <= 1.  The
// factor m is then applied inside the butterfly's own additions (a + m y is one fma), so a
// twiddle costs two instructions instead of four.  Exact rotations (1, -1, +-i) keep form 0
// with m = 1.  k4[t] = m(3t) / m(t) lets the radix-4 butterfly fold input 3 onto input 1.
// Values are computed in double and rounded once.
template <int N, int S>
struct TwFold {
    float m[N], q[N], k4[N];
    int form[N];
    constexpr TwFold() : m(), q(), k4(), form()
    {
        double md[N] = {};
        for (int t = 0; t < N; ++t) {
            double c = 0, s = 0;
            cos_sin_2pi(t, N, c, s);
            if (S < 0) s = -s;
            if (t == 0 || 4 * t == N || 2 * t == N || 4 * t == 3 * N) {
                form[t] = 0;
                md[t] = 1.0;
                q[t] = 0.0f;
            } else if ((c < 0 ? -c : c) >= (s < 0 ? -s : s)) {
                form[t] = 1;
                md[t] = c;
                q[t] = (float)(s / c);
            } else {
                form[t] = 2;
                md[t] = s;
                q[t] = (float)(c / s);
            }
            m[t] = (float)md[t];
        }
        for (int t = 0; t < N; ++t) k4[t] = (float)(md[(3 * t) % N] / md[t]);
    }
};

template <int N, int S>
struct TwF {
    static constexpr TwFold<N, S> f{};
};

// (re, im) := (re, im) W_N^t / m(t) with W = exp(S 2 pi i t / N); returns m(t)
template <int N, int S>
__device__ __forceinline__ float twiddle_folded(float &re, float &im, int t)
{
    t %= N;
    if (TwF<N, S>::f.form[t] == 0) {
        twiddle<N, S>(re, im, t);
        return 1.0f;
    }
    const float q = TwF<N, S>::f.q[t], r = re;
    if (TwF<N, S>::f.form[t] == 1) {  // (1 + i q)
        re = __builtin_fmaf(-q, im, r);
        im = __builtin_fmaf(q, r, im);
    } else {  // (q + i)
        re = __builtin_fmaf(q, r, -im);
        im = __builtin_fmaf(q, im, r);
    }
    return TwF<N, S>::f.m[t];
}

// Radix-4 butterfly on inputs twiddled by W_N^{r t} (r = 0..3), the twiddle scales folded
// into the butterfly's additions: 22 instructions instead of 28 with three general twiddles.
template <int N, int S>
__device__ __forceinline__ void dft4_twiddled(float (&vr)[4], float (&vi)[4], int t)
{
    t %= N;
    const float m1 = twiddle_folded<N, S>(vr[1], vi[1], t);
    const float m2 = twiddle_folded<N, S>(vr[2], vi[2], 2 * t);
    (void)twiddle_folded<N, S>(vr[3], vi[3], 3 * t);
    const float k = TwF<N, S>::f.k4[t];  // m(3t) / m(t)
    const float ar = __builtin_fmaf(m2, vr[2], vr[0]), ai = __builtin_fmaf(m2, vi[2], vi[0]);
    const float br = __builtin_fmaf(-m2, vr[2], vr[0]), bi = __builtin_fmaf(-m2, vi[2], vi[0]);
    // v1 + v3 = m1 (y1 + k y3), v1 - v3 = m1 (y1 - k y3)
    const float cr = __builtin_fmaf(k, vr[3], vr[1]), ci = __builtin_fmaf(k, vi[3], vi[1]);
    const float er = __builtin_fmaf(-k, vr[3], vr[1]), ei = __builtin_fmaf(-k, vi[3], vi[1]);
    // W_4 = S i
    const float fr = S < 0 ? ei : -ei;
    const float fi = S < 0 ? -er : er;
    vr[0] = __builtin_fmaf(m1, cr, ar);
    vi[0] = __builtin_fmaf(m1, ci, ai);
    vr[1] = __builtin_fmaf(m1, fr, br);
    vi[1] = __builtin_fmaf(m1, fi, bi);
    vr[2] = __builtin_fmaf(-m1, cr, ar);
    vi[2] = __builtin_fmaf(-m1, ci, ai);
    vr[3] = __builtin_fmaf(-m1, fr, br);
    vi[3] = __builtin_fmaf(-m1, fi, bi);
}

// Radix-2 butterfly on inputs twiddled by W_N^{r t}: 6 instructions instead of 8.
template <int N, int S>
__device__ __forceinline__ void dft2_twiddled(float (&vr)[2], float (&vi)[2], int t)
{
    const float m1 = twiddle_folded<N, S>(vr[1], vi[1], t);
    const float ar = vr[0], ai = vi[0];
    vr[0] = __builtin_fmaf(m1, vr[1], ar);
    vi[0] = __builtin_fmaf(m1, vi[1], ai);
    vr[1] = __builtin_fmaf(-m1, vr[1], ar);
    vi[1] = __builtin_fmaf(-m1, vi[1], ai);
}

template <int R, int S>
__device__ __forceinline__ void dft(float (&vr)[R], float (&vi)[R])
{
    if constexpr (R == 2) {
        const float ar = vr[0], ai = vi[0];
        vr[0] = ar + vr[1];
        vi[0] = ai + vi[1];
        vr[1] = ar - vr[1];
        vi[1] = ai - vi[1];
    } else if constexpr (R == 4) {
        const float ar = vr[0] + vr[2], ai = vi[0] + vi[2];
        const float br = vr[0] - vr[2], bi = vi[0] - vi[2];
        const float cr = vr[1] + vr[3], ci = vi[1] + vi[3];
        const float er = vr[1] - vr[3], ei = vi[1] - vi[3];
        // W_4 = S*i ; W*e
        const float fr = S < 0 ? ei : -ei;
        const float fi = S < 0 ? -er : er;
        vr[0] = ar + cr;
        vi[0] = ai + ci;
        vr[1] = br + fr;
        vi[1] = bi + fi;
        vr[2] = ar - cr;
        vi[2] = ai - ci;
        vr[3] = br - fr;
        vi[3] = bi - fi;
    } else if constexpr (R == 3) {
        constexpr float h = 0.86602540378443864676f;
        const float sr = vr[1] + vr[2], si = vi[1] + vi[2];
        const float dr = vr[1] - vr[2], di = vi[1] - vi[2];
        const float mr = __builtin_fmaf(-0.5f, sr, vr[0]);
        const float mi = __builtin_fmaf(-0.5f, si, vi[0]);
        const float qr = S < 0 ? h * di : -h * di;
        const float qi = S < 0 ? -h * dr : h * dr;
        vr[0] = vr[0] + sr;
        vi[0] = vi[0] + si;
        vr[1] = mr + qr;
        vi[1] = mi + qi;
        vr[2] = mr - qr;
        vi[2] = mi - qi;
    } else if constexpr (R >= 5 && R % 2 == 1) {
        // odd prime R from the symmetric pairs j, R - j: with s_j = v_j + v_{R-j} and
        // d_j = v_j - v_{R-j},  Y_m = v_0 + sum_j cos(2 pi j m / R) s_j + S i sum_j
        // sin(2 pi j m / R) d_j,  and Y_{R-m} the same with -S
        constexpr int H = R / 2;
        float sr[H + 1], si[H + 1], dr[H + 1], di[H + 1];
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            sr[j] = vr[j] + vr[R - j];
            si[j] = vi[j] + vi[R - j];
            dr[j] = vr[j] - vr[R - j];
            di[j] = vi[j] - vi[R - j];
        }
        float y0r = vr[0], y0i = vi[0];
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            y0r += sr[j];
            y0i += si[j];
        }
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            float ar = vr[0], ai = vi[0], br = 0.0f, bi = 0.0f;
#pragma unroll
            for (int j = 1; j <= H; ++j) {
                const float c = Tw<R>::t.c[(j * m) % R], sn = Tw<R>::t.s[(j * m) % R];
                ar = __builtin_fmaf(c, sr[j], ar);
                ai = __builtin_fmaf(c, si[j], ai);
                br = __builtin_fmaf(sn, dr[j], br);
                bi = __builtin_fmaf(sn, di[j], bi);
            }
            // S i B = (-S bi, S br)
            const float qr = S < 0 ? bi : -bi, qi = S < 0 ? -br : br;
            vr[m] = ar + qr;
            vi[m] = ai + qi;
            vr[R - m] = ar - qr;
            vi[R - m] = ai - qi;
        }
        vr[0] = y0r;
        vi[0] = y0i;
    } else {
        static_assert(R == 2 || R == 3 || R == 4 || (R >= 5 && R % 2 == 1), "radix");
    }
}

// One Stockham stage of radix R with NS = product of the radices already applied.
template <int N, int S, int NS, int REM>
__device__ __forceinline__ void stockham(float (&xr)[N], float (&xi)[N])
{
    if constexpr (REM > 1) {
        constexpr int R = pick_radix(REM);
        float yr[N], yi[N];
#pragma unroll
        for (int j = 0; j < N / R; ++j) {
            float vr[R], vi[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                vr[r] = xr[j + r * (N / R)];
                vi[r] = xi[j + r * (N / R)];
            }
            const int jm = j % NS;
            if constexpr (R == 4) {
                dft4_twiddled<N, S>(vr, vi, jm * (N / (NS * R)));
            } else if constexpr (R == 2) {
                dft2_twiddled<N, S>(vr, vi, jm * (N / (NS * R)));
            } else {
#pragma unroll
                for (int r = 1; r < R; ++r) twiddle<N, S>(vr[r], vi[r], r * jm * (N / (NS * R)));
                dft<R, S>(vr, vi);
            }
            const int d = (j / NS) * NS * R + jm;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                yr[d + r * NS] = vr[r];
                yi[d + r * NS] = vi[r];
            }
            // large transforms: keep the scheduler from interleaving too many butterflies
            // (each in-flight butterfly holds extra registers)
            if constexpr (N >= 64)
                if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
        stockham<N, S, NS * R, REM / R>(yr, yi);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            xr[k] = yr[k];
            xi[k] = yi[k];
        }
    }
}

// Unnormalised complex DFT of length N in place: S = -1 forward, +1 inverse.
template <int N, int S>
__device__ __forceinline__ void fft(float (&xr)[N], float (&xi)[N])
{
    static_assert(supported(N), "N must have prime factors of at most 31");
    stockham<N, S, 1, N>(xr, xi);
}

}  // namespace pivfft

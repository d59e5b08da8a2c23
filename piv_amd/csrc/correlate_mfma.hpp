// Reduced-precision correlation pass: the 2-D DFTs as dense DFT-matrix GEMMs on the matrix
// cores (v_mfma_f32_16x16x32_{f16,bf16}: 16-bit operands, FP32 accumulation).
//
// Same pass structure as k_correlate (origins and mask from the placement kernel, peak
// records to the finishing kernel, which does the sub-pixel fit in FP64 for this path); only
// the transform differs.  Per window pair (correlate.rs:99-139 restated):
//   1. ingest (shared with the FP32 kernel): bilinear samples in raw units, mean removed;
//      window energies from the samples (FP32); each window scaled to unit RMS and stored
//      as 16-bit values in LDS, column-major;
//   2. DFT along y, then along x, of A and of B:  F_A = F A F,  F_B = F B F;
//   3. conj(X) = F_A conj(F_B) with the DC bin cleared;
//   4. correlation N^2 c = Re(F conj(X) F)  (the inverse DFT as a forward one on conj(X));
//   5. find_peak on the FP32 plane (shared with the FP32 kernel).
//
// GEMM chaining.  An N x N complex matrix is held as (N/16)^2 tiles in the MFMA result
// layout: lane l = 16 g + n, register j holds entry (16 p + 4 g + j, 16 q + n) of tile (p, q).
// The A operand of v_mfma_f32_16x16x32 holds row l % 16 and k-slots 8 (l / 16) .. + 7, so a
// result tile (real and imaginary part: 4 + 4 registers) IS the A operand of Z^T C with k-slots
// (re, 4 g + j) and (im, 4 g + j): every stage is D = Z^T C with the constant DFT matrix C as
// the B operand (from LDS), and the lane dimension alternates between stages with no data
// movement.  After four stages the plane is back in (y, x) orientation.
#pragma once

#include "correlate.hpp"

namespace piv {

typedef float fv4 __attribute__((ext_vector_type(4)));

template <typename E>
struct Mfma;

template <>
struct Mfma<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    typedef _Float16 v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ fv4 mma(v8 a, v8 b, fv4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    typedef __bf16 v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ fv4 mma(v8 a, v8 b, fv4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

template <int N>
struct MfCfg {
    static_assert(N == 16 || N == 32 || N == 64, "DFT-GEMM window sizes");
    static constexpr int T = N / 16;           // tiles per side
    static constexpr int WPW = 64 / N;         // windows per wave (lane per column)
    // the DFT matrix as B operands: in registers for N <= 32 (2 T^2 x 4 VGPRs), in LDS shared
    // by four waves for N = 64
    static constexpr bool BREG = T <= 2;
    static constexpr int WAVES = BREG ? 1 : 4;
    static constexpr int WPB = WAVES * WPW;    // windows per workgroup
    static constexpr int BLOCK = 64 * WAVES;
    static constexpr int P = N + 1, PL = N * P;
    // power-of-two scale of each DFT stage keeping 16-bit intermediates near unit RMS (a DFT
    // of length N multiplies the RMS by sqrt(N), a cross spectrum squares it); exact, so it
    // only rescales the plane
    static constexpr float S = N == 16 ? 0.25f : (N == 32 ? 0.25f : 0.125f);
};

// B operands of tile (p, t) of C = F (forward DFT, F[x][k] = exp(-2 pi i x k / N)) for this
// lane, k-slots (re, 4g+j) and (im, 4g+j) of rows x = 16p + 4g + j:  B_re = [Fr; -Fi],
// B_im = [Fi; Fr].  The hardware sine and cosine take the angle in turns; x k mod N / N is
// exact, and the 16-bit rounding of the entries dominates their error.
template <int N, typename E>
__device__ __forceinline__ void mf_bconst(int p, int t, int lane, typename Mfma<E>::v8 &br,
                                          typename Mfma<E>::v8 &bi)
{
    // the stage scale S (a power of two) is folded into the matrix: exact, and the stage
    // results come out ready to pack
    constexpr float S = MfCfg<N>::S;
    const int g = lane >> 4, n = lane & 15, k = 16 * t + n;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int x = 16 * p + 4 * g + j;
        const float turns = (float)((x * k) & (N - 1)) * (1.0f / (float)N);
        const float fr = S * __builtin_amdgcn_cosf(turns), fi = -S * __builtin_amdgcn_sinf(turns);
        br[j] = (E)fr;
        br[4 + j] = (E)(-fi);
        bi[j] = (E)fi;
        bi[4 + j] = (E)fr;
    }
}

template <typename E>
__device__ __forceinline__ typename Mfma<E>::v8 mf_pack(fv4 re, fv4 im)
{
    typename Mfma<E>::v8 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = (E)re[j];
        v[4 + j] = (E)im[j];
    }
    return v;
}

// B operand source: registers (small N) or LDS
template <int T, typename E, bool REG>
struct MfB {
    using V8 = typename Mfma<E>::v8;
    const V8 (*rre)[T];
    const V8 (*rim)[T];
    const V8 *lre, *lim;
    int lane;
    __device__ __forceinline__ V8 re(int p, int t) const
    {
        if constexpr (REG) return rre[p][t];
        else return lre[(p * T + t) * 64 + lane];
    }
    __device__ __forceinline__ V8 im(int p, int t) const
    {
        if constexpr (REG) return rim[p][t];
        else return lim[(p * T + t) * 64 + lane];
    }
};

// Output column t of a stage D = Z^T C for every row of tiles q:
//   (re[q], im[q]) = sum_p A[q][p] * B[p][t]   (each B operand read once per column)
template <int T, typename E, class BS>
__device__ __forceinline__ void mf_col(const typename Mfma<E>::v8 (&A)[T][T], const BS &b, int t,
                                       fv4 (&re)[T], fv4 (&im)[T])
{
#pragma unroll
    for (int q = 0; q < T; ++q) {
        re[q] = fv4{0.f, 0.f, 0.f, 0.f};
        im[q] = fv4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int p = 0; p < T; ++p) {
        const auto br = b.re(p, t), bi = b.im(p, t);
#pragma unroll
        for (int q = 0; q < T; ++q) {
            re[q] = Mfma<E>::mma(A[q][p], br, re[q]);
            im[q] = Mfma<E>::mma(A[q][p], bi, im[q]);
        }
    }
}

template <int N, int FMT, typename E, int DEF>
__global__ __launch_bounds__(MfCfg<N>::BLOCK) void k_correlate_mf(const PassPair *__restrict__ pairs,
                                                                  uint32_t bpp, uint32_t nblocks)
{
    using M = MfCfg<N>;
    using V8 = typename Mfma<E>::v8;
    using V4 = typename Mfma<E>::v4;
    constexpr int T = M::T, WPB = M::WPB, P = M::P, PL = M::PL;
    __shared__ float s_pl[WPB * PL];      // per window: 16-bit input planes, then the FP32 plane
    // window B's samples in the deformed loader (N = 64 has no room left: one plane, twice)
    constexpr bool PL2 = DEF != 0 && N < 64;
    __shared__ float s_pl2[PL2 ? WPB * PL : 1];
    __shared__ float s_red[WPB * N];
    __shared__ float s_pkv[WPB * N];
    __shared__ int s_pki[WPB * N];
    __shared__ V8 s_bre[M::BREG ? 1 : T * T * 64], s_bim[M::BREG ? 1 : T * T * 64];

    const int tid = threadIdx.x, wlane = tid & 63;
    V8 r_bre[T][T], r_bim[T][T];  // unused (and removed) when the operands live in LDS
    if constexpr (M::BREG) {
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int t = 0; t < T; ++t) mf_bconst<N, E>(p, t, wlane, r_bre[p][t], r_bim[p][t]);
    } else {
        for (int e = tid; e < T * T * 64; e += M::BLOCK) {
            const int pt = e >> 6;
            mf_bconst<N, E>(pt / T, pt % T, e & 63, s_bre[e], s_bim[e]);
        }
    }
    MfB<T, E, M::BREG> B{r_bre, r_bim, s_bre, s_bim, wlane};

    // ---- XCD-aware block order (as k_correlate) ----
    const uint32_t L = blockIdx.x;
    const uint32_t qq = nblocks >> 3, rem = nblocks & 7, xcd = L & 7;
    const uint32_t Lr = xcd * qq + min(xcd, rem) + (L >> 3);
    const uint32_t pidx = Lr / bpp, blk = Lr - pidx * bpp;
    const PassPair &pp = pairs[pidx];
    const GridDev g = pp.grid;
    const uint32_t cells = (uint32_t)(g.nx * g.ny);
    const int wl = tid / N, lane = tid - wl * N;
    const uint32_t k = blk * WPB + wl;
    const bool active = k < cells;
    float *pr = s_pl + wl * PL;
    float *red = s_red + wl * N;

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
    float zr[N], zi[N];
    ingest<N, FMT, DEF>(pp, sig, ax, ay, bx, by, lane, red, pr, PL2 ? s_pl2 + wl * PL : nullptr, zr, zi);
    float ea = 0.f, eb = 0.f;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        ea = __builtin_fmaf(zr[r], zr[r], ea);
        eb = __builtin_fmaf(zi[r], zi[r], eb);
    }
    ea = window_sum<N>(ea, red, lane);
    eb = window_sum<N>(eb, red, lane);
    const double floor_e = energy_floor<FMT>();
    const bool has_peak = sig && !((double)ea <= floor_e || (double)eb <= floor_e);
    {
        // unit-RMS 16-bit planes, column-major: lane = column x, N consecutive rows
        const float sa = has_peak ? 1.0f / sqrtf(ea / (float)(N * N)) : 0.0f;
        const float sb = has_peak ? 1.0f / sqrtf(eb / (float)(N * N)) : 0.0f;
        E *pa = reinterpret_cast<E *>(pr), *pb = pa + N * N;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            pa[lane * N + r] = (E)(zr[r] * sa);
            pb[lane * N + r] = (E)(zi[r] * sb);
        }
    }
    __syncthreads();

    // ---- the DFT-GEMM chain, one window of this wave at a time (all 64 lanes) ----
    const int wave = tid >> 6, gq = wlane >> 4, nn = wlane & 15;
#pragma unroll 1
    for (int w = 0; w < M::WPW; ++w) {
        float *pw = s_pl + (wave * M::WPW + w) * PL;
        const E *pa = reinterpret_cast<const E *>(pw), *pb = pa + N * N;
        V8 A2a[T][T], A2b[T][T];
        // stage 1: DFT along y of each real window (imaginary k-slots zero); result tile
        // (q, t) is the stage-2 operand A2[t][q]
#pragma unroll
        for (int img = 0; img < 2; ++img) {
            const E *pl = img ? pb : pa;
            V8 A1[T][T];
#pragma unroll
            for (int q = 0; q < T; ++q)
#pragma unroll
                for (int p = 0; p < T; ++p) {
                    const V4 h = *reinterpret_cast<const V4 *>(pl + (16 * q + nn) * N + 16 * p + 4 * gq);
                    V8 v;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        v[j] = h[j];
                        v[4 + j] = (E)0.0f;
                    }
                    A1[q][p] = v;
                }
#pragma unroll
            for (int t = 0; t < T; ++t) {
                fv4 re[T], im[T];
                mf_col<T, E>(A1, B, t, re, im);
#pragma unroll
                for (int q = 0; q < T; ++q) {
                    const V8 v = mf_pack<E>(re[q], im[q]);
                    if (img) A2b[t][q] = v;
                    else A2a[t][q] = v;
                }
            }
        }
        // stage 2: DFT along x -> spectra; cross spectrum conj(X) = F_A conj(F_B), DC cleared
        V8 A3[T][T];
#pragma unroll
        for (int u = 0; u < T; ++u) {
            fv4 ar[T], ai[T], br[T], bi[T];
            mf_col<T, E>(A2a, B, u, ar, ai);
            mf_col<T, E>(A2b, B, u, br, bi);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                fv4 xr, xi;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    xr[j] = __builtin_fmaf(ar[t][j], br[t][j], ai[t][j] * bi[t][j]);
                    xi[j] = __builtin_fmaf(ai[t][j], br[t][j], -(ar[t][j] * bi[t][j]));
                }
                if (t == 0 && u == 0 && wlane == 0) {
                    xr[0] = 0.0f;
                    xi[0] = 0.0f;
                }
                A3[u][t] = mf_pack<E>(xr, xi);
            }
        }
        // stage 3: along k1; stage 4: along k2, real part only -> N^2 c[y][x] (scaled)
        V8 A4[T][T];
#pragma unroll
        for (int v = 0; v < T; ++v) {
            fv4 re[T], im[T];
            mf_col<T, E>(A3, B, v, re, im);
#pragma unroll
            for (int u = 0; u < T; ++u) A4[v][u] = mf_pack<E>(re[u], im[u]);
        }
#pragma unroll
        for (int x = 0; x < T; ++x) {
            fv4 c[T];
#pragma unroll
            for (int v = 0; v < T; ++v) c[v] = fv4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u = 0; u < T; ++u) {
                const V8 br = B.re(u, x);
#pragma unroll
                for (int v = 0; v < T; ++v) c[v] = Mfma<E>::mma(A4[v][u], br, c[v]);
            }
#pragma unroll
            for (int v = 0; v < T; ++v)
#pragma unroll
                for (int j = 0; j < 4; ++j) pw[(16 * v + 4 * gq + j) * P + 16 * x + nn] = c[v][j];
        }
    }
    __syncthreads();
    find_peak_plane<N, FMT>(pr, lane, wlane, s_pkv + wl * N, s_pki + wl * N, active, has_peak,
                            pp.peaks + k);
}

template <int N, int FMT, typename E, int DEF>
hipError_t launch_corr_mf_t(const PassPair *pairs, int npairs, uint32_t max_cells, hipStream_t st)
{
    using M = MfCfg<N>;
    const uint32_t bpp = (max_cells + M::WPB - 1) / M::WPB;
    const uint32_t blocks = bpp * (uint32_t)npairs;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_correlate_mf<N, FMT, E, DEF>), dim3(blocks), dim3(M::BLOCK), 0, st, pairs, bpp,
                       blocks);
    return hipGetLastError();
}

template <int N, typename E, int DEF>
hipError_t launch_corr_mf_f(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                            hipStream_t st)
{
    switch (fmt) {
    case FMT_U8: return launch_corr_mf_t<N, FMT_U8, E, DEF>(pairs, npairs, max_cells, st);
    case FMT_U16: return launch_corr_mf_t<N, FMT_U16, E, DEF>(pairs, npairs, max_cells, st);
    default: return launch_corr_mf_t<N, FMT_F32, E, DEF>(pairs, npairs, max_cells, st);
    }
}

// precision: 1 = fp16, 2 = bf16; deform: the pass's deformation mode (0 rigid)
template <int N>
hipError_t launch_corr_mf_n(int fmt, int precision, int deform, const PassPair *pairs, int npairs,
                            uint32_t max_cells, hipStream_t st)
{
    if (precision == 1) {
        if (deform == 1) return launch_corr_mf_f<N, _Float16, 1>(fmt, pairs, npairs, max_cells, st);
        if (deform == 2) return launch_corr_mf_f<N, _Float16, 2>(fmt, pairs, npairs, max_cells, st);
        return launch_corr_mf_f<N, _Float16, 0>(fmt, pairs, npairs, max_cells, st);
    }
    if (deform == 1) return launch_corr_mf_f<N, __bf16, 1>(fmt, pairs, npairs, max_cells, st);
    if (deform == 2) return launch_corr_mf_f<N, __bf16, 2>(fmt, pairs, npairs, max_cells, st);
    return launch_corr_mf_f<N, __bf16, 0>(fmt, pairs, npairs, max_cells, st);
}

}  // namespace piv

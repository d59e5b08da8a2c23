// Post-analysis stages on the device: per-pair flow statistics of validated fields
// (piv-core/src/flowstats.rs:31-137, piv-io/src/flow.rs:13-34) and the time-series filters
// of a run (piv-core/src/smooth.rs:44-320).
//
// All of it is HBM-bound elementwise or stencil work over [pairs][ny][nx] f32 fields:
//   k_flow_cells   one thread per cell: usable flag, velocity (px -> px/s), vorticity by the
//                  reference's central / one-sided differences, kinetic energy, and the
//                  neighbour-median residual of each component (an 8-key sorting network);
//                  per-block f64 partial sums in a fixed tree order;
//   k_flow_finish  one workgroup per pair: partials summed in block order, then the median
//                  absolute residual by an exact radix select over the residual keys;
//   k_temporal     one thread per (t, y, x): moving mean (f32, offset order) or median (a
//                  bitonic network over register keys);
//   k_spatial      one thread per (t, y, x): normalized Gaussian convolution, f32 sums in
//                  the reference's (row, column) order, the tile and its halo staged in LDS;
//   metrics        keys of hypot distances (total-order keys) + one-workgroup radix select;
//                  stream compaction of the cross-validation errors in index order.
// Arithmetic follows the reference operation by operation (no contraction: -ffp-contract=off),
// so every output is bit-identical to the oracle except the f64 sums of the flow statistics,
// whose association differs (their f32 results agree to rounding).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "post_kernels.hpp"

namespace piv {

namespace {

constexpr uint32_t NO_KEY = 0xffffffffu;  // beyond every key of a non-NaN value

// f32::total_cmp order as an unsigned key (and back)
__device__ __forceinline__ uint32_t total_key(float x)
{
    const uint32_t b = __float_as_uint(x);
    const uint32_t t = b ^ ((uint32_t)((int32_t)b >> 31) >> 1);
    return t ^ 0x80000000u;
}

__device__ __forceinline__ float key_value(uint32_t k)
{
    const uint32_t t = k ^ 0x80000000u;
    return __uint_as_float(t ^ ((uint32_t)((int32_t)t >> 31) >> 1));
}

__device__ __forceinline__ bool finite(float x) { return __builtin_isfinite(x); }

// hypotf as the reference's libm computes it: the exact squares summed in f64, rounded once
__device__ __forceinline__ float hypot_ref(float a, float b)
{
    if (__builtin_isinf(a) || __builtin_isinf(b)) return __builtin_inff();
    if (__builtin_isnan(a) || __builtin_isnan(b)) return a + b;
    return (float)__builtin_sqrt((double)a * (double)a + (double)b * (double)b);
}

__device__ __forceinline__ bool usable_cell(float dx, uint8_t fl)
{
    return finite(dx) && ((fl & INVALID) == 0 || (fl & REPLACED) != 0);
}

__device__ __forceinline__ void cswap(uint32_t &a, uint32_t &b)
{
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Batcher odd-even merge sort of 8 keys (19 compare-exchanges)
__device__ __forceinline__ void sort8(uint32_t (&k)[8])
{
    cswap(k[0], k[1]); cswap(k[2], k[3]); cswap(k[4], k[5]); cswap(k[6], k[7]);
    cswap(k[0], k[2]); cswap(k[1], k[3]); cswap(k[4], k[6]); cswap(k[5], k[7]);
    cswap(k[1], k[2]); cswap(k[5], k[6]);
    cswap(k[0], k[4]); cswap(k[1], k[5]); cswap(k[2], k[6]); cswap(k[3], k[7]);
    cswap(k[2], k[4]); cswap(k[3], k[5]);
    cswap(k[1], k[2]); cswap(k[3], k[4]); cswap(k[5], k[6]);
}

// The k-th smallest key (k < number of keys below NO_KEY) of keys[0..n): three radix passes
// over bits 31-21, 20-10 and 9-0, from the top; the whole workgroup calls it.  Counts are
// order-independent, so the result is exact and deterministic.
constexpr int SEL_BINS = 2048;

// Inclusive prefix sum of v over the workgroup (blockDim.x a multiple of 64, at most 1024);
// wsum: LDS [16].
__device__ uint32_t block_scan(uint32_t v, uint32_t *wsum)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    uint32_t before = 0;
    for (int q = 0; q < wv; ++q) before += wsum[q];
    __syncthreads();
    return v + before;
}

__device__ uint32_t block_select(const uint32_t *__restrict__ keys, uint32_t n, uint32_t k,
                                 uint32_t *hist /* LDS [SEL_BINS] */, uint32_t *sh /* LDS [18] */)
{
    constexpr int shifts[3] = {21, 10, 0};
    constexpr uint32_t widths[3] = {11, 11, 10};
    uint32_t prefix = 0, mask = 0;
    for (int pass = 0; pass < 3; ++pass) {
        const int shift = shifts[pass];
        const uint32_t nb = 1u << widths[pass], bm = nb - 1;
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t key = keys[i];
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & bm], 1u);
        }
        __syncthreads();
        // the bin holding the k-th key: block-wide prefix sum over the bins
        const uint32_t per = (nb + blockDim.x - 1) / blockDim.x, b0 = threadIdx.x * per;
        uint32_t local = 0;
        for (uint32_t j = 0; j < per; ++j) local += b0 + j < nb ? hist[b0 + j] : 0u;
        const uint32_t incl = block_scan(local, sh + 2), excl = incl - local;
        if (excl <= k && k < incl) {
            uint32_t cum = excl;
            for (uint32_t j = 0; j < per; ++j) {
                const uint32_t h = hist[b0 + j];
                if (cum + h > k) {
                    sh[0] = b0 + j;
                    sh[1] = k - cum;
                    break;
                }
                cum += h;
            }
        }
        __syncthreads();
        prefix |= sh[0] << shift;
        mask |= bm << shift;
        k = sh[1];
        __syncthreads();
    }
    return prefix;
}

// number of keys below NO_KEY
__device__ uint32_t block_count_keys(const uint32_t *__restrict__ keys, uint32_t n, uint32_t *sh)
{
    if (threadIdx.x == 0) sh[0] = 0;
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) c += keys[i] != NO_KEY;
    atomicAdd(&sh[0], c);
    __syncthreads();
    const uint32_t r = sh[0];
    __syncthreads();
    return r;
}

// smooth.rs:312-320 median of the first n keys in order (n >= 1)
__device__ float block_median(const uint32_t *__restrict__ keys, uint32_t total, uint32_t n,
                              uint32_t *hist, uint32_t *sh)
{
    const float hi = key_value(block_select(keys, total, n / 2, hist, sh));
    if (n % 2 == 1) return hi;
    const float lo = key_value(block_select(keys, total, n / 2 - 1, hist, sh));
    return 0.5f * (lo + hi);
}

constexpr int FLOW_BLOCK = 256;

struct FlowPartial {
    double ws, wa, e;
    uint32_t nw, nu, nres, pad;
};

// ---- flow statistics -------------------------------------------------------------------
__global__ __launch_bounds__(FLOW_BLOCK) void k_flow_cells(const float *__restrict__ dx,
                                                           const float *__restrict__ dy,
                                                           const uint8_t *__restrict__ flags,
                                                           int ny, int nx, float k, float step,
                                                           FlowPartial *__restrict__ part,
                                                           uint32_t *__restrict__ resid)
{
    const int cells = ny * nx;
    const int p = blockIdx.y;
    const long base = (long)p * cells;
    dx += base;
    dy += base;
    flags += base;
    resid += 2 * base;
    const int i = blockIdx.x * FLOW_BLOCK + threadIdx.x;
    double ws = 0.0, wa = 0.0, e = 0.0;
    uint32_t nw = 0, nu = 0, nres = 0;
    if (i < cells) {
        const int iy = i / nx, ix = i - iy * nx;
        // the 3x3 neighbourhood (q = 3 (dy + 1) + dx + 1), loaded together
        float ndx[9], ndy[9];
        bool nok[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int jy = iy + q / 3 - 1, jx = ix + q % 3 - 1;
            const bool in = jy >= 0 && jy < ny && jx >= 0 && jx < nx;
            const int j = in ? jy * nx + jx : i;
            ndx[q] = dx[j];
            ndy[q] = dy[j];
            nok[q] = in && usable_cell(ndx[q], flags[j]);
        }
        // flowstats.rs:31-66 on u = dx * k along y and v = dy * k along x
        const float ul = ndx[1] * k, uc = ndx[4] * k, uh = ndx[7] * k;
        const float vl = ndy[3] * k, vc = ndy[4] * k, vh = ndy[5] * k;
        const bool cu = nok[4] && finite(uc), lu = nok[1] && finite(ul), hu_ = nok[7] && finite(uh);
        const bool cv = nok[4] && finite(vc), lv = nok[3] && finite(vl), hv_ = nok[5] && finite(vh);
        bool da = false, db = false;
        float a = 0.0f, b = 0.0f;
        if (cu) {
            if (lu && hu_) { a = (uh - ul) / (2.0f * step); da = true; }
            else if (lu) { a = (uc - ul) / step; da = true; }
            else if (hu_) { a = (uh - uc) / step; da = true; }
        }
        if (cv) {
            if (lv && hv_) { b = (vh - vl) / (2.0f * step); db = true; }
            else if (lv) { b = (vc - vl) / step; db = true; }
            else if (hv_) { b = (vh - vc) / step; db = true; }
        }
        if (da && db) {
            const float w = a - b;
            if (finite(w)) {
                ws = (double)w;
                wa = (double)fabsf(w);
                nw = 1;
            }
        }
        // flowstats.rs:83-89: kinetic energy of usable, finite velocities
        if (nok[4] && finite(uc) && finite(vc)) {
            e = 0.5 * ((double)uc * (double)uc + (double)vc * (double)vc);
            nu = 1;
        }
        // flowstats.rs:106-130: residual to the median (element m/2 in total order) of the
        // usable finite neighbours, px units; absent neighbours sort last as NO_KEY
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const float *nv = c == 0 ? ndx : ndy;
            uint32_t key = NO_KEY;
            if (nok[4] && finite(nv[4])) {
                uint32_t nb[8];
                int m = 0;
#pragma unroll
                for (int q = 0, t = 0; q < 9; ++q) {
                    if (q == 4) continue;
                    const bool use = nok[q] && finite(nv[q]);
                    nb[t++] = use ? total_key(nv[q]) : NO_KEY;
                    m += use;
                }
                if (m >= 5) {
                    sort8(nb);
                    uint32_t med = nb[0];
#pragma unroll
                    for (int q = 1; q < 8; ++q) med = (q == m / 2) ? nb[q] : med;
                    key = total_key(fabsf(nv[4] - key_value(med)));
                    ++nres;
                }
            }
            resid[(long)c * cells + i] = key;
        }
    }
    // fixed-order tree reduction of the block's partial sums
    __shared__ double s_d[3][FLOW_BLOCK];
    __shared__ uint32_t s_u[3][FLOW_BLOCK];
    const int t = threadIdx.x;
    s_d[0][t] = ws;
    s_d[1][t] = wa;
    s_d[2][t] = e;
    s_u[0][t] = nw;
    s_u[1][t] = nu;
    s_u[2][t] = nres;
    __syncthreads();
    for (int h = FLOW_BLOCK / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                s_d[q][t] += s_d[q][t + h];
                s_u[q][t] += s_u[q][t + h];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        FlowPartial r;
        r.ws = s_d[0][0];
        r.wa = s_d[1][0];
        r.e = s_d[2][0];
        r.nw = s_u[0][0];
        r.nu = s_u[1][0];
        r.nres = s_u[2][0];
        r.pad = 0;
        part[(long)p * gridDim.x + blockIdx.x] = r;
    }
}

constexpr int FINISH_BLOCK = 1024;

__global__ __launch_bounds__(FINISH_BLOCK) void k_flow_finish(const FlowPartial *__restrict__ part,
                                                              int nblk,
                                                              const uint32_t *__restrict__ resid,
                                                              int cells, float step,
                                                              FlowStatsDev *__restrict__ out)
{
    __shared__ uint32_t hist[SEL_BINS];
    __shared__ uint32_t sh[18];
    __shared__ uint32_t s_nres;
    const int p = blockIdx.x;
    part += (long)p * nblk;
    // the block partials, loaded in parallel, summed in block order by one thread
    __shared__ FlowPartial s_part[FINISH_BLOCK];
    double ws = 0.0, wa = 0.0, e = 0.0;
    uint32_t nw = 0, nu = 0, nr = 0;
    for (int b0 = 0; b0 < nblk; b0 += FINISH_BLOCK) {
        const int cnt = min(FINISH_BLOCK, nblk - b0);
        __syncthreads();
        if ((int)threadIdx.x < cnt) s_part[threadIdx.x] = part[b0 + threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int b = 0; b < cnt; ++b) {
                ws += s_part[b].ws;
                wa += s_part[b].wa;
                e += s_part[b].e;
                nw += s_part[b].nw;
                nu += s_part[b].nu;
                nr += s_part[b].nres;
            }
    }
    if (threadIdx.x == 0) {
        // flowstats.rs:90-100
        FlowStatsDev r;
        r.vorticity_mean = nw > 0 ? (float)(ws / (double)nw) : __builtin_nanf("");
        r.vorticity_abs_mean = nw > 0 ? (float)(wa / (double)nw) : __builtin_nanf("");
        r.ke_mean = nu > 0 ? (float)(e / (double)nu) : __builtin_nanf("");
        r.ke_sum = (float)(e * ((double)step * (double)step));
        r.coverage = (float)nu / (float)(cells > 0 ? cells : 1);
        r.noise_px = __builtin_nanf("");
        r.n_usable = nu;
        r.n_vorticity = nw;
        out[p] = r;
        s_nres = nr;
    }
    __syncthreads();
    const uint32_t nres = s_nres;
    if (nres >= 20) {  // flowstats.rs:131-136
        const float mad = key_value(block_select(resid + 2l * p * cells, 2u * (uint32_t)cells,
                                                 nres / 2, hist, sh));
        if (threadIdx.x == 0)
            out[p].noise_px = 1.4826f * mad / sqrtf(1.0f + 3.14159265358979323846f / 16.0f);
    }
}

// ---- smoothing ---------------------------------------------------------------------------
// Bitonic sorting network over N register keys (N a power of two), fully unrolled.
template <int N>
__device__ __forceinline__ void bitonic(uint32_t (&k)[N])
{
#pragma unroll
    for (int size = 2; size <= N; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int j = i ^ stride;
                if (j > i) {
                    if ((i & size) == 0) cswap(k[i], k[j]);
                    else cswap(k[j], k[i]);
                }
            }
}

template <int N>
__device__ __forceinline__ uint32_t pick(const uint32_t (&k)[N], int idx)
{
    uint32_t r = k[0];
#pragma unroll
    for (int q = 1; q < N; ++q) r = q == idx ? k[q] : r;
    return r;
}

// NK: 0 for the mean; else the median over at most NK samples, sorted in registers (absent
// samples are NO_KEY and sort last)
template <int NK>
__global__ __launch_bounds__(256) void k_temporal(TemporalArgs args, const float *__restrict__ dx,
                                                  const float *__restrict__ dy,
                                                  const uint8_t *__restrict__ usable,
                                                  const uint8_t *__restrict__ gate,
                                                  float *__restrict__ odx, float *__restrict__ ody,
                                                  uint16_t *__restrict__ sup)
{
    const long plane = (long)args.ny * args.nx;
    const long o = (long)blockIdx.x * 256 + threadIdx.x;
    if (o >= plane * args.nt) return;
    const int t = (int)(o / plane);
    const long p = o - (long)t * plane;
    const float nan = __builtin_nanf("");
    if (!gate[o]) {
        odx[o] = nan;
        ody[o] = nan;
        sup[o] = 0;
        return;
    }
    float ru = nan, rv = nan;
    int m = 0;
    if constexpr (NK == 0) {
        // smooth.rs:303-305 mean: f32 sum from -0.0 in push (offset) order
        float su = -0.0f, sv = -0.0f;
        for (int q = 0; q < args.noff; ++q) {
            const int s = t + args.off[q];
            if (s < 0 || s >= args.nt || !usable[(long)s * plane + p]) continue;
            su += dx[(long)s * plane + p];
            sv += dy[(long)s * plane + p];
            ++m;
        }
        if (m >= args.need) {
            ru = su / (float)m;
            rv = sv / (float)m;
        }
    } else {
        // smooth.rs:312-320 median: total_cmp order, middle or mean of the two middles
        uint32_t ku[NK], kv[NK];
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            const int s = q < args.noff ? t + args.off[q] : -1;
            const bool use = s >= 0 && s < args.nt && usable[(long)s * plane + p];
            ku[q] = use ? total_key(dx[(long)s * plane + p]) : NO_KEY;
            kv[q] = use ? total_key(dy[(long)s * plane + p]) : NO_KEY;
            m += use;
        }
        if (m >= args.need) {
            bitonic(ku);
            bitonic(kv);
            if (m % 2 == 1) {
                ru = key_value(pick(ku, m / 2));
                rv = key_value(pick(kv, m / 2));
            } else {
                ru = 0.5f * (key_value(pick(ku, m / 2 - 1)) + key_value(pick(ku, m / 2)));
                rv = 0.5f * (key_value(pick(kv, m / 2 - 1)) + key_value(pick(kv, m / 2)));
            }
        }
    }
    odx[o] = ru;
    ody[o] = rv;
    sup[o] = (uint16_t)m;
}

__global__ __launch_bounds__(256) void k_spatial(SpatialArgs args, const float *__restrict__ w,
                                                 const float *__restrict__ dx,
                                                 const float *__restrict__ dy,
                                                 const uint8_t *__restrict__ gate,
                                                 float *__restrict__ odx, float *__restrict__ ody,
                                                 uint16_t *__restrict__ sup)
{
    __shared__ float s_w[SPATIAL_MAXK * SPATIAL_MAXK];
    const int r = args.r, kw = 2 * r + 1;
    for (int q = threadIdx.x; q < kw * kw; q += 256) s_w[q] = w[q];
    __syncthreads();
    const long plane = (long)args.ny * args.nx;
    const long o = (long)blockIdx.x * 256 + threadIdx.x;
    if (o >= plane * args.nt) return;
    const int t = (int)(o / plane);
    const int rem = (int)(o - (long)t * plane);
    const int y = rem / args.nx, x = rem - y * args.nx;
    const float nan = __builtin_nanf("");
    if (!gate[o] || args.none) {
        odx[o] = nan;
        ody[o] = nan;
        sup[o] = 0;
        return;
    }
    const float *u = dx + (long)t * plane, *v = dy + (long)t * plane;
    float su = 0.0f, sv = 0.0f, sw = 0.0f;
    uint32_t n = 0;
    for (int j = -r; j <= r; ++j) {
        const int yy = y + j;
        if (yy < 0 || yy >= args.ny) continue;
        for (int i = -r; i <= r; ++i) {
            const int xx = x + i;
            if (xx < 0 || xx >= args.nx) continue;
            if (args.leave >= 0 && (i < 0 ? -i : i) <= args.leave && (j < 0 ? -j : j) <= args.leave)
                continue;
            const float a = u[(long)yy * args.nx + xx], b = v[(long)yy * args.nx + xx];
            if (finite(a) && finite(b)) {
                const float wt = s_w[(j + r) * kw + (i + r)];
                su += wt * a;
                sv += wt * b;
                sw += wt;
                ++n;
            }
        }
    }
    sup[o] = (uint16_t)n;
    if (sw >= args.min_full && sw > 0.0f) {
        odx[o] = su / sw;
        ody[o] = sv / sw;
    } else {
        odx[o] = nan;
        ody[o] = nan;
    }
}

// k_spatial with a 16x16 tile of outputs of one pair and its halo staged in LDS.  Cells
// outside the grid are staged as NaN: the reference skips them before the finiteness test, and
// a non-finite value is skipped the same way, so the sums see the same terms in the same order.
constexpr int SP_TILE = 16;

__host__ __device__ constexpr size_t spatial_tile_floats(int r)
{
    return (size_t)(2 * r + 1) * (2 * r + 1) + 2 * (size_t)(SP_TILE + 2 * r) * (SP_TILE + 2 * r);
}

__global__ __launch_bounds__(256) void k_spatial_tiled(SpatialArgs args, const float *__restrict__ w,
                                                       const float *__restrict__ dx,
                                                       const float *__restrict__ dy,
                                                       const uint8_t *__restrict__ gate,
                                                       float *__restrict__ odx,
                                                       float *__restrict__ ody,
                                                       uint16_t *__restrict__ sup, int tiles_x,
                                                       int tiles)
{
    extern __shared__ float sm[];
    const int r = args.r, kw = 2 * r + 1, T = SP_TILE + 2 * r;
    float *s_w = sm, *tu = sm + kw * kw, *tv = tu + T * T;
    const int t = blockIdx.x / tiles, tile = blockIdx.x - t * tiles;
    const int ty0 = (tile / tiles_x) * SP_TILE, tx0 = (tile % tiles_x) * SP_TILE;
    const long plane = (long)args.ny * args.nx;
    const float *u = dx + (long)t * plane, *v = dy + (long)t * plane;
    const float nan = __builtin_nanf("");
    for (int q = threadIdx.x; q < kw * kw; q += 256) s_w[q] = w[q];
    for (int q = threadIdx.x; q < T * T; q += 256) {
        const int yy = ty0 - r + q / T, xx = tx0 - r + q % T;
        const bool in = yy >= 0 && yy < args.ny && xx >= 0 && xx < args.nx;
        tu[q] = in ? u[(long)yy * args.nx + xx] : nan;
        tv[q] = in ? v[(long)yy * args.nx + xx] : nan;
    }
    __syncthreads();
    const int ly = threadIdx.x / SP_TILE, lx = threadIdx.x % SP_TILE;
    const int y = ty0 + ly, x = tx0 + lx;
    if (y >= args.ny || x >= args.nx) return;
    const long o = (long)t * plane + (long)y * args.nx + x;
    if (!gate[o] || args.none) {
        odx[o] = nan;
        ody[o] = nan;
        sup[o] = 0;
        return;
    }
    float su = 0.0f, sv = 0.0f, sw = 0.0f;
    uint32_t n = 0;
    for (int j = -r; j <= r; ++j) {
        const float *ru = tu + (ly + r + j) * T + lx + r, *rv = tv + (ly + r + j) * T + lx + r;
        const float *rw = s_w + (j + r) * kw + r;
        const bool row_in = args.leave >= 0 && (j < 0 ? -j : j) <= args.leave;
        for (int i = -r; i <= r; ++i) {
            if (row_in && (i < 0 ? -i : i) <= args.leave) continue;
            const float a = ru[i], b = rv[i];
            if (finite(a) && finite(b)) {
                const float wt = rw[i];
                su += wt * a;
                sv += wt * b;
                sw += wt;
                ++n;
            }
        }
    }
    sup[o] = (uint16_t)n;
    if (sw >= args.min_full && sw > 0.0f) {
        odx[o] = su / sw;
        ody[o] = sv / sw;
    } else {
        odx[o] = nan;
        ody[o] = nan;
    }
}

// smooth.rs:303-310
__global__ void k_mask(long n, const float *__restrict__ v, const uint8_t *__restrict__ ok,
                       float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ok[i] ? v[i] : __builtin_nanf("");
}

// ---- metrics ------------------------------------------------------------------------------
constexpr int SCAN_BLOCK = 1024;

// per-block counts of usable entries (cross_validation_errors keeps index order)
__global__ __launch_bounds__(SCAN_BLOCK) void k_cv_count(long n, const uint8_t *__restrict__ ok,
                                                         uint32_t *__restrict__ counts)
{
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const long i = (long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    const bool f = i < n && ok[i];
    const uint64_t b = __ballot(f);
    if ((threadIdx.x & 63) == 0) atomicAdd(&s, (uint32_t)__popcll(b));
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s;
}

// exclusive scan of the block counts in one workgroup; total in counts[nblk] position of offs
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_counts(const uint32_t *__restrict__ counts,
                                                            uint32_t nblk,
                                                            uint64_t *__restrict__ offs)
{
    __shared__ uint64_t s_run;
    __shared__ uint64_t s_part[SCAN_BLOCK];
    if (threadIdx.x == 0) s_run = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < nblk; b0 += SCAN_BLOCK) {
        const uint32_t b = b0 + threadIdx.x;
        s_part[threadIdx.x] = b < nblk ? counts[b] : 0;
        __syncthreads();
        for (int d = 1; d < SCAN_BLOCK; d <<= 1) {
            const uint64_t add = threadIdx.x >= (unsigned)d ? s_part[threadIdx.x - d] : 0;
            __syncthreads();
            s_part[threadIdx.x] += add;
            __syncthreads();
        }
        const uint64_t incl = s_part[threadIdx.x];
        const uint64_t before = s_run;
        if (b < nblk) offs[b] = before + incl - counts[b];
        __syncthreads();
        if (threadIdx.x == SCAN_BLOCK - 1) s_run = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) offs[nblk] = s_run;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_cv_scatter(long n, const float *__restrict__ rdx,
                                                           const float *__restrict__ rdy,
                                                           const uint8_t *__restrict__ ok,
                                                           const float *__restrict__ pdx,
                                                           const float *__restrict__ pdy,
                                                           const uint64_t *__restrict__ offs,
                                                           float *__restrict__ out)
{
    __shared__ uint32_t s_wave[SCAN_BLOCK / 64];
    const long i = (long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    const bool f = i < n && ok[i];
    const uint64_t b = __ballot(f);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) s_wave[wv] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t before = 0;
    for (int q = 0; q < wv; ++q) before += s_wave[q];
    const uint32_t in_wave = (uint32_t)__popcll(lane ? (b & ((1ull << lane) - 1)) : 0ull);
    if (f) out[offs[blockIdx.x] + before + in_wave] = hypot_ref(rdx[i] - pdx[i], rdy[i] - pdy[i]);
}

// keys of the finite errors kept by `keep` (smooth.rs:236-249)
__global__ void k_keys_errors(long n, const float *__restrict__ e, const uint8_t *__restrict__ keep,
                              uint32_t *__restrict__ keys)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = e[i];
    keys[i] = (finite(x) && (!keep || keep[i])) ? total_key(x) : NO_KEY;
}

// |d(t+1) - d(t)| keys (smooth.rs:252-274): n = (nt - 1) * plane entries
__global__ void k_keys_change(long n, long plane, const float *__restrict__ dx,
                              const float *__restrict__ dy, uint32_t *__restrict__ keys)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long j = i + plane;
    const float h = hypot_ref(dx[j] - dx[i], dy[j] - dy[i]);
    keys[i] = finite(h) ? total_key(h) : NO_KEY;
}

// |d - mean of 4 neighbours| keys (smooth.rs:277-300), one entry per cell
__global__ void k_keys_rough(int nt, int ny, int nx, const float *__restrict__ dx,
                             const float *__restrict__ dy, uint32_t *__restrict__ keys)
{
    const long plane = (long)ny * nx;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= plane * nt) return;
    const int t = (int)(o / plane);
    const int rem = (int)(o - (long)t * plane);
    const int y = rem / nx, x = rem - y * nx;
    uint32_t key = NO_KEY;
    if (y >= 1 && y + 1 < ny && x >= 1 && x + 1 < nx) {
        const float *u = dx + (long)t * plane, *v = dy + (long)t * plane;
        const long c = (long)y * nx + x, up = c - nx, dn = c + nx, lf = c - 1, rt = c + 1;
        const long pts[5] = {c, up, dn, lf, rt};
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 5; ++q) ok = ok && finite(u[pts[q]]) && finite(v[pts[q]]);
        if (ok) {
            const float mu = (u[up] + u[dn] + u[lf] + u[rt]) / 4.0f;
            const float mv = (v[up] + v[dn] + v[lf] + v[rt]) / 4.0f;
            key = total_key(hypot_ref(u[c] - mu, v[c] - mv));
        }
    }
    keys[o] = key;
}

__global__ __launch_bounds__(FINISH_BLOCK) void k_median_keys(const uint32_t *__restrict__ keys,
                                                              uint32_t n, float *__restrict__ med,
                                                              uint64_t *__restrict__ count)
{
    __shared__ uint32_t hist[SEL_BINS];
    __shared__ uint32_t sh[18];
    const uint32_t m = block_count_keys(keys, n, sh);
    float r = __builtin_nanf("");
    if (m > 0) r = block_median(keys, n, m, hist, sh);
    if (threadIdx.x == 0) {
        *med = r;
        if (count) *count = m;
    }
}

inline unsigned nblocks(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t launch_flow_stats(const float *dx, const float *dy, const uint8_t *flags, int npairs,
                             int ny, int nx, float k, float step, void *part, uint32_t *resid,
                             FlowStatsDev *out, hipStream_t st)
{
    const int cells = ny * nx;
    const unsigned nb = nblocks(cells, FLOW_BLOCK);
    hipLaunchKernelGGL(k_flow_cells, dim3(nb, (unsigned)npairs), dim3(FLOW_BLOCK), 0, st, dx, dy,
                       flags, ny, nx, k, step, static_cast<FlowPartial *>(part), resid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_flow_finish, dim3((unsigned)npairs), dim3(FINISH_BLOCK), 0, st,
                       static_cast<const FlowPartial *>(part), (int)nb, resid, cells, step, out);
    return hipGetLastError();
}

size_t flow_partial_bytes(int npairs, int ny, int nx)
{
    return sizeof(FlowPartial) * (size_t)npairs * nblocks((long)ny * nx, FLOW_BLOCK);
}

hipError_t launch_temporal(const TemporalArgs &args, const float *dx, const float *dy,
                           const uint8_t *usable, const uint8_t *gate, float *odx, float *ody,
                           uint16_t *sup, hipStream_t st)
{
    const long n = (long)args.nt * args.ny * args.nx;
    if (n == 0) return hipSuccess;
    const dim3 g(nblocks(n, 256)), b(256);
    if (args.kernel == 0)
        hipLaunchKernelGGL(k_temporal<0>, g, b, 0, st, args, dx, dy, usable, gate, odx, ody, sup);
    else if (args.noff <= 8)
        hipLaunchKernelGGL(k_temporal<8>, g, b, 0, st, args, dx, dy, usable, gate, odx, ody, sup);
    else if (args.noff <= 16)
        hipLaunchKernelGGL(k_temporal<16>, g, b, 0, st, args, dx, dy, usable, gate, odx, ody, sup);
    else if (args.noff <= 32)
        hipLaunchKernelGGL(k_temporal<32>, g, b, 0, st, args, dx, dy, usable, gate, odx, ody, sup);
    else
        hipLaunchKernelGGL(k_temporal<64>, g, b, 0, st, args, dx, dy, usable, gate, odx, ody, sup);
    return hipGetLastError();
}

hipError_t launch_spatial(const SpatialArgs &args, const float *w, const float *dx, const float *dy,
                          const uint8_t *gate, float *odx, float *ody, uint16_t *sup,
                          hipStream_t st)
{
    const long n = (long)args.nt * args.ny * args.nx;
    if (n == 0) return hipSuccess;
    const size_t lds = spatial_tile_floats(args.r) * sizeof(float);
    if (lds <= 64 * 1024) {
        const int tiles_x = (args.nx + SP_TILE - 1) / SP_TILE, tiles_y = (args.ny + SP_TILE - 1) / SP_TILE;
        const long blocks = (long)tiles_x * tiles_y * args.nt;
        hipLaunchKernelGGL(k_spatial_tiled, dim3((unsigned)blocks), dim3(256), lds, st, args, w, dx,
                           dy, gate, odx, ody, sup, tiles_x, tiles_x * tiles_y);
    } else {
        hipLaunchKernelGGL(k_spatial, dim3(nblocks(n, 256)), dim3(256), 0, st, args, w, dx, dy,
                           gate, odx, ody, sup);
    }
    return hipGetLastError();
}

hipError_t launch_mask(long n, const float *v, const uint8_t *ok, float *out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mask, dim3(nblocks(n, 256)), dim3(256), 0, st, n, v, ok, out);
    return hipGetLastError();
}

size_t cv_scratch_bytes(long n)
{
    const long nb = (long)nblocks(n, SCAN_BLOCK);
    return (size_t)nb * 4 + (size_t)(nb + 1) * 8 + 16;
}

hipError_t launch_cv_errors(long n, const float *rdx, const float *rdy, const uint8_t *ok,
                            const float *pdx, const float *pdy, void *scratch, float *out,
                            uint64_t *count_dev, hipStream_t st)
{
    const unsigned nb = nblocks(n, SCAN_BLOCK);
    uint32_t *counts = static_cast<uint32_t *>(scratch);
    uint64_t *offs = reinterpret_cast<uint64_t *>(static_cast<char *>(scratch) +
                                                  ((size_t)nb * 4 + 7) / 8 * 8);
    if (nb > 0) {
        hipLaunchKernelGGL(k_cv_count, dim3(nb), dim3(SCAN_BLOCK), 0, st, n, ok, counts);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(SCAN_BLOCK), 0, st, counts, nb, offs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (nb > 0) {
        hipLaunchKernelGGL(k_cv_scatter, dim3(nb), dim3(SCAN_BLOCK), 0, st, n, rdx, rdy, ok, pdx,
                           pdy, offs, out);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipMemcpyAsync(count_dev, offs + nb, 8, hipMemcpyDeviceToDevice, st);
}

hipError_t launch_median_errors(long n, const float *e, const uint8_t *keep, uint32_t *keys,
                                float *med, uint64_t *count, hipStream_t st)
{
    if (n > 0) {
        hipLaunchKernelGGL(k_keys_errors, dim3(nblocks(n, 256)), dim3(256), 0, st, n, e, keep, keys);
        hipError_t r = hipGetLastError();
        if (r != hipSuccess) return r;
    }
    hipLaunchKernelGGL(k_median_keys, dim3(1), dim3(FINISH_BLOCK), 0, st, keys, (uint32_t)n, med,
                       count);
    return hipGetLastError();
}

hipError_t launch_pair_change(int nt, int ny, int nx, const float *dx, const float *dy,
                              uint32_t *keys, float *med, hipStream_t st)
{
    const long plane = (long)ny * nx, n = nt > 1 ? (long)(nt - 1) * plane : 0;
    if (n > 0) {
        hipLaunchKernelGGL(k_keys_change, dim3(nblocks(n, 256)), dim3(256), 0, st, n, plane, dx, dy,
                           keys);
        hipError_t r = hipGetLastError();
        if (r != hipSuccess) return r;
    }
    hipLaunchKernelGGL(k_median_keys, dim3(1), dim3(FINISH_BLOCK), 0, st, keys, (uint32_t)n, med,
                       nullptr);
    return hipGetLastError();
}

hipError_t launch_roughness(int nt, int ny, int nx, const float *dx, const float *dy,
                            uint32_t *keys, float *med, hipStream_t st)
{
    const long n = (long)nt * ny * nx;
    if (n > 0) {
        hipLaunchKernelGGL(k_keys_rough, dim3(nblocks(n, 256)), dim3(256), 0, st, nt, ny, nx, dx,
                           dy, keys);
        hipError_t r = hipGetLastError();
        if (r != hipSuccess) return r;
    }
    hipLaunchKernelGGL(k_median_keys, dim3(1), dim3(FINISH_BLOCK), 0, st, keys, (uint32_t)n, med,
                       nullptr);
    return hipGetLastError();
}

}  // namespace piv

/*
 * CPU oracle for the post-analysis stages: per-pair flow statistics and the time-series
 * filters.  TEST INFRASTRUCTURE ONLY (see piv_oracle.h): the checker that the HIP kernels of
 * the same stages are compared against.
 *
 * Restated from the Rust reference, function by function:
 *   piv-core/src/flowstats.rs  (derivative, vorticity, flow_stats, neighbour_noise)
 *   piv-io/src/flow.rs         (usable, pair_stats)
 *   piv-core/src/smooth.rs     (temporal, spatial, cross_validation_errors, median_error,
 *                               pair_to_pair_change, spatial_roughness, mask, mean, median)
 * Rust semantics kept: f32 arithmetic unless the reference casts to f64, no fused
 * multiply-add, `Iterator::sum` of f32 folds from -0.0 (Rust >= 1.85, the reference's
 * rust-version), `total_cmp` ordering for sorts and selections, libm expf / hypotf.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "piv_oracle.h"

/* f32::total_cmp key: sign-magnitude bits turned into a two's-complement order */
static int32_t total_key(float x)
{
    int32_t b;
    memcpy(&b, &x, 4);
    return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
}

static int cmp_total(const void *a, const void *b)
{
    const int32_t ka = total_key(*(const float *)a), kb = total_key(*(const float *)b);
    return (ka > kb) - (ka < kb);
}

/* select_nth_unstable_by(k, total_cmp): the element at index k of the sorted order */
static float nth_total(float *v, int64_t n, int64_t k)
{
    qsort(v, (size_t)n, sizeof(float), cmp_total);
    return v[k];
}

/* smooth.rs:312-320 median: sort by total_cmp, middle or mean of the two middles */
float or_post_median(float *v, int64_t n)
{
    qsort(v, (size_t)n, sizeof(float), cmp_total);
    if (n % 2 == 1) return v[n / 2];
    return 0.5f * (v[n / 2 - 1] + v[n / 2]);
}

/* flow.rs:13-17: finite and (accepted or REPLACED) */
void or_flow_usable(int64_t n, const float *dx, const uint8_t *flags, uint8_t *out)
{
    for (int64_t i = 0; i < n; ++i)
        out[i] = isfinite(dx[i]) && ((flags[i] & OR_INVALID) == 0 || (flags[i] & OR_REPLACED) != 0);
}

/* flowstats.rs:31-44: central where both neighbours are usable, else one-sided; the centre
 * must be usable.  have[] marks usable samples of the line. Returns 1 with a value. */
static int derivative(int32_t i, int32_t n, float h, const float *val, const uint8_t *have,
                      int64_t stride, float *out)
{
    if (!have[i * stride]) return 0;
    const float c = val[i * stride];
    const int lo = i >= 1 && have[(i - 1) * stride];
    const int hi = i + 1 < n && have[(i + 1) * stride];
    if (lo && hi) *out = (val[(i + 1) * stride] - val[(i - 1) * stride]) / (2.0f * h);
    else if (lo) *out = (c - val[(i - 1) * stride]) / h;
    else if (hi) *out = (val[(i + 1) * stride] - c) / h;
    else return 0;
    return 1;
}

/* flowstats.rs:48-66: w = du/dy - dv/dx (image axes, y down), NaN where undefined */
void or_vorticity(int32_t ny, int32_t nx, const float *u, const float *v, const uint8_t *usable,
                  float step, float *w)
{
    const int64_t n = (int64_t)ny * nx;
    uint8_t *hu = malloc((size_t)n), *hv = malloc((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        hu[i] = usable[i] && isfinite(u[i]);
        hv[i] = usable[i] && isfinite(v[i]);
    }
    for (int32_t iy = 0; iy < ny; ++iy)
        for (int32_t ix = 0; ix < nx; ++ix) {
            float a, b;
            const int ka = derivative(iy, ny, step, u + ix, hu + ix, nx, &a);
            const int kb = derivative(ix, nx, step, v + (int64_t)iy * nx, hv + (int64_t)iy * nx, 1, &b);
            w[(int64_t)iy * nx + ix] = (ka && kb) ? a - b : NAN;
        }
    free(hu);
    free(hv);
}

/* flowstats.rs:69-101 */
void or_flow_stats(int32_t ny, int32_t nx, const float *u, const float *v, const uint8_t *usable,
                   float step, or_flowstat *out)
{
    const int64_t n = (int64_t)ny * nx;
    float *w = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    or_vorticity(ny, nx, u, v, usable, step, w);
    double ws = 0.0, wa = 0.0, e = 0.0;
    int64_t nw = 0, nu = 0;
    for (int64_t i = 0; i < n; ++i)
        if (isfinite(w[i])) {
            ws += (double)w[i];
            wa += (double)fabsf(w[i]);
            ++nw;
        }
    for (int64_t i = 0; i < n; ++i)
        if (usable[i] && isfinite(u[i]) && isfinite(v[i])) {
            e += 0.5 * ((double)u[i] * (double)u[i] + (double)v[i] * (double)v[i]);
            ++nu;
        }
    free(w);
    out->vorticity_mean = nw > 0 ? (float)(ws / (double)nw) : NAN;
    out->vorticity_abs_mean = nw > 0 ? (float)(wa / (double)nw) : NAN;
    out->ke_mean = nu > 0 ? (float)(e / (double)nu) : NAN;
    out->ke_sum = (float)(e * ((double)step * (double)step));
    out->coverage = (float)nu / (float)(n > 0 ? n : 1);
    out->n_usable = nu;
    out->n_vorticity = nw;
}

/* flow.rs:20-34: displacements (px) -> px/s with k = f32(1 / dt), then flow_stats */
void or_pair_stats(int32_t ny, int32_t nx, const float *dx, const float *dy,
                   const uint8_t *usable, int32_t step, double dt_s, or_flowstat *out)
{
    const int64_t n = (int64_t)ny * nx;
    const float k = (float)(1.0 / dt_s);
    float *u = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    float *v = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    for (int64_t i = 0; i < n; ++i) {
        u[i] = dx[i] * k;
        v[i] = dy[i] * k;
    }
    or_flow_stats(ny, nx, u, v, usable, (float)step, out);
    free(u);
    free(v);
}

/* flowstats.rs:106-137 */
float or_neighbour_noise(int32_t ny, int32_t nx, const float *u, const float *v,
                         const uint8_t *usable)
{
    const int64_t n = (int64_t)ny * nx;
    float *resid = malloc((size_t)(2 * n > 0 ? 2 * n : 1) * sizeof(float));
    int64_t nr = 0;
    const float *comp[2] = {u, v};
    for (int c = 0; c < 2; ++c) {
        const float *a = comp[c];
        for (int32_t iy = 0; iy < ny; ++iy)
            for (int32_t ix = 0; ix < nx; ++ix) {
                const int64_t i = (int64_t)iy * nx + ix;
                if (!usable[i] || !isfinite(a[i])) continue;
                float nb[8];
                int m = 0;
                for (int32_t jy = iy > 0 ? iy - 1 : 0; jy < (iy + 2 < ny ? iy + 2 : ny); ++jy)
                    for (int32_t jx = ix > 0 ? ix - 1 : 0; jx < (ix + 2 < nx ? ix + 2 : nx); ++jx) {
                        const int64_t j = (int64_t)jy * nx + jx;
                        if ((jy != iy || jx != ix) && usable[j] && isfinite(a[j])) nb[m++] = a[j];
                    }
                if (m < 5) continue;
                const float med = nth_total(nb, m, m / 2);
                resid[nr++] = fabsf(a[i] - med);
            }
    }
    float out = NAN;
    if (nr >= 20) {
        const float mad = nth_total(resid, nr, nr / 2);
        out = 1.4826f * mad / sqrtf(1.0f + 3.14159265358979323846f / 16.0f);
    }
    free(resid);
    return out;
}

/* smooth.rs:44-130 TemporalKernel Mean / Median at every grid point.  leave_out: n_leave
 * offsets to skip.  Outputs NaN / support 0 where gate is false. */
void or_temporal(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy,
                 const uint8_t *usable, const uint8_t *gate, int32_t kernel, int32_t window,
                 int32_t min_support, const int32_t *leave_out, int32_t n_leave, float *out_dx,
                 float *out_dy, uint16_t *support)
{
    const int32_t half = window / 2;
    int32_t *off = malloc((size_t)(window + 1) * sizeof(int32_t));
    int32_t no = 0;
    int zero_left = 0;
    for (int32_t q = 0; q < n_leave; ++q) zero_left |= leave_out[q] == 0;
    if (!zero_left) off[no++] = 0;
    for (int sign = -1; sign <= 1; sign += 2) {
        int32_t k = 1, taken = 0;
        while (taken < half) {
            int skip = 0;
            for (int32_t q = 0; q < n_leave; ++q) skip |= leave_out[q] == sign * k;
            if (!skip) {
                off[no++] = sign * k;
                ++taken;
            }
            ++k;
        }
    }
    const int64_t plane = (int64_t)ny * nx, tot = plane * nt;
    for (int64_t i = 0; i < tot; ++i) {
        out_dx[i] = NAN;
        out_dy[i] = NAN;
        support[i] = 0;
    }
    float *bu = malloc((size_t)(no > 0 ? no : 1) * sizeof(float));
    float *bv = malloc((size_t)(no > 0 ? no : 1) * sizeof(float));
    const int32_t need = min_support > 1 ? min_support : 1;
    for (int64_t p = 0; p < plane; ++p)
        for (int32_t t = 0; t < nt; ++t) {
            const int64_t o = (int64_t)t * plane + p;
            if (!gate[o]) continue;
            int32_t m = 0;
            for (int32_t q = 0; q < no; ++q) {
                const int64_t s = (int64_t)t + off[q];
                if (s < 0 || s >= nt || !usable[s * plane + p]) continue;
                bu[m] = dx[s * plane + p];
                bv[m] = dy[s * plane + p];
                ++m;
            }
            support[o] = (uint16_t)m;
            if (m >= need) {
                if (kernel == 0) {
                    float su = -0.0f, sv = -0.0f;
                    for (int32_t q = 0; q < m; ++q) {
                        su += bu[q];
                        sv += bv[q];
                    }
                    out_dx[o] = su / (float)m;
                    out_dy[o] = sv / (float)m;
                } else {
                    out_dx[o] = or_post_median(bu, m);
                    out_dy[o] = or_post_median(bv, m);
                }
            }
        }
    free(bu);
    free(bv);
    free(off);
}

/* smooth.rs:137-208 normalized Gaussian convolution per pair; leave_out < 0: none */
void or_spatial(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy,
                const uint8_t *gate, float sigma, float min_weight, int32_t leave_out,
                float *out_dx, float *out_dy, uint16_t *support)
{
    const int64_t plane = (int64_t)ny * nx, tot = plane * nt;
    for (int64_t i = 0; i < tot; ++i) {
        out_dx[i] = NAN;
        out_dy[i] = NAN;
        support[i] = 0;
    }
    int32_t r = (int32_t)ceilf(2.5f * sigma);
    if (r < 1) r = 1;
    const int32_t kw = 2 * r + 1;
    float *ker = malloc((size_t)kw * kw * sizeof(float));
    for (int32_t j = -r; j <= r; ++j)
        for (int32_t i = -r; i <= r; ++i)
            ker[(j + r) * kw + (i + r)] = expf(-((float)(i * i + j * j)) / (2.0f * sigma * sigma));
#define EXCL(i, j) (leave_out >= 0 && ((i) < 0 ? -(i) : (i)) <= leave_out && ((j) < 0 ? -(j) : (j)) <= leave_out)
    float full = -0.0f, all = -0.0f;
    for (int32_t j = -r; j <= r; ++j)
        for (int32_t i = -r; i <= r; ++i) {
            if (!EXCL(i, j)) full += ker[(j + r) * kw + (i + r)];
            all += ker[(j + r) * kw + (i + r)];
        }
    if (full < 0.5f * all) {
        free(ker);
        return;
    }
    for (int32_t t = 0; t < nt; ++t) {
        const float *u = dx + t * plane, *v = dy + t * plane;
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t x = 0; x < nx; ++x) {
                const int64_t o = t * plane + (int64_t)y * nx + x;
                if (!gate[o]) continue;
                float su = 0.0f, sv = 0.0f, sw = 0.0f;
                uint16_t n = 0;
                for (int32_t j = -r; j <= r; ++j) {
                    const int32_t yy = y + j;
                    if (yy < 0 || yy >= ny) continue;
                    for (int32_t i = -r; i <= r; ++i) {
                        const int32_t xx = x + i;
                        if (xx < 0 || xx >= nx || EXCL(i, j)) continue;
                        const float a = u[(int64_t)yy * nx + xx], b = v[(int64_t)yy * nx + xx];
                        if (isfinite(a) && isfinite(b)) {
                            const float w = ker[(j + r) * kw + (i + r)];
                            su += w * a;
                            sv += w * b;
                            sw += w;
                            ++n;
                        }
                    }
                }
                support[o] = n;
                if (sw >= min_weight * full && sw > 0.0f) {
                    out_dx[o] = su / sw;
                    out_dy[o] = sv / sw;
                }
            }
    }
#undef EXCL
    free(ker);
}

/* smooth.rs:213-232: hypot of (raw - prediction) at usable vectors, in iteration order */
int64_t or_cross_validation_errors(int64_t n, const float *raw_dx, const float *raw_dy,
                                   const uint8_t *usable, const float *pred_dx,
                                   const float *pred_dy, float *out)
{
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i)
        if (usable[i]) out[m++] = hypotf(raw_dx[i] - pred_dx[i], raw_dy[i] - pred_dy[i]);
    return m;
}

/* smooth.rs:236-249: median of the finite errors where keep (NULL: all) */
float or_median_error(const float *errors, int64_t n, const uint8_t *keep, int64_t *count)
{
    float *e = malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i)
        if (isfinite(errors[i]) && (!keep || keep[i])) e[m++] = errors[i];
    *count = m;
    const float r = m ? or_post_median(e, m) : NAN;
    free(e);
    return r;
}

/* smooth.rs:252-274 */
float or_pair_to_pair_change(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy)
{
    const int64_t plane = (int64_t)ny * nx;
    float *d = malloc((size_t)(plane * (nt > 1 ? nt - 1 : 1) + 1) * sizeof(float));
    int64_t m = 0;
    for (int32_t t = 1; t < nt; ++t)
        for (int64_t i = 0; i < plane; ++i) {
            const float h = hypotf(dx[t * plane + i] - dx[(t - 1) * plane + i],
                                   dy[t * plane + i] - dy[(t - 1) * plane + i]);
            if (isfinite(h)) d[m++] = h;
        }
    const float r = m ? or_post_median(d, m) : NAN;
    free(d);
    return r;
}

/* smooth.rs:277-300 */
float or_spatial_roughness(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy)
{
    const int64_t plane = (int64_t)ny * nx;
    float *d = malloc((size_t)(plane * nt + 1) * sizeof(float));
    int64_t m = 0;
    for (int32_t t = 0; t < nt; ++t) {
        const float *u = dx + t * plane, *v = dy + t * plane;
        for (int32_t y = 1; y + 1 < ny; ++y)
            for (int32_t x = 1; x + 1 < nx; ++x) {
                const int64_t c = (int64_t)y * nx + x, up = c - nx, dn = c + nx, lf = c - 1, rt = c + 1;
                const int64_t pts[5] = {c, up, dn, lf, rt};
                int ok = 1;
                for (int q = 0; q < 5; ++q) ok &= isfinite(u[pts[q]]) && isfinite(v[pts[q]]);
                if (!ok) continue;
                const float mu = (u[up] + u[dn] + u[lf] + u[rt]) / 4.0f;
                const float mv = (v[up] + v[dn] + v[lf] + v[rt]) / 4.0f;
                d[m++] = hypotf(u[c] - mu, v[c] - mv);
            }
    }
    const float r = m ? or_post_median(d, m) : NAN;
    free(d);
    return r;
}

/* smooth.rs:303-310 */
void or_mask(int64_t n, const float *values, const uint8_t *usable, float *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = usable[i] ? values[i] : NAN;
}

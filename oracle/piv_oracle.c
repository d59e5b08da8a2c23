/*
 * CPU oracle for the PIV per-pair path -- TEST INFRASTRUCTURE, never the product.
 * Restates piv-core (agoodman/piv, rust/crates/piv-core/src) function by function; each
 * function cites the reference lines it follows.  Built with -ffp-contract=off so every
 * float expression rounds exactly as the Rust source does (Rust never contracts to FMA).
 */
#include "piv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- FFT (own mixed-radix; the reference uses realfft/rustfft) ---------------------- */
#define T float
#define FN(x) x##_f32_impl
#include "fft_impl.inc"
#undef T
#undef FN
#define T double
#define FN(x) x##_f64_impl
#include "fft_impl.inc"
#undef T
#undef FN

void or_fft_f32(float *data, int32_t n, int32_t sign) { or_fft_f32_impl(data, n, sign); }
void or_fft_f64(double *data, int32_t n, int32_t sign) { or_fft_f64_impl(data, n, sign); }

/* Rust f32::clamp (core/src/num/f32.rs): if x < min { min } ; if x > max { max } */
static float clamp_rs(float x, float lo, float hi)
{
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* ---- grid.rs:20-44 Grid::new ------------------------------------------------------- */
int or_grid_new(int32_t h, int32_t w, int32_t window, float overlap, or_grid *g)
{
    if (!(overlap >= 0.0f && overlap < 1.0f)) return 1;
    if (window == 0 || window > w || window > h) return 2;
    float s = roundf((float)window * (1.0f - overlap));
    int32_t step = s > 0.0f ? (int32_t)s : 0; /* `as usize` saturates at 0 */
    if (step < 1) step = 1;
    g->window = window;
    g->step = step;
    g->nx = (w - window) / step + 1;
    g->ny = (h - window) / step + 1;
    g->x0 = (w - window - (g->nx - 1) * step) / 2;
    g->y0 = (h - window - (g->ny - 1) * step) / 2;
    return 0;
}

/* grid.rs:55-66 center_x / center_y */
static float center(int32_t o0, int32_t i, int32_t step, int32_t window)
{
    return (float)(o0 + i * step) + ((float)window - 1.0f) / 2.0f;
}

/* ---- field.rs:56-81 VectorField::sample ------------------------------------------- */
void or_field_sample(const or_grid *g, const float *dx, const float *dy, float x, float y,
                     float *u, float *v)
{
    float fx = clamp_rs((x - center(g->x0, 0, g->step, g->window)) / (float)g->step, 0.0f,
                        (float)(g->nx - 1));
    float fy = clamp_rs((y - center(g->y0, 0, g->step, g->window)) / (float)g->step, 0.0f,
                        (float)(g->ny - 1));
    int32_t ix = (int32_t)floorf(fx), iy = (int32_t)floorf(fy);
    float tx = fx - (float)ix, ty = fy - (float)iy;
    int32_t ix1 = ix + 1 < g->nx - 1 ? ix + 1 : g->nx - 1;
    int32_t iy1 = iy + 1 < g->ny - 1 ? iy + 1 : g->ny - 1;
    const int32_t rr[4] = {iy, iy, iy1, iy1};
    const int32_t cc[4] = {ix, ix1, ix, ix1};
    const float ww[4] = {(1.0f - tx) * (1.0f - ty), tx * (1.0f - ty), (1.0f - tx) * ty, tx * ty};
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int k = 0; k < 4; k++) {
        float su = dx[rr[k] * g->nx + cc[k]], sv = dy[rr[k] * g->nx + cc[k]];
        if (ww[k] > 0.0f && isfinite(su) && isfinite(sv)) {
            a0 = a0 + ww[k] * su;
            a1 = a1 + ww[k] * sv;
            a2 = a2 + ww[k];
        }
    }
    if (a2 > 0.0f) {
        *u = a0 / a2;
        *v = a1 / a2;
    } else {
        *u = 0.0f;
        *v = 0.0f;
    }
}

/* ---- multipass.rs:207-231 window placement ----------------------------------------- */
static void place(int32_t origin, float shift, float max, float *a, float *b)
{
    shift = clamp_rs(shift, -max, max);
    float lo = fmaxf(-shift, 0.0f);
    float hi = fminf(max - shift, max);
    float a0 = clamp_rs((float)origin - shift / 2.0f, lo, hi);
    *a = a0;
    *b = a0 + shift;
}

void or_place_windows(const or_grid *g, int32_t h, int32_t w, const or_grid *prev,
                      const float *prev_dx, const float *prev_dy, float *origins)
{
    const int32_t n = g->window;
    const float max_x = (float)(w - n > 1 ? w - n - 1 : 0); /* saturating_sub(1) */
    const float max_y = (float)(h - n > 1 ? h - n - 1 : 0);
    for (int32_t k = 0; k < g->nx * g->ny; k++) {
        int32_t ix = k % g->nx, iy = k / g->nx;
        float px = 0.0f, py = 0.0f;
        if (prev)
            or_field_sample(prev, prev_dx, prev_dy, center(g->x0, ix, g->step, n),
                            center(g->y0, iy, g->step, n), &px, &py);
        float ax, bx, ay, by;
        place(g->x0 + ix * g->step, px, max_x, &ax, &bx);
        place(g->y0 + iy * g->step, py, max_y, &ay, &by);
        origins[4 * k + 0] = ax;
        origins[4 * k + 1] = ay;
        origins[4 * k + 2] = bx;
        origins[4 * k + 3] = by;
    }
}

/* ---- preprocess.rs:84-124 IntegralImage ------------------------------------------- */
void or_sat_build(const float *img, int32_t h, int32_t w, double *sum, double *sumsq)
{
    const int64_t s = (int64_t)w + 1;
    memset(sum, 0, sizeof(double) * (size_t)(s * (h + 1)));
    memset(sumsq, 0, sizeof(double) * (size_t)(s * (h + 1)));
    for (int32_t y = 0; y < h; y++) {
        double rs = 0.0, rq = 0.0;
        for (int32_t x = 0; x < w; x++) {
            double v = (double)img[(int64_t)y * w + x];
            rs += v;
            rq += v * v;
            int64_t i = (int64_t)(y + 1) * s + x + 1;
            sum[i] = sum[i - s] + rs;
            sumsq[i] = sumsq[i - s] + rq;
        }
    }
}

static double rect(const double *t, int64_t s, int32_t x0, int32_t y0, int32_t x1, int32_t y1)
{
    return t[y1 * s + x1] - t[y0 * s + x1] - t[y1 * s + x0] + t[y0 * s + x0];
}

void or_sat_stats(const double *sum, const double *sumsq, int32_t w, int32_t x, int32_t y,
                  int32_t bw, int32_t bh, float *mean, float *std_)
{
    const int64_t s = (int64_t)w + 1;
    const double n = (double)((int64_t)bw * bh);
    double m = rect(sum, s, x, y, x + bw, y + bh) / n;
    double var = rect(sumsq, s, x, y, x + bw, y + bh) / n - m * m;
    *mean = (float)m;
    *std_ = (float)sqrt(var > 0.0 ? var : 0.0); /* f64::max(var, 0.0) -- NaN -> 0 */
}

/* ---- preprocess.rs:161-195 gaussian_blur ------------------------------------------ */
int32_t or_gauss_kernel(float sigma, float *k)
{
    float rf = ceilf(3.0f * sigma);
    if (rf < 1.0f) rf = 1.0f;
    int32_t r = (int32_t)rf;
    float total = 0.0f;
    for (int32_t i = -r; i <= r; i++) {
        k[i + r] = expf(-(float)(i * i) / (2.0f * sigma * sigma));
    }
    for (int32_t i = 0; i <= 2 * r; i++) total += k[i]; /* Sum: fold from 0.0 in order */
    for (int32_t i = 0; i <= 2 * r; i++) k[i] /= total;
    return r;
}

void or_gaussian_blur(const float *in, int32_t h, int32_t w, float sigma, float *out)
{
    float k[512];
    const int32_t r = or_gauss_kernel(sigma, k);
    float *tmp = (float *)malloc(sizeof(float) * (size_t)h * w);
    for (int32_t y = 0; y < h; y++) {
        const float *row = in + (int64_t)y * w;
        for (int32_t x = 0; x < w; x++) {
            float acc = 0.0f;
            for (int32_t j = 0; j <= 2 * r; j++) {
                int32_t sx = x + j - r;
                sx = sx < 0 ? 0 : (sx > w - 1 ? w - 1 : sx);
                acc += k[j] * row[sx];
            }
            tmp[(int64_t)y * w + x] = acc;
        }
    }
    for (int64_t i = 0; i < (int64_t)h * w; i++) out[i] = 0.0f;
    for (int32_t y = 0; y < h; y++) {
        for (int32_t j = 0; j <= 2 * r; j++) {
            int32_t sy = y + j - r;
            sy = sy < 0 ? 0 : (sy > h - 1 ? h - 1 : sy);
            const float *row = tmp + (int64_t)sy * w;
            float *o = out + (int64_t)y * w;
            for (int32_t x = 0; x < w; x++) o[x] += k[j] * row[x];
        }
    }
    free(tmp);
}

/* ---- preprocess.rs:199-216 local_normalize ---------------------------------------- */
void or_local_normalize(const float *in, int32_t h, int32_t w, int32_t radius, float eps,
                        float *out)
{
    const size_t cells = (size_t)(w + 1) * (size_t)(h + 1);
    double *sum = (double *)malloc(sizeof(double) * cells);
    double *sq = (double *)malloc(sizeof(double) * cells);
    or_sat_build(in, h, w, sum, sq);
    for (int32_t y = 0; y < h; y++) {
        int32_t y0 = y > radius ? y - radius : 0;
        int32_t y1 = y + radius + 1 < h ? y + radius + 1 : h;
        for (int32_t x = 0; x < w; x++) {
            int32_t x0 = x > radius ? x - radius : 0;
            int32_t x1 = x + radius + 1 < w ? x + radius + 1 : w;
            float mean, sd;
            or_sat_stats(sum, sq, w, x0, y0, x1 - x0, y1 - y0, &mean, &sd);
            out[(int64_t)y * w + x] = (in[(int64_t)y * w + x] - mean) / (sd + eps);
        }
    }
    free(sum);
    free(sq);
}

/* ---- preprocess.rs:137-153 PreparedFrame::new ------------------------------------- */
void or_prepare(const float *raw, int32_t h, int32_t w, const float *bg, float sigma,
                int32_t radius, float eps, float *image, float *base)
{
    const int64_t count = (int64_t)h * w;
    for (int64_t i = 0; i < count; i++) {
        float v = raw[i];
        if (bg) {
            v = v - bg[i];
            v = v > 0.0f ? v : 0.0f; /* f32::max(v, 0.0): NaN -> 0, -0 -> +0 */
        }
        base[i] = v;
    }
    float *cur = (float *)malloc(sizeof(float) * (size_t)count);
    if (sigma > 0.0f)
        or_gaussian_blur(base, h, w, sigma, cur);
    else
        memcpy(cur, base, sizeof(float) * (size_t)count);
    if (radius > 0)
        or_local_normalize(cur, h, w, radius, eps, image);
    else
        memcpy(image, cur, sizeof(float) * (size_t)count);
    free(cur);
}

/* ---- piv-io/src/decode.rs:227-234 ------------------------------------------------- */
void or_normalize_u8(const uint8_t *in, int64_t count, float *out)
{
    for (int64_t i = 0; i < count; i++) out[i] = (float)in[i] / 255.0f;
}
void or_normalize_u16(const uint16_t *in, int64_t count, float *out)
{
    for (int64_t i = 0; i < count; i++) out[i] = (float)in[i] / 65535.0f;
}

/* ---- correlate.rs:182-210 load_window --------------------------------------------- */
double or_load_window(float *dst, const float *img, int32_t h, int32_t w, float x, float y,
                      int32_t n)
{
    (void)h;
    const int32_t x0 = (int32_t)floorf(x), y0 = (int32_t)floorf(y);
    const float tx = x - (float)x0, ty = y - (float)y0;
    if (tx < 1e-4f && ty < 1e-4f) {
        for (int32_t r = 0; r < n; r++)
            for (int32_t c = 0; c < n; c++) dst[r * n + c] = img[(int64_t)(y0 + r) * w + x0 + c];
    } else {
        const float w00 = (1.0f - tx) * (1.0f - ty), w01 = tx * (1.0f - ty);
        const float w10 = (1.0f - tx) * ty, w11 = tx * ty;
        for (int32_t r = 0; r < n; r++) {
            const float *top = img + (int64_t)(y0 + r) * w + x0;
            const float *bot = top + w;
            for (int32_t c = 0; c < n; c++)
                dst[r * n + c] = w00 * top[c] + w01 * top[c + 1] + w10 * bot[c] + w11 * bot[c + 1];
        }
    }
    double s = 0.0;
    for (int32_t i = 0; i < n * n; i++) s += (double)dst[i];
    const double mean = s / (double)(n * n);
    double energy = 0.0;
    for (int32_t i = 0; i < n * n; i++) {
        dst[i] -= (float)mean;
        energy += (double)dst[i] * (double)dst[i];
    }
    return energy;
}

/* FP64 form of load_window: same integer/bilinear decision, arithmetic in double */
static double load_window64(double *dst, const float *img, int32_t w, float x, float y, int32_t n)
{
    const int32_t x0 = (int32_t)floorf(x), y0 = (int32_t)floorf(y);
    const float txf = x - (float)x0, tyf = y - (float)y0;
    if (txf < 1e-4f && tyf < 1e-4f) {
        for (int32_t r = 0; r < n; r++)
            for (int32_t c = 0; c < n; c++)
                dst[r * n + c] = (double)img[(int64_t)(y0 + r) * w + x0 + c];
    } else {
        const double tx = (double)x - (double)x0, ty = (double)y - (double)y0;
        const double w00 = (1.0 - tx) * (1.0 - ty), w01 = tx * (1.0 - ty);
        const double w10 = (1.0 - tx) * ty, w11 = tx * ty;
        for (int32_t r = 0; r < n; r++) {
            const float *top = img + (int64_t)(y0 + r) * w + x0;
            const float *bot = top + w;
            for (int32_t c = 0; c < n; c++)
                dst[r * n + c] = w00 * top[c] + w01 * top[c + 1] + w10 * bot[c] + w11 * bot[c + 1];
        }
    }
    double s = 0.0;
    for (int32_t i = 0; i < n * n; i++) s += dst[i];
    const double mean = s / (double)(n * n);
    double energy = 0.0;
    for (int32_t i = 0; i < n * n; i++) {
        dst[i] -= mean;
        energy += dst[i] * dst[i];
    }
    return energy;
}

/* ---- window deformation ------------------------------------------------------------
 * No counterpart in the reference, which moves windows rigidly by the predictor at the
 * window centre (multipass.rs:207-231).  The definition restated here is the library's
 * (include/piv.h, piv_deformation): the rigid placement stays; pixel (c, r) of window A moves
 * by -h and of window B by +h, h = (D(p) - D(centre)) / 2, where D is the lerp-form bilinear
 * interpolant of the previous validated field (VectorField::sample's NaN-skipping weighted
 * form, field.rs:56-81, wherever a neighbour is not finite) and p = (mx + c, my + r) the
 * pixel's position midway between its A and B samples (mx = (ax + bx) / 2).  Each pixel is
 * resampled from its own fractional position: bilinear with the weights of load_window
 * (correlate.rs:189-200), or 4x4 Lagrange (cubic) weights, indices clamped into the frame. */
void or_deform_sample(const or_grid *g, const float *dx, const float *dy, float x, float y,
                      float *u, float *v)
{
    float fx = clamp_rs((x - center(g->x0, 0, g->step, g->window)) / (float)g->step, 0.0f,
                        (float)(g->nx - 1));
    float fy = clamp_rs((y - center(g->y0, 0, g->step, g->window)) / (float)g->step, 0.0f,
                        (float)(g->ny - 1));
    int32_t ix = (int32_t)floorf(fx), iy = (int32_t)floorf(fy);
    float tx = fx - (float)ix, ty = fy - (float)iy;
    int32_t ix1 = ix + 1 < g->nx - 1 ? ix + 1 : g->nx - 1;
    int32_t iy1 = iy + 1 < g->ny - 1 ? iy + 1 : g->ny - 1;
    const float u00 = dx[iy * g->nx + ix], u01 = dx[iy * g->nx + ix1];
    const float u10 = dx[iy1 * g->nx + ix], u11 = dx[iy1 * g->nx + ix1];
    const float v00 = dy[iy * g->nx + ix], v01 = dy[iy * g->nx + ix1];
    const float v10 = dy[iy1 * g->nx + ix], v11 = dy[iy1 * g->nx + ix1];
    if (isfinite(u00) && isfinite(u01) && isfinite(u10) && isfinite(u11) && isfinite(v00) &&
        isfinite(v01) && isfinite(v10) && isfinite(v11)) {
        const float ut = u00 + tx * (u01 - u00), ub = u10 + tx * (u11 - u10);
        const float vt = v00 + tx * (v01 - v00), vb = v10 + tx * (v11 - v10);
        *u = ut + ty * (ub - ut);
        *v = vt + ty * (vb - vt);
    } else {
        or_field_sample(g, dx, dy, x, y, u, v);
    }
}

/* Lagrange weights at nodes -1, 0, 1, 2 (reproduce cubics exactly) */
static void lagrange4(float t, float l[4])
{
    const float tp1 = t + 1.0f, tm1 = t - 1.0f, tm2 = t - 2.0f;
    l[0] = -(t * tm1 * tm2) / 6.0f;
    l[1] = (tp1 * tm1 * tm2) / 2.0f;
    l[2] = -(tp1 * t * tm2) / 2.0f;
    l[3] = (tp1 * t * tm1) / 6.0f;
}

static void lagrange4_64(double t, double l[4])
{
    const double tp1 = t + 1.0, tm1 = t - 1.0, tm2 = t - 2.0;
    l[0] = -(t * tm1 * tm2) / 6.0;
    l[1] = (tp1 * tm1 * tm2) / 2.0;
    l[2] = -(tp1 * t * tm2) / 2.0;
    l[3] = (tp1 * t * tm1) / 6.0;
}

static float px_clamped(const float *img, int32_t h, int32_t w, int32_t x, int32_t y)
{
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    return img[(int64_t)y * w + x];
}

/* The per-pixel shift h of the window pair with rigid origins (ax, ay), (bx, by). */
void or_deform_shifts(const or_grid *prev, const float *pdx, const float *pdy, int32_t n,
                      float ax, float ay, float bx, float by, float *hx, float *hy)
{
    const float mx = (ax + bx) * 0.5f, my = (ay + by) * 0.5f;
    const float half = ((float)n - 1.0f) / 2.0f;
    float uc, vc;
    or_deform_sample(prev, pdx, pdy, mx + half, my + half, &uc, &vc);
    for (int32_t r = 0; r < n; r++)
        for (int32_t c = 0; c < n; c++) {
            float u, v;
            or_deform_sample(prev, pdx, pdy, mx + (float)c, my + (float)r, &u, &v);
            hx[r * n + c] = (u - uc) * 0.5f;
            hy[r * n + c] = (v - vc) * 0.5f;
        }
}

/* load_window (correlate.rs:182-210) of a deformed window: sgn -1 for A (moves by -h), +1 for
 * B; mode 1 bilinear, 2 Lagrange.  With h = 0 and mode 1 every sample is load_window's. */
double or_load_window_deformed(float *dst, const float *img, int32_t h, int32_t w, float x,
                               float y, int32_t n, int32_t mode, const float *hx,
                               const float *hy, int32_t sgn)
{
    const int32_t x0 = (int32_t)floorf(x), y0 = (int32_t)floorf(y);
    const float tx = x - (float)x0, ty = y - (float)y0;
    for (int32_t r = 0; r < n; r++)
        for (int32_t c = 0; c < n; c++) {
            const int32_t k = r * n + c;
            const float fx = sgn < 0 ? tx - hx[k] : tx + hx[k];
            const float fy = sgn < 0 ? ty - hy[k] : ty + hy[k];
            const float ixf = floorf(fx), iyf = floorf(fy);
            const float wx = fx - ixf, wy = fy - iyf;
            const int32_t xi = x0 + c + (int32_t)ixf, yi = y0 + r + (int32_t)iyf;
            float v;
            if (wx < 1e-4f && wy < 1e-4f) {
                v = px_clamped(img, h, w, xi, yi);
            } else if (mode == 1) {
                const float w00 = (1.0f - wx) * (1.0f - wy), w01 = wx * (1.0f - wy);
                const float w10 = (1.0f - wx) * wy, w11 = wx * wy;
                v = w00 * px_clamped(img, h, w, xi, yi) + w01 * px_clamped(img, h, w, xi + 1, yi) +
                    w10 * px_clamped(img, h, w, xi, yi + 1) + w11 * px_clamped(img, h, w, xi + 1, yi + 1);
            } else {
                float lx[4], ly[4];
                lagrange4(wx, lx);
                lagrange4(wy, ly);
                v = 0.0f;
                for (int32_t j = 0; j < 4; j++) {
                    float hsum = 0.0f;
                    for (int32_t i = 0; i < 4; i++)
                        hsum = hsum + lx[i] * px_clamped(img, h, w, xi - 1 + i, yi - 1 + j);
                    v = v + ly[j] * hsum;
                }
            }
            dst[k] = v;
        }
    double s = 0.0;
    for (int32_t i = 0; i < n * n; i++) s += (double)dst[i];
    const double mean = s / (double)(n * n);
    double energy = 0.0;
    for (int32_t i = 0; i < n * n; i++) {
        dst[i] -= (float)mean;
        energy += (double)dst[i] * (double)dst[i];
    }
    return energy;
}

/* FP64 form: positions as in the FP32 form (placement is f32), weights and sums in double */
double or_load_window_deformed64(double *dst, const float *img, int32_t h, int32_t w, float x,
                                 float y, int32_t n, int32_t mode, const float *hx,
                                 const float *hy, int32_t sgn)
{
    const int32_t x0 = (int32_t)floorf(x), y0 = (int32_t)floorf(y);
    const float tx = x - (float)x0, ty = y - (float)y0;
    for (int32_t r = 0; r < n; r++)
        for (int32_t c = 0; c < n; c++) {
            const int32_t k = r * n + c;
            const float fx = sgn < 0 ? tx - hx[k] : tx + hx[k];
            const float fy = sgn < 0 ? ty - hy[k] : ty + hy[k];
            const float ixf = floorf(fx), iyf = floorf(fy);
            const int32_t xi = x0 + c + (int32_t)ixf, yi = y0 + r + (int32_t)iyf;
            const double wx = (double)fx - (double)ixf, wy = (double)fy - (double)iyf;
            double v;
            if ((float)wx < 1e-4f && (float)wy < 1e-4f) {
                v = px_clamped(img, h, w, xi, yi);
            } else if (mode == 1) {
                const double w00 = (1.0 - wx) * (1.0 - wy), w01 = wx * (1.0 - wy);
                const double w10 = (1.0 - wx) * wy, w11 = wx * wy;
                v = w00 * px_clamped(img, h, w, xi, yi) + w01 * px_clamped(img, h, w, xi + 1, yi) +
                    w10 * px_clamped(img, h, w, xi, yi + 1) + w11 * px_clamped(img, h, w, xi + 1, yi + 1);
            } else {
                double lx[4], ly[4];
                lagrange4_64(wx, lx);
                lagrange4_64(wy, ly);
                v = 0.0;
                for (int32_t j = 0; j < 4; j++) {
                    double hsum = 0.0;
                    for (int32_t i = 0; i < 4; i++)
                        hsum += lx[i] * px_clamped(img, h, w, xi - 1 + i, yi - 1 + j);
                    v += ly[j] * hsum;
                }
            }
            dst[k] = v;
        }
    double s = 0.0;
    for (int32_t i = 0; i < n * n; i++) s += dst[i];
    const double mean = s / (double)(n * n);
    double energy = 0.0;
    for (int32_t i = 0; i < n * n; i++) {
        dst[i] -= mean;
        energy += dst[i] * dst[i];
    }
    return energy;
}

/* ---- correlate.rs:215-303 find_peak ----------------------------------------------- */
#define FIND_PEAK_BODY(TT, LOG, PEAK_T)                                                      \
    const int32_t half = n / 2;                                                              \
    TT best = -INFINITY;                                                                     \
    int32_t by = 0, bx = 0;                                                                  \
    for (int32_t i = 0; i < n * n; i++) {                                                    \
        if (corr[i] > best) {                                                                \
            best = corr[i];                                                                  \
            by = i / n;                                                                      \
            bx = i % n;                                                                      \
        }                                                                                    \
    }                                                                                        \
    int32_t hw = 0;                                                                          \
    const int32_t steps[4][2] = {{1, 0}, {n - 1, 0}, {0, 1}, {0, n - 1}};                    \
    for (int d = 0; d < 4; d++) {                                                            \
        int32_t k = 0;                                                                       \
        while (k < half && AT(by + k * steps[d][0], bx + k * steps[d][1]) > (TT)0.5 * best)  \
            k++;                                                                             \
        if (k > hw) hw = k;                                                                  \
    }                                                                                        \
    int32_t exclude = (int32_t)ceilf(1.5f * (float)hw);                                      \
    if (exclude < 2) exclude = 2;                                                            \
    TT second = -INFINITY;                                                                   \
    for (int32_t y = 0; y < n; y++) {                                                        \
        for (int32_t x = 0; x < n; x++) {                                                    \
            const TT v = AT(y, x);                                                           \
            int32_t dy_ = abs(y - by), dx_ = abs(x - bx);                                    \
            dy_ = dy_ < n - dy_ ? dy_ : n - dy_;                                             \
            dx_ = dx_ < n - dx_ ? dx_ : n - dx_;                                             \
            const int32_t cd = dy_ > dx_ ? dy_ : dx_;                                        \
            if (v > second && cd > exclude) {                                                \
                int lm = 1;                                                                  \
                for (int32_t oy = -1; oy <= 1 && lm; oy++)                                   \
                    for (int32_t ox = -1; ox <= 1; ox++) {                                   \
                        if (oy == 0 && ox == 0) continue;                                    \
                        if (!(AT(y + n + oy, x + n + ox) <= v)) {                            \
                            lm = 0;                                                          \
                            break;                                                           \
                        }                                                                    \
                    }                                                                        \
                if (lm) second = v;                                                          \
            }                                                                                \
        }                                                                                    \
    }                                                                                        \
    TT sub[2];                                                                               \
    for (int axis = 0; axis < 2; axis++) {                                                   \
        TT m = axis == 0 ? AT(by, bx + n - 1) : AT(by + n - 1, bx);                          \
        TT c = best;                                                                         \
        TT p = axis == 0 ? AT(by, bx + 1) : AT(by + 1, bx);                                  \
        TT d;                                                                                \
        if (m > 0 && c > 0 && p > 0) {                                                       \
            TT lm_ = LOG(m), lc = LOG(c), lp = LOG(p);                                       \
            d = (lm_ - lp) / ((TT)2.0 * (lm_ - (TT)2.0 * lc + lp));                          \
        } else {                                                                             \
            d = (m - p) / ((TT)2.0 * (m - (TT)2.0 * c + p));                                 \
        }                                                                                    \
        if (isfinite(d)) {                                                                   \
            if (d < (TT)-0.5) d = (TT)-0.5;                                                  \
            if (d > (TT)0.5) d = (TT)0.5;                                                    \
        } else {                                                                             \
            d = 0;                                                                           \
        }                                                                                    \
        sub[axis] = d;                                                                       \
    }                                                                                        \
    TT wx = bx >= half ? (TT)bx - (TT)n : (TT)bx;                                            \
    TT wy = by >= half ? (TT)by - (TT)n : (TT)by;                                            \
    TT ppr;                                                                                  \
    if (second > 0) {                                                                        \
        ppr = best / second;                                                                 \
        if (ppr > (TT)1000.0) ppr = (TT)1000.0; /* f32::min(NaN, 1000) = 1000 */             \
        if (ppr != ppr) ppr = (TT)1000.0;                                                    \
    } else {                                                                                 \
        ppr = (TT)1000.0;                                                                    \
    }                                                                                        \
    out->dx = wx + sub[0];                                                                   \
    out->dy = wy + sub[1];                                                                   \
    out->height = best;                                                                      \
    out->ppr = ppr;                                                                          \
    out->iy = by;                                                                            \
    out->ix = bx;

#define AT(yy, xx) corr[(((yy) % n) * n) + ((xx) % n)]

/* f32 ln: correctly rounded through double (glibc's logf is also correctly rounded to
 * within its documented bound; the two agree except for rare boundary cases) */
static float ln32(float x) { return (float)log((double)x); }

void or_find_peak(const float *corr, int32_t n, or_peak *out) { FIND_PEAK_BODY(float, ln32, or_peak) }

void or_find_peak64(const double *corr, int32_t n, or_peak64 *out)
{
    FIND_PEAK_BODY(double, log, or_peak64)
}
#undef AT

/* ---- correlate.rs:99-178 correlate / forward / inverse ---------------------------- */
/* forward: per row R2C (first m bins) stored column-major out[c*n+y], then a complex
 * FFT down each of the m columns (correlate.rs:141-160) */
#define CORR_FWD(TT, FFT)                                                                   \
    static void fwd_##TT(const TT *win, TT *out, TT *row, int32_t n)                         \
    {                                                                                        \
        const int32_t m = n / 2 + 1;                                                         \
        for (int32_t y = 0; y < n; y++) {                                                    \
            for (int32_t x = 0; x < n; x++) {                                                \
                row[2 * x] = win[y * n + x];                                                 \
                row[2 * x + 1] = 0;                                                          \
            }                                                                                \
            FFT(row, n, -1);                                                                 \
            for (int32_t c = 0; c < m; c++) {                                                \
                out[2 * (c * n + y)] = row[2 * c];                                           \
                out[2 * (c * n + y) + 1] = row[2 * c + 1];                                   \
            }                                                                                \
        }                                                                                    \
        for (int32_t c = 0; c < m; c++) FFT(out + 2 * c * n, n, -1);                         \
    }                                                                                        \
    static void inv_##TT(TT *ta, TT *row, TT *corr, int32_t n)                               \
    {                                                                                        \
        const int32_t m = n / 2 + 1;                                                         \
        for (int32_t c = 0; c < m; c++) FFT(ta + 2 * c * n, n, +1);                          \
        for (int32_t y = 0; y < n; y++) {                                                    \
            for (int32_t c = 0; c < m; c++) {                                                \
                row[2 * c] = ta[2 * (c * n + y)];                                            \
                row[2 * c + 1] = ta[2 * (c * n + y) + 1];                                    \
            }                                                                                \
            row[1] = 0;                                                                      \
            row[2 * (m - 1) + 1] = 0;                                                        \
            for (int32_t c = m; c < n; c++) {                                                \
                row[2 * c] = row[2 * (n - c)];                                               \
                row[2 * c + 1] = -row[2 * (n - c) + 1];                                      \
            }                                                                                \
            FFT(row, n, +1);                                                                 \
            for (int32_t x = 0; x < n; x++) corr[y * n + x] = row[2 * x];                    \
        }                                                                                    \
    }

CORR_FWD(float, or_fft_f32_impl)
CORR_FWD(double, or_fft_f64_impl)

/* deform: NULL (rigid windows, the reference) or {mode, hx, hy} of or_deform_shifts */
typedef struct {
    int32_t mode;
    const float *hx, *hy;
} deform_t;

static int correlate_f32(const float *a, const float *b, int32_t h, int32_t w, int32_t n,
                         float ax, float ay, float bx, float by, or_peak *out, float *corr_out,
                         const deform_t *def);
static int correlate_f64(const float *a, const float *b, int32_t h, int32_t w, int32_t n,
                         float ax, float ay, float bx, float by, or_peak64 *out,
                         double *corr_out, const deform_t *def);

int or_correlate_f32(const float *a, const float *b, int32_t h, int32_t w, int32_t n, float ax,
                     float ay, float bx, float by, or_peak *out, float *corr_out)
{
    return correlate_f32(a, b, h, w, n, ax, ay, bx, by, out, corr_out, NULL);
}

int or_correlate_f64(const float *a, const float *b, int32_t h, int32_t w, int32_t n, float ax,
                     float ay, float bx, float by, or_peak64 *out, double *corr_out)
{
    return correlate_f64(a, b, h, w, n, ax, ay, bx, by, out, corr_out, NULL);
}

static int correlate_f32(const float *a, const float *b, int32_t h, int32_t w, int32_t n,
                         float ax, float ay, float bx, float by, or_peak *out, float *corr_out,
                         const deform_t *def)
{
    const int32_t m = n / 2 + 1;
    float *wa = (float *)malloc(sizeof(float) * (size_t)n * n * 3);
    float *wb = wa + n * n;
    float *corr = wb + n * n;
    float *ta = (float *)malloc(sizeof(float) * 2 * (size_t)m * n * 2);
    float *tb = ta + 2 * m * n;
    float *row = (float *)malloc(sizeof(float) * 2 * (size_t)n);
    int ok = 0;
    const double ea = def ? or_load_window_deformed(wa, a, h, w, ax, ay, n, def->mode, def->hx, def->hy, -1)
                          : or_load_window(wa, a, h, w, ax, ay, n);
    const double eb = def ? or_load_window_deformed(wb, b, h, w, bx, by, n, def->mode, def->hx, def->hy, 1)
                          : or_load_window(wb, b, h, w, bx, by, n);
    if (!(ea <= 1e-12 || eb <= 1e-12)) {
        fwd_float(wa, ta, row, n);
        fwd_float(wb, tb, row, n);
        for (int32_t i = 0; i < m * n; i++) {
            /* fa.conj() * fb  (num_complex Mul) */
            const float ar = ta[2 * i], ai = -ta[2 * i + 1];
            const float br = tb[2 * i], bi = tb[2 * i + 1];
            ta[2 * i] = ar * br - ai * bi;
            ta[2 * i + 1] = ar * bi + ai * br;
        }
        inv_float(ta, row, corr, n);
        /* `1.0 / (X) as f32`: `as` binds tighter than `/`, so the division is in f32 */
        const float denom = (float)((double)(n * n) * sqrt(ea * eb));
        const float scale = 1.0f / denom;
        for (int32_t i = 0; i < n * n; i++) corr[i] *= scale;
        or_find_peak(corr, n, out);
        if (corr_out) memcpy(corr_out, corr, sizeof(float) * (size_t)n * n);
        ok = 1;
    }
    free(wa);
    free(ta);
    free(row);
    return ok;
}

static int correlate_f64(const float *a, const float *b, int32_t h, int32_t w, int32_t n,
                         float ax, float ay, float bx, float by, or_peak64 *out,
                         double *corr_out, const deform_t *def)
{
    const int32_t m = n / 2 + 1;
    double *wa = (double *)malloc(sizeof(double) * (size_t)n * n * 3);
    double *wb = wa + n * n;
    double *corr = wb + n * n;
    double *ta = (double *)malloc(sizeof(double) * 2 * (size_t)m * n * 2);
    double *tb = ta + 2 * m * n;
    double *row = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    int ok = 0;
    const double ea = def ? or_load_window_deformed64(wa, a, h, w, ax, ay, n, def->mode, def->hx, def->hy, -1)
                          : load_window64(wa, a, w, ax, ay, n);
    const double eb = def ? or_load_window_deformed64(wb, b, h, w, bx, by, n, def->mode, def->hx, def->hy, 1)
                          : load_window64(wb, b, w, bx, by, n);
    if (!(ea <= 1e-12 || eb <= 1e-12)) {
        fwd_double(wa, ta, row, n);
        fwd_double(wb, tb, row, n);
        for (int32_t i = 0; i < m * n; i++) {
            const double ar = ta[2 * i], ai = -ta[2 * i + 1];
            const double br = tb[2 * i], bi = tb[2 * i + 1];
            ta[2 * i] = ar * br - ai * bi;
            ta[2 * i + 1] = ar * bi + ai * br;
        }
        inv_double(ta, row, corr, n);
        const double scale = 1.0 / ((double)(n * n) * sqrt(ea * eb));
        for (int32_t i = 0; i < n * n; i++) corr[i] *= scale;
        or_find_peak64(corr, n, out);
        if (corr_out) memcpy(corr_out, corr, sizeof(double) * (size_t)n * n);
        ok = 1;
    }
    free(wa);
    free(ta);
    free(row);
    return ok;
}

/* ---- multipass.rs:194-261 PivAnalyzer::pass (given origins) ----------------------- */
void or_correlate_pass(const float *img_a, const float *img_b, const double *sat_a_sum,
                       const double *sat_a_sq, const double *sat_b_sum, const double *sat_b_sq,
                       int32_t h, int32_t w, const or_grid *g, const float *origins,
                       float min_mean, float min_std, int32_t precision, int32_t threads,
                       float *dx, float *dy, float *ppr, uint8_t *flags, int32_t *peak_iy,
                       int32_t *peak_ix, double *dx64, double *dy64, int32_t deform,
                       const or_grid *prev, const float *prev_dx, const float *prev_dy)
{
    const int32_t n = g->window;
    const int32_t count = g->nx * g->ny;
    const int deformed = deform != 0 && prev != NULL;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
#else
    (void)threads;
#endif
    for (int32_t k = 0; k < count; k++) {
        const float ax = origins[4 * k], ay = origins[4 * k + 1];
        const float bx = origins[4 * k + 2], by = origins[4 * k + 3];
        int sig = 1;
        if (sat_a_sum) {
            float mean, sd;
            or_sat_stats(sat_a_sum, sat_a_sq, w, (int32_t)roundf(ax), (int32_t)roundf(ay), n, n,
                         &mean, &sd);
            if (!(mean >= min_mean && sd >= min_std)) sig = 0;
            if (sig) {
                or_sat_stats(sat_b_sum, sat_b_sq, w, (int32_t)roundf(bx), (int32_t)roundf(by), n,
                             n, &mean, &sd);
                if (!(mean >= min_mean && sd >= min_std)) sig = 0;
            }
        }
        int ok = 0;
        float rdx = NAN, rdy = NAN, rppr = 0.0f;
        int32_t piy = -1, pix = -1;
        float *shifts = NULL;
        deform_t def = {deform, NULL, NULL};
        if (sig && deformed) {
            shifts = (float *)malloc(sizeof(float) * 2 * (size_t)n * n);
            or_deform_shifts(prev, prev_dx, prev_dy, n, ax, ay, bx, by, shifts, shifts + n * n);
            def.hx = shifts;
            def.hy = shifts + n * n;
        }
        if (sig) {
            if (precision == 64) {
                or_peak64 p;
                ok = correlate_f64(img_a, img_b, h, w, n, ax, ay, bx, by, &p, NULL, shifts ? &def : NULL);
                if (ok) {
                    double ddx = ((double)bx - (double)ax) + p.dx;
                    double ddy = ((double)by - (double)ay) + p.dy;
                    rdx = (float)ddx;
                    rdy = (float)ddy;
                    rppr = (float)p.ppr;
                    piy = p.iy;
                    pix = p.ix;
                    if (dx64) dx64[k] = ddx;
                    if (dy64) dy64[k] = ddy;
                }
            } else {
                or_peak p;
                ok = correlate_f32(img_a, img_b, h, w, n, ax, ay, bx, by, &p, NULL, shifts ? &def : NULL);
                if (ok) {
                    rdx = bx - ax + p.dx;
                    rdy = by - ay + p.dy;
                    rppr = p.ppr;
                    piy = p.iy;
                    pix = p.ix;
                }
            }
        }
        free(shifts);
        dx[k] = ok ? rdx : NAN;
        dy[k] = ok ? rdy : NAN;
        ppr[k] = ok ? rppr : 0.0f;
        flags[k] = ok ? 0 : (uint8_t)OR_NO_SIGNAL;
        if (peak_iy) peak_iy[k] = piy;
        if (peak_ix) peak_ix[k] = pix;
        if (!ok && dx64) dx64[k] = NAN;
        if (!ok && dy64) dy64[k] = NAN;
    }
}

/* ---- validate.rs:51-171 ----------------------------------------------------------- */
/* f32::total_cmp key: flip all bits of negatives, the sign bit of positives */
static int32_t total_key(float v)
{
    int32_t b;
    memcpy(&b, &v, 4);
    return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
}

static int cmp_total(const void *pa, const void *pb)
{
    int32_t a = total_key(*(const float *)pa), b = total_key(*(const float *)pb);
    return (a > b) - (a < b);
}

float or_median(float *vals, int32_t n)
{
    qsort(vals, (size_t)n, sizeof(float), cmp_total);
    if (n % 2 == 1) return vals[n / 2];
    return 0.5f * (vals[n / 2 - 1] + vals[n / 2]);
}

static int collect(const float *dx, const float *dy, const uint8_t *snap, int32_t ny, int32_t nx,
                   int32_t ix, int32_t iy, int32_t radius, float *nu, float *nv)
{
    int cnt = 0;
    int32_t ylo = iy > radius ? iy - radius : 0, yhi = iy + radius < ny - 1 ? iy + radius : ny - 1;
    int32_t xlo = ix > radius ? ix - radius : 0, xhi = ix + radius < nx - 1 ? ix + radius : nx - 1;
    for (int32_t y = ylo; y <= yhi; y++)
        for (int32_t x = xlo; x <= xhi; x++) {
            if ((x != ix || y != iy) && (snap[y * nx + x] & OR_INVALID) == 0) {
                nu[cnt] = dx[y * nx + x];
                nv[cnt] = dy[y * nx + x];
                cnt++;
            }
        }
    return cnt;
}

static float residual(float value, float *nb, int32_t cnt, float eps)
{
    const float med = or_median(nb, cnt);
    float dev[32];
    for (int32_t i = 0; i < cnt; i++) dev[i] = fabsf(nb[i] - med);
    return fabsf(value - med) / (or_median(dev, cnt) + eps);
}

void or_validate(int32_t ny, int32_t nx, float *dx, float *dy, const float *ppr, uint8_t *flags,
                 const or_vcfg *cfg, float max_displacement, or_summary *s)
{
    memset(s, 0, sizeof(*s));
    const int32_t count = nx * ny;
    for (int32_t i = 0; i < count; i++) {
        if (flags[i] & OR_NO_SIGNAL) {
            s->no_signal++;
            continue;
        }
        if (hypotf(dx[i], dy[i]) > max_displacement) {
            flags[i] |= OR_OUT_OF_RANGE;
            s->out_of_range++;
        }
        if (ppr[i] < cfg->ppr_threshold) {
            flags[i] |= OR_LOW_PPR;
            s->low_ppr++;
        }
    }
    uint8_t *snap = (uint8_t *)malloc((size_t)count);
    memcpy(snap, flags, (size_t)count);
    float nu[32], nv[32];
    for (int32_t iy = 0; iy < ny; iy++)
        for (int32_t ix = 0; ix < nx; ix++) {
            if (snap[iy * nx + ix] & OR_INVALID) continue;
            int cnt = collect(dx, dy, snap, ny, nx, ix, iy, 1, nu, nv);
            if (cnt < 2) continue;
            float ru = residual(dx[iy * nx + ix], nu, cnt, cfg->median_eps);
            float rv = residual(dy[iy * nx + ix], nv, cnt, cfg->median_eps);
            if (sqrtf(ru * ru + rv * rv) > cfg->median_threshold) {
                flags[iy * nx + ix] |= OR_MEDIAN_OUTLIER;
                s->median_outliers++;
            }
        }
    if (cfg->replace) {
        memcpy(snap, flags, (size_t)count);
        for (int32_t iy = 0; iy < ny; iy++)
            for (int32_t ix = 0; ix < nx; ix++) {
                const uint8_t f = snap[iy * nx + ix];
                if (!((f & OR_NO_SIGNAL) == 0 && (f & OR_INVALID) != 0)) continue;
                int replaced = 0;
                for (int32_t radius = 1; radius <= 2; radius++) {
                    int cnt = collect(dx, dy, snap, ny, nx, ix, iy, radius, nu, nv);
                    if (cnt > 0) {
                        dx[iy * nx + ix] = or_median(nu, cnt);
                        dy[iy * nx + ix] = or_median(nv, cnt);
                        flags[iy * nx + ix] |= OR_REPLACED;
                        s->replaced++;
                        replaced = 1;
                        break;
                    }
                }
                if (!replaced) {
                    dx[iy * nx + ix] = NAN;
                    dy[iy * nx + ix] = NAN;
                }
            }
    }
    free(snap);
}

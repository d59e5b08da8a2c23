/*
 * CPU oracle for the PIV per-pair path.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker that the HIP product path is compared against.  Only tests/, the smoke entry
 * point and the CPU-baseline leg of bench.py may load it; the product never does.
 *
 * It is a plain-C restatement of the Rust reference crate piv-core (agoodman/piv), written
 * from the reference text, function by function (file:line citations on each function in
 * piv_oracle.c).  Two precisions:
 *   - "f32": the reference's FP32 arithmetic in the reference's operation order.  The only
 *     part that cannot follow the reference bit for bit is the FFT (realfft 3.5.0 /
 *     rustfft 6.4.1 are third-party crates, not vendored): the oracle uses its own
 *     mixed-radix FFT, so correlation planes agree to FP32 rounding, not bitwise.
 *   - "f64": the same algorithm in FP64 end to end (window load, FFT, peak fit), the
 *     ground truth for reduced-precision paths.
 *
 * Pinning (see tests/test_oracle_*.py): origins, validation and grids bit-exact against the
 * Rust-written golden fixtures; raw peaks <= 1e-4 px against them; the FFT against a direct
 * DFT; correlation against a direct circular cross-correlation; find_peak, validate, blur,
 * normalize, SAT and sample against the reference's own closed-form unit tests.
 */
#ifndef PIV_ORACLE_H
#define PIV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* flags (piv-core/src/field.rs:6-20) */
#define OR_NO_SIGNAL 1u
#define OR_LOW_PPR 2u
#define OR_MEDIAN_OUTLIER 4u
#define OR_REPLACED 8u
#define OR_OUT_OF_RANGE 16u
#define OR_INVALID (OR_NO_SIGNAL | OR_LOW_PPR | OR_MEDIAN_OUTLIER | OR_OUT_OF_RANGE)

typedef struct {
    int32_t window, step, nx, ny, x0, y0;
} or_grid;

typedef struct {
    float ppr_threshold, median_threshold, median_eps;
    int32_t replace;
} or_vcfg;

typedef struct {
    int64_t no_signal, low_ppr, median_outliers, out_of_range, replaced;
} or_summary;

/* Grid::new -- returns 0 ok, 1 config error (overlap), 2 image too small */
int or_grid_new(int32_t h, int32_t w, int32_t window, float overlap, or_grid *g);

/* VectorField::sample */
void or_field_sample(const or_grid *g, const float *dx, const float *dy, float x, float y,
                     float *u, float *v);

/* multipass.rs place closure, for all cells of a grid; prev may be NULL (first pass) */
void or_place_windows(const or_grid *g, int32_t h, int32_t w, const or_grid *prev,
                      const float *prev_dx, const float *prev_dy, float *origins /* [N*4] */);

/* IntegralImage::new / stats */
void or_sat_build(const float *img, int32_t h, int32_t w, double *sum, double *sumsq);
void or_sat_stats(const double *sum, const double *sumsq, int32_t w, int32_t x, int32_t y,
                  int32_t bw, int32_t bh, float *mean, float *std_);

/* preprocessing */
int32_t or_gauss_kernel(float sigma, float *k /* >= 2r+1 */);
void or_gaussian_blur(const float *in, int32_t h, int32_t w, float sigma, float *out);
void or_local_normalize(const float *in, int32_t h, int32_t w, int32_t radius, float eps,
                        float *out);
/* PreparedFrame::new: image (conditioned) and base (background-subtracted raw, for stats) */
void or_prepare(const float *raw, int32_t h, int32_t w, const float *bg, float sigma,
                int32_t radius, float eps, float *image, float *base);

/* decode.rs:227-234 sample normalisation */
void or_normalize_u8(const uint8_t *in, int64_t count, float *out);
void or_normalize_u16(const uint16_t *in, int64_t count, float *out);

/* correlate.rs:182-210 load_window (f32 reference order); returns energy */
double or_load_window(float *dst, const float *img, int32_t h, int32_t w, float x, float y,
                      int32_t n);

typedef struct {
    float dx, dy, height, ppr;
    int32_t iy, ix; /* integer peak (row, column) in the circular plane */
} or_peak;

typedef struct {
    double dx, dy, height, ppr;
    int32_t iy, ix;
} or_peak64;

/* correlate.rs:215-303 find_peak */
void or_find_peak(const float *corr, int32_t n, or_peak *out);
void or_find_peak64(const double *corr, int32_t n, or_peak64 *out);

/* complex FFT, in place, unnormalised; sign -1 forward, +1 inverse; data interleaved re,im */
void or_fft_f32(float *data, int32_t n, int32_t sign);
void or_fft_f64(double *data, int32_t n, int32_t sign);

/* CorrelationPlan::correlate; returns 1 with a peak, 0 for no signal. corr_out optional [n*n] */
int or_correlate_f32(const float *a, const float *b, int32_t h, int32_t w, int32_t n, float ax,
                     float ay, float bx, float by, or_peak *out, float *corr_out);
int or_correlate_f64(const float *a, const float *b, int32_t h, int32_t w, int32_t n, float ax,
                     float ay, float bx, float by, or_peak64 *out, double *corr_out);

/* PivAnalyzer::pass for given origins (multipass.rs:194-261): mask + correlate + assemble.
 * sat_* may be NULL when min_mean and min_std are both -inf (mask disabled).
 * precision 32 or 64 (64 writes the f64 displacement into dx64/dy64 when non-NULL).
 * peak_iy/peak_ix optional. threads <= 0: all available. */
void or_correlate_pass(const float *img_a, const float *img_b, const double *sat_a_sum,
                       const double *sat_a_sq, const double *sat_b_sum, const double *sat_b_sq,
                       int32_t h, int32_t w, const or_grid *g, const float *origins,
                       float min_mean, float min_std, int32_t precision, int32_t threads,
                       float *dx, float *dy, float *ppr, uint8_t *flags, int32_t *peak_iy,
                       int32_t *peak_ix, double *dx64, double *dy64, int32_t deform,
                       const or_grid *prev, const float *prev_dx, const float *prev_dy);

/* Window deformation (the library's definition; no counterpart in the reference): deform 0
 * rigid, 1 bilinear, 2 Lagrange-cubic resampling, with the previous pass's validated field
 * prev_dx / prev_dy on grid prev (NULL on a first pass: rigid).  The pieces, exposed for
 * tests: the predictor interpolant D, the per-pixel shifts h of one window pair, and the
 * deformed load_window (sgn -1: window A, +1: window B). */
void or_deform_sample(const or_grid *g, const float *dx, const float *dy, float x, float y,
                      float *u, float *v);
void or_deform_shifts(const or_grid *prev, const float *pdx, const float *pdy, int32_t n,
                      float ax, float ay, float bx, float by, float *hx, float *hy);
double or_load_window_deformed(float *dst, const float *img, int32_t h, int32_t w, float x,
                               float y, int32_t n, int32_t mode, const float *hx,
                               const float *hy, int32_t sgn);
double or_load_window_deformed64(double *dst, const float *img, int32_t h, int32_t w, float x,
                                 float y, int32_t n, int32_t mode, const float *hx,
                                 const float *hy, int32_t sgn);

/* validate.rs:51-130 */
void or_validate(int32_t ny, int32_t nx, float *dx, float *dy, const float *ppr, uint8_t *flags,
                 const or_vcfg *cfg, float max_displacement, or_summary *summary);

/* median helper (validate.rs:163-171), exposed for tests; sorts vals in place */
float or_median(float *vals, int32_t n);

/* ---- post-analysis stages (post_oracle.c) ---------------------------------------------- */
typedef struct {
    float vorticity_mean, vorticity_abs_mean, ke_mean, ke_sum, coverage;
    int64_t n_usable, n_vorticity;
} or_flowstat;

void or_flow_usable(int64_t n, const float *dx, const uint8_t *flags, uint8_t *out);
void or_vorticity(int32_t ny, int32_t nx, const float *u, const float *v, const uint8_t *usable,
                  float step, float *w);
void or_flow_stats(int32_t ny, int32_t nx, const float *u, const float *v, const uint8_t *usable,
                   float step, or_flowstat *out);
void or_pair_stats(int32_t ny, int32_t nx, const float *dx, const float *dy,
                   const uint8_t *usable, int32_t step, double dt_s, or_flowstat *out);
float or_neighbour_noise(int32_t ny, int32_t nx, const float *u, const float *v,
                         const uint8_t *usable);
float or_post_median(float *v, int64_t n);
void or_temporal(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy,
                 const uint8_t *usable, const uint8_t *gate, int32_t kernel, int32_t window,
                 int32_t min_support, const int32_t *leave_out, int32_t n_leave, float *out_dx,
                 float *out_dy, uint16_t *support);
void or_spatial(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy,
                const uint8_t *gate, float sigma, float min_weight, int32_t leave_out,
                float *out_dx, float *out_dy, uint16_t *support);
int64_t or_cross_validation_errors(int64_t n, const float *raw_dx, const float *raw_dy,
                                   const uint8_t *usable, const float *pred_dx,
                                   const float *pred_dy, float *out);
float or_median_error(const float *errors, int64_t n, const uint8_t *keep, int64_t *count);
float or_pair_to_pair_change(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy);
float or_spatial_roughness(int32_t nt, int32_t ny, int32_t nx, const float *dx, const float *dy);
void or_mask(int64_t n, const float *values, const uint8_t *usable, float *out);

#ifdef __cplusplus
}
#endif
#endif

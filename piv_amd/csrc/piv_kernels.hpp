// Launchers for the HIP kernels (piv_kernels.hip), called by the host orchestration.
#pragma once

#include <hip/hip_runtime.h>

#include "piv_device.hpp"

namespace piv {

// Largest window size with a correlation kernel; every even size from 8 up to it has one.
constexpr int kMaxWindow = 128;
bool correlate_supported(int n);
// whether a pass of window n (deformation mode `deform`) runs the generic-size kernel
// (k_correlate_any) rather than the register-resident FFT kernel of its size
bool correlate_generic(int n, int deform);
// variant: index into corr_variants(n) (variants.hpp); 0 is the default
// deform: the pass's deformation mode (0 rigid, 1 bilinear, 2 bicubic; only pairs with a
// predictor deform)
hipError_t launch_correlate(int n, int fmt, int variant, const PassPair *pairs, int npairs,
                            uint32_t max_cells, hipStream_t st, int deform = 0);
// Block maps: the per-window and per-vector kernels run ceil(max_cells / 256) blocks for every
// pair, or, with `begins` (device uint32 [npairs + 1]: pair p owns blocks begins[p] ..
// begins[p + 1] - 1, 256 windows or vectors each), exactly the blocks each pair needs:
// total_blocks = begins[npairs].
// placement + mask (one thread per window), and sub-pixel fit + assembly of the raw field
hipError_t launch_place_pass(const PassPair *pairs, int npairs, uint32_t max_cells, hipStream_t st,
                             const uint32_t *begins = nullptr, uint32_t total_blocks = 0);
hipError_t launch_peak_finish(const PassPair *pairs, int npairs, uint32_t max_cells, hipStream_t st,
                              bool f64 = false, const uint32_t *begins = nullptr,
                              uint32_t total_blocks = 0);
// DFT-GEMM correlation on the matrix cores (precision 1 = fp16, 2 = bf16), n in {16, 32, 64}
bool correlate_mf_supported(int n);
hipError_t launch_correlate_mf(int n, int fmt, int precision, const PassPair *pairs, int npairs,
                               uint32_t max_cells, hipStream_t st, int deform = 0);
// placement + mask, correlation, sub-pixel fit and assembly of one pass (three launches)
hipError_t launch_pass(int n, int fmt, int variant, const PassPair *pairs, int npairs,
                       uint32_t max_cells, hipStream_t st);
// both validation launches of one pass; summaries must be zeroed beforehand
hipError_t launch_validate(const ValidatePair *jobs, int njobs, uint32_t max_cells,
                           hipStream_t st, const uint32_t *begins = nullptr,
                           uint32_t total_blocks = 0);
hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st);
// nframes > 1: a batch of frames (raw frames raw_frame_bytes apart, base images H*W apart,
// tables sat_stride doubles apart), one launch per kernel
hipError_t launch_base(int fmt, const void *raw, int pitch, const float *bg, int H, int W,
                       float *out, hipStream_t st, int nframes = 1, long raw_frame_bytes = 0);
// SAT of normalised images; `carry` is scratch of sat_carry_bytes(H, W, nframes)
hipError_t launch_sat(const float *img, int H, int W, double *sum, double *sq, double *carry,
                      hipStream_t st, int nframes = 1, long sat_stride = 0);
// SAT straight from u8 / u16 / f32 samples (no background), a batch of frames raw_frame_bytes
// apart: the same additions in the same order as launch_sat on the normalised image
hipError_t launch_sat_raw(int fmt, const void *raw, int pitch, long raw_frame_bytes, int H, int W,
                          double *sum, double *sq, double *carry, hipStream_t st, int nframes = 1,
                          long sat_stride = 0, int screen_t = 0);
// screen_t > 0 (u16 frames, default mask): screen the frames for aligned all-zero
// screen_t x screen_t tiles first (screen_t divides 32) and build only the tables of frames
// holding one, recording in each frame's state word whether its tables were built
size_t sat_carry_bytes(int H, int W, int nframes);
// SAT straight from the samples of frames at their own addresses (imgs[i], tables with entry
// (0, 0) of the sums at sums[i], squares sat_table_elems(H, W) later), up to 64 frames per
// launch; `carry` is scratch of sat_carry_bytes(H, W, min(nframes, 64))
hipError_t launch_sat_list(int fmt, const void *const *imgs, double *const *sums, int nframes,
                           int pitch, int H, int W, double *carry, hipStream_t st, int screen_t = 0);
// The base image (normalised samples, minus the background when bg is given, clamped at 0)
// and its Gaussian blur in one pass: `out` receives the blurred image, `base_out` (optional)
// the base; bit-identical to launch_base followed by launch_blur.  Radius 1..16 (blur_fusable).
// nframes > 1: frames raw_frame_bytes apart, images H*W apart.
bool blur_fusable(int r);
hipError_t launch_blur_fused(int fmt, const void *raw, int pitch, const float *bg, int H, int W,
                             const float *taps, int r, float *base_out, float *out, hipStream_t st,
                             int nframes = 1, long raw_frame_bytes = 0);
// nframes > 1: a batch of images H*W apart (tables sat_stride doubles apart)
hipError_t launch_blur(const float *in, int H, int W, const float *taps, int r, float *tmp,
                       float *out, hipStream_t st, int nframes = 1);
hipError_t launch_local_norm(const float *in, int H, int W, int radius, float eps,
                             const double *sum, const double *sq, float *out, hipStream_t st,
                             int nframes = 1, long sat_stride = 0);
hipError_t launch_normalize(int fmt, const void *raw, long count, float *out, hipStream_t st);
// dst[i] = src[i * channels + plane] for `count` pixels of u8 / u16 / f32 samples
hipError_t launch_extract_plane(int fmt, const void *src, long count, int channels, int plane,
                                void *dst, hipStream_t st);
hipError_t launch_bg_add(int fmt, const void *raw, int pitch, int H, int W, int mode, int first,
                         float *acc, hipStream_t st);
hipError_t launch_bg_finish(float *acc, long n, float count, hipStream_t st);
hipError_t launch_place(const GridDev &g, const GridDev &prev, int has_prev, const float *pdx,
                        const float *pdy, int H, int W, float *origins, hipStream_t st);


}  // namespace piv

// Internal launchers shared between the drop-in entry points and the fused
// pipeline. All launchers enqueue on `s` and return a cbmv_status.
#pragma once

#include "common.hpp"

namespace cbmv {

// Output placement of a cost volume: element (y, x, d) at y*sy + x*sx + d*sd.
struct VolStrides {
  int64_t sy, sx, sd;
};
inline VolStrides strides_hwd(int64_t H, int64_t W, int64_t D) { return {W * D, D, 1}; }
inline VolStrides strides_dhw(int64_t H, int64_t W, int64_t D) { return {W, 1, H * W}; }
inline VolStrides strides_hdw(int64_t H, int64_t W, int64_t D) { return {D * W, 1, W}; }

// ---- matchers.hip ----
// The cost launchers take an optional `mins` (H*W doubles): when not null,
// each pixel's minimum over d of the values it writes (ascending d, invalid
// entries as RAND_MAX) goes to mins[y * W + x]. With `rm` (H*W words, all
// ones beforehand) the window kernels also merge each pixel's right-view
// minimum (over the elements (x + d, d) inside the row) into rm[y * W + x]:
// as a code for census and SAD of Sobel, as the order-preserving key of the
// value for NCC and zero-mean SAD (dkey: x >= 0 as the bits with the sign
// set, x < 0 as the complemented bits); *rm_done says whether they did
// (the generic kernels of other window sizes do not).
int census_words(int wsize);
// Per-pixel census descriptors, words = census_words(wsize) uint64 per pixel.
int launch_census_desc(const uint8_t* img, int H, int W, int wsize, uint64_t* desc, hipStream_t s);
int launch_census_cost_u8(const uint64_t* dl, const uint64_t* dr, int H, int W, int D, int wsize,
                          uint8_t* out, VolStrides st, hipStream_t s, double* mins = nullptr,
                          unsigned long long* rm = nullptr, bool* rm_done = nullptr);
int launch_census_cost_f64(const uint64_t* dl, const uint64_t* dr, int H, int W, int D, int wsize,
                           double* out, VolStrides st, hipStream_t s);
// Window sums of a u8 image at valid centres: A = sum, C = 1/sqrt(w^2*sum(x^2) - A^2),
// M = A / w^2 (any of A, C, M may be null).
int launch_window_stats(const uint8_t* img, int H, int W, int wsize, uint32_t* A, double* C,
                        double* M, hipStream_t s);
int launch_ncc_cost(const uint8_t* L, const uint8_t* R, const uint32_t* AL, const uint32_t* AR,
                    const double* CL, const double* CR, int H, int W, int D, int wsize,
                    double* out, VolStrides st, hipStream_t s, double* mins = nullptr,
                    unsigned long long* rm = nullptr, bool* rm_done = nullptr);
int launch_sobel_f64(const uint8_t* img, int H, int W, double* out, hipStream_t s);
int launch_sobel_i16(const uint8_t* img, int H, int W, int16_t* out, hipStream_t s);
int launch_sadsob_f64(const double* L, const double* R, int H, int W, int D, int wsize,
                      double* out, VolStrides st, hipStream_t s);
int launch_sadsob_u16(const int16_t* L, const int16_t* R, int H, int W, int D, int wsize,
                      uint16_t* out, VolStrides st, hipStream_t s, double* mins = nullptr,
                      unsigned long long* rm = nullptr, bool* rm_done = nullptr);
int launch_zsad(const uint8_t* L, const uint8_t* R, const double* ML, const double* MR, int H,
                int W, int D, int wsize, double* out, VolStrides st, hipStream_t s,
                double* mins = nullptr, unsigned long long* rm = nullptr,
                bool* rm_done = nullptr);

// ---- post.hip (shared by the drop-in `post` module and the pipeline) ----
int launch_cross(const float* img, int H, int W, int L1, float tau1, float* out, hipStream_t s);
int launch_median2d(const float* in, int H, int W, int ksize, float* out, hipStream_t s);
int launch_mean2d(const float* in, int H, int W, const float* kernel, int ksize, float alpha2,
                  float* out, hipStream_t s);
int launch_outlier(const float* d0, const float* d1, int H, int W, int disp_max, float* out,
                   hipStream_t s);
int launch_interp_mismatch(const float* d0, const float* outlier, int H, int W, float* out,
                           hipStream_t s);
int launch_interp_occlusion(const float* d0, const float* outlier, int H, int W, float* out,
                            hipStream_t s);
// Copy of an [H,W] image with sgm_pad_columns() columns on each side (row
// pitch W + 2 * pad) holding values far outside the image range, different
// at neighbouring cells: the right-image argument of the SGM launchers.
constexpr int kSgmPad = 520;
int sgm_pad_columns();
int launch_pad_rows(const float* img, int H, int W, float* out, hipStream_t s);
// One SGM path over every scanline of an [H,W,D] volume (strides in elements,
// d contiguous). x1 is the padded right image. path 0..3 = right, left, down, up (post.cu:184-208); 4..7
// the diagonals (dx,dy) = (1,1), (-1,-1), (1,-1), (-1,1). mode bit 0: the
// accumulator is not read (first path into a zeroed sum); bit 1: the sum is
// multiplied by `scale` when stored (last path).
int launch_sgm_path_f64(const float* x0, const float* x1, const double* in, double* out, int H,
                        int W, int D, VolStrides st, float pi1, float pi2, float tau_so,
                        float alpha1, float q1, float q2, int direction, int path, int mode,
                        double scale, hipStream_t s);
int launch_sgm_path_f32(const float* x0, const float* x1, const float* in, float* out, int H,
                        int W, int D, VolStrides st, float pi1, float pi2, float tau_so,
                        float alpha1, float q1, float q2, int direction, int path, int mode,
                        double scale, hipStream_t s);

// ---- sgm_pair.hip ----
// Two SGM paths per sweep (8-path chains, ndisp <= 256): pair 0 = E + SE, 1 = W + NW,
// 2 = S + SW, 3 = N + NE; mode and scale as for launch_sgm_path_f32. edge
// (sgm_pair_edge_floats), edge_mn (sgm_pair_edge_entries) and ctr (4
// words, zero) pass the diagonal values between workgroups; the edge words
// start as launch_sgm_pair_fill(always = true) leaves them. spin: wall-clock
// ticks after which a wait gives up (counted in ctr[2]).
int64_t sgm_pair_edge_floats(int H, int W, int D);
int64_t sgm_pair_edge_entries(int H, int W, int D);
int launch_sgm_pair_fill(float* edge, int64_t n_edge, float* edge_mn, int64_t n_mn, unsigned* ctr,
                         bool always, hipStream_t s);
int launch_sgm_pair_f32(const float* x0, const float* x1p, const float* in, float* out, int H,
                        int W, int D, float pi1, float pi2, float tau_so, float alpha1, float q1,
                        float q2, int direction, int pair, int mode, double scale, float* edge,
                        float* edge_mn, unsigned* ctr, long long spin, hipStream_t s);

}  // namespace cbmv

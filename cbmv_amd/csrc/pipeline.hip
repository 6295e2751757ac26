// Fused device pipeline: u8 stereo pair in device memory -> f32 disparity map.
//
// Stages (all on one stream, no host synchronisation inside a run):
//  1. matchers      census / NCC / SAD-of-Sobel / zero-mean SAD volumes, layout
//                   [H, D, W] with compact types (u8, f64, u16, f64)
//  2. statistics    per pixel, per matcher: minimum and sum of exponentials
//                   over d, for the left view and the right view (whose cost
//                   at (y, x', d) is the left cost at (y, x' + d, d))
//  3. classifier    20 features per voxel (features.hpp) -> random forest or
//                   MLP -> matching cost P(y, x, d) = probability of no match,
//                   written [H, W, D] float32
//  4. post          cross arms, SGM, cross-based aggregation, winner-take-all
//                   (+ subpixel), left-right check and interpolation, median
//                   and weighted-mean filters (test.py:153-449)
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>

#include "features.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace cbmv {
namespace {

constexpr int kB = 256;

// XCD-aware block order. The dispatcher hands the blocks of a grid to the
// eight XCDs (each with its own L2) in turn, block b to XCD b mod 8. A 1D
// grid of 8 * per blocks mapped through xcd_block gives each XCD a run of
// consecutive logical blocks, so neighbouring tiles (which read each other's
// columns) share an L2 and run at about the same time. Blocks with a
// logical id >= the tile count exit at once.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int per) {
  return (b % kXcds) * per + b / kXcds;
}
__host__ __device__ inline int xcd_per(int64_t nblocks) { return (int)((nblocks + kXcds - 1) / kXcds); }

// ------------------------------------------------------------------ statistics
// One thread per pixel (x fastest). For each matcher: minimum with strict <
// from RAND_MAX, then the sum of exp(-(c - m)^2 / sigma), both in ascending d
// as in featextract.cpp:322-327,373-383. The matcher kernels write the left
// view's minima into rows 0..3 of the left statistics as they fill the
// volumes, and (the window kernels) the right view's minima over the row
// into rmins, so that only the sums take a pass here; for a matcher whose
// kernel did not, the right view's minimum takes a pass of its own. Right
// view: c_R(x', d) = c(x' + d, d) for x' + d < W, else the batch's first
// element (featextract.cpp:130).
// The left and right views of a matcher run in the same loops (two
// independent loads per step), and each matcher has its own code path (the
// volume type is a template argument) so that the loops unroll into
// straight-line loads.
// Element of a volume row at a 32-bit byte offset from the row's base (a
// row of D x W elements stays below 4 GiB; the base is uniform across the
// workgroup, so the loads take the scalar-base + 32-bit-offset form).
template <typename T>
__device__ __forceinline__ T row_at(const T* row, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(row) + byte_off);
}
template <int K>
__device__ __forceinline__ double code_value(uint32_t c) {
  if constexpr (K == 0) return c == kCensusInvalid ? kRandMax : (double)c;
  else return c == kSobInvalid ? kRandMax : (double)c;
}

template <int K>
__device__ __forceinline__ double vol_at(const Vols& v, int64_t idx) {
  if constexpr (K == 0) {
    const uint8_t c = v.census[idx];
    return c == kCensusInvalid ? kRandMax : (double)c;
  } else if constexpr (K == 1) {
    return v.ncc[idx];
  } else if constexpr (K == 2) {
    const uint16_t c = v.sob[idx];
    return c == kSobInvalid ? kRandMax : (double)c;
  } else {
    return v.sad[idx];
  }
}

// exp(a) for a <= 0 in float64 by Tang's table-driven method (ACM TOMS 15(2),
// 1989): a = (256 m + j) ln2/256 + r with |r| <= ln2/512, exp(a) = 2^m
// 2^(j/256) e^r, e^r by its degree-4 Taylor polynomial (truncation below
// 4e-17 relative). ln2/256 is split Cody-Waite style: the high part has 34
// significant bits, so k * hi is exact for |k| < 2^19 (|a| <= 746). tab[j] =
// 2^(j/256). Relative error about one ulp; below -746 the result is 0.
__device__ __forceinline__ double exp_nonpos(double a, const double* tab) {
  constexpr double kInvLn2x256 = 369.3299304675746;         // 256 / ln2
  constexpr double kLn2x256Hi = 0x1.62e42fef80000p-9;      // ln2/256, 34 bits
  constexpr double kLn2x256Lo = 0x1.1cf79abc9e3b4p-44;     // ln2/256 - hi
  a = fmax(a, -746.0);
  const double kf = rint(a * kInvLn2x256);
  const int k = (int)kf;
  const double r = fma(kf, -kLn2x256Lo, fma(kf, -kLn2x256Hi, a));
  double p = fma(r, 1.0 / 24, 1.0 / 6);
  p = fma(r, p, 0.5);
  p = fma(r, p, 1.0);
  p = fma(r, p, 1.0);
  return ldexp(tab[k & 255] * p, k >> 8);
}

template <int K>
__device__ __forceinline__ double row_val(const Vols& v, int64_t rowbase, uint32_t off) {
  if constexpr (K == 1) return row_at(v.ncc + rowbase, off);
  else if constexpr (K == 3) return row_at(v.sad + rowbase, off);
  else if constexpr (K == 0) return code_value<0>(row_at(v.census + rowbase, off));
  else return code_value<2>(row_at(v.sob + rowbase, off));
}

template <int K>
__device__ __forceinline__ void stats_matcher(const Vols& v, int W, int D, int64_t HW,
                                              int64_t rowbase, int64_t pix, int x, double fill,
                                              double sigma, const double* tab,
                                              const unsigned long long* __restrict__ rmk,
                                              double* __restrict__ sl,
                                              double* __restrict__ sr) {
  constexpr uint32_t kSize = K == 0 ? 1 : K == 2 ? 2 : 8;
  // d < dmax: the right view's x + d is inside the row; beyond it, the fill
  const int dmax = min(D, W - x);
  // the left view's minimum is the matcher kernel's (written into sl); so
  // is the right view's over the row (rmk: a code for SAD of Sobel, a dkey
  // for NCC and ZSAD) when the matcher kernel formed it, else it is taken
  // along the row's diagonal here. The volumes hold no NaN: fmin is the
  // strict-< running minimum (up to the sign of a zero, which no consumer
  // of the minimum sees)
  const double mL = sl[K * HW + pix];
  double mR = dmax < D ? fill : kRandMax;
  if (rmk) {
    const unsigned long long k = rmk[pix];
    double r;
    if constexpr (K == 2) r = code_value<2>(k >= kSobInvalid ? kSobInvalid : (uint32_t)k);
    else r = dkey_value(k);
    mR = fmin(mR, r);
  } else {
#pragma unroll 8
    for (int d = 0; d < dmax; d++)
      mR = fmin(mR, row_val<K>(v, rowbase, (uint32_t)(d * (W + 1) + x) * kSize));
  }
  // exp(-t^2 / sigma) as exp(t^2 * (-1/sigma)): the product differs from
  // the quotient by at most an ulp, a relative change of at most |a| 2^-52
  // in each term
  const double nis = -1.0 / sigma;
  double sL = 0.0, sR = 0.0;
#pragma unroll 4
  for (int d = 0; d < D; d++) {
    const uint32_t off = (uint32_t)(d * W + x) * kSize;
    const double tl = row_val<K>(v, rowbase, off) - mL;
    const double crv = row_val<K>(v, rowbase, d < dmax ? off + d * kSize : off);
    const double tr = (d < dmax ? crv : fill) - mR;
    sL += exp_nonpos(tl * tl * nis, tab);
    sR += exp_nonpos(tr * tr * nis, tab);
  }
  sl[(4 + K) * HW + pix] = sL;
  sr[K * HW + pix] = mR;
  sr[(4 + K) * HW + pix] = sR;
}

// Integer-coded matchers (census u8, SAD of Sobel u16). Codes order like the
// values they stand for (the invalid code is the largest and stands for
// RAND_MAX), so the minimum is an integer minimum of codes, and a term
// exp(-(c - m)^2 / sigma) depends on the code difference t = c - m only:
// etab[t] for t < ntab, 0 beyond (etab[ntab] = 0; exp underflows to 0 there).
// An invalid c contributes exp(0) = 1 when m is invalid too, else
// exp(-(RAND_MAX - m)^2 / sigma) = 0.
template <typename T, int K>
__device__ __forceinline__ void stats_coded(const T* __restrict__ vol, int W, int D, int64_t HW,
                                            int64_t rowbase, int64_t pix, int x, uint32_t fillc,
                                            const double* etab, uint32_t ntab,
                                            const unsigned long long* __restrict__ rmk,
                                            double* __restrict__ sl, double* __restrict__ sr) {
  constexpr uint32_t kInv = (uint32_t)(T)~(T)0;
  const int dmax = min(D, W - x);
  const T* row = vol + rowbase;
  // the left view's minimum from the matcher kernel (RAND_MAX: every code
  // invalid); the right view's over the row from it too (rmk, a code) or
  // along the row's diagonal
  const double ml = sl[K * HW + pix];
  const uint32_t mL = ml == kRandMax ? kInv : (uint32_t)ml;
  uint32_t mR = dmax < D ? fillc : kInv;
  if (rmk) {
    const unsigned long long k = rmk[pix];
    mR = min(mR, k >= kInv ? kInv : (uint32_t)k);
  } else {
#pragma unroll 8
    for (int d = 0; d < dmax; d++)
      mR = min(mR, (uint32_t)row_at(row, (uint32_t)(d * (W + 1) + x) * (uint32_t)sizeof(T)));
  }
  const uint32_t zl = mL == kInv ? 0u : ntab, zr = mR == kInv ? 0u : ntab;
  double sL = 0.0, sR = 0.0;
#pragma unroll 8
  for (int d = 0; d < D; d++) {
    const uint32_t off = (uint32_t)(d * W + x) * (uint32_t)sizeof(T);
    const uint32_t cl = row_at(row, off);
    const uint32_t crv = row_at(row, d < dmax ? off + d * (uint32_t)sizeof(T) : off);
    const uint32_t cr = d < dmax ? crv : fillc;
    sL += etab[cl == kInv ? zl : min(cl - mL, ntab)];
    sR += etab[cr == kInv ? zr : min(cr - mR, ntab)];
  }
  sl[(4 + K) * HW + pix] = sL;
  sr[K * HW + pix] = mR == kInv ? kRandMax : (double)mR;
  sr[(4 + K) * HW + pix] = sR;
}

// table sizes of the coded matchers: census codes lie below 256; the SAD of
// Sobel table covers t < sqrt(745.2 sigma) (beyond it the term is 0) when
// that fits in kSobTab entries, otherwise that matcher takes the float64 path
constexpr int kCenTab = 256, kSobTab = 4096;

__global__ void __launch_bounds__(kB) k_stats(Vols v, int H, int W, int D, int rows_per_batch,
                                              FeatureSpec fs,
                                              const unsigned long long* __restrict__ rmins,
                                              uint32_t rmask, double* __restrict__ sl,
                                              double* __restrict__ sr) {
  __shared__ double tab[256];
  __shared__ double ctab[kCenTab + 1], stab[kSobTab + 1];
  // exp(-t^2 / sigma) as the reference writes it (featextract.cpp:373-383)
  const double sob_reach = sqrt(745.2 * fs.sigma[2]);
  const bool sob_coded = sob_reach < kSobTab;
  const uint32_t nsob = sob_coded ? (uint32_t)sob_reach + 1 : 0;
  for (int i = threadIdx.x; i < 256; i += blockDim.x) tab[i] = exp2(i / 256.0);
  for (int t = threadIdx.x; t <= kCenTab; t += kB)
    ctab[t] = t < kCenTab ? exp(-((double)t * t) / fs.sigma[0]) : 0.0;
  if (sob_coded)
    for (int t = threadIdx.x; t <= (int)nsob; t += kB)
      stab[t] = t < (int)nsob ? exp(-((double)t * t) / fs.sigma[2]) : 0.0;
  __syncthreads();
  // logical tile (bx, y), row-major, in XCD order: the tiles of a row share
  // an L2, where the right view's columns x + d of the next tile are read
  const int nbx = (W + kB - 1) / kB;
  const int l = xcd_block(blockIdx.x, xcd_per((int64_t)nbx * H));
  if (l >= nbx * H) return;
  const int x = (l % nbx) * kB + threadIdx.x;
  const int y = l / nbx;
  if (x >= W) return;
  const int64_t HW = (int64_t)H * W;
  const int64_t pix = (int64_t)y * W + x;
  const int64_t rowbase = (int64_t)y * D * W;
  const int64_t f = (int64_t)((y / rows_per_batch) * rows_per_batch) * D * W;
  // the right view's row minima the matcher kernels formed (bit K of rmask)
  auto rmk = [&](int k) { return (rmask >> k) & 1u ? rmins + k * HW : nullptr; };
  stats_coded<uint8_t, 0>(v.census, W, D, HW, rowbase, pix, x, v.census[f], ctab, kCenTab, rmk(0),
                          sl, sr);
  stats_matcher<1>(v, W, D, HW, rowbase, pix, x, vol_at<1>(v, f), fs.sigma[1], tab, rmk(1), sl, sr);
  if (sob_coded)
    stats_coded<uint16_t, 2>(v.sob, W, D, HW, rowbase, pix, x, v.sob[f], stab, nsob, rmk(2), sl,
                             sr);
  else
    stats_matcher<2>(v, W, D, HW, rowbase, pix, x, vol_at<2>(v, f), fs.sigma[2], tab, rmk(2), sl,
                     sr);
  stats_matcher<3>(v, W, D, HW, rowbase, pix, x, vol_at<3>(v, f), fs.sigma[3], tab, rmk(3), sl, sr);
}

// ------------------------------------------------------------------ features + forest
// Block = 64 x-values x 64 disparities of one row; a wavefront takes one x
// and the 64 disparities (lane = d - d0), so that its lanes are voxels of
// one pixel: their walks through a tree coincide more often than those of
// 64 neighbouring pixels at one disparity (the forest kernels ran 6 %
// faster so), and the [H, W, D] output row of 64 disparities is one
// contiguous store.
constexpr int kTX = 64, kTD = 64;

__global__ void __launch_bounds__(kB) k_features_rf(Vols v, Stats st, FeatureSpec fs, int H, int W,
                                                    int D, int rows_per_batch, ForestView f,
                                                    float* __restrict__ P) {
  __shared__ float xs[20][kB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x0 = blockIdx.x * kTX, d0 = blockIdx.y * kTD, y = blockIdx.z;
  const int d = d0 + lane;
  for (int j = 0; j < kTX / 4; j++) {
    const int x = x0 + wave + 4 * j;
    if (x < W && d < D) {
      voxel_features(v, st, fs, H, W, D, rows_per_batch, y, x, d,
                     [&](int i, float val) { xs[i][tid] = val; });
      const double acc = forest_sum0<2>(f, [&](uint32_t i) { return xs[i][tid]; });
      // sklearn: sum over trees, then divide by the tree count
      P[((int64_t)y * W + x) * D + d] = (float)(acc / f.n_trees);
    }
  }
}

// The same with the first kTopLevels levels of every tree in LDS (forests of
// at most kTopMaxTrees trees): 16 wavefronts per block, one block per CU
// (155 KB of LDS: the features of 1024 voxels, 80 KB, and the top levels),
// so the CU keeps the 16 wavefronts the register count allows and the top
// table is loaded once per 4096 voxels.
constexpr int kRfWaves = 16;

__global__ void __launch_bounds__(64 * kRfWaves) k_features_rf_top(
    Vols v, Stats st, FeatureSpec fs, int H, int W, int D, int rows_per_batch, ForestView f,
    ForestTop top, float* __restrict__ P) {
  __shared__ float xs[20][64 * kRfWaves];
  __shared__ GpuNode lnodes[kTopMaxTrees * kTopNodes];
  __shared__ uint32_t lexits[kTopMaxTrees * kTopExits];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < f.n_trees * kTopNodes; i += 64 * kRfWaves) lnodes[i] = top.nodes[i];
  for (int i = tid; i < f.n_trees * kTopExits; i += 64 * kRfWaves) lexits[i] = top.exits[i];
  __syncthreads();
  const int x0 = blockIdx.x * kTX, d0 = blockIdx.y * kTD, y = blockIdx.z;
  const int d = d0 + lane;
  for (int j = 0; j < kTX / kRfWaves; j++) {
    const int x = x0 + wave + kRfWaves * j;
    if (x < W && d < D) {
      voxel_features(v, st, fs, H, W, D, rows_per_batch, y, x, d,
                     [&](int i, float val) { xs[i][tid] = val; });
      const double acc =
          forest_sum0_top<2>(f, lnodes, lexits, [&](uint32_t i) { return xs[i][tid]; });
      P[((int64_t)y * W + x) * D + d] = (float)(acc / f.n_trees);
    }
  }
}

// Features of one row for tests: out[d][x][20].
__global__ void k_row_features(Vols v, Stats st, FeatureSpec fs, int H, int W, int D,
                               int rows_per_batch, int y, float* __restrict__ out) {
  const int x = blockIdx.x * kB + threadIdx.x;
  const int d = blockIdx.y;
  if (x >= W) return;
  float* o = out + ((int64_t)d * W + x) * 20;
  voxel_features(v, st, fs, H, W, D, rows_per_batch, y, x, d, [&](int i, float val) { o[i] = val; });
}

// Single-precision statistics of both views for the MLP features: minima
// rounded to float, reciprocals of the sums computed in float64.
__global__ void k_stats_f32(const double* __restrict__ sl, const double* __restrict__ sr,
                            int64_t HW, float* __restrict__ out) {
  const int64_t p = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (p >= HW) return;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    out[k * HW + p] = (float)sl[k * HW + p];
    out[(4 + k) * HW + p] = (float)(1.0 / sl[(4 + k) * HW + p]);
    out[(8 + k) * HW + p] = (float)sr[k * HW + p];
    out[(12 + k) * HW + p] = (float)(1.0 / sr[(4 + k) * HW + p]);
  }
}

// Single-precision features of listed voxels (y, x, d), for MLP training.
__global__ void k_sample_features(Vols v, StatsF st, FeatureSpec fs, int H, int W, int D,
                                  int rows_per_batch, const int32_t* __restrict__ samp, int64_t n,
                                  float* __restrict__ out) {
  const int64_t k = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (k >= n) return;
  const int y = samp[3 * k], x = samp[3 * k + 1], d = samp[3 * k + 2];
  float* o = out + k * 20;
  if (y < 0 || y >= H || x < 0 || x >= W || d < 0 || d >= D) {
    for (int i = 0; i < 20; i++) o[i] = 0.f;
    return;
  }
  voxel_features_f32(v, st, fs, H, W, D, rows_per_batch, y, x, d, [&](int i, float val) { o[i] = val; });
}

// ------------------------------------------------------------------ image normalisation
// (img - mean) / std in float32 (test.py:489-492), with numpy's float32
// reductions reproduced bit for bit: numpy sums a contiguous float32 array
// in chunks of its ufunc buffer (8192 elements); each chunk is summed
// pairwise (halves rounded down to a multiple of 8 until at most 128
// elements remain; such a block is summed with 8 interleaved accumulators
// combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then its tail in order;
// fewer than 8 elements in order from 0), and the chunk sums are added in
// order starting from 0. The mean and the variance are that sum divided by
// n in double and rounded to float (numpy 2: float32 scalar over an intp
// scalar); std is the float square root of the variance, whose input is the
// float32 array (x - mean)^2, each square rounded. Rounding a double
// quotient or square root of floats to float equals the correctly rounded
// float result (53 >= 2*24 + 2 bits), so double arithmetic is used for the
// divisions and the square root.
constexpr int kNpChunk = 8192;

// element i of the reduced array: the image value, or its squared deviation
struct NpElem {
  const uint8_t* lds;
  float mean;
  bool square;
  __device__ float operator()(int i) const {
#pragma clang fp contract(off)
    const float v = (float)lds[i];
    if (!square) return v;
    const float t = v - mean;
    return t * t;
  }
};

// pairwise_sum of numpy on elements [a, a + n), n <= 128 (one block)
__device__ float np_block_sum(const NpElem& e, int a, int n) {
#pragma clang fp contract(off)
  if (n < 8) {
    float r = 0.f;
    for (int i = 0; i < n; i++) r += e(a + i);
    return r;
  }
  float r[8];
#pragma unroll
  for (int j = 0; j < 8; j++) r[j] = e(a + j);
  int i = 8;
  for (; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] += e(a + i + j);
  }
  float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; i++) res += e(a + i);
  return res;
}

__device__ __forceinline__ int np_half(int n) {
  const int n2 = n / 2;
  return n2 - n2 % 8;
}

// One wavefront per chunk (blockIdx.x) of image blockIdx.y: lane l takes
// the node of the pairwise tree reached by the 6 bits of l (a bit of 1 at
// a node that is not split leaves the lane empty), sums it (at most one
// more split: a node at depth 6 of a chunk holds at most 142 elements),
// and the lanes combine their sums back up the tree in its order.
// SQ: the squared deviations from the image's mean (stat[2*img]).
template <bool SQ>
__global__ void __launch_bounds__(64) k_np_chunk_sums(const uint8_t* __restrict__ img0,
                                                      const uint8_t* __restrict__ img1, int64_t n,
                                                      const float* __restrict__ stat,
                                                      float* __restrict__ chunks, int nchunks) {
  __shared__ uint8_t buf[kNpChunk];
  const int lane = threadIdx.x;
  const int im = blockIdx.y;
  const uint8_t* img = im ? img1 : img0;
  const int64_t a0 = (int64_t)blockIdx.x * kNpChunk;
  const int m = (int)min((int64_t)kNpChunk, n - a0);
  for (int i = lane; i < m; i += 64) buf[i] = img[a0 + i];
  __syncthreads();
  NpElem e{buf, SQ ? stat[2 * im] : 0.f, SQ};
  int a = 0, len = m;
  bool alive = true;
  unsigned split = 0;  // bit k: the node at depth k was split
  for (int k = 0; k < 6; k++) {
    const int bit = (lane >> (5 - k)) & 1;
    if (len > 128) {
      split |= 1u << k;
      const int h = np_half(len);
      if (bit) {
        a += h;
        len -= h;
      } else {
        len = h;
      }
    } else if (bit) {
      alive = false;
    }
  }
  float v = 0.f;
  if (alive) {
    if (len > 128) {
      const int h = np_half(len);
      v = np_block_sum(e, a, h) + np_block_sum(e, a + h, len - h);
    } else {
      v = np_block_sum(e, a, len);
    }
  }
  // combine: at depth k the node's first lane adds its right child's sum
  for (int k = 5; k >= 0; k--) {
    const int s = 1 << (5 - k);
    const float other = __shfl_down(v, s, 64);
    if ((lane & (2 * s - 1)) == 0 && ((split >> k) & 1)) v = v + other;
  }
  if (lane == 0) chunks[im * nchunks + blockIdx.x] = v;
}

// The chunk sums added in order from 0, divided by n: the mean (SQ false,
// stat[2*img]) or the standard deviation (SQ true, stat[2*img+1]).
template <bool SQ>
__global__ void k_np_finish(const float* __restrict__ chunks, int nchunks, int64_t n,
                            float* __restrict__ stat) {
  if (threadIdx.x != 0) return;
  const int im = blockIdx.x;
  float s = 0.f;
  for (int c = 0; c < nchunks; c++) s += chunks[im * nchunks + c];
  const float q = (float)((double)s / (double)n);
  if (SQ)
    stat[2 * im + 1] = (float)sqrt((double)q);
  else
    stat[2 * im] = q;
}

__global__ void k_img_normalise(const uint8_t* __restrict__ img0, const uint8_t* __restrict__ img1,
                                int64_t n, const float* __restrict__ stat,
                                float* __restrict__ out0, float* __restrict__ out1) {
#pragma clang fp contract(off)
  const int im = blockIdx.y;
  const uint8_t* img = im ? img1 : img0;
  float* out = im ? out1 : out0;
  const float mean = stat[2 * im], sd = stat[2 * im + 1];
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
    const float t = (float)img[i] - mean;
    out[i] = (float)((double)t / (double)sd);
  }
}
static int np_chunks(int64_t n) { return (int)((n + kNpChunk - 1) / kNpChunk); }

// scratch of the normalisation: both images' chunk sums, then mean and std of each
static size_t normalise_scratch_bytes(int64_t n) { return ((size_t)2 * np_chunks(n) + 4) * 4; }

// Normalises both u8 images of n pixels into float32 (outL, outR).
static int launch_normalise(const uint8_t* L, const uint8_t* R, int64_t n, float* scratch,
                            float* outL, float* outR, hipStream_t s) {
  const int nc = np_chunks(n);
  float* chunks = scratch;
  float* stat = scratch + 2 * nc;
  k_np_chunk_sums<false><<<dim3(nc, 2), 64, 0, s>>>(L, R, n, stat, chunks, nc);
  k_np_finish<false><<<2, 64, 0, s>>>(chunks, nc, n, stat);
  k_np_chunk_sums<true><<<dim3(nc, 2), 64, 0, s>>>(L, R, n, stat, chunks, nc);
  k_np_finish<true><<<2, 64, 0, s>>>(chunks, nc, n, stat);
  const int grid = (int)min((int64_t)1024, (n + kB - 1) / kB);
  k_img_normalise<<<dim3(grid, 2), kB, 0, s>>>(L, R, n, stat, outL, outR);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

// ------------------------------------------------------------------ right-view cost
// PR[y, x', d] = P[y, x' + d, d] for x' + d < W, else P[0] (get_right_cost,
// featextract.cpp:115-152, applied to the whole volume at test.py:156).
// The right view's cost PR[y][x][d] = P[y][x+d][d], or P[0] where x + d is
// outside the image (get_right_cost, featextract.cpp:115-152). A workgroup
// builds tx output pixels of one row in LDS (tx = 64, or 32 when D > 256 so
// that the tile stays within 64 KB): for each source pixel xs the lanes read
// the run of disparities that map into the tile (contiguous in d), then the
// tile is written out as one contiguous block.
constexpr int kRightCostThreads = 512;  // 8 wavefronts: 16 per CU at 64 KB of LDS per block

__global__ void __launch_bounds__(kRightCostThreads) k_right_cost(const float* __restrict__ P, int H,
                                                                  int W, int D, int tx,
                                                                  float* __restrict__ PR) {
  constexpr int NW = kRightCostThreads / 64;
  extern __shared__ float tile[];  // [tx][D]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x0 = blockIdx.x * tx, y = blockIdx.y;
  const int64_t row = (int64_t)y * W * D;
  const float fill = P[0];
  const int x = x0 + lane;  // output pixel of this lane (lanes >= tx idle)
  // source pixels xs = x + d for d in [0, D): xs in [x0, x0 + tx - 1 + D - 1],
  // column xs = x0 + wave + NW (u + 8 j); eight loads in flight per
  // wavefront (unconditional, clamped addresses)
  const int xend = x0 + tx + D - 1;
  for (int xb = x0 + wave; xb < xend; xb += 8 * NW) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int xs = xb + NW * u;
      const int d = min(max(xs - x, 0), D - 1);
      v[u] = P[row + (int64_t)min(xs, W - 1) * D + d];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int xs = xb + NW * u;
      const int d = xs - x;
      if (xs < xend && lane < tx && d >= 0 && d < D) tile[lane * D + d] = xs < W ? v[u] : fill;
    }
  }
  __syncthreads();
  const int nx = min(tx, W - x0);
  float* out = PR + row + (int64_t)x0 * D;
  for (int e = tid; e < nx * D; e += kRightCostThreads) out[e] = tile[e];
}

// ------------------------------------------------------------------ CBCA on [H, W, D]
// Pass 1: per (row, x, d) the float sum and count over the horizontal arm at
// row y of the intersection of the two crosses (post.cu:418-426). The arm
// bounds lie within x +- (L1 - 1), so the lanes of a pixel read one pixel's
// disparities at a time (coalesced) under a per-lane mask, in ascending x.
__global__ void k_cbca_h(const float* __restrict__ vol, const float* __restrict__ lc,
                         const float* __restrict__ rc, int H, int W, int D, int L1, int dir,
                         float* __restrict__ hsum, uint8_t* __restrict__ hcnt) {
  const int d = threadIdx.x;
  const int x = blockIdx.x, y = blockIdx.y;
  if (d >= D) return;
  const int64_t HW = (int64_t)H * W;
  const int64_t pix = (int64_t)y * W + x;
  const int dd = d * dir;
  const int xr = x + dd;
  float sum = 0.f;
  int cnt = 0;
  if (xr >= 0 && xr < W) {
    const int xs = (int)fmaxf(lc[pix], rc[pix + dd] - dd);
    const int xt = (int)fminf(lc[HW + pix], rc[HW + pix + dd] - dd);
    const int lo = max(x - L1 + 1, 0), hi = min(x + L1 - 1, W - 1);
    const float* row = vol + (int64_t)y * W * D + d;
    for (int xx = lo; xx <= hi; xx++)
      if (xx > xs && xx < xt) {
        sum += row[(int64_t)xx * D];
        cnt++;
      }
  }
  const int64_t o = pix * D + d;
  hsum[o] = sum;
  hcnt[o] = (uint8_t)cnt;
}

// Pass 2: sum of the row sums over the vertical arm (post.cu:415-417), mean.
// Voxels whose matching pixel is outside the image keep their value
// (post.cu:409-410). With WTA set, the block (one pixel, thread = d) also
// takes the first minimum over d and the subpixel parabola (post.cu:453-468),
// writing only the two disparity maps.
template <bool WTA>
__global__ void k_cbca_v(const float* __restrict__ vol, const float* __restrict__ hsum,
                         const uint8_t* __restrict__ hcnt, const float* __restrict__ lc,
                         const float* __restrict__ rc, int H, int W, int D, int L1, int dir,
                         float* __restrict__ out, float* __restrict__ wta,
                         float* __restrict__ sub) {
  __shared__ float vals[1024];
  __shared__ float rmin[16];
  __shared__ int ridx[16];
  const int d = threadIdx.x;
  const int x = blockIdx.x, y = blockIdx.y;
  const int64_t HW = (int64_t)H * W;
  const int64_t pix = (int64_t)y * W + x;
  float res = INFINITY;
  if (d < D) {
    const int dd = d * dir;
    const int xr = x + dd;
    const int64_t o = pix * D + d;
    if (xr < 0 || xr >= W) {
      res = vol[o];
    } else {
      const int ys = (int)fmaxf(lc[2 * HW + pix], rc[2 * HW + pix + dd]);
      const int yt = (int)fminf(lc[3 * HW + pix], rc[3 * HW + pix + dd]);
      const int lo = max(y - L1 + 1, 0), hi = min(y + L1 - 1, H - 1);
      float sum = 0.f;
      int cnt = 0;
      for (int yy = lo; yy <= hi; yy++)
        if (yy > ys && yy < yt) {
          const int64_t q = ((int64_t)yy * W + x) * D + d;
          sum += hsum[q];
          cnt += hcnt[q];
        }
      res = sum / cnt;
    }
    if (!WTA) out[o] = res;
  }
  if (WTA) {
    vals[d] = res;
    // first minimum over d (numpy argmin, test.py:409)
    float bv = res;
    int bi = d < D ? d : 0x7fffffff;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const float ov = __shfl_xor(bv, m, 64);
      const int oi = __shfl_xor(bi, m, 64);
      if (ov < bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if ((d & 63) == 0) {
      rmin[d >> 6] = bv;
      ridx[d >> 6] = bi;
    }
    __syncthreads();
    if (d == 0) {
      const int nw = (blockDim.x + 63) >> 6;
      float b = rmin[0];
      int bidx = ridx[0];
      for (int w = 1; w < nw; w++)
        if (rmin[w] < b || (rmin[w] == b && ridx[w] < bidx)) {
          b = rmin[w];
          bidx = ridx[w];
        }
      wta[pix] = (float)bidx;
      float o = (float)bidx;
      if (1 <= bidx && bidx < D - 1) {
        const float cn = vals[bidx - 1], cz = vals[bidx], cp = vals[bidx + 1];
        const float denom = 2 * (cp + cn - 2 * cz);
        if ((double)denom > 1e-5) {
          const double r = fmin(1.0, fmax(-1.0, (double)((cp - cn) / denom)));
          o = (float)((double)bidx - r);
        }
      }
      sub[pix] = o;
    }
  }
}

// x clamped to [0, 1]: one instruction, the clamp modifier on the operation
// that produces x.
__device__ __forceinline__ float sat01(float x) { return __builtin_amdgcn_fmed3f(x, 0.f, 1.f); }

// ---- fused CBCA: row sums and the column mean in one pass ----
// A workgroup takes TX columns x 64 disparities (wavefront w = column x0 + w,
// lane = disparity; TX = 4, 8 or 16, a tuning choice) and walks down the
// image. Each row entering the
// 2*L1-1 row window is staged once in LDS (the 8 columns plus L1-1 halo
// columns on each side, double-buffered: one barrier per row); every thread
// forms that row's masked sum over its horizontal range from the staged row
// (ascending x, as post.cu:420-426) and keeps the window of row sums and
// counts in registers, then writes the mean over its vertical range. The
// arithmetic is that of the separate row and column walkers (row sums in
// ascending x, then added in ascending y), without their row-sum volumes in
// memory. Loads for the next PF rows sit in a register ring. With WTA, the
// pass also keeps each pixel's first minimum over d (test.py:409): per
// output row, every wavefront (one column, 64 disparities) finds its
// minimum and the lowest disparity holding it, and one 64-bit atomic minimum
// of (ordered value bits, d) merges the disparity blocks into wkey, whose
// low word is then the winner (ties go to the smaller d, as in argmin).
template <int L1, int PF, int TX, bool WTA>
__global__ void __launch_bounds__(TX * 64, 1024 / (TX * 64)) k_cbca_fused(
    const float* __restrict__ vol, const float* __restrict__ lc, const float* __restrict__ rc,
    int H, int W, int D, int dir, int chunk_rows, float* __restrict__ out,
    unsigned long long* __restrict__ wkey) {
  constexpr int K = 2 * L1 - 1;
  constexpr int R = L1 - 1;
  constexpr int NT = TX * 64;           // threads
  constexpr int NC = TX + 2 * R;        // staged columns
  constexpr int NE = NC * 64;           // staged elements per row
  constexpr int NL = (NE + NT - 1) / NT;  // staged elements per thread
  static_assert(NT > K * K, "reciprocal table is filled by one pass of the threads");
  // arm values staged per row: the left image's four planes at the tile's
  // TX columns, the right image's at the TX + 63 columns that the tile's
  // voxels match
  constexpr int NRC = TX + 63;
  constexpr int NAE = 4 * TX + 4 * NRC;
  constexpr int NA = (NAE + NT - 1) / NT;  // staged arm values per thread
  __shared__ float rowbuf[2][NC][64];
  __shared__ float armbuf[2][NAE];
  __shared__ float rtab[K * K + 1];  // 1/c rounded to nearest (rtab[0] unused)
  // logical tile (gx, gy, gz) = (column strip, disparity block, row chunk),
  // column strips fastest, in XCD order: neighbouring strips, which stage
  // each other's halo columns, share an L2
  const int ngx = (W + TX - 1) / TX, ngy = (D + 63) / 64, ngz = (H + chunk_rows - 1) / chunk_rows;
  const int l = xcd_block(blockIdx.x, xcd_per((int64_t)ngx * ngy * ngz));
  if (l >= ngx * ngy * ngz) return;  // whole workgroup, before any barrier
  const int gx = l % ngx, gy = (l / ngx) % ngy, gz = l / (ngx * ngy);
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (tid <= K * K) rtab[tid] = tid ? 1.0f / (float)tid : 0.f;
  const int x0 = gx * TX, d0 = gy * 64;
  const int x = x0 + w, d = d0 + lane;
  const bool ok = x < W && d < D;
  const int dd = d * dir;
  const int xr = x + dd;
  const bool inside = ok && xr >= 0 && xr < W;
  const int xc = min(x, W - 1), dc = min(d, D - 1);
  // Volume accesses are buffer accesses: a per-row descriptor and a 32-bit
  // per-lane byte offset that stays fixed along the walk (the host keeps a
  // volume row below 4 GiB).
  const uint32_t rowD = (uint32_t)W * (uint32_t)D * 4u;  // bytes per volume row
  const char* const vb = reinterpret_cast<const char*>(vol);
  // Arm values, staged in LDS with each row (one load per thread for
  // TX = 8, instead of eight per thread): element e < 4 TX is the left
  // image's plane e / TX at column x0 + e % TX; element 4 TX + e' the right
  // image's plane e' / NRC at column xbase + e' % NRC, where the voxel
  // (x, d) of the tile reads column x + d dir = xbase + (xr - xbase) with
  // 0 <= xr - xbase < NRC. Columns are clamped to the image as the volume's
  // are. Planes 0, 1 are read at the entering row, planes 2, 3 at the
  // output row.
  const int xbase = dir < 0 ? x0 - d0 - 63 : x0 + d0;
  const int xo = xr - xbase;
  const int64_t HW = (int64_t)H * W;
  const float* aptr[NA];
  bool arow[NA];  // the value belongs to the output row (planes 2, 3)
#pragma unroll
  for (int k = 0; k < NA; k++) {
    const int e = min(tid + NT * k, NAE - 1);
    const bool left = e < 4 * TX;
    const int ei = left ? e : e - 4 * TX;
    const int plane = left ? ei / TX : ei / NRC;
    const int col = left ? min(x0 + ei % TX, W - 1) : min(max(xbase + ei % NRC, 0), W - 1);
    aptr[k] = (left ? lc : rc) + plane * HW + col;
    arow[k] = plane >= 2;
  }
  // staged element e = c * 64 + l: column x0 - R + c (clamped), disparity
  // d0 + l (clamped); columns outside the image are staged as 0
  uint32_t soff[NL];
  bool sin[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int e = tid + NT * k;
    const int cx = x0 - R + e / 64;
    sin[k] = cx >= 0 && cx < W;
    const int c = min(max(cx, 0), W - 1);
    const int l = min(d0 + e % 64, D - 1);
    soff[k] = e < NE ? (uint32_t)(c * D + l) * 4u : 0u;
  }
  // only voxels whose match is outside the image copy their cost; the
  // others read the row's first element, one cache line for the wavefront
  const uint32_t cpb = inside ? 0u : (uint32_t)(xc * D + dc) * 4u;
  const uint32_t ob = (uint32_t)(xc * D + dc) * 4u;
  // taps x-1 and x+1 belong to every row's range except at the borders of
  // either image (a distance-1 arm step is always accepted, post.cu:336-396):
  // per-thread 0/1 weights
  const float cwl = (x >= 1 && xr >= 1) ? 1.f : 0.f;
  const float cwr = (x <= W - 2 && xr <= W - 2) ? 1.f : 0.f;
  const float ccen = cwl + 1.f + cwr;  // pixels of the centre three taps
  // arm ends relative to tap 0 (column x - R), integer-valued floats:
  // a = end + (R - x) for the left image's arm, the right image's arm also
  // shifted by -dd
  const float xoff = (float)(R - x), xoffr = (float)(R - x - dd);
  // ring of raw loaded values (no arithmetic before use)
  float rs[PF][NL], ra[PF][NA], rcp[PF];
  // iteration i brings row r = i into the window and writes row y = i - R
  auto load = [&](int slot, int i) {
    const int r = min(i, H - 1);
    const int y = min(max(i - R, 0), H - 1);
    const __amdgpu_buffer_rsrc_t vr = buffer_rsrc(vb + (int64_t)r * rowD, rowD);
    const __amdgpu_buffer_rsrc_t vy = buffer_rsrc(vb + (int64_t)y * rowD, rowD);
#pragma unroll
    for (int k = 0; k < NL; k++) rs[slot][k] = buffer_ldf(vr, soff[k], 0);
#pragma unroll
    for (int k = 0; k < NA; k++) ra[slot][k] = aptr[k][(int64_t)(arow[k] ? y : r) * W];
    rcp[slot] = buffer_ldf(vy, cpb, 0);
  };
  // window of row sums and their pixel counts (counts as floats, exact)
  float ws[K], wc[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    ws[q] = 0.f;
    wc[q] = 0.f;
  }
  // this workgroup writes rows [ya, yb); its walk starts R rows above ya
  // (the window fills there) and ends when row yb - 1 is written
  const int ya = gz * chunk_rows, yb = min(H, ya + chunk_rows);
  const int s0 = max(ya - R, 0), n = yb + R;
#pragma unroll
  for (int k = 0; k < PF; k++) load(k, s0 + k);
  // one row of the walk; ring slot k holds the inputs of iteration i
  auto row_step = [&](int k, int i) {
    // nothing of this iteration is scheduled into the previous one (it
    // would wait early for loads still in flight)
    __builtin_amdgcn_sched_barrier(0);
    const int buf = i & 1;
    // staged element e of this row sits at rowbuf[buf][0][0] + e; thread
    // (w, lane) stages e = tid + NT q and reads its taps at column w + q,
    // disparity lane, i.e. at e = tid + 64 q: one base address per row
    float* const rt = &rowbuf[buf][0][0] + tid;
#pragma unroll
    for (int q = 0; q < NL; q++) {
      const int e = tid + NT * q;
      if (q < NL - 1 || e < NE) rt[NT * q] = sin[q] ? rs[k][q] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < NA; q++) {
      const int e = tid + NT * q;
      if (q < NA - 1 || e < NAE) armbuf[buf][e] = ra[k][q];
    }
    __syncthreads();
    // this voxel's arms: the left image's at column x (the same for the
    // wavefront), the right image's at column xr
    const float* const al = &armbuf[buf][w];
    const float* const ar = &armbuf[buf][4 * TX + xo];
    const float rh0 = al[0], rh1 = al[TX], rv2 = al[2 * TX], rv3 = al[3 * TX];
    const float rq0 = ar[0], rq1 = ar[NRC], rw2 = ar[2 * NRC], rw3 = ar[3 * NRC];
    // The entering row's sum over [xs+1, xt-1] within x-R .. x+R, in
    // ascending x. Tap q (column x-R+q) is in iff a < q < b with
    // a = xs + R - x and b = xt + R - x; the left taps (q < R-1) only need
    // the first test, the right ones (q > R+1) the second, and the centre
    // three are in (x+-1 weighted by cwl / cwr). All are integers, so the
    // weight sat(q - a) or sat(b - q) is exactly 0 or 1 (one subtraction
    // with the clamp modifier) and fma(0, v, s) = s, fma(1, v, s) = s + v:
    // the sum is the masked sum. Adding an outside-image 0 is exact.
    const bool hrow = inside && i < H;
    const float a = fmaxf(rh0 + xoff, rq0 + xoffr);
    const float b = fminf(rh1 + xoff, rq1 + xoffr);
    float hs = 0.f;
#pragma unroll
    for (int q = 0; q < R - 1; q++)
      hs = fmaf(sat01((float)q - a), rt[64 * q], hs);
    hs = fmaf(cwl, rt[64 * (R - 1)], hs);
    hs += rt[64 * R];
    hs = fmaf(cwr, rt[64 * (R + 1)], hs);
#pragma unroll
    for (int q = R + 2; q < K; q++)
      hs = fmaf(sat01(b - (float)q), rt[64 * q], hs);
    // pixels in the row's range: R-1 - max(a+1, 0) left, b - (R+2) right,
    // each clamped to [0, R-1]
    const float hc = __builtin_amdgcn_fmed3f((float)(R - 2) - a, 0.f, (float)(R - 1)) +
                     __builtin_amdgcn_fmed3f(b - (float)(R + 2), 0.f, (float)(R - 1)) + ccen;
#pragma unroll
    for (int q = 0; q < K - 1; q++) {
      ws[q] = ws[q + 1];
      wc[q] = wc[q + 1];
    }
    ws[K - 1] = hrow ? hs : 0.f;
    wc[K - 1] = hrow ? hc : 0.f;
    const int y = i - R;
    if (y >= ya) {
      // rows y-1, y+1 are in (or outside the image, where the window
      // holds zeros): only the outer rows need their one test, row q of
      // the window (row y-R+q) being in iff av < q < bv, as above
      const float yoff = (float)(R - y);
      const float av = fmaxf(rv2 + yoff, rw2 + yoff), bv = fminf(rv3 + yoff, rw3 + yoff);
      float sum = 0.f, cnt = 0.f;
#pragma unroll
      for (int q = 0; q < R - 1; q++) {
        const float m = sat01((float)q - av);
        sum = fmaf(m, ws[q], sum);
        cnt = fmaf(m, wc[q], cnt);
      }
#pragma unroll
      for (int q = R - 1; q <= R + 1; q++) {
        sum += ws[q];
        cnt += wc[q];
      }
#pragma unroll
      for (int q = R + 2; q < K; q++) {
        const float m = sat01(bv - (float)q);
        sum = fmaf(m, ws[q], sum);
        cnt = fmaf(m, wc[q], cnt);
      }
      // sum / cnt, correctly rounded for every quotient of at least
      // 2^-126 (one subnormal ulp below): the reciprocal rounded to
      // nearest, then one residual correction (Markstein); checked
      // exhaustively over all non-negative floats and cnt = 1 .. 127.
      // cnt >= 1 for inside voxels (the centre tap); others copy.
      const float rc1 = rtab[(int)cnt];
      const float q1 = sum * rc1;
      const float mean = fmaf(fmaf(-q1, cnt, sum), rc1, q1);
      const float res = inside ? mean : rcp[k];
      if (ok)
        buffer_stf(res, buffer_rsrc(reinterpret_cast<char*>(out) + (int64_t)y * rowD, rowD), ob, 0);
      if constexpr (WTA) {
        if (x < W) {  // uniform: the wavefront's column
          // ordered bits: non-negative floats order like their bits,
          // negative ones like their complement (+ 0.f turns a -0 into +0,
          // which compares equal to it as a float); lanes past D carry the
          // largest key. The minimum key is the float minimum, and the
          // lowest lane holding it the first disparity.
          const uint32_t u = __float_as_uint(res + 0.f);
          const uint32_t ou = d < D ? ((u & 0x80000000u) ? ~u : (u | 0x80000000u)) : ~0u;
          const uint32_t om = wave_min_u32_dpp(ou);
          const uint64_t hit = __ballot(ou == om);  // never 0: a lane holds om
          const uint32_t dw = (uint32_t)(d0 + __builtin_ctzll(hit));
          // uniform address and value: one lane's atomic (the compiler's
          // atomic optimizer elects it)
          atomicMin(&wkey[(int64_t)y * W + x], ((unsigned long long)om << 32) | dw);
        }
      }
    }
  };
  // main part: every prefetch is an iteration of the walk. Blocks of K
  // prefetch groups first: after K * PF rows the K-row window is back in
  // the registers it started in, so the unrolled block shifts it by
  // renaming rather than by register moves.
  int i0 = s0;
  for (; i0 + (K + 1) * PF <= n; i0 += K * PF) {
#pragma unroll
    for (int u = 0; u < K; u++) {
#pragma unroll
      for (int k = 0; k < PF; k++) {
        row_step(k, i0 + u * PF + k);
        load(k, i0 + u * PF + k + PF);
      }
    }
  }
  for (; i0 + 2 * PF <= n; i0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      row_step(k, i0 + k);
      load(k, i0 + k + PF);
    }
  }
  // tail (block-uniform bounds: every thread reaches the same barriers)
  for (; i0 < n; i0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      if (i0 + k < n) {
        row_step(k, i0 + k);
        load(k, i0 + k + PF);
      }
    }
  }
}

// Winner-take-all with the first minimum over d (test.py:409) and the
// subpixel parabola of post.cu:453-483, one wavefront per pixel (lane l
// holds disparities l*DPL .. l*DPL+DPL-1): the minimum by a DPP reduction,
// its first disparity by a ballot, the neighbours by readlane.
// WTA and subpixel maps from the per-pixel keys of the last aggregation pass
// (low word: the first disparity of minimum cost); the parabola is
// k_wta_wave's, from the three costs around the winner.
__global__ void k_wta_finish(const float* __restrict__ vol,
                             const unsigned long long* __restrict__ wkey, int64_t npix, int D,
                             float* __restrict__ wta, float* __restrict__ sub) {
  const int64_t pix = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (pix >= npix) return;
  const int bidx = (int)(uint32_t)wkey[pix];
  wta[pix] = (float)bidx;
  float o = (float)bidx;
  if (1 <= bidx && bidx < D - 1) {
    const float* c = vol + pix * D + bidx;
    const float cn = c[-1], cz = c[0], cp = c[1];
    const float denom = 2 * (cp + cn - 2 * cz);
    if ((double)denom > 1e-5) {
      const double r = fmin(1.0, fmax(-1.0, (double)((cp - cn) / denom)));
      o = (float)((double)bidx - r);
    }
  }
  sub[pix] = o;
}

// Fill kernel (valid under stream capture, unlike a memset node).
__global__ void k_fill_u64(unsigned long long* __restrict__ a, int64_t n, unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
    a[i] = v;
}

template <int DPL>
__global__ void __launch_bounds__(256) k_wta_wave(const float* __restrict__ vol, int64_t npix,
                                                  int D, float* __restrict__ wta,
                                                  float* __restrict__ sub) {
  const int lane = threadIdx.x & 63;
  const int64_t pix = (int64_t)__builtin_amdgcn_readfirstlane(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (pix >= npix) return;
  const float* p = vol + pix * D;
  float v[DPL];
#pragma unroll
  for (int j = 0; j < DPL; j++) {
    const int d = lane * DPL + j;
    v[j] = d < D ? p[d] : INFINITY;
  }
  float lm = v[0];
#pragma unroll
  for (int j = 1; j < DPL; j++) lm = fminf(lm, v[j]);
  const float m = wave_min_dpp(lm);
  int jf = DPL;
#pragma unroll
  for (int j = DPL - 1; j >= 0; j--) jf = v[j] == m ? j : jf;
  const uint64_t hit = __ballot(jf < DPL);
  const int l0 = (int)__builtin_ctzll(hit);
  const int bidx = l0 * DPL + __builtin_amdgcn_readlane(jf, l0);
  // cost at disparity t (uniform), read from the lane that holds it
  auto cost_at = [&](int t) {
    float c = v[0];
#pragma unroll
    for (int j = 1; j < DPL; j++) c = (t % DPL) == j ? v[j] : c;
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), t / DPL));
  };
  if (lane == 0) {
    wta[pix] = (float)bidx;
  }
  float o = (float)bidx;
  if (1 <= bidx && bidx < D - 1) {
    const float cn = cost_at(bidx - 1), cz = cost_at(bidx), cp = cost_at(bidx + 1);
    const float denom = 2 * (cp + cn - 2 * cz);
    if ((double)denom > 1e-5) {
      const double r = fmin(1.0, fmax(-1.0, (double)((cp - cn) / denom)));
      o = (float)((double)bidx - r);
    }
  }
  if (lane == 0) sub[pix] = o;
}


// Gaussian of Testing.__gaussian (test.py:108-118), including its (+1, +1) offset.
std::vector<float> gaussian_kernel(float sigma, int& ks) {
  const int kr = (int)std::ceil((double)sigma * 3);
  ks = 2 * kr + 1;
  std::vector<float> k((size_t)ks * ks);
  for (int i = 0; i < ks; i++)
    for (int j = 0; j < ks; j++) {
      const double yy = (i - 1) - kr, xx = (j - 1) - kr;
      k[(size_t)i * ks + j] =
          (float)std::exp(-(xx * xx + yy * yy) / (2.0 * (double)sigma * (double)sigma));
    }
  return k;
}

}  // namespace
}  // namespace cbmv

using namespace cbmv;

#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

struct cbmv_pipeline {
  int H = 0, W = 0, D = 0;
  int capH = 0, capW = 0, capD = 0;  // size the buffers were allocated for
  cbmv_params p{};
  int rows_per_batch = 1;
  // matchers
  DevBuf descL, descR, AL, AR, CL, CR, ML, MR, sobL, sobR;
  DevBuf census, ncc, sob, sad;
  DevBuf statsL, statsR;
  // the right view's minima per pixel over the row, merged by the matcher
  // kernels (codes or float64 keys, [4][H*W]); bit k of rmin_mask: matcher
  // k formed them in the last matcher stage
  DevBuf rmins;
  uint32_t rmin_mask = 0;
  DevBuf statsF;      // single-precision statistics for the MLP [16][H*W]
  // classifier
  DevBuf forest_nodes, forest_leaf0, forest_leaf1, forest_roots;
  // the forest's top levels for the LDS walk (forests of <= kTopMaxTrees
  // trees); CBMV_FOREST_TOP=0 at creation keeps the global-only walk
  DevBuf forest_top_nodes, forest_top_exits;
  bool forest_no_top = false;
  int forest_trees = 0;
  bool have_forest = false;
  MlpDevice mlp;
  bool have_mlp = false;
  DevBuf pmax;        // MLP best p(match) per pixel [H,W]
  DevBuf gate;        // per-pixel gate flags / list
  // post
  DevBuf imgL, imgR, imgRp, part, lcross, rcross;
  DevBuf cost, costR, volA, volB, hsum, hcnt;
  DevBuf* agg = nullptr;  // volA or volB: the last chain's final volume (fused aggregation)
  DevBuf wkey;            // per-pixel first-minimum keys of the last aggregation pass
  // A side stream (default priority) runs independent work concurrently,
  // forked and joined by events: two of the four matchers, and, with the LR
  // check and the fused aggregation, the right view's chain (with its own
  // ping-pong volumes and keys, volC / volD / wkeyR).
  DevBuf volC, volD, wkeyR;
  // paired SGM sweeps (8 paths, D <= 256): per view, the edge values passed
  // between workgroups, their minima and the sweep counters
  bool pairs = false;
  long long pair_spin = 0;
  DevBuf pairEdge[2], pairMn[2], pairCtr[2];
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  cbmv_pipeline() = default;
  cbmv_pipeline(const cbmv_pipeline&) = delete;
  cbmv_pipeline& operator=(const cbmv_pipeline&) = delete;
  ~cbmv_pipeline() {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
  }
  DevBuf wtaL, wtaR, subL, subR, outl, tmp1, tmp2, gauss;
  int gauss_k = 0;
};

static ForestView pipeline_forest(const cbmv_pipeline* pl) {
  return ForestView{pl->forest_nodes.as<GpuNode>(), pl->forest_leaf0.as<double>(),
                    pl->forest_leaf1.as<double>(), pl->forest_roots.as<int32_t>(),
                    pl->forest_trees, 20};
}

static FeatureSpec feature_spec() {
  // ratio epsilons and likelihood sigmas of test.py:71-73,548-556
  FeatureSpec fs{};
  fs.e[0] = 0.01;
  fs.e[1] = 1.01;
  fs.e[2] = 0.01;
  fs.e[3] = 0.01;
  fs.sigma[0] = 2.0 * 8 * 8;
  fs.sigma[1] = 0.02;
  fs.sigma[2] = 2.0 * 100 * 100;
  fs.sigma[3] = 2.0 * 100 * 100;
  return fs;
}

static Vols pipeline_vols(const cbmv_pipeline* pl) {
  return Vols{pl->census.as<uint8_t>(), pl->ncc.as<double>(), pl->sob.as<uint16_t>(),
              pl->sad.as<double>()};
}

static Stats pipeline_stats(const cbmv_pipeline* pl) {
  return Stats{pl->statsL.as<double>(), pl->statsR.as<double>()};
}

// stage 1
static int run_matchers(cbmv_pipeline* pl, const uint8_t* L, const uint8_t* R, hipStream_t s) {
  const int H = pl->H, W = pl->W, D = pl->D;
  const cbmv_params& p = pl->p;
  const VolStrides st = strides_hdw(H, W, D);
  const int64_t HW = (int64_t)H * W;
  // the right view's minima start at all ones (the matcher kernels merge
  // into them with atomic minima)
  unsigned long long* const rmins = pl->rmins.as<unsigned long long>();
  k_fill_u64<<<grid_1d(4 * HW, kB), kB, 0, s>>>(rmins, 4 * HW, ~0ull);
  CBMV_LAUNCH_CHECK();
  // NCC and zero-mean SAD on the side stream, census and SAD of Sobel here:
  // the four matchers share only their (read-only) input images
  hipStream_t s2 = pl->side ? pl->side : s;
  if (pl->side) {
    CBMV_HIP(hipEventRecord(pl->ev_fork, s));
    CBMV_HIP(hipStreamWaitEvent(s2, pl->ev_fork, 0));
  }
  RC(launch_window_stats(L, H, W, p.ncc_w, pl->AL.as<uint32_t>(), pl->CL.as<double>(), nullptr, s2));
  RC(launch_window_stats(R, H, W, p.ncc_w, pl->AR.as<uint32_t>(), pl->CR.as<double>(), nullptr, s2));
  // each matcher writes its left-view minima into row K of the left
  // statistics and its right-view minima into plane K of rmins
  double* const mins = pl->statsL.as<double>();
  bool done[4] = {false, false, false, false};
  RC(launch_ncc_cost(L, R, pl->AL.as<uint32_t>(), pl->AR.as<uint32_t>(), pl->CL.as<double>(),
                     pl->CR.as<double>(), H, W, D, p.ncc_w, pl->ncc.as<double>(), st, s2,
                     mins + 1 * HW, rmins + 1 * HW, &done[1]));
  RC(launch_window_stats(L, H, W, p.sad_w, nullptr, nullptr, pl->ML.as<double>(), s2));
  RC(launch_window_stats(R, H, W, p.sad_w, nullptr, nullptr, pl->MR.as<double>(), s2));
  RC(launch_zsad(L, R, pl->ML.as<double>(), pl->MR.as<double>(), H, W, D, p.sad_w,
                 pl->sad.as<double>(), st, s2, mins + 3 * HW, rmins + 3 * HW, &done[3]));
  RC(launch_census_desc(L, H, W, p.census_w, pl->descL.as<uint64_t>(), s));
  RC(launch_census_desc(R, H, W, p.census_w, pl->descR.as<uint64_t>(), s));
  RC(launch_census_cost_u8(pl->descL.as<uint64_t>(), pl->descR.as<uint64_t>(), H, W, D, p.census_w,
                           pl->census.as<uint8_t>(), st, s, mins, rmins, &done[0]));
  RC(launch_sobel_i16(L, H, W, pl->sobL.as<int16_t>(), s));
  RC(launch_sobel_i16(R, H, W, pl->sobR.as<int16_t>(), s));
  RC(launch_sadsob_u16(pl->sobL.as<int16_t>(), pl->sobR.as<int16_t>(), H, W, D, p.sob_w,
                       pl->sob.as<uint16_t>(), st, s, mins + 2 * HW, rmins + 2 * HW, &done[2]));
  pl->rmin_mask = 0;
  for (int k = 0; k < 4; k++) pl->rmin_mask |= done[k] ? 1u << k : 0u;
  if (pl->side) {
    CBMV_HIP(hipEventRecord(pl->ev_join, s2));
    CBMV_HIP(hipStreamWaitEvent(s, pl->ev_join, 0));
  }
  return CBMV_OK;
}

// stage 2
static int run_stats(cbmv_pipeline* pl, hipStream_t s) {
  const int H = pl->H, W = pl->W, D = pl->D;
  // k_stats addresses a volume row (D x W elements of up to 8 bytes, plus
  // the right view's shift) with 32-bit byte offsets
  if ((int64_t)D * (W + 1) * 8 > (int64_t)UINT32_MAX) {
    set_error("statistics: a volume row (W*D*8 bytes) must stay below 4 GiB");
    return CBMV_EINVAL;
  }
  const FeatureSpec fs = feature_spec();
  k_stats<<<kXcds * xcd_per((int64_t)((W + kB - 1) / kB) * H), kB, 0, s>>>(
      pipeline_vols(pl), H, W, D, pl->rows_per_batch, fs,
      pl->rmins.as<const unsigned long long>(), pl->rmin_mask, pl->statsL.as<double>(),
      pl->statsR.as<double>());
  const int64_t HW = (int64_t)H * W;
  k_stats_f32<<<grid_1d(HW, kB), kB, 0, s>>>(pl->statsL.as<double>(), pl->statsR.as<double>(), HW,
                                              pl->statsF.as<float>());
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

// stage 3
static int run_classifier(cbmv_pipeline* pl, hipStream_t s) {
  const int H = pl->H, W = pl->W, D = pl->D;
  const int cls = pl->p.classifier;
  if (cls == 0) {
    if (!pl->have_forest) {
      set_error("pipeline: no forest set (cbmv_pipeline_set_forest)");
      return CBMV_ESTATE;
    }
    dim3 grid((W + kTX - 1) / kTX, (D + kTD - 1) / kTD, H);
    if (pl->forest_top_exits.p && !pl->forest_no_top) {
      const ForestTop top{pl->forest_top_nodes.as<GpuNode>(), pl->forest_top_exits.as<uint32_t>()};
      k_features_rf_top<<<grid, 64 * kRfWaves, 0, s>>>(pipeline_vols(pl), pipeline_stats(pl),
                                                      feature_spec(), H, W, D, pl->rows_per_batch,
                                                      pipeline_forest(pl), top,
                                                      pl->cost.as<float>());
    } else {
      k_features_rf<<<grid, kB, 0, s>>>(pipeline_vols(pl), pipeline_stats(pl), feature_spec(), H,
                                        W, D, pl->rows_per_batch, pipeline_forest(pl),
                                        pl->cost.as<float>());
    }
    CBMV_LAUNCH_CHECK();
    return CBMV_OK;
  }
  if (!pl->have_mlp) {
    set_error("pipeline: no MLP set (cbmv_pipeline_set_mlp)");
    return CBMV_ESTATE;
  }
  if (cls == 2 && !pl->have_forest) {
    set_error("pipeline: the gated MLP needs a forest for its fallback");
    return CBMV_ESTATE;
  }
  MlpRun r{};
  r.vols = pipeline_vols(pl);
  r.stats = pipeline_stats(pl);
  r.statsf = StatsF{pl->statsF.as<float>()};
  r.fs = feature_spec();
  r.H = H;
  r.W = W;
  r.D = D;
  r.rows_per_batch = pl->rows_per_batch;
  r.cost = pl->cost.as<float>();
  r.pmax = pl->pmax.as<float>();
  r.gate = cls == 2;
  r.gate_threshold = pl->p.gate_threshold;
  r.gate_flags = pl->gate.as<int32_t>();
  r.forest = pipeline_forest(pl);
  return mlp_classify(pl->mlp, r, s);
}

// One fused aggregation pass of an [H,W,D] volume (arm length 6).
// wkey (per-pixel keys, all ones on entry) non-null: the pass also keeps
// each pixel's first minimum over d for k_wta_finish.
static int launch_cbca_fused(const float* in, const float* arms_a, const float* arms_b, int H,
                             int W, int D, int dir, int tile, float* out,
                             unsigned long long* wkey, hipStream_t s) {
  const unsigned gy = (unsigned)((D + 63) / 64);
  // 32-bit byte offsets within a volume row and across the four arm planes
  if ((int64_t)W * D * 4 > (int64_t)UINT32_MAX || (int64_t)H * W * 16 > (int64_t)UINT32_MAX) {
    set_error("aggregation: W*D*4 and H*W*16 bytes must stay below 4 GiB");
    return CBMV_EINVAL;
  }
  if (tile != 4 && tile != 16) tile = 8;
  // Row chunks: a column strip's walk is split into chunks of rows (each
  // repeats 2R = 10 rows of window fill) so that the workgroups fill the
  // chip in whole rounds; the chunk count minimises rounds x walk length.
  const unsigned gx = (unsigned)((W + tile - 1) / tile);
  const int64_t slots = (int64_t)std::max(device_cus(), 1) * (1024 / (tile * 64));
  int best = 1;
  int64_t best_cost = INT64_MAX;
  for (int nch = 1; nch <= 32 && nch <= H; nch++) {
    const int rows = (H + nch - 1) / nch;
    const int64_t rounds = ((int64_t)gx * gy * nch + slots - 1) / slots;
    const int64_t cost = rounds * (rows + 10);
    if (cost < best_cost) {
      best_cost = cost;
      best = nch;
    }
  }
  const int chunk = (H + best - 1) / best;
  const unsigned gz = (unsigned)((H + chunk - 1) / chunk);
  const unsigned nb = (unsigned)(kXcds * xcd_per((int64_t)gx * gy * gz));
#define CBMV_CBCA_GO(TX, NT, WTA)                                                     \
  k_cbca_fused<6, 3, TX, WTA><<<nb, NT, 0, s>>>(in, arms_a, arms_b, H, W, D, dir, chunk, out, \
                                                 wkey)
  if (wkey != nullptr) {
    switch (tile) {
      case 4: CBMV_CBCA_GO(4, 256, true); break;
      case 16: CBMV_CBCA_GO(16, 1024, true); break;
      default: CBMV_CBCA_GO(8, 512, true); break;
    }
  } else {
    switch (tile) {
      case 4: CBMV_CBCA_GO(4, 256, false); break;
      case 16: CBMV_CBCA_GO(16, 1024, false); break;
      default: CBMV_CBCA_GO(8, 512, false); break;
    }
  }
#undef CBMV_CBCA_GO
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

// WTA and subpixel maps of an [H,W,D] volume (k_wta_wave, lane width by D).
static int launch_wta_sub(const float* vol, int H, int W, int D, float* wta, float* sub,
                          hipStream_t s) {
  const int64_t np = (int64_t)H * W;
  const unsigned g = (unsigned)((np + 3) / 4);
  if (D <= 64)
    k_wta_wave<1><<<g, 256, 0, s>>>(vol, np, D, wta, sub);
  else if (D <= 128)
    k_wta_wave<2><<<g, 256, 0, s>>>(vol, np, D, wta, sub);
  else if (D <= 256)
    k_wta_wave<4><<<g, 256, 0, s>>>(vol, np, D, wta, sub);
  else if (D <= 512)
    k_wta_wave<8><<<g, 256, 0, s>>>(vol, np, D, wta, sub);
  else {
    set_error("WTA supports ndisp <= 512 (got %d)", D);
    return CBMV_EINVAL;
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

// SGM iterations on vol (in place through volB), then CBCA; returns the final
// [H,W,D] volume in *final (or WTA directly when fuse_wta).
// va, vb: the chain's ping-pong volumes; wk: its WTA keys; left: the chain
// whose final volume buffer 13 exposes.
static int run_chain(cbmv_pipeline* pl, const float* cost_in, const float* arms_a,
                     const float* arms_b, int dir, float* wta, float* sub, DevBuf& va, DevBuf& vb,
                     DevBuf& wk, bool left, int view, hipStream_t s) {
  const int H = pl->H, W = pl->W, D = pl->D;
  const cbmv_params& p = pl->p;
  const int64_t n = (int64_t)H * W * D;
  const VolStrides st = strides_hwd(H, W, D);
  const dim3 pgrid(W, H), pblock((D + 63) / 64 * 64);
  float* A = va.as<float>();
  float* B = vb.as<float>();
  const float* cur = cost_in;
  // cbca_i1 aggregations before SGM (test.py:182-184, 239-241)
  for (int i = 0; i < p.cbca_i1; i++) {
    float* dst = (cur == A) ? B : A;
    if (p.L1 == 6) {
      RC(launch_cbca_fused(cur, arms_a, arms_b, H, W, D, dir, p.cbca_tile, dst, nullptr, s));
    } else {
      k_cbca_h<<<pgrid, pblock, 0, s>>>(cur, arms_a, arms_b, H, W, D, p.L1, dir,
                                        pl->hsum.as<float>(), pl->hcnt.as<uint8_t>());
      k_cbca_v<false><<<pgrid, pblock, 0, s>>>(cur, pl->hsum.as<float>(), pl->hcnt.as<uint8_t>(),
                                               arms_a, arms_b, H, W, D, p.L1, dir, dst, nullptr,
                                               nullptr);
    }
    CBMV_LAUNCH_CHECK();
    cur = dst;
  }
  // SGM iterations: zeroed sum, paths, divide by the path count
  // (test.py:378-383); the first path writes instead of adding to zero and
  // the last one scales, so no separate clearing or dividing pass.
  const bool pairs = pl->pairs && p.sgm_paths == 8 && p.sgm_iters > 0;
  float* const pe = pl->pairEdge[view].as<float>();
  float* const pm = pl->pairMn[view].as<float>();
  unsigned* const pc = pl->pairCtr[view].as<unsigned>();
  if (pairs)  // refills the edge buffer only after a sweep that timed out
    RC(launch_sgm_pair_fill(pe, (int64_t)(pl->pairEdge[view].bytes / 4), pm,
                            (int64_t)(pl->pairMn[view].bytes / 4), pc, false, s));
  for (int it = 0; it < p.sgm_iters && pairs; it++) {
    // E + SE, W + NW, S + SW, N + NE: each sweep reads the volume and the
    // sum once for two paths
    float* dst = (cur == A) ? B : A;
    for (int pair = 0; pair < 4; pair++) {
      const int mode = (pair == 0 ? 1 : 0) | (pair == 3 ? 2 : 0);
      RC(launch_sgm_pair_f32(pl->imgL.as<float>(), pl->imgRp.as<float>(), cur, dst, H, W, D, p.pi1,
                             p.pi2, p.tau_so, p.alpha1, p.sgm_q1, p.sgm_q2, dir, pair, mode,
                             1.0 / p.sgm_paths, pe, pm, pc, pl->pair_spin, s));
    }
    cur = dst;
  }
  for (int it = 0; it < p.sgm_iters && !pairs; it++) {
    float* dst = (cur == A) ? B : A;
    for (int path = 0; path < p.sgm_paths; path++) {
      const int mode = (path == 0 ? 1 : 0) | (path == p.sgm_paths - 1 ? 2 : 0);
      RC(launch_sgm_path_f32(pl->imgL.as<float>(), pl->imgRp.as<float>(), cur, dst, H, W, D, st,
                             p.pi1, p.pi2, p.tau_so, p.alpha1, p.sgm_q1, p.sgm_q2, dir, path, mode,
                             1.0 / p.sgm_paths, s));
    }
    cur = dst;
  }
  (void)n;
  if (p.cbca_i2 == 0) {
    if (left && (cur == A || cur == B)) pl->agg = cur == A ? &va : &vb;
    return launch_wta_sub(cur, H, W, D, wta, sub, s);
  }
  const bool sliding = p.L1 == 6;
  for (int i = 0; i < p.cbca_i2; i++) {
    const bool last = i == p.cbca_i2 - 1;
    float* dst = (cur == A) ? B : A;
    if (sliding) {
      // one pass per iteration; the last one also keeps each pixel's first
      // minimum over d, and k_wta_finish turns it into the WTA and
      // subpixel maps (reading three costs per pixel, not the volume)
      unsigned long long* wkey = nullptr;
      if (last) {
        wkey = wk.as<unsigned long long>();
        const int64_t np = (int64_t)H * W;
        k_fill_u64<<<grid_1d(np, kB), kB, 0, s>>>(wkey, np, ~0ull);
        CBMV_LAUNCH_CHECK();
      }
      RC(launch_cbca_fused(cur, arms_a, arms_b, H, W, D, dir, p.cbca_tile, dst, wkey, s));
      cur = dst;
      if (last) {
        if (left) pl->agg = cur == A ? &va : &vb;
        const int64_t np = (int64_t)H * W;
        k_wta_finish<<<grid_1d(np, kB), kB, 0, s>>>(cur, wkey, np, D, wta, sub);
        CBMV_LAUNCH_CHECK();
      }
      continue;
    }
    k_cbca_h<<<pgrid, pblock, 0, s>>>(cur, arms_a, arms_b, H, W, D, p.L1, dir, pl->hsum.as<float>(),
                                      pl->hcnt.as<uint8_t>());
    if (last)
      k_cbca_v<true><<<pgrid, pblock, 0, s>>>(cur, pl->hsum.as<float>(), pl->hcnt.as<uint8_t>(),
                                              arms_a, arms_b, H, W, D, p.L1, dir, nullptr, wta, sub);
    else
      k_cbca_v<false><<<pgrid, pblock, 0, s>>>(cur, pl->hsum.as<float>(), pl->hcnt.as<uint8_t>(),
                                               arms_a, arms_b, H, W, D, p.L1, dir, dst, nullptr,
                                               nullptr);
    CBMV_LAUNCH_CHECK();
    cur = dst;
  }
  return CBMV_OK;
}

// stage 4
static int run_post(cbmv_pipeline* pl, const uint8_t* L, const uint8_t* R, float* disp,
                    hipStream_t s) {
  const int H = pl->H, W = pl->W, D = pl->D;
  const cbmv_params& p = pl->p;
  const int64_t np = (int64_t)H * W;
  RC(launch_normalise(L, R, np, pl->part.as<float>(), pl->imgL.as<float>(), pl->imgR.as<float>(), s));
  RC(launch_pad_rows(pl->imgR.as<float>(), H, W, pl->imgRp.as<float>(), s));
  RC(launch_cross(pl->imgL.as<float>(), H, W, p.L1, p.tau1, pl->lcross.as<float>(), s));
  RC(launch_cross(pl->imgR.as<float>(), H, W, p.L1, p.tau1, pl->rcross.as<float>(), s));
  const int direction = -1;  // do_post (test.py:499,501)
  // the right view's chain on the side stream when there is one (it shares
  // only read-only inputs with the left view's)
  const bool side = p.lr_check && pl->side != nullptr && pl->volC.p != nullptr;
  if (p.lr_check) {
    hipStream_t sr = side ? pl->side : s;
    if (side) {
      CBMV_HIP(hipEventRecord(pl->ev_fork, s));
      CBMV_HIP(hipStreamWaitEvent(sr, pl->ev_fork, 0));
    }
    const int tx = D <= 256 ? 64 : 32;
    k_right_cost<<<dim3((W + tx - 1) / tx, H), kRightCostThreads, (size_t)tx * D * 4, sr>>>(
        pl->cost.as<float>(), H, W, D, tx, pl->costR.as<float>());
    CBMV_LAUNCH_CHECK();
    RC(run_chain(pl, pl->costR.as<float>(), pl->rcross.as<float>(), pl->lcross.as<float>(),
                 -direction, pl->wtaR.as<float>(), pl->subR.as<float>(), side ? pl->volC : pl->volA,
                 side ? pl->volD : pl->volB, side ? pl->wkeyR : pl->wkey, false, 1, sr));
    if (side) CBMV_HIP(hipEventRecord(pl->ev_join, sr));
  }
  RC(run_chain(pl, pl->cost.as<float>(), pl->lcross.as<float>(), pl->rcross.as<float>(), direction,
               pl->wtaL.as<float>(), pl->subL.as<float>(), pl->volA, pl->volB, pl->wkey, true, 0, s));
  if (side) CBMV_HIP(hipStreamWaitEvent(s, pl->ev_join, 0));
  const float* cur;
  if (p.lr_check) {
    // the subpixel result is overwritten by interpolate_mismatch (test.py:293-300)
    RC(launch_outlier(pl->wtaL.as<float>(), pl->wtaR.as<float>(), H, W, D, pl->outl.as<float>(), s));
    RC(launch_interp_mismatch(pl->wtaL.as<float>(), pl->outl.as<float>(), H, W, pl->tmp1.as<float>(), s));
    RC(launch_interp_occlusion(pl->tmp1.as<float>(), pl->outl.as<float>(), H, W, pl->tmp2.as<float>(), s));
    cur = pl->tmp2.as<float>();
  } else {
    cur = pl->subL.as<float>();
  }
  for (int i = 0; i < p.median_iters; i++) {
    float* dst = (cur == pl->tmp1.as<float>()) ? pl->tmp2.as<float>() : pl->tmp1.as<float>();
    RC(launch_median2d(cur, H, W, p.median_w, dst, s));
    cur = dst;
  }
  RC(launch_mean2d(cur, H, W, pl->gauss.as<float>(), pl->gauss_k, p.blur_t, disp, s));
  return CBMV_OK;
}

extern "C" {

void cbmv_params_default(cbmv_params* p) {
  // Testing.__init__ defaults (test.py:44-47) and the SAD-of-Sobel window (test.py:131)
  p->census_w = 11;
  p->ncc_w = 3;
  p->sad_w = 5;
  p->sob_w = 5;
  p->pi1 = 1.3f;
  p->pi2 = 10.f;
  p->sgm_q1 = 2.2f;
  p->sgm_q2 = 3.f;
  p->alpha1 = 1.4f;
  p->tau_so = 0.12f;
  p->sgm_iters = 2;
  p->L1 = 6;
  p->tau1 = 0.07f;
  p->cbca_i1 = 0;
  p->cbca_i2 = 4;
  p->median_iters = 1;
  p->median_w = 3;
  p->blur_sigma = 2.2f;
  p->blur_t = 1.3f;
  p->sgm_paths = 4;
  p->lr_check = 0;
  p->test_batch = 100000;
  p->classifier = 0;
  p->gate_threshold = 0.5f;
  p->cbca_tile = 0;
}

cbmv_pipeline* cbmv_pipeline_create(int64_t H, int64_t W, int32_t D, const cbmv_params* params) {
  clear_error();
  cbmv_params p;
  if (params)
    p = *params;
  else
    cbmv_params_default(&p);
  auto fail = [](const char* msg) -> cbmv_pipeline* {
    set_error("pipeline: %s", msg);
    return nullptr;
  };
  if (H < 16 || W < 16 || H > 65535 || W > 65535) return fail("image size must be in [16, 65535]");
  if (D < 2 || D > 512) return fail("ndisp must be in [2, 512]");
  if (p.census_w < 1 || p.census_w > 15 || p.census_w * p.census_w > 254)
    return fail("census window must be in [1, 15]");
  if (p.sob_w < 1 || (int64_t)p.sob_w * p.sob_w * 2040 >= 65535)
    return fail("SAD-of-Sobel window must be in [1, 5]");
  if (p.ncc_w < 1 || p.ncc_w > 31 || p.sad_w < 1 || p.sad_w > 31) return fail("bad window size");
  if (p.sgm_paths != 4 && p.sgm_paths != 8) return fail("sgm_paths must be 4 or 8");
  if (p.L1 < 1 || p.L1 > 120) return fail("L1 must be in [1, 120]");
  if (p.median_w < 1 || p.median_w > 11) return fail("median window must be in [1, 11]");
  if (p.test_batch < W) return fail("test_batch must be >= W (rows per batch >= 1)");
  if (p.classifier < 0 || p.classifier > 2) return fail("classifier must be 0, 1 or 2");
  if (p.cbca_tile != 0 && p.cbca_tile != 4 && p.cbca_tile != 8 && p.cbca_tile != 16)
    return fail("cbca_tile must be 0, 4, 8 or 16");
  auto* pl = new cbmv_pipeline();
  pl->H = pl->capH = (int)H;
  pl->W = pl->capW = (int)W;
  pl->D = pl->capD = D;
  pl->p = p;
  pl->rows_per_batch = p.test_batch / (int)W;
  const size_t np = (size_t)H * W, nv = np * D;
  const int words = census_words(p.census_w);
  int rc = 0;
  rc |= pl->descL.alloc(np * words * 8);
  rc |= pl->descR.alloc(np * words * 8);
  rc |= pl->AL.alloc(np * 4);
  rc |= pl->AR.alloc(np * 4);
  rc |= pl->CL.alloc(np * 8);
  rc |= pl->CR.alloc(np * 8);
  rc |= pl->ML.alloc(np * 8);
  rc |= pl->MR.alloc(np * 8);
  rc |= pl->sobL.alloc(np * 2);
  rc |= pl->sobR.alloc(np * 2);
  rc |= pl->census.alloc(nv);
  rc |= pl->ncc.alloc(nv * 8);
  rc |= pl->sob.alloc(nv * 2);
  rc |= pl->sad.alloc(nv * 8);
  rc |= pl->statsL.alloc(np * 8 * 8);
  rc |= pl->rmins.alloc(np * 4 * 8);
  rc |= pl->statsR.alloc(np * 8 * 8);
  rc |= pl->statsF.alloc(np * 16 * 4);
  rc |= pl->imgL.alloc(np * 4);
  rc |= pl->imgR.alloc(np * 4);
  rc |= pl->imgRp.alloc((size_t)H * (W + 2 * sgm_pad_columns()) * 4);
  rc |= pl->part.alloc(normalise_scratch_bytes(np));
  rc |= pl->lcross.alloc(4 * np * 4);
  rc |= pl->rcross.alloc(4 * np * 4);
  rc |= pl->cost.alloc(nv * 4);
  if (p.lr_check) rc |= pl->costR.alloc(nv * 4);
  rc |= pl->volA.alloc(nv * 4);
  rc |= pl->volB.alloc(nv * 4);
  if (p.L1 != 6) {
    // row sums and counts of the two-pass aggregation (the fused pass for
    // L1 = 6 keeps them on chip)
    rc |= pl->hsum.alloc(nv * 4);
    rc |= pl->hcnt.alloc(nv);
  }
  if (p.L1 == 6) rc |= pl->wkey.alloc(np * 8);
  if (p.L1 == 6 && p.lr_check) {
    rc |= pl->volC.alloc(nv * 4);
    rc |= pl->volD.alloc(nv * 4);
    rc |= pl->wkeyR.alloc(np * 8);
  }
  // Paired sweeps for 8-path chains with D <= 256, by default when the two
  // views' chains run concurrently (LR check): there they take less time
  // than one sweep per path (half the memory traffic for the two chains
  // that share the device); a chain alone is faster with one sweep per path
  // (fewer instructions per voxel). CBMV_SGM_PAIRS=1 / 0 forces them on / off.
  const char* top_env = getenv("CBMV_FOREST_TOP");
  pl->forest_no_top = top_env && top_env[0] == '0';
  const char* pairs_env = getenv("CBMV_SGM_PAIRS");
  const bool two_chains = p.lr_check && p.L1 == 6;
  pl->pairs = p.sgm_paths == 8 && D <= 256 &&
              (pairs_env && pairs_env[0] ? pairs_env[0] == '1' : two_chains);
  if (pl->pairs) {
    for (int v = 0; v < (p.lr_check ? 2 : 1); v++) {
      rc |= pl->pairEdge[v].alloc((size_t)sgm_pair_edge_floats((int)H, (int)W, D) * 4);
      rc |= pl->pairMn[v].alloc((size_t)sgm_pair_edge_entries((int)H, (int)W, D) * 4);
      rc |= pl->pairCtr[v].alloc(16);
      if (!rc && (hipMemsetAsync(pl->pairCtr[v].p, 0, 16, lib_stream()) != hipSuccess ||
                  launch_sgm_pair_fill(pl->pairEdge[v].as<float>(),
                                       (int64_t)(pl->pairEdge[v].bytes / 4),
                                       pl->pairMn[v].as<float>(), (int64_t)(pl->pairMn[v].bytes / 4),
                                       pl->pairCtr[v].as<unsigned>(), true, lib_stream()))) {
        set_error("pipeline: SGM sweep buffers could not be initialised");
        rc = CBMV_EHIP;
      }
    }
    // a wait between sweep workgroups gives up after 0.5 s of wall clock
    int dev = 0, rate_khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        rate_khz <= 0)
      rate_khz = 100000;
    pl->pair_spin = (long long)rate_khz * 500;
  }
  if (!rc && (hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming) != hipSuccess)) {
    set_error("pipeline: side stream or events could not be created");
    rc = CBMV_EHIP;
  }
  rc |= pl->wtaL.alloc(np * 4);
  rc |= pl->wtaR.alloc(np * 4);
  rc |= pl->subL.alloc(np * 4);
  rc |= pl->subR.alloc(np * 4);
  rc |= pl->outl.alloc(np * 4);
  rc |= pl->tmp1.alloc(np * 4);
  rc |= pl->tmp2.alloc(np * 4);
  if (p.classifier != 0) {
    rc |= pl->pmax.alloc(np * 4);
    rc |= pl->gate.alloc(np * 4 + 64);
  }
  std::vector<float> g = gaussian_kernel(p.blur_sigma, pl->gauss_k);
  rc |= to_device(pl->gauss, g.data(), g.size() * 4);
  if (rc || sync()) {
    delete pl;
    return nullptr;
  }
  return pl;
}

int cbmv_pipeline_set_forest(cbmv_pipeline* pl, const cbmv_forest* f) {
  clear_error();
  CBMV_REQUIRE(pl && f, "null pipeline or forest");
  CBMV_REQUIRE(forest_n_features(f) == 20, "the pipeline's forest must take 20 features");
  const auto& nodes = forest_host_nodes(f);
  const auto& leaf0 = forest_host_leaf0(f);
  const auto& roots = forest_host_roots(f);
  std::vector<double> leaf1(leaf0.size(), 0.0);
  RC(to_device(pl->forest_nodes, nodes.data(), nodes.size() * sizeof(GpuNode)));
  RC(to_device(pl->forest_leaf0, leaf0.data(), leaf0.size() * 8));
  RC(to_device(pl->forest_leaf1, leaf1.data(), leaf1.size() * 8));
  RC(to_device(pl->forest_roots, roots.data(), roots.size() * 4));
  const int T = forest_n_trees(f);
  pl->forest_top_nodes.release();
  pl->forest_top_exits.release();
  if (T <= kTopMaxTrees) {
    // heap-ordered copies of each tree's first kTopLevels levels (a leaf with
    // its global index in the threshold's bits) and the global indices of
    // the positions one level below; positions under a leaf are never
    // reached (the walk stays on the leaf)
    std::vector<GpuNode> tn((size_t)T * kTopNodes, GpuNode{0.f, kLeafFeature});
    std::vector<uint32_t> tx((size_t)T * kTopExits, 0u);
    std::vector<uint32_t> g((size_t)kTopNodes + kTopExits);  // global index by heap position
    for (int t = 0; t < T; t++) {
      g[0] = (uint32_t)roots[t];
      for (int h = 0; h < kTopNodes; h++) {
        GpuNode nd = nodes[g[h]];
        if ((nd.meta & 31u) == kLeafFeature) {
          g[2 * h + 1] = g[2 * h + 2] = g[h];
          uint32_t gi = g[h];
          std::memcpy(&nd.thr, &gi, 4);
        } else {
          g[2 * h + 1] = g[h] + 1;
          g[2 * h + 2] = nd.meta >> 5;
        }
        tn[(size_t)t * kTopNodes + h] = nd;
      }
      for (int e = 0; e < kTopExits; e++) tx[(size_t)t * kTopExits + e] = g[kTopNodes + e];
    }
    RC(to_device(pl->forest_top_nodes, tn.data(), tn.size() * sizeof(GpuNode)));
    RC(to_device(pl->forest_top_exits, tx.data(), tx.size() * 4));
  }
  RC(sync());
  pl->forest_trees = forest_n_trees(f);
  pl->have_forest = true;
  return CBMV_OK;
}

int cbmv_pipeline_set_mlp(cbmv_pipeline* pl, const float* mean, const float* inv_std,
                          const float* w1, const float* b1, const float* w2, const float* b2,
                          const float* w3, const float* b3, int32_t activation) {
  clear_error();
  CBMV_REQUIRE(pl, "null pipeline");
  RC(mlp_upload(pl->mlp, mean, inv_std, w1, b1, w2, b2, w3, b3, activation));
  pl->have_mlp = true;
  return CBMV_OK;
}

int cbmv_pipeline_reshape(cbmv_pipeline* pl, int64_t H, int64_t W, int32_t D) {
  clear_error();
  CBMV_REQUIRE(pl, "null pipeline");
  CBMV_REQUIRE(H >= 16 && W >= 16 && D >= 2, "pair size %lldx%lld, ndisp %d too small",
               (long long)W, (long long)H, D);
  CBMV_REQUIRE(H <= pl->capH && W <= pl->capW && D <= pl->capD,
               "pair %lldx%lld, ndisp %d exceeds the pipeline's %dx%d, ndisp %d", (long long)W,
               (long long)H, D, pl->capW, pl->capH, pl->capD);
  CBMV_REQUIRE(pl->p.test_batch >= W, "test_batch must be >= W");
  pl->H = (int)H;
  pl->W = (int)W;
  pl->D = D;
  pl->rows_per_batch = pl->p.test_batch / (int)W;
  return CBMV_OK;
}

int cbmv_pipeline_set_cbca_tile(cbmv_pipeline* pl, int32_t tile) {
  clear_error();
  CBMV_REQUIRE(pl, "null pipeline");
  CBMV_REQUIRE(tile == 0 || tile == 4 || tile == 8 || tile == 16,
               "cbca_tile must be 0, 4, 8 or 16 (got %d)", tile);
  pl->p.cbca_tile = tile;
  return CBMV_OK;
}

int cbmv_pipeline_set_sgm_pairs(cbmv_pipeline* pl, int32_t on) {
  clear_error();
  CBMV_REQUIRE(pl, "null pipeline");
  CBMV_REQUIRE(on == 0 || on == 1, "sgm_pairs must be 0 or 1 (got %d)", on);
  CBMV_REQUIRE(!on || pl->pairEdge[0].p != nullptr,
               "paired SGM sweeps were not set up for this pipeline (they need 8 paths, "
               "ndisp <= 256 and the LR check, or CBMV_SGM_PAIRS=1 at creation)");
  pl->pairs = on != 0;
  return CBMV_OK;
}

int cbmv_pipeline_run(cbmv_pipeline* pl, const uint8_t* d_left, const uint8_t* d_right,
                      float* d_disp, void* stream) {
  clear_error();
  CBMV_REQUIRE(pl && d_left && d_right && d_disp, "null argument");
  hipStream_t s = (hipStream_t)stream;
  RC(run_matchers(pl, d_left, d_right, s));
  RC(run_stats(pl, s));
  RC(run_classifier(pl, s));
  RC(run_post(pl, d_left, d_right, d_disp, s));
  return CBMV_OK;
}

int cbmv_pipeline_run_stage(cbmv_pipeline* pl, int32_t stage, const uint8_t* d_left,
                            const uint8_t* d_right, float* d_disp, void* stream) {
  clear_error();
  CBMV_REQUIRE(pl, "null pipeline");
  hipStream_t s = (hipStream_t)stream;
  switch (stage) {
    case 0:
      CBMV_REQUIRE(d_left && d_right, "null image");
      return run_matchers(pl, d_left, d_right, s);
    case 1: return run_stats(pl, s);
    case 2: return run_classifier(pl, s);
    case 3:
      CBMV_REQUIRE(d_left && d_right && d_disp, "null argument");
      return run_post(pl, d_left, d_right, d_disp, s);
    default: set_error("unknown stage %d", stage); return CBMV_EINVAL;
  }
}

int cbmv_pipeline_buffer(cbmv_pipeline* pl, int32_t which, void** ptr, int64_t* nbytes) {
  clear_error();
  CBMV_REQUIRE(pl && ptr && nbytes, "null argument");
  DevBuf* b = nullptr;
  switch (which) {
    case 0: b = &pl->census; break;
    case 1: b = &pl->ncc; break;
    case 2: b = &pl->sob; break;
    case 3: b = &pl->sad; break;
    case 4: b = &pl->statsL; break;
    case 5: b = &pl->statsR; break;
    case 6: b = &pl->cost; break;
    case 7: b = &pl->wtaL; break;
    case 8: b = &pl->wtaR; break;
    case 9: b = &pl->outl; break;
    case 10: b = &pl->pmax; break;
    case 11: b = &pl->gate; break;
    case 12: b = &pl->costR; break;
    case 13:
      if (!pl->agg) {
        set_error("no aggregated volume: run a chain with the fused aggregation (L1 = 6) first");
        return CBMV_EINVAL;
      }
      b = pl->agg;
      break;
    case 14: b = &pl->imgL; break;
    case 15: b = &pl->imgR; break;
    default: set_error("unknown buffer %d", which); return CBMV_EINVAL;
  }
  *ptr = b->p;
  *nbytes = (int64_t)b->bytes;
  return CBMV_OK;
}

int cbmv_pipeline_sgm_timeouts(cbmv_pipeline* pl, int64_t* count) {
  clear_error();
  CBMV_REQUIRE(pl && count, "null argument");
  *count = 0;
  CBMV_HIP(hipDeviceSynchronize());
  for (int v = 0; v < 2; v++) {
    if (!pl->pairCtr[v].p) continue;
    unsigned c[4];
    CBMV_HIP(hipMemcpy(c, pl->pairCtr[v].p, sizeof(c), hipMemcpyDeviceToHost));
    *count += c[2];
  }
  return CBMV_OK;
}

int cbmv_pipeline_row_features(cbmv_pipeline* pl, int64_t y, float* d_out, void* stream) {
  clear_error();
  CBMV_REQUIRE(pl && d_out, "null argument");
  CBMV_REQUIRE(y >= 0 && y < pl->H, "row out of range");
  k_row_features<<<dim3((pl->W + kB - 1) / kB, pl->D), kB, 0, (hipStream_t)stream>>>(
      pipeline_vols(pl), pipeline_stats(pl), feature_spec(), pl->H, pl->W, pl->D,
      pl->rows_per_batch, (int)y, d_out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int cbmv_pipeline_sample_features(cbmv_pipeline* pl, const int32_t* d_samples, int64_t N,
                                  float* d_out, void* stream) {
  clear_error();
  CBMV_REQUIRE(pl && d_samples && d_out && N >= 0, "bad arguments");
  if (N == 0) return CBMV_OK;
  k_sample_features<<<grid_1d(N, kB), kB, 0, (hipStream_t)stream>>>(
      pipeline_vols(pl), StatsF{pl->statsF.as<float>()}, feature_spec(), pl->H, pl->W, pl->D,
      pl->rows_per_batch, d_samples, N, d_out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

void cbmv_pipeline_destroy(cbmv_pipeline* pl) {
  if (pl) (void)hipDeviceSynchronize();
  delete pl;
}

}  // extern "C"

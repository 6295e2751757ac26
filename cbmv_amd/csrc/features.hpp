// The 20-dimensional CBMV feature vector of one voxel (y, x, d), computed on
// the fly from the four matcher volumes and the per-pixel statistics.
//
// Order and formulas of Testing.__extract_features_lr (test.py:521-595):
//   [0..3]   raw cost of census, NCC, SAD-of-Sobel, zero-mean SAD
//   [4..7]   ratio      (m + e) / (c + e), 0 if m is the invalid value
//            (extract_pkrn_test, featextract.cpp:303-342)
//   [8..11]  likelihood exp(-((c - m)^2 / sigma)) / sum_k exp(-(c_k - m)^2 / sigma)
//            (extract_aml_testing, featextract.cpp:402-450)
//   [12..19] the same two features of the right view, mapped back to the left
//            pixel: at (y, x, d) they use the right view's statistics of pixel
//            x - d (the right view's cost there is this voxel's cost,
//            featextract.cpp:141); for x < d the reference fills with the
//            batch's first element (featextract.cpp:467), i.e. the right-view
//            feature of (row0, 0, 0) with row0 the first row of the
//            floor(100000 / W)-row batch (test.py:608).
// All arithmetic is float64 like the reference; the caller casts each feature
// to float32, which is what scikit-learn does before traversing trees.
#pragma once

#include "common.hpp"

namespace cbmv {

// Pipeline volumes, layout [H, D, W] (x fastest).
struct Vols {
  const uint8_t* census;  // 255 = invalid
  const double* ncc;      // RAND_MAX = invalid
  const uint16_t* sob;    // 65535 = invalid
  const double* sad;      // RAND_MAX = invalid
};

// Per-pixel statistics, structure of arrays: [k][H*W] with k = 0..3 the
// minimum of matcher k and k = 4..7 the sum of its exponentials.
struct Stats {
  const double* left;
  const double* right;
};

struct FeatureSpec {
  double e[4];      // ratio epsilon per matcher (test.py:548-551)
  double sigma[4];  // likelihood sigma per matcher (test.py:71-73,553-556)
};

__device__ __forceinline__ double vol_load(const Vols& v, int k, int64_t idx) {
  switch (k) {
    case 0: {
      const uint8_t c = v.census[idx];
      return c == kCensusInvalid ? kRandMax : (double)c;
    }
    case 1: return v.ncc[idx];
    case 2: {
      const uint16_t c = v.sob[idx];
      return c == kSobInvalid ? kRandMax : (double)c;
    }
    default: return v.sad[idx];
  }
}

__device__ __forceinline__ double ratio_feature(double m, double c, double e) {
  return (m == kRandMax) ? 0.0 : (m + e) / (c + e);
}

__device__ __forceinline__ double aml_feature(double m, double den, double c, double sigma) {
  const double t = c - m;
  return (m == kRandMax) ? 0.0 : exp(-((t * t) / sigma)) / den;
}

// Single-precision variant for the MLP, whose inputs are bf16 anyway:
// hardware reciprocal and exp2 (v_rcp_f32, v_exp_f32), 1/sigma * log2(e)
// folded into one constant.
__device__ __forceinline__ float ratio_feature_f(float m, float c, float e) {
  return (m == (float)kRandMax) ? 0.f : (m + e) * __builtin_amdgcn_rcpf(c + e);
}

__device__ __forceinline__ float aml_feature_f(float m, float rden, float c, float nl2s) {
  const float t = c - m;
  return (m == (float)kRandMax) ? 0.f : __builtin_amdgcn_exp2f((t * t) * nl2s) * rden;
}

// Single-precision per-pixel statistics for the MLP path, [16][H*W]:
// rows 0..3 the left minima, 4..7 the reciprocals of the left sums,
// 8..11 and 12..15 the same for the right view (k_stats_f32 derives them
// from the float64 statistics).
struct StatsF {
  const float* s;
};

// Constants of matcher k in single precision: ratio epsilon and
// -log2(e) / sigma.
struct MatcherConst {
  float e, nl2s;
};

__device__ __forceinline__ MatcherConst matcher_const(const FeatureSpec& fs, int k) {
  // selects rather than a runtime index into the by-value argument
  const double e = k == 0 ? fs.e[0] : k == 1 ? fs.e[1] : k == 2 ? fs.e[2] : fs.e[3];
  const double s = k == 0 ? fs.sigma[0] : k == 1 ? fs.sigma[1] : k == 2 ? fs.sigma[2] : fs.sigma[3];
  return MatcherConst{(float)e, (float)(-1.4426950408889634 / s)};
}

// Raw matcher values as loaded (no conversion yet), so that loads can be
// issued ahead of their use.
struct RawCost {
  uint32_t cen;    // census byte
  uint32_t sob;    // SAD-of-Sobel u16
  double ncc, sad;
};

__device__ __forceinline__ RawCost raw_load(const Vols& v, int64_t idx) {
  return RawCost{v.census[idx], v.sob[idx], v.ncc[idx], v.sad[idx]};
}

// Matcher k's cost from the raw values, invalid entries as RAND_MAX.
__device__ __forceinline__ float raw_cost_f(const RawCost& r, int k) {
  const float cen = r.cen == kCensusInvalid ? (float)kRandMax : (float)r.cen;
  const float sob = r.sob == kSobInvalid ? (float)kRandMax : (float)r.sob;
  return k == 0 ? cen : k == 1 ? (float)r.ncc : k == 2 ? sob : (float)r.sad;
}

// The five single-precision features of matcher k (raw cost, left ratio,
// left likelihood, right ratio, right likelihood = features k, 4+k, 8+k,
// 12+k, 16+k) from the voxel's cost c, the cost cr the right view sees
// (c itself when x >= d, else the batch's first element) and the four
// statistics.
__device__ __forceinline__ void matcher_features_f(float c, float cr, float mL, float rsL,
                                                   float mR, float rsR, MatcherConst mc,
                                                   float f[5]) {
  f[0] = c;
  f[1] = ratio_feature_f(mL, c, mc.e);
  f[2] = aml_feature_f(mL, rsL, c, mc.nl2s);
  f[3] = ratio_feature_f(mR, cr, mc.e);
  f[4] = aml_feature_f(mR, rsR, cr, mc.nl2s);
}

// All 20 single-precision features of voxel (y, x, d) into out(i, value),
// in the order of voxel_features; the MLP inference kernel computes the same
// values through matcher_features_f, one matcher per lane group.
template <typename Sink>
__device__ __forceinline__ void voxel_features_f32(const Vols& v, StatsF sf, const FeatureSpec& fs,
                                                   int H, int W, int D, int rows_per_batch, int y,
                                                   int x, int d, Sink out) {
  const int64_t HW = (int64_t)H * W;
  const int64_t pix = (int64_t)y * W + x;
  const bool back = x >= d;
  const int row0 = (y / rows_per_batch) * rows_per_batch;
  const int64_t rp = back ? pix - d : (int64_t)row0 * W;
  const RawCost r = raw_load(v, ((int64_t)y * D + d) * W + x);
  const RawCost r0 = raw_load(v, (int64_t)row0 * D * W);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float c = raw_cost_f(r, k);
    const float cr = back ? c : raw_cost_f(r0, k);
    float f[5];
    matcher_features_f(c, cr, sf.s[k * HW + pix], sf.s[(4 + k) * HW + pix], sf.s[(8 + k) * HW + rp],
                       sf.s[(12 + k) * HW + rp], matcher_const(fs, k), f);
#pragma unroll
    for (int s = 0; s < 5; s++) out(k + 4 * s, f[s]);
  }
}

// Computes the 20 features of voxel (y, x, d) as float32 into out(i, value).
template <typename Sink>
__device__ __forceinline__ void voxel_features(const Vols& v, const Stats& st,
                                               const FeatureSpec& fs, int H, int W, int D,
                                               int rows_per_batch, int y, int x, int d,
                                               Sink out) {
  const int64_t HW = (int64_t)H * W;
  const int64_t pix = (int64_t)y * W + x;
  const int64_t vidx = ((int64_t)y * D + d) * W + x;
  const bool back = x >= d;
  const int64_t rpix = pix - d;
  const int row0 = (y / rows_per_batch) * rows_per_batch;
  const int64_t fpix = (int64_t)row0 * W;           // (row0, x = 0)
  const int64_t fidx = (int64_t)row0 * D * W;       // volume element (row0, d = 0, x = 0)
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double c = vol_load(v, k, vidx);
    const double mL = st.left[k * HW + pix];
    const double sL = st.left[(4 + k) * HW + pix];
    out(k, (float)c);
    out(4 + k, (float)ratio_feature(mL, c, fs.e[k]));
    out(8 + k, (float)aml_feature(mL, sL, c, fs.sigma[k]));
    double r, a;
    if (back) {
      const double mR = st.right[k * HW + rpix];
      const double sR = st.right[(4 + k) * HW + rpix];
      r = ratio_feature(mR, c, fs.e[k]);
      a = aml_feature(mR, sR, c, fs.sigma[k]);
    } else {
      const double c0 = vol_load(v, k, fidx);
      const double mR = st.right[k * HW + fpix];
      const double sR = st.right[(4 + k) * HW + fpix];
      r = ratio_feature(mR, c0, fs.e[k]);
      a = aml_feature(mR, sR, c0, fs.sigma[k]);
    }
    out(12 + k, (float)r);
    out(16 + k, (float)a);
  }
}

}  // namespace cbmv

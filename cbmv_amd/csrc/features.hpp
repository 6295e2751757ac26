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

template <typename Sink>
__device__ __forceinline__ void voxel_features_f32(const Vols& v, const Stats& st,
                                                   const FeatureSpec& fs, int H, int W, int D,
                                                   int rows_per_batch, int y, int x, int d,
                                                   Sink out) {
  const int64_t HW = (int64_t)H * W;
  const int64_t pix = (int64_t)y * W + x;
  const int64_t vidx = ((int64_t)y * D + d) * W + x;
  const bool back = x >= d;
  const int row0 = (y / rows_per_batch) * rows_per_batch;
  const int64_t rp = back ? pix - d : (int64_t)row0 * W;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float c = (float)vol_load(v, k, vidx);
    const float e = (float)fs.e[k];
    const float nl2s = (float)(-1.4426950408889634 / fs.sigma[k]);
    const float mL = (float)st.left[k * HW + pix];
    const float sL = __builtin_amdgcn_rcpf((float)st.left[(4 + k) * HW + pix]);
    const float mR = (float)st.right[k * HW + rp];
    const float sR = __builtin_amdgcn_rcpf((float)st.right[(4 + k) * HW + rp]);
    const float cr = back ? c : (float)vol_load(v, k, (int64_t)row0 * D * W);
    out(k, c);
    out(4 + k, ratio_feature_f(mL, c, e));
    out(8 + k, aml_feature_f(mL, sL, c, nl2s));
    out(12 + k, ratio_feature_f(mR, cr, e));
    out(16 + k, aml_feature_f(mR, sR, cr, nl2s));
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

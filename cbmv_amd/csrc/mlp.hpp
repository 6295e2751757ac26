// BF16 MFMA multilayer perceptron (20 -> 128 -> 64 -> 1, GELU, sigmoid) used
// as the matching-probability function in place of the random forest.
#pragma once

#include "common.hpp"
#include "features.hpp"
#include "forest.hpp"

namespace cbmv {

constexpr int kMlpIn = 20, kMlpH1 = 128, kMlpH2 = 64;

struct MlpDevice {
  DevBuf params;  // packed bf16 weights + f32 biases / normalisation
  bool ready = false;
};

struct MlpRun {
  Vols vols;
  Stats stats;
  FeatureSpec fs;
  int H, W, D, rows_per_batch;
  float* cost;        // [H,W,D] 1 - p(match), or the forest's P(no match) for gated pixels
  float* prob_match;  // [H,W,D] p(match)
  bool gate;
  float gate_threshold;
  int32_t* gate_flags;  // [H*W] + counters
  ForestView forest;
};

int mlp_upload(MlpDevice& m, const float* mean, const float* inv_std, const float* w1,
               const float* b1, const float* w2, const float* b2, const float* w3,
               const float* b3);
int mlp_classify(const MlpDevice& m, const MlpRun& r, hipStream_t s);

}  // namespace cbmv

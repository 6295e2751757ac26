// BF16 MFMA multilayer perceptron (20 -> 128 -> 64 -> 1, GELU or ReLU
// hidden units, sigmoid output) used as the matching-probability function in
// place of the random forest. Inputs are the 20 CBMV features standardised as
// z = clamp((x - mean) * inv_std, -8, 8).
#pragma once

#include "common.hpp"
#include "features.hpp"
#include "forest.hpp"

namespace cbmv {

constexpr int kMlpIn = 20, kMlpInPad = 32, kMlpH1 = 128, kMlpH2 = 64;
// hidden activations (values of the C-ABI's activation argument)
constexpr int kActGelu = 0, kActRelu = 1;

// Device parameter block (one allocation), see mlp.hip for the layout.
struct MlpDevice {
  DevBuf params;
  int act = kActGelu;
  bool ready = false;
};

struct MlpRun {
  Vols vols;
  Stats stats;     // float64 statistics (forest fallback)
  StatsF statsf;   // single-precision statistics (MLP features)
  FeatureSpec fs;
  int H, W, D, rows_per_batch;
  float* cost;          // [H,W,D] 1 - p(match); gated pixels get the forest's P(no match)
  float* pmax;          // [H,W] best p(match) of each pixel (filled when gating)
  bool gate;
  float gate_threshold;
  int32_t* gate_flags;  // [H*W + 16]: per-pixel list, count at the end
  ForestView forest;
};

int mlp_upload(MlpDevice& m, const float* mean, const float* inv_std, const float* w1,
               const float* b1, const float* w2, const float* b2, const float* w3,
               const float* b3, int act);
int mlp_classify(const MlpDevice& m, const MlpRun& r, hipStream_t s);

}  // namespace cbmv

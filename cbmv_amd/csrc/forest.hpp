// Flat random-forest format for the GPU traversal kernels.
//
// Each tree is stored in preorder, so the left child of node n is n + 1.
// A node is 8 bytes: the split threshold as float32 rounded toward -inf
// (for a float32 feature x, x <= thr32 is exactly sklearn's
// (double)x <= thr64), and a packed word: bits 0..4 the feature index
// (31 marks a leaf), bits 5..31 the absolute index of the right child, or
// for a leaf the index of its class-0 fraction in `leaf0` (float64, as
// sklearn's tree_.value).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cbmv {

constexpr uint32_t kLeafFeature = 31;

struct GpuNode {
  float thr;
  uint32_t meta;
};

struct ForestView {
  const GpuNode* nodes;
  const double* leaf0;  // class-0 fraction per leaf
  const double* leaf1;  // class-1 fraction per leaf
  const int32_t* roots;
  int32_t n_trees;
  int32_t n_features;
};

// Sum over trees of the class-0 leaf fraction reached by one sample whose
// features are read through `feat(i)`; the caller divides by n_trees.
template <typename FeatFn>
__device__ __forceinline__ double forest_sum0(const ForestView& f, FeatFn feat) {
  double acc = 0.0;
  for (int t = 0; t < f.n_trees; t++) {
    uint32_t n = (uint32_t)f.roots[t];
    while (true) {
      const GpuNode nd = f.nodes[n];
      const uint32_t fi = nd.meta & 31u;
      if (fi == kLeafFeature) {
        acc += f.leaf0[nd.meta >> 5];
        break;
      }
      n = (feat(fi) <= nd.thr) ? n + 1 : (nd.meta >> 5);
    }
  }
  return acc;
}

// Same sum with G trees walked together: each iteration issues the G node
// loads back to back, so their latencies overlap instead of adding up. The
// leaf fractions are still added in tree order, like scikit-learn's
// sequential accumulation.
template <int G, typename FeatFn>
__device__ __forceinline__ double forest_sum0_ilp(const ForestView& f, FeatFn feat) {
  double acc = 0.0;
  for (int t0 = 0; t0 < f.n_trees; t0 += G) {
    uint32_t n[G], leaf[G];
    bool live[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      live[g] = t0 + g < f.n_trees;
      n[g] = live[g] ? (uint32_t)f.roots[t0 + g] : 0u;
      leaf[g] = 0u;
    }
    bool any = true;
    while (any) {
      GpuNode nd[G];
#pragma unroll
      for (int g = 0; g < G; g++) nd[g] = f.nodes[n[g]];
      any = false;
#pragma unroll
      for (int g = 0; g < G; g++) {
        if (!live[g]) continue;
        const uint32_t fi = nd[g].meta & 31u;
        if (fi == kLeafFeature) {
          live[g] = false;
          leaf[g] = nd[g].meta >> 5;
        } else {
          n[g] = (feat(fi) <= nd[g].thr) ? n[g] + 1 : (nd[g].meta >> 5);
          any = true;
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; g++)
      if (t0 + g < f.n_trees) acc += f.leaf0[leaf[g]];
  }
  return acc;
}

}  // namespace cbmv

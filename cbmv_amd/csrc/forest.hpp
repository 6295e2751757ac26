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

// Sums over trees of the leaf class fractions reached by one sample whose
// features are read through `feat(i)` (class 0 into acc0 and, with BOTH,
// class 1 into acc1); the caller divides by n_trees. G trees are walked
// together and branch-free: every lane advances each of them by a select (a
// lane on a leaf stays there) and the walk repeats while any lane of the
// wavefront is inside one of them, so the G node loads of a step overlap and
// no per-lane control flow (exec-mask bookkeeping on the scalar unit) is
// needed. The leaf fractions are added in tree order, like scikit-learn's
// sequential accumulation.
template <int G, bool BOTH, typename FeatFn>
__device__ __forceinline__ void forest_sums(const ForestView& f, FeatFn feat, double& acc0,
                                            double& acc1) {
  acc0 = 0.0;
  acc1 = 0.0;
  for (int t0 = 0; t0 < f.n_trees; t0 += G) {
    uint32_t n[G], meta[G];
#pragma unroll
    for (int g = 0; g < G; g++) n[g] = (uint32_t)f.roots[min(t0 + g, f.n_trees - 1)];
    while (true) {
      GpuNode nd[G];
#pragma unroll
      for (int g = 0; g < G; g++) nd[g] = f.nodes[n[g]];
      bool inner = false;
#pragma unroll
      for (int g = 0; g < G; g++) {
        const uint32_t fi = nd[g].meta & 31u;
        const bool leaf = fi == kLeafFeature;
        const float v = feat(leaf ? 0u : fi);
        const uint32_t nx = v <= nd[g].thr ? n[g] + 1 : (nd[g].meta >> 5);
        n[g] = leaf ? n[g] : nx;
        meta[g] = nd[g].meta;
        inner |= !leaf;
      }
      if (!__any(inner)) break;
    }
#pragma unroll
    for (int g = 0; g < G; g++)
      if (t0 + g < f.n_trees) {
        acc0 += f.leaf0[meta[g] >> 5];
        if (BOTH) acc1 += f.leaf1[meta[g] >> 5];
      }
  }
}

// The first kTopLevels levels of every tree of a forest of at most
// kTopMaxTrees trees, in heap order (children of heap position i at 2i + 1
// and 2i + 2), for a copy in LDS: `nodes` holds the entries of depths
// 0 .. kTopLevels - 1 (split nodes as in the global array; a leaf with its
// global index in place of the threshold, which a leaf does not use), `exits`
// the global index of the node at every heap position of depth kTopLevels,
// where the walk continues in global memory.
constexpr int kTopLevels = 7;
constexpr int kTopNodes = (1 << kTopLevels) - 1;
constexpr int kTopExits = 1 << kTopLevels;
constexpr int kTopMaxTrees = 50;

struct ForestTop {
  const GpuNode* nodes;   // [n_trees][kTopNodes]
  const uint32_t* exits;  // [n_trees][kTopExits]
};

// forest_sums' class-0 sum, each tree's first kTopLevels steps taken on the
// LDS copies (lnodes, lexits: ForestTop's arrays) and the rest in global
// memory; the same leaves, added in the same order. A lane that reaches a
// leaf above the last cached level stays on it (its last LDS read is that
// leaf, whose threshold bits give its global index), and its global walk
// ends at the first step.
template <int G, typename FeatFn>
__device__ __forceinline__ double forest_sum0_top(const ForestView& f, const GpuNode* lnodes,
                                                  const uint32_t* lexits, FeatFn feat) {
  double acc0 = 0.0;
  for (int t0 = 0; t0 < f.n_trees; t0 += G) {
    uint32_t n[G], meta[G], h[G], lthr[G];
    int tt[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      tt[g] = min(t0 + g, f.n_trees - 1);
      h[g] = 0;
      lthr[g] = 0;
    }
#pragma unroll
    for (int k = 0; k < kTopLevels; k++) {
#pragma unroll
      for (int g = 0; g < G; g++) {
        const GpuNode nd = lnodes[tt[g] * kTopNodes + h[g]];
        const uint32_t fi = nd.meta & 31u;
        const bool leaf = fi == kLeafFeature;
        const float v = feat(leaf ? 0u : fi);
        const uint32_t c = 2 * h[g] + (v <= nd.thr ? 1u : 2u);
        h[g] = leaf ? h[g] : c;
        lthr[g] = __float_as_uint(nd.thr);
      }
    }
#pragma unroll
    for (int g = 0; g < G; g++)
      n[g] = h[g] < (uint32_t)kTopNodes ? lthr[g]
                                        : lexits[tt[g] * kTopExits + (h[g] - kTopNodes)];
    while (true) {
      GpuNode nd[G];
#pragma unroll
      for (int g = 0; g < G; g++) nd[g] = f.nodes[n[g]];
      bool inner = false;
#pragma unroll
      for (int g = 0; g < G; g++) {
        const uint32_t fi = nd[g].meta & 31u;
        const bool leaf = fi == kLeafFeature;
        const float v = feat(leaf ? 0u : fi);
        const uint32_t nx = v <= nd[g].thr ? n[g] + 1 : (nd[g].meta >> 5);
        n[g] = leaf ? n[g] : nx;
        meta[g] = nd[g].meta;
        inner |= !leaf;
      }
      if (!__any(inner)) break;
    }
#pragma unroll
    for (int g = 0; g < G; g++)
      if (t0 + g < f.n_trees) acc0 += f.leaf0[meta[g] >> 5];
  }
  return acc0;
}

// Class-0 sum only (the cost volume's P(non-match) before the division).
template <int G, typename FeatFn>
__device__ __forceinline__ double forest_sum0(const ForestView& f, FeatFn feat) {
  double a0, a1;
  forest_sums<G, false>(f, feat, a0, a1);
  return a0;
}

}  // namespace cbmv

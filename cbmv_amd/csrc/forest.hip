// Random forest: conversion of sklearn-style node arrays to the flat
// preorder GPU format, and predict_proba on host features (the drop-in for
// RandomForestClassifier.predict_proba, test.py:620).
#include <cmath>
#include <vector>

#include "forest.hpp"
#include "internal.hpp"

namespace cbmv {
namespace {

// One sample per thread; the sample's features sit in LDS, column-major by
// thread, so the data-dependent feature reads are bank-conflict free.
constexpr int kFB = 128;
constexpr int kMaxFeatures = 31;

__global__ void __launch_bounds__(kFB) k_forest_predict(ForestView f, const float* __restrict__ X,
                                                        int64_t N, double* __restrict__ proba) {
  __shared__ float xs[kMaxFeatures][kFB];
  const int64_t i = blockIdx.x * (int64_t)kFB + threadIdx.x;
  if (i < N)
    for (int k = 0; k < f.n_features; k++) xs[k][threadIdx.x] = X[i * f.n_features + k];
  if (i >= N) return;
  const int tid = threadIdx.x;
  double acc0, acc1;
  forest_sums<2, true>(f, [&](uint32_t k) { return xs[k][tid]; }, acc0, acc1);
  // sklearn sums the per-tree probabilities and divides by the tree count
  proba[2 * i] = acc0 / f.n_trees;
  proba[2 * i + 1] = acc1 / f.n_trees;
}

}  // namespace
}  // namespace cbmv

using namespace cbmv;

struct cbmv_forest {
  int32_t n_trees = 0, n_features = 0;
  int64_t n_nodes = 0, n_leaves = 0;
  DevBuf nodes, leaf0, leaf1, roots;
  // host copies, so a pipeline can take its own device copy
  std::vector<GpuNode> h_nodes;
  std::vector<double> h_leaf0, h_leaf1;
  std::vector<int32_t> h_roots;
};

namespace cbmv {
ForestView forest_view(const cbmv_forest* f) {
  return ForestView{f->nodes.as<GpuNode>(), f->leaf0.as<double>(), f->leaf1.as<double>(),
                    f->roots.as<int32_t>(), f->n_trees, f->n_features};
}
const std::vector<GpuNode>& forest_host_nodes(const cbmv_forest* f) { return f->h_nodes; }
const std::vector<double>& forest_host_leaf0(const cbmv_forest* f) { return f->h_leaf0; }
const std::vector<int32_t>& forest_host_roots(const cbmv_forest* f) { return f->h_roots; }
int forest_n_trees(const cbmv_forest* f) { return f->n_trees; }
int forest_n_features(const cbmv_forest* f) { return f->n_features; }
}  // namespace cbmv

#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

static float round_down_f32(double t) {
  float f = (float)t;
  if ((double)f > t) f = std::nextafter(f, -INFINITY);
  return f;
}

extern "C" {

cbmv_forest* cbmv_forest_create(int32_t n_trees, const int32_t* roots, int64_t n_nodes,
                                const int32_t* feature, const double* threshold,
                                const int32_t* left, const int32_t* right, const double* value,
                                int32_t n_features) {
  clear_error();
  if (n_trees <= 0 || n_nodes <= 0 || n_features <= 0 || n_features > kMaxFeatures ||
      n_nodes >= (1ll << 27)) {
    set_error("forest: need 1..%d features, >= 1 tree and < 2^27 nodes", kMaxFeatures);
    return nullptr;
  }
  auto* f = new cbmv_forest();
  f->n_trees = n_trees;
  f->n_features = n_features;
  f->h_nodes.reserve(n_nodes);
  // preorder re-layout of every tree; explicit stack of (old index, parent slot)
  for (int t = 0; t < n_trees; t++) {
    const int32_t r = roots[t];
    if (r < 0 || r >= n_nodes) {
      set_error("forest: root %d of tree %d out of range", r, t);
      delete f;
      return nullptr;
    }
    f->h_roots.push_back((int32_t)f->h_nodes.size());
    std::vector<std::pair<int32_t, int64_t>> stack;  // (old node, new index of parent to patch or -1)
    stack.push_back({r, -1});
    int64_t guard = 0;
    while (!stack.empty()) {
      auto [o, patch] = stack.back();
      stack.pop_back();
      if (o < 0 || o >= n_nodes || ++guard > n_nodes) {
        set_error("forest: malformed tree %d", t);
        delete f;
        return nullptr;
      }
      const int64_t ni = (int64_t)f->h_nodes.size();
      if (patch >= 0) f->h_nodes[patch].meta |= (uint32_t)ni << 5;  // right child link
      GpuNode nd{};
      if (left[o] < 0) {
        nd.thr = 0.f;
        nd.meta = kLeafFeature | ((uint32_t)f->h_leaf0.size() << 5);
        f->h_leaf0.push_back(value[2 * o]);
        f->h_leaf1.push_back(value[2 * o + 1]);
        f->h_nodes.push_back(nd);
      } else {
        if (feature[o] < 0 || feature[o] >= n_features || right[o] < 0) {
          set_error("forest: bad split node %d of tree %d", o, t);
          delete f;
          return nullptr;
        }
        nd.thr = round_down_f32(threshold[o]);
        nd.meta = (uint32_t)feature[o];
        f->h_nodes.push_back(nd);
        // right subtree is emitted after the whole left subtree: push it first
        stack.push_back({right[o], ni});
        stack.push_back({left[o], -1});
      }
    }
  }
  f->n_nodes = (int64_t)f->h_nodes.size();
  f->n_leaves = (int64_t)f->h_leaf0.size();
  if (to_device(f->nodes, f->h_nodes.data(), f->h_nodes.size() * sizeof(GpuNode)) ||
      to_device(f->leaf0, f->h_leaf0.data(), f->h_leaf0.size() * 8) ||
      to_device(f->leaf1, f->h_leaf1.data(), f->h_leaf1.size() * 8) ||
      to_device(f->roots, f->h_roots.data(), f->h_roots.size() * 4) || sync()) {
    delete f;
    return nullptr;
  }
  return f;
}

int cbmv_forest_n_trees(const cbmv_forest* f) { return f ? f->n_trees : CBMV_EINVAL; }

int cbmv_forest_predict_proba(const cbmv_forest* f, const float* X, int64_t N, double* proba) {
  clear_error();
  CBMV_REQUIRE(f != nullptr, "null forest");
  CBMV_REQUIRE(N >= 0, "negative sample count");
  if (N == 0) return CBMV_OK;
  hipStream_t s = lib_stream();
  DevBuf dx, dp;
  RC(to_device(dx, X, (size_t)N * f->n_features * 4));
  RC(dp.alloc((size_t)N * 16));
  k_forest_predict<<<grid_1d(N, kFB), kFB, 0, s>>>(forest_view(f), dx.as<float>(), N, dp.as<double>());
  CBMV_LAUNCH_CHECK();
  return to_host(proba, dp, (size_t)N * 16);
}

void cbmv_forest_destroy(cbmv_forest* f) {
  if (f) (void)hipStreamSynchronize(lib_stream());
  delete f;
}

}  // extern "C"

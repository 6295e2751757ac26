// Random-forest trainer on the host: CART trees with Gini impurity and exact
// best splits, bootstrap samples and a random subset of candidate features
// per node, the classifier the reference fits with scikit-learn
// (RandomForestClassifier(min_samples_leaf=700, n_estimators=50),
// train.py:287-289), for hosts where scikit-learn is not installed. The
// output is the flat node format that cbmv_forest_create takes.
//
// Split rules follow scikit-learn's best splitter so that one tree grown on
// all features without bootstrap equals sklearn's DecisionTreeClassifier:
// * features are float32; a feature whose values in a node span no more
//   than 1e-7 is constant there;
// * candidate positions lie between consecutive distinct sorted values;
//   both children must keep min_samples_leaf samples (rows of non-zero
//   weight); the best position maximises -W_L gini_L - W_R gini_R (the
//   Gini proxy of the children), the first one found wins ties;
// * the threshold is the midpoint a/2 + b/2 of the two values, or a when
//   the midpoint rounds to b;
// * a node is a leaf when it is at max_depth, holds fewer than
//   2 min_samples_leaf (or 2) samples, is pure, or no candidate feature has
//   a valid split;
// * candidate features are drawn as scikit-learn's splitter draws them:
//   Fisher-Yates over a feature array kept from node to node, its xorshift
//   stream (values below 2^31), max_features draws (more while every one
//   was constant), features found constant in an ancestor skipped;
// * nodes are numbered depth first, left child first; a node's value is its
//   weighted class fractions.
// Bootstrap draws n rows with replacement (weights = counts) from a
// SplitMix64 stream per tree (seed, tree index), which also seeds the tree's
// splitter stream (a single tree may set that state directly:
// DecisionTreeClassifier(random_state=r) uses
// numpy.random.RandomState(r).randint(0, 2**31 - 1)). Trees are grown in
// parallel on std::threads; the forest depends on the seed only.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/cbmv_abi.h"

namespace cbmv {
void set_error(const char* fmt, ...);
void clear_error();
}  // namespace cbmv

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, n) (n >= 1; Lemire's multiply-shift, bias below 2^-32)
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

struct Tree {
  std::vector<int32_t> feature, left, right;
  std::vector<double> threshold, value;  // value: [n, 2]
};

struct Grower {
  const float* X;
  const int32_t* y;
  int64_t n;
  int nf;
  cbmv_cart_params p;
  // per tree
  std::vector<int64_t> samples;  // rows of non-zero weight, partitioned per node
  std::vector<double> w;         // row weights (bootstrap counts)
  std::vector<int> features, const_features;
  Tree t;

  struct Split {
    int f = -1;
    int64_t pos = 0;  // split index into samples[start, end)
    double thr = 0, proxy = -INFINITY;
  };

  void leaf_value(int64_t s, int64_t e, double& c0, double& c1) const {
    c0 = c1 = 0;
    for (int64_t i = s; i < e; i++) {
      const int64_t r = samples[i];
      (y[r] ? c1 : c0) += w[r];
    }
  }

  // scikit-learn's splitter random stream (xorshift, values below 2^31)
  uint32_t rstate = 1;
  uint32_t rand_r() {
    rstate ^= rstate << 13;
    rstate ^= rstate >> 17;
    rstate ^= rstate << 5;
    return rstate & 0x7FFFFFFFu;
  }
  int rand_int(int lo, int hi) { return lo + (int)(rand_r() % (uint32_t)(hi - lo)); }

  // Best split of samples[s, e) over up to max_features candidate features
  // drawn in scikit-learn's order: Fisher-Yates over the persistent
  // `features` array, whose first `nconst` entries are features found
  // constant in an ancestor (skipped without sorting); features found
  // constant here join them. On success samples[s, e) holds the node's rows
  // sorted by the winning feature, so the left child is its first `pos`
  // rows. `nconst` becomes the children's count of known constants.
  Split best_split(int64_t s, int64_t e, double tot0, double tot1, int& nconst) {
    Split best;
    const int64_t m = e - s;
    const int64_t msl = std::max<int64_t>(p.min_samples_leaf, 1);
    std::vector<int64_t> idx(m), tmp(m);
    std::vector<float> v(m);
    std::vector<std::pair<float, int64_t>> kv(m);
    int f_i = nf, visited = 0, found = 0, drawn_const = 0;
    const int known = nconst;
    int total = known;
    while (f_i > total && (visited < p.max_features || visited <= found + drawn_const)) {
      visited++;
      int f_j = rand_int(drawn_const, f_i - found);
      if (f_j < known) {  // a known constant: drawn, not evaluated
        std::swap(features[drawn_const], features[f_j]);
        drawn_const++;
        continue;
      }
      f_j += found;
      const int f = features[f_j];
      for (int64_t i = 0; i < m; i++) {
        const int64_t r = samples[s + i];
        kv[i] = {X[r * nf + f], r};
      }
      std::sort(kv.begin(), kv.end());
      for (int64_t i = 0; i < m; i++) {
        v[i] = kv[i].first;
        idx[i] = kv[i].second;
      }
      if (v[m - 1] <= v[0] + 1e-7f) {  // constant in this node
        std::swap(features[f_j], features[total]);
        found++;
        total++;
        continue;
      }
      f_i--;
      std::swap(features[f_i], features[f_j]);
      double l0 = 0, l1 = 0;
      int64_t p0 = 0;  // rows already on the left
      bool improved = false;
      for (int64_t q = 0; q + 1 < m;) {
        // advance to the last row of the run of values within 1e-7 of v[q]
        int64_t r = q;
        while (r + 1 < m && v[r + 1] <= v[r] + 1e-7f) r++;
        const int64_t pos = r + 1;  // first row of the right child
        for (; p0 < pos; p0++) (y[idx[p0]] ? l1 : l0) += w[idx[p0]];
        q = pos;
        if (pos >= m) break;
        if (pos < msl || m - pos < msl) continue;
        const double wl = l0 + l1, wr = (tot0 - l0) + (tot1 - l1);
        if (wl <= 0 || wr <= 0) continue;
        const double r0 = tot0 - l0, r1 = tot1 - l1;
        // -W_R gini_R - W_L gini_L, the children's Gini impurities as
        // 1 - sum_c n_c^2 / W^2
        const double gl = 1.0 - (l0 * l0 + l1 * l1) / (wl * wl);
        const double gr = 1.0 - (r0 * r0 + r1 * r1) / (wr * wr);
        const double proxy = -wr * gr - wl * gl;
        if (proxy > best.proxy) {
          best.proxy = proxy;
          best.f = f;
          best.pos = pos;
          double thr = (double)v[pos - 1] / 2.0 + (double)v[pos] / 2.0;
          if (thr == (double)v[pos] || std::isinf(thr)) thr = (double)v[pos - 1];
          best.thr = thr;
          improved = true;
        }
      }
      if (improved) std::copy(idx.begin(), idx.end(), tmp.begin());
    }
    if (best.f >= 0) std::copy(tmp.begin(), tmp.end(), samples.begin() + s);
    // the known constants keep their order for the siblings and children;
    // the new ones follow them
    std::copy(const_features.begin(), const_features.begin() + known, features.begin());
    std::copy(features.begin() + known, features.begin() + total, const_features.begin() + known);
    nconst = total;
    return best;
  }

  int grow(int64_t s, int64_t e, int depth, int nconst) {
    const int id = (int)t.feature.size();
    double c0, c1;
    leaf_value(s, e, c0, c1);
    const double tot = c0 + c1;
    t.feature.push_back(-1);
    t.left.push_back(-1);
    t.right.push_back(-1);
    t.threshold.push_back(-2.0);  // sklearn's TREE_UNDEFINED threshold of a leaf
    t.value.push_back(tot > 0 ? c0 / tot : 0.0);
    t.value.push_back(tot > 0 ? c1 / tot : 0.0);
    const int64_t m = e - s;
    const double gini = tot > 0 ? 1.0 - (c0 * c0 + c1 * c1) / (tot * tot) : 0.0;
    const bool leaf = (p.max_depth > 0 && depth >= p.max_depth) || m < 2 ||
                      m < 2 * (int64_t)p.min_samples_leaf || gini <= DBL_EPSILON;
    if (leaf) return id;
    const Split sp = best_split(s, e, c0, c1, nconst);
    if (sp.f < 0) return id;
    t.feature[id] = sp.f;
    t.threshold[id] = sp.thr;
    const int l = grow(s, s + sp.pos, depth + 1, nconst);
    t.left[id] = l;
    const int r = grow(s + sp.pos, e, depth + 1, nconst);
    t.right[id] = r;
    return id;
  }

  void fit(int tree_index, const double* weights) {
    Rng rng{p.seed * 0x100000001B3ull + (uint64_t)tree_index * 0x9E3779B97F4A7C15ull + 1};
    rstate = p.splitter_state && p.n_trees == 1 ? p.splitter_state
                                                : (uint32_t)(rng.next() >> 33) | 1u;
    w.assign(n, 0.0);
    if (weights) {
      for (int64_t i = 0; i < n; i++) w[i] = weights[i];
    } else if (p.bootstrap) {
      for (int64_t i = 0; i < n; i++) w[(int64_t)(((rng.next() >> 11) * (double)n) * 0x1.0p-53)] += 1.0;
    } else {
      for (int64_t i = 0; i < n; i++) w[i] = 1.0;
    }
    samples.clear();
    for (int64_t i = 0; i < n; i++)
      if (w[i] != 0.0) samples.push_back(i);
    features.resize(nf);
    for (int i = 0; i < nf; i++) features[i] = i;
    const_features.assign(nf, 0);
    t = Tree();
    if (!samples.empty()) grow(0, (int64_t)samples.size(), 0, 0);
  }
};

}  // namespace

struct cbmv_cart_model {
  int nf = 0;
  std::vector<Tree> trees;
};

extern "C" {

void cbmv_cart_params_default(cbmv_cart_params* p) {
  // RandomForestClassifier(min_samples_leaf=700, n_estimators=50) (train.py:287-289)
  // with its defaults: sqrt(20) = 4 candidate features, bootstrap, no depth limit
  p->n_trees = 50;
  p->max_features = 4;
  p->min_samples_leaf = 700;
  p->max_depth = 0;
  p->bootstrap = 1;
  p->n_threads = 0;
  p->seed = 0;
  p->splitter_state = 0;
}

cbmv_cart_model* cbmv_cart_fit(const float* X, const int32_t* y, int64_t n, int32_t n_features,
                               const double* sample_weight, const cbmv_cart_params* params) {
  cbmv::clear_error();
  cbmv_cart_params p;
  if (params)
    p = *params;
  else
    cbmv_cart_params_default(&p);
  if (!X || !y || n < 1 || n_features < 1) {
    cbmv::set_error("cart: need X[n, n_features] and y[n] with n >= 1");
    return nullptr;
  }
  if (p.n_trees < 1 || p.min_samples_leaf < 1 || p.max_depth < 0 || p.max_features < 1) {
    cbmv::set_error("cart: n_trees, min_samples_leaf and max_features must be >= 1, "
                    "max_depth >= 0 (0 = unlimited)");
    return nullptr;
  }
  if (sample_weight && p.n_trees != 1) {
    cbmv::set_error("cart: explicit sample weights fit a single tree (n_trees = 1)");
    return nullptr;
  }
  for (int64_t i = 0; i < n; i++)
    if (y[i] != 0 && y[i] != 1) {
      cbmv::set_error("cart: labels must be 0 or 1 (row %lld has %d)", (long long)i, y[i]);
      return nullptr;
    }
  if (sample_weight)
    for (int64_t i = 0; i < n; i++)
      if (!(sample_weight[i] >= 0.0) || std::isinf(sample_weight[i])) {
        cbmv::set_error("cart: sample weights must be finite and >= 0");
        return nullptr;
      }
  p.max_features = std::min(p.max_features, n_features);
  auto model = std::make_unique<cbmv_cart_model>();
  model->nf = n_features;
  model->trees.resize(p.n_trees);
  int nt = p.n_threads > 0 ? p.n_threads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, p.n_trees));
  std::atomic<int> next{0};
  auto work = [&]() {
    Grower g{X, y, n, n_features, p, {}, {}, {}, {}, {}};
    for (int k = next++; k < p.n_trees; k = next++) {
      g.fit(k, sample_weight);
      model->trees[k] = std::move(g.t);
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < nt; i++) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  return model.release();
}

int64_t cbmv_cart_n_nodes(const cbmv_cart_model* m) {
  if (!m) return -1;
  int64_t s = 0;
  for (const Tree& t : m->trees) s += (int64_t)t.feature.size();
  return s;
}

int32_t cbmv_cart_n_trees(const cbmv_cart_model* m) { return m ? (int32_t)m->trees.size() : -1; }

int cbmv_cart_export(const cbmv_cart_model* m, int32_t* roots, int32_t* feature, double* threshold,
                     int32_t* left, int32_t* right, double* value) {
  cbmv::clear_error();
  if (!m || !roots || !feature || !threshold || !left || !right || !value) {
    cbmv::set_error("cart_export: null argument");
    return CBMV_EINVAL;
  }
  int64_t off = 0;
  for (size_t k = 0; k < m->trees.size(); k++) {
    const Tree& t = m->trees[k];
    roots[k] = (int32_t)off;
    for (size_t i = 0; i < t.feature.size(); i++) {
      feature[off + i] = t.feature[i];
      threshold[off + i] = t.threshold[i];
      left[off + i] = t.left[i] >= 0 ? (int32_t)(t.left[i] + off) : -1;
      right[off + i] = t.right[i] >= 0 ? (int32_t)(t.right[i] + off) : -1;
      value[2 * (off + i)] = t.value[2 * i];
      value[2 * (off + i) + 1] = t.value[2 * i + 1];
    }
    off += (int64_t)t.feature.size();
  }
  return CBMV_OK;
}

void cbmv_cart_destroy(cbmv_cart_model* m) { delete m; }

}  // extern "C"

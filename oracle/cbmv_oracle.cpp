// CPU oracle for the CBMV cost-volume path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, in plain C++/OpenMP, the
// arithmetic of the reference's matchers (cpp/matchers/matchers.cpp),
// feature extraction (cpp/featextract/featextract.cpp) and post-processing
// kernels (cpp/post/post.cu). It is the checker the HIP kernels are compared
// with; nothing in the product links or calls it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Each function cites the reference lines it follows. Where the reference
// mixes float and double, the same types are used here, in the same order of
// operations, so that the values are bit-identical to what the reference
// computes. Build with -ffp-contract=off so no multiply-add is fused.
//
// Layout conventions (same as the reference's return values):
//   census       -> [H, W, D]
//   ncc/sadsob/zsad -> [D, H, W]
//   post sgm works on [H, W, D]; cbca and subpixel on [D, H, W].

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kInvalid = 2147483647.0;  // RAND_MAX in glibc, the reference's fill

inline int64_t idx3(int64_t a, int64_t b, int64_t c, int64_t nb, int64_t nc) {
  return (a * nb + b) * nc + c;
}

// Selection sort with the same comparison as the device helper `sort`
// (post.cu:46-59). Only the sorted result matters, but keep it identical.
void selection_sort(float* x, int n) {
  for (int i = 0; i < n - 1; i++) {
    int mn = i;
    for (int j = i + 1; j < n; j++)
      if (x[j] < x[mn]) mn = j;
    float t = x[mn];
    x[mn] = x[i];
    x[i] = t;
  }
}

}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---------------------------------------------------------------------------
// Matchers
// ---------------------------------------------------------------------------

// census, matchers.cpp:209-350. Bit k of a pixel's descriptor is
// (centre < neighbour_k) over the wsize*wsize window in row-major order; the
// cost is the number of differing bits (the reference counts equal 16-bit
// lanes over the zero-padded vector and subtracts from the padded length, so
// padding lanes always compare equal). Valid centres: y in [wc, H-wsize-1+wc],
// x in [wc, W-wsize-1+wc], d <= x - wc (matchers.cpp:278-279,312-316).
void orc_census(const uint8_t* L, const uint8_t* R, int H, int W, int D, int wsize,
                double* out /* [H,W,D] */) {
  const int64_t total = (int64_t)H * W * D;
  for (int64_t i = 0; i < total; i++) out[i] = kInvalid;
  const int wc = wsize / 2;
  const int nbits = wsize * wsize;
  const int words = (nbits + 63) / 64;
  std::vector<uint64_t> dl((size_t)H * W * words, 0), dr((size_t)H * W * words, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H - wsize; i++) {
    for (int j = 0; j < W - wsize; j++) {
      const int cy = i + wc, cx = j + wc;
      const int cl = L[(int64_t)cy * W + cx], cr = R[(int64_t)cy * W + cx];
      uint64_t* pl = &dl[((int64_t)cy * W + cx) * words];
      uint64_t* pr = &dr[((int64_t)cy * W + cx) * words];
      for (int k = 0; k < nbits; k++) {
        const int wy = k / wsize, wx = k % wsize;
        const int64_t p = (int64_t)(i + wy) * W + (j + wx);
        if (cl < (int)L[p]) pl[k >> 6] |= (uint64_t)1 << (k & 63);
        if (cr < (int)R[p]) pr[k >> 6] |= (uint64_t)1 << (k & 63);
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H - wsize; i++) {
    for (int j = 0; j < W - wsize; j++) {
      const int cy = i + wc, cx = j + wc;
      const int end = std::min(D, j + 1);
      const uint64_t* pl = &dl[((int64_t)cy * W + cx) * words];
      for (int d = 0; d < end; d++) {
        const uint64_t* pr = &dr[((int64_t)cy * W + (cx - d)) * words];
        int ham = 0;
        for (int w = 0; w < words; w++) ham += __builtin_popcountll(pl[w] ^ pr[w]);
        out[idx3(cy, cx, d, W, D)] = (double)ham;
      }
    }
  }
}

// nccNister, matchers.cpp:24-205. Window sums are exact integers; the
// finish is -(wsize^2 * sum(LR) - sum(L) sum(R)) * Cl * Cr with
// C = 1/sqrt(wsize^2 * sum(X^2) - sum(X)^2) in double, evaluated left to
// right. A non-finite C (constant window) gives the value 1 (:178-181).
// Valid: y in [wc, H-wsize-1+wc], x in [d+wc, W-wsize-1+wc] (:168-173).
void orc_ncc(const uint8_t* L, const uint8_t* R, int H, int W, int D, int wsize,
             double* out /* [D,H,W] */) {
  const int64_t total = (int64_t)H * W * D;
  for (int64_t i = 0; i < total; i++) out[i] = kInvalid;
  const int wc = wsize / 2;
  const int sq = wsize * wsize;
  std::vector<uint64_t> Al((size_t)H * W, 0), Ar((size_t)H * W, 0);
  std::vector<double> Cl((size_t)H * W, 0.0), Cr((size_t)H * W, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H - wsize; i++) {
    for (int j = 0; j < W - wsize; j++) {
      uint64_t sl = 0, sr = 0, ql = 0, qr = 0;
      for (int a = 0; a < wsize; a++)
        for (int b = 0; b < wsize; b++) {
          const int64_t p = (int64_t)(i + a) * W + (j + b);
          sl += L[p];
          sr += R[p];
          ql += (uint64_t)L[p] * L[p];
          qr += (uint64_t)R[p] * R[p];
        }
      const int64_t c = (int64_t)(i + wc) * W + (j + wc);
      Al[c] = sl;
      Ar[c] = sr;
      // (double)(sq*B) - (double)A*A, then 1/sqrt (matchers.cpp:127-128)
      Cl[c] = 1 / (std::sqrt((double)((uint64_t)sq * ql) - (double)sl * (double)sl));
      Cr[c] = 1 / (std::sqrt((double)((uint64_t)sq * qr) - (double)sr * (double)sr));
    }
  }
#pragma omp parallel for schedule(dynamic)
  for (int d = 0; d < D; d++) {
    for (int i = 0; i < H - wsize; i++) {
      const int cy = i + wc;
      for (int j = d; j < W - wsize; j++) {
        uint64_t lr = 0;
        for (int a = 0; a < wsize; a++)
          for (int b = 0; b < wsize; b++) {
            const int64_t p = (int64_t)(i + a) * W + (j + b);
            lr += (uint64_t)L[p] * R[p - d];
          }
        const int64_t c = (int64_t)cy * W + (j + wc);
        const int64_t cRr = (int64_t)cy * W + (j - d + wc);
        double v;
        if (std::isfinite(Cl[c]) && std::isfinite(Cr[cRr])) {
          // (double)(sq*lD - Al*Ar): both products exact, difference exact.
          const double num = (double)sq * (double)lr - (double)(Al[c] * Ar[cRr]);
          v = -num * Cl[c] * Cr[cRr];
        } else {
          v = 1;
        }
        out[(int64_t)d * H * W + c] = v;
      }
    }
  }
}

// sobel, matchers.cpp:515-555: horizontal Sobel, written at
// (y,x) in [1,H-3] x [1,W-3], zero elsewhere.
void orc_sobel(const uint8_t* img, int H, int W, double* out) {
  for (int64_t i = 0; i < (int64_t)H * W; i++) out[i] = 0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H - 3; i++)
    for (int j = 0; j < W - 3; j++) {
      const uint8_t* r0 = img + (int64_t)i * W + j;
      const uint8_t* r1 = r0 + W;
      const uint8_t* r2 = r1 + W;
      const int v = -r0[0] + r0[2] - 2 * r1[0] + 2 * r1[2] - r2[0] + r2[2];
      out[(int64_t)(i + 1) * W + (j + 1)] = (double)v;
    }
}

// sadsob, matchers.cpp:353-436: wsize x wsize SAD of two double images.
// The reference sums through a per-d integral image; for integer-valued
// inputs (Sobel output) every partial sum is an exact integer, so the direct
// sum is identical. Valid: y in [wc, H-wsize-1+wc], x in [d+wc, W-wsize-1+wc].
void orc_sadsob(const double* Ls, const double* Rs, int H, int W, int D, int wsize,
                double* out /* [D,H,W] */) {
  const int64_t total = (int64_t)H * W * D;
  for (int64_t i = 0; i < total; i++) out[i] = kInvalid;
  const int wc = wsize / 2;
#pragma omp parallel for schedule(dynamic)
  for (int d = 0; d < D; d++)
    for (int i = 0; i < H - wsize; i++)
      for (int j = d; j < W - wsize; j++) {
        double s = 0;
        for (int a = 0; a < wsize; a++)
          for (int b = 0; b < wsize; b++) {
            const int64_t p = (int64_t)(i + a) * W + (j + b);
            s += std::fabs(Ls[p] - Rs[p - d]);
          }
        out[(int64_t)d * H * W + (int64_t)(i + wc) * W + (j + wc)] = s;
      }
}

// zsad, matchers.cpp:440-512. Means are double sums / wsize^2 (:471-483);
// each tap is fabs(((l - ml) - r) + mr) accumulated in double in window
// row-major order (:499-505).
void orc_zsad(const uint8_t* L, const uint8_t* R, int H, int W, int D, int wsize,
              double* out /* [D,H,W] */) {
  const int64_t total = (int64_t)H * W * D;
  for (int64_t i = 0; i < total; i++) out[i] = kInvalid;
  const int wc = wsize / 2;
  const int sq = wsize * wsize;
  std::vector<double> ml((size_t)H * W, 0.0), mr((size_t)H * W, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H - wsize; i++)
    for (int j = 0; j < W - wsize; j++) {
      double sl = 0, sr = 0;
      for (int a = 0; a < wsize; a++)
        for (int b = 0; b < wsize; b++) {
          const int64_t p = (int64_t)(i + a) * W + (j + b);
          sl += L[p];
          sr += R[p];
        }
      const int64_t c = (int64_t)(i + wc) * W + (j + wc);
      ml[c] = sl / sq;
      mr[c] = sr / sq;
    }
#pragma omp parallel for schedule(dynamic)
  for (int d = 0; d < D; d++)
    for (int i = 0; i < H - wsize; i++)
      for (int j = d; j < W - wsize; j++) {
        const int64_t c = (int64_t)(i + wc) * W + (j + wc);
        const double mL = ml[c];
        const double mR = mr[c - d];
        double s = 0;
        for (int a = 0; a < wsize; a++)
          for (int b = 0; b < wsize; b++) {
            const int64_t p = (int64_t)(i + a) * W + (j + b);
            s += std::fabs((double)L[p] - mL - (double)R[p - d] + mR);
          }
        out[(int64_t)d * H * W + c] = s;
      }
}

// ---------------------------------------------------------------------------
// Feature extraction (featextract.cpp)
// ---------------------------------------------------------------------------

// get_right_cost, featextract.cpp:115-152: out[y,x,d] = in[y,x+d,d] for
// x < W-d, else in[0] (the first element of the passed array).
void orc_right_cost(const double* in, int H, int W, int D, double* out) {
  const double fill = in[0];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++)
      for (int d = 0; d < D; d++)
        out[idx3(i, j, d, W, D)] = (j < W - d) ? in[idx3(i, j + d, d, W, D)] : fill;
}

// get_left_cost_from_right, featextract.cpp:452-488: out[y,x,d] = in[y,x-d,d]
// for x >= d, else in[0].
void orc_left_cost(const double* in, int H, int W, int D, double* out) {
  const double fill = in[0];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < H; i++)
    for (int j = 0; j < W; j++)
      for (int d = 0; d < D; d++)
        out[idx3(i, j, d, W, D)] = (j >= d) ? in[idx3(i, j - d, d, W, D)] : fill;
}

static inline double row_min(const double* v, int D) {
  double m = kInvalid;  // strict < from RAND_MAX (featextract.cpp:322-327)
  for (int k = 0; k < D; k++)
    if (v[k] < m) m = v[k];
  return m;
}

// extract_pkrn_test, featextract.cpp:303-342 (2-argument extract_ratio).
void orc_ratio(const double* vol, int64_t N, int D, double e, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const double* v = vol + i * D;
    const double m = row_min(v, D);
    for (int j = 0; j < D; j++) out[i * D + j] = (m == kInvalid) ? 0.0 : (m + e) / (v[j] + e);
  }
}

// extract_pkrn, featextract.cpp:254-300 (3-argument extract_ratio).
void orc_ratio_samples(const double* vol, int64_t N, int D, const int32_t* samp, double e,
                       double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const double* v = vol + i * D;
    const double m = row_min(v, D);
    for (int k = 0; k < 3; k++)
      out[3 * i + k] = (m == kInvalid) ? 0.0 : (m + e) / (v[samp[3 * i + k]] + e);
  }
}

static inline double aml_denom(const double* v, int D, double m, double sigma) {
  double den = 0;
  for (int k = 0; k < D; k++) {
    const double num = v[k] - m;
    den += std::exp(-(num * num) / sigma);
  }
  return den;
}

// extract_aml_testing, featextract.cpp:402-450 (2-argument extract_likelihood).
void orc_aml(const double* vol, int64_t N, int D, double sigma, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const double* v = vol + i * D;
    const double m = row_min(v, D);
    const double den = aml_denom(v, D, m, sigma);
    for (int j = 0; j < D; j++) {
      const double t = v[j] - m;
      out[i * D + j] = (m == kInvalid) ? 0.0 : std::exp(-((t * t) / sigma)) / den;
    }
  }
}

// extract_aml, featextract.cpp:345-399 (3-argument extract_likelihood).
void orc_aml_samples(const double* vol, int64_t N, int D, const int32_t* samp, double sigma,
                     double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const double* v = vol + i * D;
    const double m = row_min(v, D);
    const double den = aml_denom(v, D, m, sigma);
    for (int k = 0; k < 3; k++) {
      const double t = v[samp[3 * i + k]] - m;
      out[3 * i + k] = (m == kInvalid) ? 0.0 : std::exp(-((t * t) / sigma)) / den;
    }
  }
}

// get_samples, featextract.cpp:217-252.
void orc_get_samples(const double* vol, int64_t N, int D, const int32_t* samp, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++)
    for (int k = 0; k < 3; k++) out[3 * i + k] = vol[i * D + samp[3 * i + k]];
}

// ---------------------------------------------------------------------------
// Random forest (scikit-learn predict_proba semantics: the sample is cast to
// float32, a node sends it left when (double)x <= threshold, the per-tree
// probability is the leaf's stored class fraction, the forest sums them in
// tree order in double and divides by the number of trees).
// Nodes of all trees are concatenated; children are absolute node indices,
// -1 marks a leaf. value0 is the class-0 (non-match) fraction of each node.
// ---------------------------------------------------------------------------
void orc_forest_predict(const float* X, int64_t N, int F, const int32_t* feature,
                        const double* threshold, const int32_t* left, const int32_t* right,
                        const double* value0, const int32_t* roots, int n_trees,
                        double* out /* [N] P(class 0) */) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; i++) {
    const float* x = X + i * F;
    double acc = 0;
    for (int t = 0; t < n_trees; t++) {
      int n = roots[t];
      while (left[n] >= 0) n = ((double)x[feature[n]] <= threshold[n]) ? left[n] : right[n];
      acc += value0[n];
    }
    out[i] = acc / n_trees;
  }
}

// ---------------------------------------------------------------------------
// Post-processing (post.cu)
// ---------------------------------------------------------------------------

// cross, post.cu:336-396. out[dir,y,x] for dir 0..3 = (-x, +x, -y, +y):
// the first coordinate along the arm that is rejected (colour difference
// >= tau1, arm length >= L1, or outside the image). Distance 1 is always
// accepted.
void orc_cross(const float* img, int H, int W, int L1, float tau1, float* out /* [4,H,W] */) {
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)4 * H * W; id++) {
    int64_t t = id;
    const int x = (int)(t % W);
    t /= W;
    const int y = (int)(t % H);
    const int dir = (int)(t / H);
    const int dx = dir == 0 ? -1 : (dir == 1 ? 1 : 0);
    const int dy = dir == 2 ? -1 : (dir == 3 ? 1 : 0);
    const float c = img[(int64_t)y * W + x];
    int xx = x + dx, yy = y + dy;
    for (;; xx += dx, yy += dy) {
      if (xx < 0 || xx >= W || yy < 0 || yy >= H) break;
      const int dist = std::max(std::abs(xx - x), std::abs(yy - y));
      if (dist == 1) continue;
      if (std::fabs(c - img[(int64_t)yy * W + xx]) >= tau1) break;
      if (dist >= L1) break;
    }
    out[id] = (float)(dir <= 1 ? xx : yy);
  }
}

// cbca, post.cu:399-450, volume layout [D,H,W]. The support region is the
// intersection of the left-image cross at x and the right-image cross at
// x + d*direction; values are read as float and summed in float in row-major
// order; the mean is sum/cnt in float. Voxels whose matching pixel is
// outside the image are copied.
void orc_cbca(const float* lc, const float* rc, const double* vol, double* out, int D, int H,
              int W, int direction) {
  const int64_t HW = (int64_t)H * W;
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)D * HW; id++) {
    int64_t t = id;
    const int x = (int)(t % W);
    t /= W;
    const int y = (int)(t % H);
    const int d = (int)(t / H);
    const int xr = x + d * direction;
    if (xr < 0 || xr >= W) {
      out[id] = vol[id];
      continue;
    }
    float sum = 0;
    int cnt = 0;
    const int yy_s = (int)std::max(lc[(2 * (int64_t)H + y) * W + x], rc[(2 * (int64_t)H + y) * W + xr]);
    const int yy_t = (int)std::min(lc[(3 * (int64_t)H + y) * W + x], rc[(3 * (int64_t)H + y) * W + xr]);
    for (int yy = yy_s + 1; yy < yy_t; yy++) {
      const int xx_s = (int)std::max(lc[(0 * (int64_t)H + yy) * W + x],
                                     rc[(0 * (int64_t)H + yy) * W + xr] - d * direction);
      const int xx_t = (int)std::min(lc[(1 * (int64_t)H + yy) * W + x],
                                     rc[(1 * (int64_t)H + yy) * W + xr] - d * direction);
      for (int xx = xx_s + 1; xx < xx_t; xx++) {
        const float v = (float)vol[(int64_t)d * HW + (int64_t)yy * W + xx];
        sum += v;
        cnt++;
      }
    }
    out[id] = sum / cnt;
  }
}

// One SGM path, post.cu:178-261 (sgm2<dir>), run over every scanline.
// path 0..3 = right, left, down, up as in the reference; 4..7 are the
// diagonal extension (dx,dy) = (1,1), (-1,-1), (1,-1), (-1,1), where paths
// with dy = +1 use the 'down' penalty asymmetry and dy = -1 the 'up' one.
// in/out layout [H,W,D]; out accumulates (+=) the path values in double.
static void sgm_path(const float* x0, const float* x1, const double* in, double* out, int H,
                     int W, int D, float pi1, float pi2, float tau_so, float alpha1, float q1,
                     float q2, int direction, int path) {
  static const int kdx[8] = {1, -1, 0, 0, 1, -1, 1, -1};
  static const int kdy[8] = {0, 0, 1, -1, 1, -1, -1, 1};
  const int dx = kdx[path], dy = kdy[path];
  // Scanlines: each is identified by its start pixel; walk until leaving.
  std::vector<std::pair<int, int>> starts;
  if (dy == 0) {
    for (int y = 0; y < H; y++) starts.push_back({dx > 0 ? 0 : W - 1, y});
  } else if (dx == 0) {
    for (int x = 0; x < W; x++) starts.push_back({x, dy > 0 ? 0 : H - 1});
  } else {
    const int y0 = dy > 0 ? 0 : H - 1;
    const int x0s = dx > 0 ? 0 : W - 1;
    for (int x = 0; x < W; x++) starts.push_back({x, y0});
    for (int y = 0; y < H; y++)
      if (y != y0) starts.push_back({x0s, y});
  }
  const bool down_rule = dy > 0;
  const bool up_rule = dy < 0;
#pragma omp parallel
  {
    std::vector<float> prev(D), cur(D);
#pragma omp for schedule(dynamic)
    for (size_t s = 0; s < starts.size(); s++) {
      int x = starts[s].first, y = starts[s].second;
      bool first = true;
      for (; x >= 0 && x < W && y >= 0 && y < H; x += dx, y += dy) {
        const int64_t pix = (int64_t)y * W + x;
        if (first) {
          for (int d = 0; d < D; d++) {
            const float val = (float)in[pix * D + d];
            out[pix * D + d] += val;
            cur[d] = val;
          }
          first = false;
        } else {
          float mn = prev[0];
          for (int d = 1; d < D; d++)
            if (prev[d] < mn) mn = prev[d];
          const int64_t q = (int64_t)(y - dy) * W + (x - dx);
          const float D1 = std::fabs(x0[pix] - x0[q]);
          for (int d = 0; d < D; d++) {
            const int xx = x + d * direction;
            float D2;
            if (xx < 0 || xx >= W || xx - dx < 0 || xx - dx >= W)
              D2 = 10;
            else
              D2 = std::fabs(x1[pix + d * direction] - x1[q + d * direction]);
            float P1, P2;
            if (D1 < tau_so && D2 < tau_so) {
              P1 = pi1;
              P2 = pi2;
            } else if (D1 > tau_so && D2 > tau_so) {
              P1 = pi1 / (q1 * q2);
              P2 = pi2 / (q1 * q2);
            } else {
              P1 = pi1 / q1;
              P2 = pi2 / q1;
            }
            float c = std::min(prev[d], mn + P2);
            if (d - 1 >= 0) c = std::min(c, prev[d - 1] + (down_rule ? P1 / alpha1 : P1));
            if (d + 1 < D) c = std::min(c, prev[d + 1] + (up_rule ? P1 / alpha1 : P1));
            const double val = (in[pix * D + d] + c - mn);
            out[pix * D + d] += val;
            cur[d] = (float)val;
          }
        }
        std::swap(prev, cur);
      }
    }
  }
}

// sgm_w, post.cu:263-333: the paths in order right, left, down, up (and the
// four diagonals when npaths == 8). `out` must be zeroed by the caller
// (test.py:379 set_to_zero) and is accumulated in double.
void orc_sgm(const float* x0, const float* x1, const double* in, double* out, int H, int W,
             int D, float pi1, float pi2, float tau_so, float alpha1, float q1, float q2,
             int direction, int npaths) {
  for (int p = 0; p < npaths; p++)
    sgm_path(x0, x1, in, out, H, W, D, pi1, pi2, tau_so, alpha1, q1, q2, direction, p);
}

// One path alone (0..7, numbered as in sgm_path), accumulated into `out`;
// lets the tests check each direction's recurrence on its own.
void orc_sgm_path(const float* x0, const float* x1, const double* in, double* out, int H, int W,
                  int D, float pi1, float pi2, float tau_so, float alpha1, float q1, float q2,
                  int direction, int path) {
  sgm_path(x0, x1, in, out, H, W, D, pi1, pi2, tau_so, alpha1, q1, q2, direction, path);
}

// subpixel_enchancement, post.cu:453-483. cost layout [D,H,W].
void orc_subpixel(const float* d0, const double* cost, int D, int H, int W, float* out) {
  const int64_t HW = (int64_t)H * W;
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < HW; id++) {
    const int d = (int)d0[id];
    out[id] = (float)d;
    if (1 <= d && d < D - 1) {
      const float cn = (float)cost[(int64_t)(d - 1) * HW + id];
      const float cz = (float)cost[(int64_t)d * HW + id];
      const float cp = (float)cost[(int64_t)(d + 1) * HW + id];
      const float denom = 2 * (cp + cn - 2 * cz);
      if (denom > 1e-5) {
        const double r = std::min(1.0, std::max(-1.0, (double)((cp - cn) / denom)));
        out[id] = (float)(d - r);
      }
    }
  }
}

// median2d, post.cu:486-518: median (element n/2 of the sorted in-bounds
// window) over a ksize x ksize window.
void orc_median2d(const float* in, int H, int W, int ksize, float* out) {
  const int r = ksize / 2;
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)H * W; id++) {
    const int x = (int)(id % W), y = (int)(id / W);
    float xs[121];
    int n = 0;
    for (int xx = x - r; xx <= x + r; xx++)
      for (int yy = y - r; yy <= y + r; yy++)
        if (0 <= xx && xx < W && 0 <= yy && yy < H) xs[n++] = in[(int64_t)yy * W + xx];
    selection_sort(xs, n);
    out[id] = xs[n / 2];
  }
}

// mean2d, post.cu:520-566: weighted mean over neighbours whose value differs
// by less than alpha2; kernel element i counts xx in the outer loop and yy in
// the inner loop, i.e. kernel[(xx-x+r)*k + (yy-y+r)]. The reference builds
// post.cu with nvcc's default --fmad=true (cpp/post/makefile:27), which fuses
// `sum += img * kernel` (post.cu:533) into one fused multiply-add: so does
// this restatement. It is the only multiply-add in post.cu whose fusion
// changes a result (2*cz in post.cu:462 is exact).
void orc_mean2d(const float* in, int H, int W, const float* kernel, int ksize, float alpha2,
                float* out) {
  const int r = ksize / 2;
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)H * W; id++) {
    const int x = (int)(id % W), y = (int)(id / W);
    float sum = 0, cnt = 0;
    int i = 0;
    for (int xx = x - r; xx <= x + r; xx++)
      for (int yy = y - r; yy <= y + r; yy++, i++)
        if (0 <= xx && xx < W && 0 <= yy && yy < H &&
            std::fabs(in[(int64_t)yy * W + xx] - in[id]) < alpha2) {
          sum = std::fma(in[(int64_t)yy * W + xx], kernel[i], sum);
          cnt += kernel[i];
        }
    out[id] = sum / cnt;
  }
}

// outlier_detection, post.cu:717-754: 0 = match, 1 = occlusion, 2 = mismatch.
void orc_outlier(const float* d0, const float* d1, int H, int W, int disp_max, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)H * W; id++) {
    const int x = (int)(id % W);
    const int d0i = (int)d0[id];
    if (x - d0i < 0) {
      out[id] = 1;
    } else if (std::fabs(d0[id] - d1[id - d0i]) < 1.1) {
      out[id] = 0;
    } else {
      out[id] = 1;
      for (int d = 0; d < disp_max; d++)
        if (x - d >= 0 && std::fabs((float)d - d1[id - d]) < 1.1) {
          out[id] = 2;
          break;
        }
    }
  }
}

// interpolate_mismatch, post.cu:757-827: for mismatch pixels, march along 16
// directions (rounding half away from zero) to the first non-mismatch pixel
// and take element n/2 of the sorted values found. With no value found the
// input value is kept (the reference asserts instead).
void orc_interp_mismatch(const float* d0, const float* outlier, int H, int W, float* out) {
  static const float dirs[32] = {0, 1,  -0.5, 1,  -1, 1,  -1, 0.5, -1, 0,  -1,  -0.5, -1, -1, -0.5, -1,
                                 0, -1, 0.5,  -1, 1,  -1, 1,  -0.5, 1, 0,  1,   0.5,  1,  1,  0.5,  1};
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t id = 0; id < (int64_t)H * W; id++) {
    if (outlier[id] != 2) {
      out[id] = d0[id];
      continue;
    }
    float vals[16];
    int n = 0;
    const int x = (int)(id % W), y = (int)(id / W);
    for (int k = 0; k < 16; k++) {
      const float dx = dirs[2 * k], dy = dirs[2 * k + 1];
      float xx = (float)x, yy = (float)y;
      int xi = (int)std::round(xx), yi = (int)std::round(yy);
      while (0 <= yi && yi < H && 0 <= xi && xi < W && outlier[(int64_t)yi * W + xi] == 2) {
        xx += dx;
        yy += dy;
        xi = (int)std::round(xx);
        yi = (int)std::round(yy);
      }
      if (0 <= yi && yi < H && 0 <= xi && xi < W) vals[n++] = d0[(int64_t)yi * W + xi];
    }
    if (n == 0) {
      out[id] = d0[id];
      continue;
    }
    selection_sort(vals, n);
    out[id] = vals[n / 2];
  }
}

// interpolate_occlusion, post.cu:829-869: nearest match pixel to the left,
// else to the right, else unchanged.
void orc_interp_occlusion(const float* d0, const float* outlier, int H, int W, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)H * W; id++) {
    if (outlier[id] != 1) {
      out[id] = d0[id];
      continue;
    }
    const int x = (int)(id % W);
    int dx = 0;
    while (x + dx >= 0 && outlier[id + dx] != 0) dx--;
    if (x + dx < 0) {
      dx = 0;
      while (x + dx < W && outlier[id + dx] != 0) dx++;
    }
    out[id] = (x + dx < W) ? d0[id + dx] : d0[id];
  }
}

}  // extern "C"

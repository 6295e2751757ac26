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
#include <cmath>
#include <vector>

#include "features.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace cbmv {
namespace {

constexpr int kB = 256;

// ------------------------------------------------------------------ statistics
// One thread per pixel (x fastest). For each matcher: minimum with strict <
// from RAND_MAX, then the sum of exp(-(c - m)^2 / sigma), both in ascending d
// as in featextract.cpp:322-327,373-383. Right view: c_R(x', d) = c(x' + d, d)
// for x' + d < W, else the batch's first element (featextract.cpp:130).
__global__ void __launch_bounds__(kB) k_stats(Vols v, int H, int W, int D, int rows_per_batch,
                                              FeatureSpec fs, double* __restrict__ sl,
                                              double* __restrict__ sr) {
  const int x = blockIdx.x * kB + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W) return;
  const int64_t HW = (int64_t)H * W;
  const int64_t pix = (int64_t)y * W + x;
  const int64_t rowbase = (int64_t)y * D * W;
  const int row0 = (y / rows_per_batch) * rows_per_batch;
#pragma unroll 1
  for (int k = 0; k < 4; k++) {
    const double sigma = fs.sigma[k];
    // left view
    double m = kRandMax;
    for (int d = 0; d < D; d++) {
      const double c = vol_load(v, k, rowbase + (int64_t)d * W + x);
      if (c < m) m = c;
    }
    double den = 0.0;
    for (int d = 0; d < D; d++) {
      const double t = vol_load(v, k, rowbase + (int64_t)d * W + x) - m;
      den += exp(-(t * t) / sigma);
    }
    sl[k * HW + pix] = m;
    sl[(4 + k) * HW + pix] = den;
    // right view
    const double fill = vol_load(v, k, (int64_t)row0 * D * W);
    m = kRandMax;
    for (int d = 0; d < D; d++) {
      const double c = (x + d < W) ? vol_load(v, k, rowbase + (int64_t)d * W + x + d) : fill;
      if (c < m) m = c;
    }
    den = 0.0;
    for (int d = 0; d < D; d++) {
      const double c = (x + d < W) ? vol_load(v, k, rowbase + (int64_t)d * W + x + d) : fill;
      const double t = c - m;
      den += exp(-(t * t) / sigma);
    }
    sr[k * HW + pix] = m;
    sr[(4 + k) * HW + pix] = den;
  }
}

// ------------------------------------------------------------------ features + forest
// Block = 64 x-values x 64 disparities of one row; thread (lane, wave) handles
// x = x0 + lane and d = d0 + wave + 4 j. The cost tile is transposed through
// LDS so that the [H, W, D] output is written in contiguous 256-byte rows.
constexpr int kTX = 64, kTD = 64;

__global__ void __launch_bounds__(kB) k_features_rf(Vols v, Stats st, FeatureSpec fs, int H, int W,
                                                    int D, int rows_per_batch, ForestView f,
                                                    float* __restrict__ P) {
  __shared__ float xs[20][kB];
  __shared__ float tile[kTX][kTD + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x0 = blockIdx.x * kTX, d0 = blockIdx.y * kTD, y = blockIdx.z;
  const int x = x0 + lane;
  for (int j = 0; j < kTD / 4; j++) {
    const int dd = wave + 4 * j;
    const int d = d0 + dd;
    if (x < W && d < D) {
      voxel_features(v, st, fs, H, W, D, rows_per_batch, y, x, d,
                     [&](int i, float val) { xs[i][tid] = val; });
      const double acc = forest_sum0_ilp<4>(f, [&](uint32_t i) { return xs[i][tid]; });
      // sklearn: sum over trees, then divide by the tree count
      tile[lane][dd] = (float)(acc / f.n_trees);
    }
  }
  __syncthreads();
  for (int j = 0; j < kTX / 4; j++) {
    const int r = wave + 4 * j;
    const int xx = x0 + r, d = d0 + lane;
    if (xx < W && d < D) P[((int64_t)y * W + xx) * D + d] = tile[r][lane];
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
// (img - mean) / std in float32 with mean and population std from float64 sums
// (test.py:489-492).
__global__ void k_img_partials(const uint8_t* __restrict__ img, int64_t n, double* __restrict__ part) {
  __shared__ double s1[kB / 64], s2[kB / 64];
  double a = 0, b = 0;
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
    const double v = img[i];
    a += v;
    b += v * v;
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    s1[threadIdx.x >> 6] = a;
    s2[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int w = 0; w < kB / 64; w++) {
      ta += s1[w];
      tb += s2[w];
    }
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
  }
}

__global__ void k_img_normalise(const uint8_t* __restrict__ img, int64_t n,
                                const double* __restrict__ part, int nparts,
                                float* __restrict__ out) {
  __shared__ float ms[2];
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int i = 0; i < nparts; i++) {
      a += part[2 * i];
      b += part[2 * i + 1];
    }
    const double mean = a / (double)n;
    const double var = fmax(b / (double)n - mean * mean, 0.0);
    ms[0] = (float)mean;
    ms[1] = (float)sqrt(var);
  }
  __syncthreads();
  const float mean = ms[0], sd = ms[1];
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
    out[i] = ((float)img[i] - mean) / sd;
}

// ------------------------------------------------------------------ right-view cost
// PR[y, x', d] = P[y, x' + d, d] for x' + d < W, else P[0] (get_right_cost,
// featextract.cpp:115-152, applied to the whole volume at test.py:156).
__global__ void k_right_cost(const float* __restrict__ P, int H, int W, int D,
                             float* __restrict__ PR) {
  const int d = threadIdx.x;
  const int x = blockIdx.x, y = blockIdx.y;
  if (d >= D) return;
  const int64_t o = ((int64_t)y * W + x) * D + d;
  PR[o] = (x + d < W) ? P[o + (int64_t)d * D] : P[0];
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

// ---- sliding-window CBCA (the pipeline's default for its arm length L1) ----
// Row pass: a wavefront walks one row (y, 64 disparities) along x keeping the
// 2*L1-1 costs around x in registers, so every cost is loaded once; the
// masked sum runs in ascending x like post.cu:420-426. The next PF steps'
// inputs (new cost and the arm bounds) sit in a register ring.
template <int L1, int PF>
__global__ void __launch_bounds__(256) k_cbca_rows(const float* __restrict__ vol,
                                                   const float* __restrict__ lc,
                                                   const float* __restrict__ rc, int H, int W,
                                                   int D, int dir, float* __restrict__ hsum,
                                                   uint8_t* __restrict__ hcnt, int nwork) {
  constexpr int K = 2 * L1 - 1;
  const int lane = threadIdx.x & 63;
  const int work = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (work >= nwork) return;
  const int nchunk = (D + 63) / 64;
  const int y = work / nchunk;
  const int d = (work % nchunk) * 64 + lane;
  const bool dok = d < D;
  const int dd = d * dir;
  const int64_t HW = (int64_t)H * W;
  const float* row = vol + (int64_t)y * W * D + d;
  const float* lrow = lc + (int64_t)y * W;
  const float* rrow = rc + (int64_t)y * W;
  float win[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int xx = k - L1;
    win[k] = (dok && xx >= 0 && xx < W) ? row[(int64_t)xx * D] : 0.f;
  }
  // ring of raw loaded values (clamped addresses, no arithmetic before use)
  float rv[PF], rl0[PF], rl1[PF], rr0[PF], rr1[PF];
  const int dcl = min(d, D - 1);
  const float* rowc = vol + (int64_t)y * W * D + dcl;
  auto load = [&](int slot, int x) {
    x = min(x, W - 1);
    const int xn = min(x + L1 - 1, W - 1);
    rv[slot] = rowc[(int64_t)xn * D];
    const int xr = min(max(x + dd, 0), W - 1);
    rl0[slot] = lrow[x];
    rl1[slot] = lrow[HW + x];
    rr0[slot] = rrow[xr];
    rr1[slot] = rrow[HW + xr];
  };
#pragma unroll
  for (int k = 0; k < PF; k++) load(k, k);
  for (int x0 = 0; x0 < W; x0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const int x = x0 + k;
      if (x >= W) break;
#pragma unroll
      for (int q = 0; q < K - 1; q++) win[q] = win[q + 1];
      win[K - 1] = (dok && x + L1 - 1 < W) ? rv[k] : 0.f;
      const int xr = x + dd;
      const bool vox = dok && xr >= 0 && xr < W;
      const int xs = vox ? (int)fmaxf(rl0[k], rr0[k] - dd) : 0x7fffffff;  // empty range
      const int xt = vox ? (int)fminf(rl1[k], rr1[k] - dd) : 0;
      float sum = 0.f;
      int cnt = 0;
#pragma unroll
      for (int q = 0; q < K; q++) {
        const int xx = x - (L1 - 1) + q;
        if (xx > xs && xx < xt) {
          sum += win[q];
          cnt++;
        }
      }
      if (dok) {
        const int64_t o = ((int64_t)y * W + x) * D + d;
        hsum[o] = sum;
        hcnt[o] = (uint8_t)cnt;
      }
      load(k, x + PF);
    }
  }
}

// Column pass: a workgroup (one wavefront per 64 disparities) walks one
// column down y keeping the 2*L1-1 row sums around y in registers; the mean
// over the vertical arm (post.cu:415-417). With WTA the workgroup also takes
// the first minimum over d of each pixel and the subpixel parabola, writing
// only the disparity maps.
template <int L1, int PF, bool WTA>
__global__ void __launch_bounds__(512) k_cbca_cols(const float* __restrict__ vol,
                                                   const float* __restrict__ hsum,
                                                   const uint8_t* __restrict__ hcnt,
                                                   const float* __restrict__ lc,
                                                   const float* __restrict__ rc, int H, int W,
                                                   int D, int dir, float* __restrict__ out,
                                                   float* __restrict__ wta,
                                                   float* __restrict__ sub) {
  constexpr int K = 2 * L1 - 1;
  __shared__ float vals[2][512];
  __shared__ float rmin[2][8];
  __shared__ int ridx[2][8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = blockDim.x >> 6;
  const int x = blockIdx.x;
  const int d = tid;
  const bool dok = d < D;
  const int dd = d * dir;
  const int xr = x + dd;
  const bool inside = dok && xr >= 0 && xr < W;
  const int64_t HW = (int64_t)H * W;
  const int64_t colstride = (int64_t)W * D;
  const float* hcol = hsum + (int64_t)x * D + d;
  const uint8_t* ccol = hcnt + (int64_t)x * D + d;
  float ws[K];
  int wc[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int yy = k - L1;
    const bool ok = inside && yy >= 0 && yy < H;
    ws[k] = ok ? hcol[yy * colstride] : 0.f;
    wc[k] = ok ? (int)ccol[yy * colstride] : 0;
  }
  // ring of raw loaded values (clamped addresses, no arithmetic before use)
  float rvs[PF], rl2[PF], rl3[PF], rr2[PF], rr3[PF], rcopy[PF];
  uint8_t rvc[PF];
  const int dcl = min(d, D - 1);
  const int xrc = min(max(xr, 0), W - 1);
  const float* hcolc = hsum + (int64_t)x * D + dcl;
  const uint8_t* ccolc = hcnt + (int64_t)x * D + dcl;
  auto load = [&](int slot, int y) {
    y = min(y, H - 1);
    const int yn = min(y + L1 - 1, H - 1);
    rvs[slot] = hcolc[yn * colstride];
    rvc[slot] = ccolc[yn * colstride];
    const int64_t prow = (int64_t)y * W;
    rl2[slot] = lc[2 * HW + prow + x];
    rl3[slot] = lc[3 * HW + prow + x];
    rr2[slot] = rc[2 * HW + prow + xrc];
    rr3[slot] = rc[3 * HW + prow + xrc];
    // only voxels whose match is outside the image copy their cost; the others
    // all read element 0, one cache line for the whole wavefront
    rcopy[slot] = vol[inside ? (int64_t)0 : (prow + x) * D + dcl];
  };
#pragma unroll
  for (int k = 0; k < PF; k++) load(k, k);
  int buf = 0;
  for (int y0 = 0; y0 < H; y0 += PF) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const int y = y0 + k;
      if (y >= H) break;
#pragma unroll
      for (int q = 0; q < K - 1; q++) {
        ws[q] = ws[q + 1];
        wc[q] = wc[q + 1];
      }
      const bool nok = inside && y + L1 - 1 < H;
      ws[K - 1] = nok ? rvs[k] : 0.f;
      wc[K - 1] = nok ? (int)rvc[k] : 0;
      float res;
      if (inside) {
        const int ys = (int)fmaxf(rl2[k], rr2[k]), yt = (int)fminf(rl3[k], rr3[k]);
        float sum = 0.f;
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < K; q++) {
          const int yy = y - (L1 - 1) + q;
          if (yy > ys && yy < yt) {
            sum += ws[q];
            cnt += wc[q];
          }
        }
        res = sum / cnt;
      } else {
        res = dok ? rcopy[k] : INFINITY;
      }
      const int64_t pix = (int64_t)y * W + x;
      if (!WTA) {
        if (dok) out[pix * D + d] = res;
      } else {
        vals[buf][d] = res;
        float bv = res;
        int bi = dok ? d : 0x7fffffff;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
          const float ov = __shfl_xor(bv, m, 64);
          const int oi = __shfl_xor(bi, m, 64);
          if (ov < bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
          }
        }
        if (lane == 0) {
          rmin[buf][wave] = bv;
          ridx[buf][wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
          float b = rmin[buf][0];
          int bidx = ridx[buf][0];
          for (int w = 1; w < nw; w++)
            if (rmin[buf][w] < b || (rmin[buf][w] == b && ridx[buf][w] < bidx)) {
              b = rmin[buf][w];
              bidx = ridx[buf][w];
            }
          wta[pix] = (float)bidx;
          float o = (float)bidx;
          if (1 <= bidx && bidx < D - 1) {
            const float cn = vals[buf][bidx - 1], cz = vals[buf][bidx], cp = vals[buf][bidx + 1];
            const float denom = 2 * (cp + cn - 2 * cz);
            if ((double)denom > 1e-5) {
              const double r = fmin(1.0, fmax(-1.0, (double)((cp - cn) / denom)));
              o = (float)((double)bidx - r);
            }
          }
          sub[pix] = o;
        }
        buf ^= 1;
      }
      load(k, y + PF);
    }
  }
}

// WTA (+ subpixel) of a volume without aggregation (cbca_i2 = 0).
__global__ void k_wta(const float* __restrict__ vol, int H, int W, int D, float* __restrict__ wta,
                      float* __restrict__ sub) {
  __shared__ float vals[1024];
  __shared__ float rmin[16];
  __shared__ int ridx[16];
  const int d = threadIdx.x;
  const int64_t pix = (int64_t)blockIdx.y * W + blockIdx.x;
  const float res = d < D ? vol[pix * D + d] : INFINITY;
  vals[d] = res;
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

// SGM output of the last path is divided by the number of paths (test.py:382).
__global__ void k_scale(float* __restrict__ a, int64_t n, float s) {
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
    a[i] *= s;
}

__global__ void k_zero(float* __restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
    a[i] = 0.f;
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
  cbmv_params p{};
  int rows_per_batch = 1;
  // matchers
  DevBuf descL, descR, AL, AR, CL, CR, ML, MR, sobL, sobR;
  DevBuf census, ncc, sob, sad;
  DevBuf statsL, statsR;
  DevBuf statsF;      // single-precision statistics for the MLP [16][H*W]
  // classifier
  DevBuf forest_nodes, forest_leaf0, forest_leaf1, forest_roots;
  int forest_trees = 0;
  bool have_forest = false;
  MlpDevice mlp;
  bool have_mlp = false;
  DevBuf pmax;        // MLP best p(match) per pixel [H,W]
  DevBuf gate;        // per-pixel gate flags / list
  // post
  DevBuf imgL, imgR, imgRp, part, lcross, rcross;
  DevBuf cost, costR, volA, volB, hsum, hcnt;
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
  RC(launch_census_desc(L, H, W, p.census_w, pl->descL.as<uint64_t>(), s));
  RC(launch_census_desc(R, H, W, p.census_w, pl->descR.as<uint64_t>(), s));
  RC(launch_census_cost_u8(pl->descL.as<uint64_t>(), pl->descR.as<uint64_t>(), H, W, D, p.census_w,
                           pl->census.as<uint8_t>(), st, s));
  RC(launch_window_stats(L, H, W, p.ncc_w, pl->AL.as<uint32_t>(), pl->CL.as<double>(), nullptr, s));
  RC(launch_window_stats(R, H, W, p.ncc_w, pl->AR.as<uint32_t>(), pl->CR.as<double>(), nullptr, s));
  RC(launch_ncc_cost(L, R, pl->AL.as<uint32_t>(), pl->AR.as<uint32_t>(), pl->CL.as<double>(),
                     pl->CR.as<double>(), H, W, D, p.ncc_w, pl->ncc.as<double>(), st, s));
  RC(launch_sobel_i16(L, H, W, pl->sobL.as<int16_t>(), s));
  RC(launch_sobel_i16(R, H, W, pl->sobR.as<int16_t>(), s));
  RC(launch_sadsob_u16(pl->sobL.as<int16_t>(), pl->sobR.as<int16_t>(), H, W, D, p.sob_w,
                       pl->sob.as<uint16_t>(), st, s));
  RC(launch_window_stats(L, H, W, p.sad_w, nullptr, nullptr, pl->ML.as<double>(), s));
  RC(launch_window_stats(R, H, W, p.sad_w, nullptr, nullptr, pl->MR.as<double>(), s));
  RC(launch_zsad(L, R, pl->ML.as<double>(), pl->MR.as<double>(), H, W, D, p.sad_w,
                 pl->sad.as<double>(), st, s));
  return CBMV_OK;
}

// stage 2
static int run_stats(cbmv_pipeline* pl, hipStream_t s) {
  const int H = pl->H, W = pl->W, D = pl->D;
  k_stats<<<dim3((W + kB - 1) / kB, H), kB, 0, s>>>(pipeline_vols(pl), H, W, D, pl->rows_per_batch,
                                                    feature_spec(), pl->statsL.as<double>(),
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
    k_features_rf<<<grid, kB, 0, s>>>(pipeline_vols(pl), pipeline_stats(pl), feature_spec(), H, W,
                                      D, pl->rows_per_batch, pipeline_forest(pl),
                                      pl->cost.as<float>());
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

// SGM iterations on vol (in place through volB), then CBCA; returns the final
// [H,W,D] volume in *final (or WTA directly when fuse_wta).
static int run_chain(cbmv_pipeline* pl, const float* cost_in, const float* arms_a,
                     const float* arms_b, int dir, float* wta, float* sub, hipStream_t s) {
  const int H = pl->H, W = pl->W, D = pl->D;
  const cbmv_params& p = pl->p;
  const int64_t n = (int64_t)H * W * D;
  const VolStrides st = strides_hwd(H, W, D);
  const dim3 pgrid(W, H), pblock((D + 63) / 64 * 64);
  float* A = pl->volA.as<float>();
  float* B = pl->volB.as<float>();
  const float* cur = cost_in;
  // cbca_i1 aggregations before SGM (test.py:182-184, 239-241)
  for (int i = 0; i < p.cbca_i1; i++) {
    float* dst = (cur == A) ? B : A;
    k_cbca_h<<<pgrid, pblock, 0, s>>>(cur, arms_a, arms_b, H, W, D, p.L1, dir, pl->hsum.as<float>(),
                                      pl->hcnt.as<uint8_t>());
    k_cbca_v<false><<<pgrid, pblock, 0, s>>>(cur, pl->hsum.as<float>(), pl->hcnt.as<uint8_t>(),
                                             arms_a, arms_b, H, W, D, p.L1, dir, dst, nullptr, nullptr);
    CBMV_LAUNCH_CHECK();
    cur = dst;
  }
  // SGM iterations: zeroed sum, paths, divide by the path count
  // (test.py:378-383); the first path writes instead of adding to zero and
  // the last one scales, so no separate clearing or dividing pass.
  for (int it = 0; it < p.sgm_iters; it++) {
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
    k_wta<<<pgrid, pblock, 0, s>>>(cur, H, W, D, wta, sub);
    CBMV_LAUNCH_CHECK();
    return CBMV_OK;
  }
  const bool sliding = p.L1 == 6;
  for (int i = 0; i < p.cbca_i2; i++) {
    const bool last = i == p.cbca_i2 - 1;
    float* dst = (cur == A) ? B : A;
    if (sliding) {
      const int nwork = H * ((D + 63) / 64);
      k_cbca_rows<6, 8><<<(nwork + 3) / 4, 256, 0, s>>>(cur, arms_a, arms_b, H, W, D, dir,
                                                         pl->hsum.as<float>(),
                                                         pl->hcnt.as<uint8_t>(), nwork);
      if (last)
        k_cbca_cols<6, 4, true><<<W, pblock, 0, s>>>(cur, pl->hsum.as<float>(),
                                                     pl->hcnt.as<uint8_t>(), arms_a, arms_b, H, W,
                                                     D, dir, nullptr, wta, sub);
      else
        k_cbca_cols<6, 4, false><<<W, pblock, 0, s>>>(cur, pl->hsum.as<float>(),
                                                      pl->hcnt.as<uint8_t>(), arms_a, arms_b, H,
                                                      W, D, dir, dst, nullptr, nullptr);
      CBMV_LAUNCH_CHECK();
      cur = dst;
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
  const int nparts = 256;
  k_img_partials<<<nparts, kB, 0, s>>>(L, np, pl->part.as<double>());
  k_img_normalise<<<1024, kB, 0, s>>>(L, np, pl->part.as<double>(), nparts, pl->imgL.as<float>());
  k_img_partials<<<nparts, kB, 0, s>>>(R, np, pl->part.as<double>() + 2 * nparts);
  k_img_normalise<<<1024, kB, 0, s>>>(R, np, pl->part.as<double>() + 2 * nparts, nparts,
                                      pl->imgR.as<float>());
  CBMV_LAUNCH_CHECK();
  RC(launch_pad_rows(pl->imgR.as<float>(), H, W, pl->imgRp.as<float>(), s));
  RC(launch_cross(pl->imgL.as<float>(), H, W, p.L1, p.tau1, pl->lcross.as<float>(), s));
  RC(launch_cross(pl->imgR.as<float>(), H, W, p.L1, p.tau1, pl->rcross.as<float>(), s));
  const int direction = -1;  // do_post (test.py:499,501)
  if (p.lr_check) {
    k_right_cost<<<dim3(W, H), (D + 63) / 64 * 64, 0, s>>>(pl->cost.as<float>(), H, W, D,
                                                          pl->costR.as<float>());
    CBMV_LAUNCH_CHECK();
    RC(run_chain(pl, pl->costR.as<float>(), pl->rcross.as<float>(), pl->lcross.as<float>(),
                 -direction, pl->wtaR.as<float>(), pl->subR.as<float>(), s));
  }
  RC(run_chain(pl, pl->cost.as<float>(), pl->lcross.as<float>(), pl->rcross.as<float>(), direction,
               pl->wtaL.as<float>(), pl->subL.as<float>(), s));
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
  auto* pl = new cbmv_pipeline();
  pl->H = (int)H;
  pl->W = (int)W;
  pl->D = D;
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
  rc |= pl->statsR.alloc(np * 8 * 8);
  rc |= pl->statsF.alloc(np * 16 * 4);
  rc |= pl->imgL.alloc(np * 4);
  rc |= pl->imgR.alloc(np * 4);
  rc |= pl->imgRp.alloc((size_t)H * (W + 2 * sgm_pad_columns()) * 4);
  rc |= pl->part.alloc(4 * 256 * 8);
  rc |= pl->lcross.alloc(4 * np * 4);
  rc |= pl->rcross.alloc(4 * np * 4);
  rc |= pl->cost.alloc(nv * 4);
  if (p.lr_check) rc |= pl->costR.alloc(nv * 4);
  rc |= pl->volA.alloc(nv * 4);
  rc |= pl->volB.alloc(nv * 4);
  rc |= pl->hsum.alloc(nv * 4);
  rc |= pl->hcnt.alloc(nv);
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
    default: set_error("unknown buffer %d", which); return CBMV_EINVAL;
  }
  *ptr = b->p;
  *nbytes = (int64_t)b->bytes;
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

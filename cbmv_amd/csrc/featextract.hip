// Drop-in feature-extraction entry points (libfeatextract,
// cpp/featextract/featextract.cpp:519-540) on HIP: host float64 arrays in and
// out, one kernel per call.
//
// Per-row reductions (minimum, sum of exponentials) run one thread per row
// walking d in ascending order, the reference's order, so the float64 sums
// round the same way (featextract.cpp:322-327,373-383).
#include "ops.hpp"

namespace cbmv {
namespace {

constexpr int kTile = 64;

// out[c][r] = in[r][c] for an rows x cols matrix, through a padded LDS tile.
__global__ void k_transpose_f64(const double* __restrict__ in, double* __restrict__ out,
                                int64_t rows, int64_t cols) {
  __shared__ double tile[kTile][kTile + 1];
  const int64_t ntc = (cols + kTile - 1) / kTile;
  const int64_t c0 = ((int64_t)blockIdx.x % ntc) * kTile, r0 = ((int64_t)blockIdx.x / ntc) * kTile;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 64 x 4
  for (int k = ty; k < kTile; k += 4) {
    const int64_t r = r0 + k, c = c0 + tx;
    if (r < rows && c < cols) tile[k][tx] = in[r * cols + c];
  }
  __syncthreads();
  for (int k = ty; k < kTile; k += 4) {
    const int64_t c = c0 + k, r = r0 + tx;
    if (r < rows && c < cols) out[c * rows + r] = tile[tx][k];
  }
}

__global__ void k_get_cost(const double* __restrict__ in, double* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[2 * i];
}

// sign = +1: out[y,x,d] = in[y,x+d,d] for x+d < W (get_right_cost);
// sign = -1: out[y,x,d] = in[y,x-d,d] for x >= d (get_left_cost_from_right);
// otherwise in[0].
__global__ void k_reindex(const double* __restrict__ in, double* __restrict__ out, int64_t H,
                          int64_t W, int64_t D, int sign) {
  const double fill = in[0];
  const int64_t n = H * W * D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = i % D;
    const int64_t x = (i / D) % W;
    const int64_t xs = x + sign * d;
    out[i] = (xs >= 0 && xs < W) ? in[i + (xs - x) * D] : fill;
  }
}

__device__ __forceinline__ double row_min(const double* v, int64_t D) {
  double m = kRandMax;
  for (int64_t k = 0; k < D; k++)
    if (v[k] < m) m = v[k];
  return m;
}

__device__ __forceinline__ double row_den(const double* v, int64_t D, double m, double sigma) {
  double den = 0;
  for (int64_t k = 0; k < D; k++) {
    const double num = v[k] - m;
    den += exp(-(num * num) / sigma);
  }
  return den;
}

__global__ void k_ratio(const double* __restrict__ vol, int64_t N, int64_t D, double e,
                        double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* v = vol + i * D;
  const double m = row_min(v, D);
  for (int64_t j = 0; j < D; j++) out[i * D + j] = (m == kRandMax) ? 0.0 : (m + e) / (v[j] + e);
}

__global__ void k_ratio_samples(const double* __restrict__ vol, int64_t N, int64_t D,
                                const int32_t* __restrict__ samp, double e,
                                double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* v = vol + i * D;
  const double m = row_min(v, D);
  for (int k = 0; k < 3; k++)
    out[3 * i + k] = (m == kRandMax) ? 0.0 : (m + e) / (v[samp[3 * i + k]] + e);
}

__global__ void k_aml(const double* __restrict__ vol, int64_t N, int64_t D, double sigma,
                      double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* v = vol + i * D;
  const double m = row_min(v, D);
  const double den = row_den(v, D, m, sigma);
  for (int64_t j = 0; j < D; j++) {
    const double t = v[j] - m;
    out[i * D + j] = (m == kRandMax) ? 0.0 : exp(-((t * t) / sigma)) / den;
  }
}

__global__ void k_aml_samples(const double* __restrict__ vol, int64_t N, int64_t D,
                              const int32_t* __restrict__ samp, double sigma,
                              double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* v = vol + i * D;
  const double m = row_min(v, D);
  const double den = row_den(v, D, m, sigma);
  for (int k = 0; k < 3; k++) {
    const double t = v[samp[3 * i + k]] - m;
    out[3 * i + k] = (m == kRandMax) ? 0.0 : exp(-((t * t) / sigma)) / den;
  }
}

__global__ void k_get_samples(const double* __restrict__ vol, int64_t N, int64_t D,
                              const int32_t* __restrict__ samp, double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int k = 0; k < 3; k++) out[3 * i + k] = vol[i * D + samp[3 * i + k]];
}

__global__ void k_labels(int64_t n, int32_t* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[3 * i] = 1;
  out[3 * i + 1] = 0;
  out[3 * i + 2] = 0;
}

__device__ __forceinline__ uint32_t draw(uint64_t seed, uint64_t ctr) {
  // splitmix64 finaliser on a (seed, counter) key
  uint64_t z = seed + (ctr + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (uint32_t)(z >> 33);  // 31 random bits, like rand()
}

// featextract.cpp:183-209 with the draws keyed by (seed, 2*s + k).
__global__ void k_d_indices(const int32_t* __restrict__ gt, int64_t n, int32_t maxd, int32_t th,
                            uint64_t seed, int32_t* __restrict__ out, int* __restrict__ bad) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int g = gt[s];
  if (g < 0 || g > maxd) {
    atomicOr(bad, 1);
    return;
  }
  const bool low = !(g - th <= 0);
  const bool high = !(g + th >= maxd);
  const int lo_n = g - th;                  // [0, g - th)
  const int hi_0 = g + th + 1;              // [g + th + 1, maxd)
  const int hi_n = maxd - hi_0;
  int d1, d2;
  if (low) {
    d1 = (int)(draw(seed, 2 * s) % (uint32_t)lo_n);
  } else {
    if (hi_n <= 0) { atomicOr(bad, 2); return; }
    d1 = (int)(draw(seed, 2 * s) % (uint32_t)hi_n) + hi_0;
  }
  if (high) {
    if (hi_n <= 0) { atomicOr(bad, 2); return; }
    d2 = (int)(draw(seed, 2 * s + 1) % (uint32_t)hi_n) + hi_0;
  } else {
    if (lo_n <= 0) { atomicOr(bad, 2); return; }
    d2 = (int)(draw(seed, 2 * s + 1) % (uint32_t)lo_n);
  }
  out[3 * s] = g;
  out[3 * s + 1] = d1;
  out[3 * s + 2] = d2;
}

constexpr int kB = 256;

}  // namespace
}  // namespace cbmv

using namespace cbmv;

#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

static int transpose(const double* in, int64_t rows, int64_t cols, double* out) {
  CBMV_REQUIRE(rows > 0 && cols > 0, "empty array");
  hipStream_t s = lib_stream();
  const size_t n = (size_t)rows * cols * sizeof(double);
  DevBuf di, dout;
  RC(to_device(di, in, n));
  RC(dout.alloc(n));
  const int64_t tiles = ((cols + kTile - 1) / kTile) * ((rows + kTile - 1) / kTile);
  CBMV_REQUIRE(tiles < 0x7fffffff, "array too large for the transpose grid");
  k_transpose_f64<<<dim3((unsigned)tiles), dim3(kTile, 4), 0, s>>>(di.as<double>(), dout.as<double>(), rows, cols);
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, n);
}

static int reindex(const double* in, int64_t H, int64_t W, int64_t D, int sign, double* out) {
  CBMV_REQUIRE(H > 0 && W > 0 && D > 0, "empty volume");
  hipStream_t s = lib_stream();
  const size_t n = (size_t)H * W * D;
  DevBuf di, dout;
  RC(to_device(di, in, n * 8));
  RC(dout.alloc(n * 8));
  k_reindex<<<grid_1d(n, kB), kB, 0, s>>>(di.as<double>(), dout.as<double>(), H, W, D, sign);
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, n * 8);
}

static int check_samples(const int32_t* samp, int64_t N, int64_t D) {
  for (int64_t i = 0; i < 3 * N; i++)
    CBMV_REQUIRE(samp[i] >= 0 && samp[i] < D, "sample index %d outside [0, %lld)", samp[i],
                 (long long)D);
  return CBMV_OK;
}

extern "C" {

int cbmv_swap_axes(const double* in_dhw, int64_t D, int64_t H, int64_t W, double* out_hwd) {
  clear_error();
  return transpose(in_dhw, D, H * W, out_hwd);
}

int cbmv_swap_axes_back(const double* in_hwd, int64_t H, int64_t W, int64_t D, double* out_dhw) {
  clear_error();
  return transpose(in_hwd, H * W, D, out_dhw);
}

int cbmv_get_cost(const double* prob, int64_t H, int64_t W, int64_t D, double* out) {
  clear_error();
  CBMV_REQUIRE(H > 0 && W > 0 && D > 0, "empty volume");
  hipStream_t s = lib_stream();
  const size_t n = (size_t)H * W * D;
  DevBuf di, dout;
  RC(to_device(di, prob, 2 * n * 8));
  RC(dout.alloc(n * 8));
  k_get_cost<<<grid_1d(n, kB), kB, 0, s>>>(di.as<double>(), dout.as<double>(), (int64_t)n);
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, n * 8);
}

int cbmv_get_right_cost(const double* in, int64_t H, int64_t W, int64_t D, double* out) {
  clear_error();
  return reindex(in, H, W, D, +1, out);
}

int cbmv_get_left_cost(const double* in, int64_t H, int64_t W, int64_t D, double* out) {
  clear_error();
  return reindex(in, H, W, D, -1, out);
}

int cbmv_generate_d_indices(const int32_t* gt, int64_t n, int32_t maxd, int32_t thresh,
                            uint64_t seed, int32_t* out_n3) {
  clear_error();
  CBMV_REQUIRE(n >= 0, "negative sample count");
  if (n == 0) return CBMV_OK;
  hipStream_t s = lib_stream();
  DevBuf dg, dout, dbad;
  RC(to_device(dg, gt, n * 4));
  RC(dout.alloc(n * 12));
  RC(dbad.alloc(4));
  CBMV_HIP(hipMemsetAsync(dbad.p, 0, 4, s));
  k_d_indices<<<grid_1d(n, kB), kB, 0, s>>>(dg.as<int32_t>(), n, maxd, thresh, seed,
                                            dout.as<int32_t>(), dbad.as<int>());
  CBMV_LAUNCH_CHECK();
  int bad = 0;
  RC(to_host(&bad, dbad, 4));
  CBMV_REQUIRE(!(bad & 1), "ground-truth disparity outside [0, maxd=%d]", maxd);
  CBMV_REQUIRE(!(bad & 2),
               "empty negative-sample range (e.g. gt == maxd-2 with thresh 1, where the "
               "reference divides by zero)");
  return to_host(out_n3, dout, n * 12);
}

int cbmv_get_samples(const double* vol, int64_t N, int64_t D, const int32_t* samp, double* out) {
  clear_error();
  CBMV_REQUIRE(N > 0 && D > 0, "empty volume");
  RC(check_samples(samp, N, D));
  hipStream_t s = lib_stream();
  DevBuf dv, ds, dout;
  RC(to_device(dv, vol, N * D * 8));
  RC(to_device(ds, samp, N * 12));
  RC(dout.alloc(N * 24));
  k_get_samples<<<grid_1d(N, kB), kB, 0, s>>>(dv.as<double>(), N, D, ds.as<int32_t>(), dout.as<double>());
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, N * 24);
}

int cbmv_extract_ratio(const double* vol, int64_t N, int64_t D, double e, double* out) {
  clear_error();
  CBMV_REQUIRE(N > 0 && D > 0, "empty volume");
  hipStream_t s = lib_stream();
  DevBuf dv, dout;
  RC(to_device(dv, vol, N * D * 8));
  RC(dout.alloc(N * D * 8));
  k_ratio<<<grid_1d(N, 64), 64, 0, s>>>(dv.as<double>(), N, D, e, dout.as<double>());
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, N * D * 8);
}

int cbmv_extract_ratio_samples(const double* vol, int64_t N, int64_t D, const int32_t* samp,
                               double e, double* out) {
  clear_error();
  CBMV_REQUIRE(N > 0 && D > 0, "empty volume");
  RC(check_samples(samp, N, D));
  hipStream_t s = lib_stream();
  DevBuf dv, ds, dout;
  RC(to_device(dv, vol, N * D * 8));
  RC(to_device(ds, samp, N * 12));
  RC(dout.alloc(N * 24));
  k_ratio_samples<<<grid_1d(N, 64), 64, 0, s>>>(dv.as<double>(), N, D, ds.as<int32_t>(), e, dout.as<double>());
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, N * 24);
}

int cbmv_extract_likelihood(const double* vol, int64_t N, int64_t D, double sigma, double* out) {
  clear_error();
  CBMV_REQUIRE(N > 0 && D > 0, "empty volume");
  hipStream_t s = lib_stream();
  DevBuf dv, dout;
  RC(to_device(dv, vol, N * D * 8));
  RC(dout.alloc(N * D * 8));
  k_aml<<<grid_1d(N, 64), 64, 0, s>>>(dv.as<double>(), N, D, sigma, dout.as<double>());
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, N * D * 8);
}

int cbmv_extract_likelihood_samples(const double* vol, int64_t N, int64_t D,
                                    const int32_t* samp, double sigma, double* out) {
  clear_error();
  CBMV_REQUIRE(N > 0 && D > 0, "empty volume");
  RC(check_samples(samp, N, D));
  hipStream_t s = lib_stream();
  DevBuf dv, ds, dout;
  RC(to_device(dv, vol, N * D * 8));
  RC(to_device(ds, samp, N * 12));
  RC(dout.alloc(N * 24));
  k_aml_samples<<<grid_1d(N, 64), 64, 0, s>>>(dv.as<double>(), N, D, ds.as<int32_t>(), sigma, dout.as<double>());
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, N * 24);
}

int cbmv_generate_labels(int64_t n, int32_t* out) {
  clear_error();
  CBMV_REQUIRE(n >= 0, "negative sample count");
  if (n == 0) return CBMV_OK;
  hipStream_t s = lib_stream();
  DevBuf dout;
  RC(dout.alloc(n * 12));
  k_labels<<<grid_1d(n, kB), kB, 0, s>>>(n, dout.as<int32_t>());
  CBMV_LAUNCH_CHECK();
  return to_host(out, dout, n * 12);
}

}  // extern "C"

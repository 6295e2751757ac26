// Post-processing (cpp/post/post.cu rewritten for CDNA): cross arms, SGM,
// cross-based aggregation, subpixel refinement, median and weighted-mean
// filters, left-right check and the two interpolations; and the drop-in
// `post` module's device arrays.
//
// Arithmetic follows the reference's types: float images and float path
// state in SGM with the path value formed in double (post.cu:217-260), float
// accumulation in CBCA (post.cu:412-429), float filters.
#include <cmath>

#include "ops.hpp"

namespace cbmv {
namespace {

constexpr int kB = 256;

// ------------------------------------------------------------------ cross
// post.cu:336-382. out[dir][y][x], dir 0..3 = -x, +x, -y, +y.
__global__ void k_cross(const float* __restrict__ img, int H, int W, int L1, float tau1,
                        float* __restrict__ out) {
  const int64_t n = (int64_t)4 * H * W;
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= n) return;
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
    const int dist = max(abs(xx - x), abs(yy - y));
    if (dist == 1) continue;                                     // always accepted
    if (fabsf(c - img[(int64_t)yy * W + xx]) >= tau1) break;     // rule 1
    if (dist >= L1) break;                                       // rule 2
  }
  out[id] = (float)(dir <= 1 ? xx : yy);
}

// ------------------------------------------------------------------ order statistics
// k-th smallest (0-based) of vals[0..N), where entries beyond the valid ones
// are +inf; equals element k of the ascending sort (post.cu:46-59 sort + xs[n/2]).
template <int N>
__device__ __forceinline__ float kth_smallest(const float (&v)[N], int k) {
  float res = v[0];
#pragma unroll
  for (int i = 0; i < N; i++) {
    int lt = 0, le = 0;
#pragma unroll
    for (int j = 0; j < N; j++) {
      lt += v[j] < v[i];
      le += v[j] <= v[i];
    }
    if (lt <= k && k < le) res = v[i];
  }
  return res;
}

// post.cu:486-505.
template <int R>
__global__ void k_median(const float* __restrict__ in, int H, int W, float* __restrict__ out) {
  constexpr int K = 2 * R + 1;
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)H * W) return;
  const int x = (int)(id % W), y = (int)(id / W);
  float v[K * K];
  int n = 0;
#pragma unroll
  for (int a = 0; a < K; a++)
#pragma unroll
    for (int b = 0; b < K; b++) {
      const int xx = x - R + a, yy = y - R + b;
      const bool ok = 0 <= xx && xx < W && 0 <= yy && yy < H;
      v[a * K + b] = ok ? in[(int64_t)yy * W + xx] : INFINITY;
      n += ok;
    }
  out[id] = kth_smallest<K * K>(v, n / 2);
}

// post.cu:520-540: kernel element i runs xx in the outer loop, yy inner.
__global__ void k_mean2d(const float* __restrict__ img, const float* __restrict__ kernel, int H,
                         int W, int r, float alpha2, float* __restrict__ out) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)H * W) return;
  const int x = (int)(id % W), y = (int)(id / W);
  const float c = img[id];
  float sum = 0, cnt = 0;
  int i = 0;
  for (int xx = x - r; xx <= x + r; xx++)
    for (int yy = y - r; yy <= y + r; yy++, i++) {
      if (0 <= xx && xx < W && 0 <= yy && yy < H) {
        const float v = img[(int64_t)yy * W + xx];
        if (fabsf(v - c) < alpha2) {
          sum = fmaf(v, kernel[i], sum);  // fused, as nvcc compiles post.cu:533
          cnt += kernel[i];
        }
      }
    }
  out[id] = sum / cnt;
}

// post.cu:717-739: 0 match, 1 occlusion, 2 mismatch.
__global__ void k_outlier(const float* __restrict__ d0, const float* __restrict__ d1, int H,
                          int W, int disp_max, float* __restrict__ out) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)H * W) return;
  const int x = (int)(id % W);
  const int d0i = (int)d0[id];
  float o;
  if (x - d0i < 0 || x - d0i >= W) {  // (>= W only for negative input disparities)
    o = 1;
  } else if ((double)fabsf(d0[id] - d1[id - d0i]) < 1.1) {
    o = 0;
  } else {
    o = 1;
    for (int d = 0; d < disp_max; d++)
      if (x - d >= 0 && (double)fabsf((float)d - d1[id - d]) < 1.1) {
        o = 2;
        break;
      }
  }
  out[id] = o;
}

// post.cu:757-814. With no non-mismatch pixel in any direction the input is
// kept (the reference asserts).
__global__ void k_interp_mismatch(const float* __restrict__ d0, const float* __restrict__ outl,
                                  int H, int W, float* __restrict__ out) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)H * W) return;
  if (outl[id] != 2) {
    out[id] = d0[id];
    return;
  }
  // direction table of post.cu:759-776
  const float dirs[32] = {0, 1,  -0.5f, 1,  -1, 1,  -1, 0.5f, -1, 0,  -1, -0.5f, -1, -1, -0.5f, -1,
                          0, -1, 0.5f,  -1, 1,  -1, 1,  -0.5f, 1, 0,  1,  0.5f,  1,  1,  0.5f,  1};
  const int x = (int)(id % W), y = (int)(id / W);
  float vals[16];
  int n = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const float dx = dirs[2 * k], dy = dirs[2 * k + 1];
    float xx = (float)x, yy = (float)y;
    int xi = (int)roundf(xx), yi = (int)roundf(yy);
    while (0 <= yi && yi < H && 0 <= xi && xi < W && outl[(int64_t)yi * W + xi] == 2) {
      xx += dx;
      yy += dy;
      xi = (int)roundf(xx);
      yi = (int)roundf(yy);
    }
    const bool ok = 0 <= yi && yi < H && 0 <= xi && xi < W;
    vals[k] = ok ? d0[(int64_t)yi * W + xi] : INFINITY;
    n += ok;
  }
  out[id] = n == 0 ? d0[id] : kth_smallest<16>(vals, n / 2);
}

// post.cu:829-855.
__global__ void k_interp_occlusion(const float* __restrict__ d0, const float* __restrict__ outl,
                                   int H, int W, float* __restrict__ out) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)H * W) return;
  if (outl[id] != 1) {
    out[id] = d0[id];
    return;
  }
  const int x = (int)(id % W);
  int dx = 0;
  while (x + dx >= 0 && outl[id + dx] != 0) dx--;
  if (x + dx < 0) {
    dx = 0;
    while (x + dx < W && outl[id + dx] != 0) dx++;
  }
  out[id] = (x + dx < W) ? d0[id + dx] : d0[id];
}

// ------------------------------------------------------------------ SGM padding
// Right image with kSgmPad columns on both sides. A padding cell holds
// 1e20 * (1 + (c & 1) + 2 (y & 1)): two cells at a horizontal, vertical or
// diagonal step from each other differ by at least 1e20, and so does a
// padding cell from an image value, so an SGM step whose right-image pair
// leaves the image sees a colour difference above any threshold (the
// reference's D2 = 10, post.cu:229-234) without testing the columns.
__global__ void k_pad_rows(const float* __restrict__ img, int H, int W, float* __restrict__ out) {
  const int64_t pitch = (int64_t)W + 2 * kSgmPad;
  const int64_t n = (int64_t)H * pitch;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t y = i / pitch, c = i % pitch - kSgmPad;
    out[i] = (c >= 0 && c < W) ? img[y * W + c]
                               : 1e20f * (float)(1 + (int)(c & 1) + 2 * (int)(y & 1));
  }
}

// ------------------------------------------------------------------ CBCA (drop-in, exact)
// post.cu:399-433 on a [D,H,W] volume: the mean over the support region,
// values read as float and summed in float in row-major order.
template <typename T>
__global__ void k_cbca_dhw(const float* __restrict__ lc, const float* __restrict__ rc,
                           const T* __restrict__ vol, T* __restrict__ out, int D, int H, int W,
                           int direction) {
  const int64_t HW = (int64_t)H * W;
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= (int64_t)D * HW) return;
  int64_t t = id;
  const int x = (int)(t % W);
  t /= W;
  const int y = (int)(t % H);
  const int d = (int)(t / H);
  const int dd = d * direction;
  const int xr = x + dd;
  if (xr < 0 || xr >= W) {
    out[id] = vol[id];
    return;
  }
  // arm ends produced by cross() lie in [-1, H] / [-1, W]; clamping only
  // matters for arrays that did not come from cross()
  const int yy_s = max((int)fmaxf(lc[(2 * (int64_t)H + y) * W + x], rc[(2 * (int64_t)H + y) * W + xr]), -1);
  const int yy_t = min((int)fminf(lc[(3 * (int64_t)H + y) * W + x], rc[(3 * (int64_t)H + y) * W + xr]), H);
  float sum = 0;
  int cnt = 0;
  for (int yy = yy_s + 1; yy < yy_t; yy++) {
    const int xx_s = max((int)fmaxf(lc[(int64_t)yy * W + x], rc[(int64_t)yy * W + xr] - dd), -1);
    const int xx_t = min((int)fminf(lc[(HW + (int64_t)yy * W) + x], rc[(HW + (int64_t)yy * W) + xr] - dd), W);
    const T* row = vol + (int64_t)d * HW + (int64_t)yy * W;
    for (int xx = xx_s + 1; xx < xx_t; xx++) {
      sum += (float)row[xx];
      cnt++;
    }
  }
  out[id] = (T)(sum / cnt);
}

// post.cu:453-468 on a [D,H,W] cost.
template <typename T>
__global__ void k_subpixel_dhw(const float* __restrict__ d0, const T* __restrict__ c2,
                               float* __restrict__ out, int64_t HW, int D) {
  const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (id >= HW) return;
  const int d = (int)d0[id];
  float o = (float)d;
  if (1 <= d && d < D - 1) {
    const float cn = (float)c2[(int64_t)(d - 1) * HW + id];
    const float cz = (float)c2[(int64_t)d * HW + id];
    const float cp = (float)c2[(int64_t)(d + 1) * HW + id];
    const float denom = 2 * (cp + cn - 2 * cz);
    if ((double)denom > 1e-5) {
      const double r = fmin(1.0, fmax(-1.0, (double)((cp - cn) / denom)));
      o = (float)((double)d - r);
    }
  }
  out[id] = o;
}

template <typename T>
__global__ void k_copy(const T* __restrict__ in, T* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

__global__ void k_divide(double* __restrict__ a, double val, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] /= val;
}

// out[c][r] = in[r][c] (rows x cols), padded LDS tile; 1-D grid of tiles.
__global__ void k_transpose_dev_f64(const double* __restrict__ in, double* __restrict__ out,
                                    int64_t rows, int64_t cols) {
  __shared__ double tile[64][65];
  const int64_t ntc = (cols + 63) / 64;
  const int64_t c0 = ((int64_t)blockIdx.x % ntc) * 64, r0 = ((int64_t)blockIdx.x / ntc) * 64;
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int k = ty; k < 64; k += 4) {
    const int64_t r = r0 + k, c = c0 + tx;
    if (r < rows && c < cols) tile[k][tx] = in[r * cols + c];
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int64_t c = c0 + k, r = r0 + tx;
    if (r < rows && c < cols) out[c * rows + r] = tile[tx][k];
  }
}

}  // namespace

// ------------------------------------------------------------------ launchers
int launch_cross(const float* img, int H, int W, int L1, float tau1, float* out, hipStream_t s) {
  const int64_t n = (int64_t)4 * H * W;
  k_cross<<<grid_1d(n, kB), kB, 0, s>>>(img, H, W, L1, tau1, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_median2d(const float* in, int H, int W, int ksize, float* out, hipStream_t s) {
  const int r = ksize / 2;
  const int64_t n = (int64_t)H * W;
  switch (r) {
    case 0: k_median<0><<<grid_1d(n, kB), kB, 0, s>>>(in, H, W, out); break;
    case 1: k_median<1><<<grid_1d(n, kB), kB, 0, s>>>(in, H, W, out); break;
    case 2: k_median<2><<<grid_1d(n, kB), kB, 0, s>>>(in, H, W, out); break;
    case 3: k_median<3><<<grid_1d(n, kB), kB, 0, s>>>(in, H, W, out); break;
    case 4: k_median<4><<<grid_1d(n, kB), kB, 0, s>>>(in, H, W, out); break;
    case 5: k_median<5><<<grid_1d(n, kB), kB, 0, s>>>(in, H, W, out); break;
    default: set_error("median window %d larger than 11 (post.cu:493)", ksize); return CBMV_EINVAL;
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_mean2d(const float* in, int H, int W, const float* kernel, int ksize, float alpha2,
                  float* out, hipStream_t s) {
  const int64_t n = (int64_t)H * W;
  k_mean2d<<<grid_1d(n, kB), kB, 0, s>>>(in, kernel, H, W, ksize / 2, alpha2, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_outlier(const float* d0, const float* d1, int H, int W, int disp_max, float* out,
                   hipStream_t s) {
  k_outlier<<<grid_1d((int64_t)H * W, kB), kB, 0, s>>>(d0, d1, H, W, disp_max, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_interp_mismatch(const float* d0, const float* outlier, int H, int W, float* out,
                           hipStream_t s) {
  k_interp_mismatch<<<grid_1d((int64_t)H * W, kB), kB, 0, s>>>(d0, outlier, H, W, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_interp_occlusion(const float* d0, const float* outlier, int H, int W, float* out,
                            hipStream_t s) {
  k_interp_occlusion<<<grid_1d((int64_t)H * W, kB), kB, 0, s>>>(d0, outlier, H, W, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_pad_rows(const float* img, int H, int W, float* out, hipStream_t s) {
  const int64_t n = (int64_t)H * (W + 2 * kSgmPad);
  k_pad_rows<<<grid_1d(n, kB) > 8192 ? 8192 : grid_1d(n, kB), kB, 0, s>>>(img, H, W, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int sgm_pad_columns() { return kSgmPad; }

}  // namespace cbmv

// ===========================================================================
// Drop-in `post` module: device arrays and one entry point per function.
// ===========================================================================
using namespace cbmv;

struct cbmv_array {
  int32_t dtype;
  int32_t ndims;
  int64_t shape[4];
  int64_t count;  // elements allocated
  void* d;
};

#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

static size_t esize(const cbmv_array* a) { return a->dtype == CBMV_F64 ? 8 : 4; }

static int live(const cbmv_array* a, const char* what) {
  CBMV_REQUIRE(a != nullptr, "%s: null array", what);
  if (!a->d) {
    set_error("%s: array memory was freed", what);
    return CBMV_ESTATE;
  }
  return CBMV_OK;
}

static int need(const cbmv_array* a, int dtype, int ndims, const char* what) {
  RC(live(a, what));
  CBMV_REQUIRE(a->dtype == dtype, "%s: expected a %s array", what,
               dtype == CBMV_F64 ? "float64 (CUDA_double_array)" : "float32 (CUDA_float_array)");
  CBMV_REQUIRE(ndims == 0 || a->ndims == ndims, "%s: expected %d dimensions, got %d", what, ndims,
               a->ndims);
  return CBMV_OK;
}

static int64_t nel(const cbmv_array* a) {
  int64_t n = 1;
  for (int i = 0; i < a->ndims; i++) n *= a->shape[i];
  return n;
}

extern "C" {

cbmv_array* cbmv_array_create(int32_t dtype, int32_t ndims, const int64_t* shape,
                              const void* host_data) {
  clear_error();
  if ((dtype != CBMV_F32 && dtype != CBMV_F64) || ndims < 1 || ndims > 4) {
    set_error("array: dtype must be float32/float64 and 1..4 dimensions");
    return nullptr;
  }
  cbmv_array* a = new cbmv_array{dtype, ndims, {0, 0, 0, 0}, 1, nullptr};
  for (int i = 0; i < ndims; i++) {
    if (shape[i] < 0) {
      delete a;
      set_error("array: negative dimension");
      return nullptr;
    }
    a->shape[i] = shape[i];
    a->count *= shape[i];
  }
  const size_t bytes = (size_t)a->count * esize(a);
  if (hipMalloc(&a->d, bytes ? bytes : 1) != hipSuccess) {
    set_error("array: hipMalloc of %zu bytes failed", bytes);
    delete a;
    return nullptr;
  }
  if (bytes && host_data &&
      hipMemcpy(a->d, host_data, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("array: host-to-device copy failed");
    (void)hipFree(a->d);
    delete a;
    return nullptr;
  }
  return a;
}

int cbmv_array_dtype(const cbmv_array* a) { return a ? a->dtype : CBMV_EINVAL; }
int cbmv_array_ndims(const cbmv_array* a) { return a ? a->ndims : CBMV_EINVAL; }

int cbmv_array_shape(const cbmv_array* a, int64_t* shape_out) {
  CBMV_REQUIRE(a != nullptr, "null array");
  for (int i = 0; i < a->ndims; i++) shape_out[i] = a->shape[i];
  return CBMV_OK;
}

int cbmv_array_get(const cbmv_array* a, void* host_out) {
  clear_error();
  RC(live(a, "get_array"));
  RC(sync());
  const int64_t n = nel(a);
  CBMV_REQUIRE(n <= a->count, "get_array: logical shape exceeds the allocation");
  CBMV_HIP(hipMemcpy(host_out, a->d, (size_t)n * esize(a), hipMemcpyDeviceToHost));
  return CBMV_OK;
}

int cbmv_array_set_to_zero(cbmv_array* a) {
  clear_error();
  RC(live(a, "set_to_zero"));
  CBMV_HIP(hipMemsetAsync(a->d, 0, (size_t)a->count * esize(a), lib_stream()));
  return sync();
}

int cbmv_array_free(cbmv_array* a) {
  clear_error();
  CBMV_REQUIRE(a != nullptr, "null array");
  if (a->d) {
    RC(sync());
    (void)hipFree(a->d);
    a->d = nullptr;
  }
  return CBMV_OK;
}

void cbmv_array_destroy(cbmv_array* a) {
  if (!a) return;
  if (a->d) {
    (void)hipStreamSynchronize(lib_stream());
    (void)hipFree(a->d);
  }
  delete a;
}

int cbmv_post_sgm(cbmv_array* left, cbmv_array* right, cbmv_array* in_cost, cbmv_array* out_cost,
                  cbmv_array* tmp, float pi1, float pi2, float tau_so, float alpha1, float q1,
                  float q2, int32_t direction) {
  (void)tmp;  // the path state lives in LDS
  clear_error();
  RC(need(left, CBMV_F32, 2, "sgm"));
  RC(need(right, CBMV_F32, 2, "sgm"));
  RC(need(in_cost, CBMV_F64, 3, "sgm"));
  RC(need(out_cost, CBMV_F64, 3, "sgm"));
  const int H = (int)in_cost->shape[0], W = (int)in_cost->shape[1], D = (int)in_cost->shape[2];
  CBMV_REQUIRE(left->shape[0] == H && left->shape[1] == W && right->shape[0] == H &&
                   right->shape[1] == W,
               "sgm: image and cost shapes differ");
  CBMV_REQUIRE(out_cost->count >= nel(in_cost), "sgm: output array too small");
  hipStream_t s = lib_stream();
  DevBuf padded;
  RC(padded.alloc((size_t)H * (W + 2 * kSgmPad) * 4));
  RC(launch_pad_rows((const float*)right->d, H, W, padded.as<float>(), s));
  // every path adds into dout, like output[...] += val in post.cu:212,259
  for (int p = 0; p < 4; p++)
    RC(launch_sgm_path_f64((const float*)left->d, padded.as<float>(), (const double*)in_cost->d,
                           (double*)out_cost->d, H, W, D, strides_hwd(H, W, D), pi1, pi2, tau_so,
                           alpha1, q1, q2, direction, p, 0, 1.0, s));
  return sync();
}

int cbmv_post_cross(cbmv_array* in, cbmv_array* out, int32_t L1, float tau1) {
  clear_error();
  RC(need(in, CBMV_F32, 2, "cross"));
  RC(need(out, CBMV_F32, 3, "cross"));
  const int H = (int)out->shape[1], W = (int)out->shape[2];
  CBMV_REQUIRE(out->shape[0] == 4 && in->shape[0] == H && in->shape[1] == W,
               "cross: output must be [4,H,W] of an [H,W] image");
  RC(launch_cross((const float*)in->d, H, W, L1, tau1, (float*)out->d, lib_stream()));
  return sync();
}

int cbmv_post_cbca(cbmv_array* leftc, cbmv_array* rightc, cbmv_array* vol, cbmv_array* vol_out,
                   int32_t direction) {
  clear_error();
  RC(need(leftc, CBMV_F32, 3, "cbca"));
  RC(need(rightc, CBMV_F32, 3, "cbca"));
  RC(need(vol, CBMV_F64, 3, "cbca"));
  RC(need(vol_out, CBMV_F64, 3, "cbca"));
  CBMV_REQUIRE(direction == -1 || direction == 1, "cbca: direction must be -1 or 1");
  const int D = (int)vol->shape[0], H = (int)vol->shape[1], W = (int)vol->shape[2];
  CBMV_REQUIRE(leftc->shape[1] == H && leftc->shape[2] == W && rightc->shape[1] == H &&
                   rightc->shape[2] == W,
               "cbca: cross arrays do not match the [D,H,W] volume");
  CBMV_REQUIRE(vol_out->count >= nel(vol), "cbca: output array too small");
  const int64_t n = (int64_t)D * H * W;
  k_cbca_dhw<double><<<grid_1d(n, 256), 256, 0, lib_stream()>>>(
      (const float*)leftc->d, (const float*)rightc->d, (const double*)vol->d, (double*)vol_out->d, D,
      H, W, direction);
  CBMV_LAUNCH_CHECK();
  return sync();
}

int cbmv_post_subpixel(cbmv_array* d, cbmv_array* cost, cbmv_array* out) {
  clear_error();
  RC(need(d, CBMV_F32, 2, "subpixel_enchancement"));
  RC(need(cost, CBMV_F64, 3, "subpixel_enchancement"));
  RC(need(out, CBMV_F32, 2, "subpixel_enchancement"));
  const int64_t HW = d->shape[0] * d->shape[1];
  CBMV_REQUIRE(cost->shape[1] * cost->shape[2] == HW, "subpixel: cost is not [D,H,W]");
  k_subpixel_dhw<double><<<grid_1d(HW, 256), 256, 0, lib_stream()>>>(
      (const float*)d->d, (const double*)cost->d, (float*)out->d, HW, (int)cost->shape[0]);
  CBMV_LAUNCH_CHECK();
  return sync();
}

int cbmv_post_median2d(cbmv_array* in, cbmv_array* out, int32_t ksize) {
  clear_error();
  RC(need(in, CBMV_F32, 2, "median2d"));
  RC(need(out, CBMV_F32, 0, "median2d"));
  CBMV_REQUIRE(ksize >= 1 && ksize <= 11, "median2d: kernel size must be in [1, 11]");
  CBMV_REQUIRE(out->count >= nel(in), "median2d: output array too small");
  RC(launch_median2d((const float*)in->d, (int)in->shape[0], (int)in->shape[1], ksize,
                     (float*)out->d, lib_stream()));
  return sync();
}

int cbmv_post_mean2d(cbmv_array* in, cbmv_array* out, const float* kernel, int32_t kh, int32_t kw,
                     float alpha2) {
  clear_error();
  RC(need(in, CBMV_F32, 2, "mean2d"));
  RC(need(out, CBMV_F32, 0, "mean2d"));
  CBMV_REQUIRE(kh == kw && (kh % 2) == 1, "mean2d: the kernel must be square with odd size");
  CBMV_REQUIRE(out->count >= nel(in), "mean2d: output array too small");
  DevBuf dk;
  RC(to_device(dk, kernel, (size_t)kh * kw * 4));
  RC(launch_mean2d((const float*)in->d, (int)in->shape[0], (int)in->shape[1], dk.as<float>(), kh,
                   alpha2, (float*)out->d, lib_stream()));
  return sync();
}

static int swap(cbmv_array* in, cbmv_array* out, bool back) {
  RC(need(in, CBMV_F64, 3, back ? "swap_axis_back" : "swap_axis"));
  RC(need(out, CBMV_F64, 3, back ? "swap_axis_back" : "swap_axis"));
  CBMV_REQUIRE(out->count >= nel(in), "swap: output array too small");
  // swap_axis: [D,H,W] -> [H,W,D] = transpose of [D, HW]; back: [HW, D] -> [D, HW]
  const int64_t rows = back ? in->shape[0] * in->shape[1] : in->shape[0];
  const int64_t cols = back ? in->shape[2] : in->shape[1] * in->shape[2];
  const int64_t tiles = ((cols + 63) / 64) * ((rows + 63) / 64);
  k_transpose_dev_f64<<<dim3((unsigned)tiles), dim3(64, 4), 0, lib_stream()>>>(
      (const double*)in->d, (double*)out->d, rows, cols);
  CBMV_LAUNCH_CHECK();
  const int64_t a = in->shape[0], b = in->shape[1], c = in->shape[2];
  if (!back) {  // post.cu:612-614
    out->shape[0] = b;
    out->shape[1] = c;
    out->shape[2] = a;
  } else {      // post.cu:663-665
    out->shape[0] = c;
    out->shape[1] = a;
    out->shape[2] = b;
  }
  return sync();
}

int cbmv_post_swap_axis(cbmv_array* in, cbmv_array* out) {
  clear_error();
  return swap(in, out, false);
}

int cbmv_post_swap_axis_back(cbmv_array* in, cbmv_array* out) {
  clear_error();
  return swap(in, out, true);
}

int cbmv_post_copy(cbmv_array* in, cbmv_array* out) {
  clear_error();
  RC(live(in, "copy"));
  RC(live(out, "copy"));
  CBMV_REQUIRE(in->dtype == out->dtype, "copy: arrays of different dtypes");
  const int64_t n = nel(in);
  CBMV_REQUIRE(out->count >= n, "copy: output array too small");
  if (in->dtype == CBMV_F64) {
    k_copy<double><<<grid_1d(n, 256), 256, 0, lib_stream()>>>((const double*)in->d, (double*)out->d, n);
    // CUDA_copy_double copies the logical shape as well (post.cu:702-704)
    out->ndims = in->ndims;
    for (int i = 0; i < 4; i++) out->shape[i] = in->shape[i];
  } else {
    k_copy<float><<<grid_1d(n, 256), 256, 0, lib_stream()>>>((const float*)in->d, (float*)out->d, n);
  }
  CBMV_LAUNCH_CHECK();
  return sync();
}

int cbmv_post_divide(cbmv_array* in, double val) {
  clear_error();
  RC(need(in, CBMV_F64, 0, "CUDA_divide_double"));
  const int64_t n = in->count;
  k_divide<<<grid_1d(n, 256), 256, 0, lib_stream()>>>((double*)in->d, val, n);
  CBMV_LAUNCH_CHECK();
  return sync();
}

int cbmv_post_outlier_detection(cbmv_array* d0, cbmv_array* d1, cbmv_array* out,
                                int32_t disp_max) {
  clear_error();
  RC(need(d0, CBMV_F32, 2, "outlier_detection"));
  RC(need(d1, CBMV_F32, 2, "outlier_detection"));
  RC(need(out, CBMV_F32, 2, "outlier_detection"));
  CBMV_REQUIRE(nel(d1) >= nel(d0) && nel(out) >= nel(d0), "outlier_detection: shape mismatch");
  RC(launch_outlier((const float*)d0->d, (const float*)d1->d, (int)d0->shape[0], (int)d0->shape[1],
                    disp_max, (float*)out->d, lib_stream()));
  return sync();
}

int cbmv_post_interpolate_mismatch(cbmv_array* d0, cbmv_array* outlier, cbmv_array* out) {
  clear_error();
  RC(need(d0, CBMV_F32, 2, "interpolate_mismatch"));
  RC(need(outlier, CBMV_F32, 2, "interpolate_mismatch"));
  RC(need(out, CBMV_F32, 2, "interpolate_mismatch"));
  CBMV_REQUIRE(nel(outlier) >= nel(d0) && nel(out) >= nel(d0), "interpolate_mismatch: shape mismatch");
  RC(launch_interp_mismatch((const float*)d0->d, (const float*)outlier->d, (int)d0->shape[0],
                            (int)d0->shape[1], (float*)out->d, lib_stream()));
  return sync();
}

int cbmv_post_interpolate_occlusion(cbmv_array* d0, cbmv_array* outlier, cbmv_array* out) {
  clear_error();
  RC(need(d0, CBMV_F32, 2, "interpolate_occlusion"));
  RC(need(outlier, CBMV_F32, 2, "interpolate_occlusion"));
  RC(need(out, CBMV_F32, 2, "interpolate_occlusion"));
  CBMV_REQUIRE(nel(outlier) >= nel(d0) && nel(out) >= nel(d0), "interpolate_occlusion: shape mismatch");
  RC(launch_interp_occlusion((const float*)d0->d, (const float*)outlier->d, (int)d0->shape[0],
                             (int)d0->shape[1], (float*)out->d, lib_stream()));
  return sync();
}

}  // extern "C"

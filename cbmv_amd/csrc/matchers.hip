// Matching costs: census, NCC (Nister), SAD of Sobel-x, zero-mean SAD.
//
// One thread per pixel (y, x) with x fastest across the wavefront; each
// thread walks the disparities d = 0..D-1 of its pixel. With the pipeline's
// [H,D,W] volume layout the stores of a wavefront at a given d are 64
// consecutive elements, and the right-image reads at x - d are consecutive too.
//
// Validity follows the reference's loop bounds: with wc = w/2, a voxel is
// computed for y - wc in [0, H-w) and j = x - wc in [d, W-w)
// (matchers.cpp:278-279,312-316 for census, :168-173 NCC, :412-417 SAD of
// Sobel, :495-497 zero-mean SAD); other voxels hold the invalid value.
//
// Arithmetic matches the reference exactly: census and SAD-of-Sobel are
// integers; NCC's window sums are integers and its finish is the reference's
// double expression; zero-mean SAD is evaluated in double in the reference's
// order ((l - ml) - r) + mr, summed in window row-major order.
#include "ops.hpp"

#include <mutex>
#include <type_traits>

namespace cbmv {

namespace {

// one row of 256 pixels per block: at each disparity a block's stores to an
// [H, D, W] volume are 2 KB contiguous (64 x 4 blocks wrote 512 B to each
// of four rows; NCC 1.45 -> 1.24 ms at 1500x1000x256)
constexpr int kBX = 256, kBY = 1;

__device__ __forceinline__ bool voxel_valid(int y, int x, int d, int H, int W, int w) {
  const int wc = w / 2;
  const int i = y - wc, j = x - wc;
  return i >= 0 && i < H - w && j >= d && j < W - w;
}

__device__ __forceinline__ bool centre_valid(int y, int x, int H, int W, int w) {
  const int wc = w / 2;
  const int i = y - wc, j = x - wc;
  return i >= 0 && i < H - w && j >= 0 && j < W - w;
}

// Right-image values that move one lane per disparity step. In the window
// kernels lane i of a wavefront is column x0 + i, and what its voxel at
// disparity d reads at the right image (window column x - wc - d, pixel
// x - d) is what lane i-1 read at d-1. So each step shifts the N dwords
// one lane up (DPP wave_shr) and lane 0 takes the next entry of a feed
// that each lane k loads for step dn + k, 64 steps at a time; the only
// per-step memory access left is the store.
template <int N>
struct LaneFeed {
  uint32_t cur[N];   // this step's values
  uint32_t feed[N];  // lane 0 holds the next step's
  int used = 64;     // entries of the feed taken
  // load(dn, feed): lane k's entries for step dn + k
  template <typename Load>
  __device__ __forceinline__ void advance(int d, Load load) {
    if (used == 64) {
      load(d, feed);
      used = 0;
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
      cur[k] = (uint32_t)__builtin_amdgcn_update_dpp((int)feed[k], (int)cur[k], 0x138, 0xf, 0xf,
                                                     false);  // wave_shr:1, lane 0 keeps feed
      feed[k] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)feed[k], 0x130, 0xf, 0xf,
                                                      true);  // wave_shl:1
    }
    used++;
  }
};

// The right view's minimum per pixel (featextract.cpp:130 with the
// statistics' minimum, featextract.cpp:322-327), formed while the volume is
// written. The right view of pixel p at disparity d is the volume element
// (p + d, d), which lane p + d of a wavefront computes at step d; so a
// pixel's running minimum moves one lane up per step and, where it leaves
// the wavefront (lane 63) or the disparities end, is merged into rm[p] with
// a 64-bit atomic minimum (a pixel's right view spans up to D columns,
// several wavefronts). Each step rotates the running values one lane up
// (DPP wave_ror): lane 0 takes the identity for the pixel x0 - d that
// starts there, and the value that left lane 63 is shifted into lane 0 of
// an outbox register (DPP wave_shr), which holds the values of the last
// steps in its lanes, the newest in lane 0. The kernels flush the outbox
// (one atomic per pending lane) after a group of steps, before its oldest
// value would leave it. Lanes of columns outside the image carry the
// invalid value, which lowers no minimum. Coded matchers keep codes (the
// invalid code is the largest), float64 ones the value, merged as its
// dkey; rm starts at all ones.
template <bool F64>
struct RightMin {
  // identity: all ones (a code above every code) / the bits of +inf
  static constexpr uint32_t kIdLo = F64 ? 0u : 0xffffffffu, kIdHi = F64 ? 0x7ff00000u : 0u;
  uint32_t lo = kIdLo, hi = kIdHi;  // running minimum: a code, or the halves of a float64
  uint32_t olo = kIdLo, ohi = kIdHi;  // outbox
  int fl = 0;  // steps up to fl have been flushed (step 0 leaves nothing)
  __device__ __forceinline__ static uint32_t ror(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x13c, 0xf, 0xf, false);  // wave_ror:1
  }
  __device__ __forceinline__ static uint32_t shr(uint32_t v, uint32_t first) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)first, (int)v, 0x138, 0xf, 0xf,
                                                 false);  // wave_shr:1, lane 0 takes first
  }
  __device__ __forceinline__ uint64_t key_of(uint32_t l, uint32_t h) const {
    if constexpr (F64) return dkey(__longlong_as_double((long long)(((uint64_t)h << 32) | l)));
    else return l;
  }
  template <typename V>
  __device__ __forceinline__ void step(V v, int lane) {
    const uint32_t tl = ror(lo);  // lane 0: the value leaving lane 63
    olo = shr(olo, tl);
    if constexpr (F64) {
      const uint32_t th = ror(hi);
      ohi = shr(ohi, th);
      const double s = __longlong_as_double((long long)(((uint64_t)(lane == 0 ? kIdHi : th) << 32) |
                                                        (lane == 0 ? kIdLo : tl)));
      // the volumes hold no NaN: the strict-< minimum (up to the sign of a zero)
      const double r = (double)v < s ? (double)v : s;
      const uint64_t b = (uint64_t)__double_as_longlong(r);
      lo = (uint32_t)b;
      hi = (uint32_t)(b >> 32);
    } else {
      lo = min(lane == 0 ? kIdLo : tl, (uint32_t)v);
    }
  }
  // after step d: outbox lane j holds the value that left at step d - j
  // (the pixel x0 + 64 - (d - j)); flush the pending ones when more than
  // 64 - g are, g being the most steps before the next call
  __device__ __forceinline__ void sync(unsigned long long* rm, int64_t prow, int x0, int lane, int d,
                                       int W, int g) {
    if (d - fl > 64 - g) {
      const int p = x0 + 64 - (d - lane);
      if (lane < d - fl && p >= 0 && p < W)
        atomicMin(rm + prow + p, (unsigned long long)key_of(olo, ohi));
      fl = d;
    }
  }
  // after the last step (d = D - 1): the outbox and every lane's pixel
  __device__ __forceinline__ void end(unsigned long long* rm, int64_t prow, int x0, int x, int lane,
                                      int D, int W) {
    sync(rm, prow, x0, lane, D - 1, W, 64);
    const int p = x - (D - 1);
    if (p >= 0 && p < W) atomicMin(rm + prow + p, (unsigned long long)key_of(lo, hi));
  }
};

// Window column c (clamped to the image) of an image, one byte per window
// row (rows r0 .. r0 + WS - 1), packed four rows to a dword.
template <int WS>
__device__ __forceinline__ void window_column(const uint8_t* __restrict__ img, int64_t row0, int W,
                                              int c, uint32_t* pk) {
  c = min(max(c, 0), W - 1);
#pragma unroll
  for (int k = 0; k < (WS + 3) / 4; k++) pk[k] = 0;
#pragma unroll
  for (int a = 0; a < WS; a++)
    pk[a / 4] |= (uint32_t)img[row0 + (int64_t)a * W + c] << (8 * (a % 4));
}

// ---------------------------------------------------------------- census
template <int WORDS>
__global__ void k_census_desc(const uint8_t* __restrict__ img, int H, int W, int w,
                              uint64_t* __restrict__ desc) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H || !centre_valid(y, x, H, W, w)) return;
  const int wc = w / 2;
  const int c = img[(int64_t)y * W + x];
  uint64_t bits[WORDS];
#pragma unroll
  for (int k = 0; k < WORDS; k++) bits[k] = 0;
  int k = 0;
  for (int a = 0; a < w; a++) {
    const uint8_t* row = img + (int64_t)(y - wc + a) * W + (x - wc);
    for (int b = 0; b < w; b++, k++) {
      // bit = centre < neighbour (matchers.cpp:292-293)
      const uint64_t bit = (c < (int)row[b]) ? 1ull : 0ull;
#pragma unroll
      for (int q = 0; q < WORDS; q++)
        if ((k >> 6) == q) bits[q] |= bit << (k & 63);
    }
  }
  uint64_t* o = desc + ((int64_t)y * W + x) * WORDS;
#pragma unroll
  for (int q = 0; q < WORDS; q++) o[q] = bits[q];
}

template <int WORDS, typename OutT>
__global__ void k_census_cost(const uint64_t* __restrict__ dl, const uint64_t* __restrict__ dr,
                              int H, int W, int D, int w, OutT* __restrict__ out, VolStrides st,
                              OutT invalid, double* __restrict__ mins,
                              unsigned long long* __restrict__ rm) {
  const int lane = threadIdx.x & 63;
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int x0 = x - lane;  // the wavefront's first column
  // every lane of a wavefront with a column in the image stays (the right
  // descriptors move between its lanes, as in k_ncc_win)
  if (x0 >= W || y >= H) return;
  const bool inx = x < W;
  const int xc = min(x, W - 1);
  OutT* o = out + y * st.sy + xc * st.sx;
  OutT m = invalid;
  const int wc = w / 2;
  if (y - wc < 0 || y - wc >= H - w) {  // the whole row is invalid
    for (int d = 0; d < D; d++)
      if (inx) o[d * st.sd] = invalid;
    if (mins && inx) mins[(int64_t)y * W + x] = kRandMax;
    if constexpr (std::is_integral<OutT>::value)
      if (rm && inx) rm[(int64_t)y * W + x] = (uint32_t)invalid;
    return;
  }
  // voxel (x, d) is valid when the centre is and d <= x - wc
  const int dmax = inx && centre_valid(y, xc, H, W, w) ? min(D - 1, x - wc) : -1;
  const int64_t prow = (int64_t)y * W;
  uint64_t l[WORDS];
#pragma unroll
  for (int q = 0; q < WORDS; q++) l[q] = dl[(prow + xc) * WORDS + q];
  // the right descriptor of pixel x - d, moving one lane per step
  LaneFeed<2 * WORDS> mv;
#pragma unroll
  for (int q = 0; q < WORDS; q++) {
    const uint64_t r = dr[(prow + xc) * WORDS + q];
    mv.cur[2 * q] = (uint32_t)r;
    mv.cur[2 * q + 1] = (uint32_t)(r >> 32);
  }
  auto load = [&](int dn, uint32_t* f) {
    const int c = min(max(x0 - (dn + lane), 0), W - 1);
#pragma unroll
    for (int q = 0; q < WORDS; q++) {
      const uint64_t r = dr[(prow + c) * WORDS + q];
      f[2 * q] = (uint32_t)r;
      f[2 * q + 1] = (uint32_t)(r >> 32);
    }
  };
  RightMin<false> rmv;
  for (int d = 0; d < D; d++) {
    if (d > 0) mv.advance(d, load);
    int ham = 0;
#pragma unroll
    for (int q = 0; q < WORDS; q++)
      ham += __popcll(l[q] ^ (((uint64_t)mv.cur[2 * q + 1] << 32) | mv.cur[2 * q]));
    const OutT v = d <= dmax ? (OutT)ham : invalid;
    if (inx) o[d * st.sd] = v;
    m = v < m ? v : m;
    if constexpr (std::is_integral<OutT>::value)
      if (rm) {
        rmv.step((uint32_t)v, lane);
        rmv.sync(rm, prow, x0, lane, d, W, 1);
      }
  }
  if constexpr (std::is_integral<OutT>::value)
    if (rm) rmv.end(rm, prow, x0, x, lane, D, W);
  if (mins && inx) mins[(int64_t)y * W + x] = m == invalid ? kRandMax : (double)m;
}

// ---------------------------------------------------------------- window stats
__global__ void k_window_stats(const uint8_t* __restrict__ img, int H, int W, int w,
                               uint32_t* __restrict__ A, double* __restrict__ C,
                               double* __restrict__ M) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H || !centre_valid(y, x, H, W, w)) return;
  const int wc = w / 2;
  uint32_t s = 0;
  uint64_t q = 0;
  for (int a = 0; a < w; a++) {
    const uint8_t* row = img + (int64_t)(y - wc + a) * W + (x - wc);
    for (int b = 0; b < w; b++) {
      const uint32_t v = row[b];
      s += v;
      q += v * v;
    }
  }
  const int64_t p = (int64_t)y * W + x;
  const int sq = w * w;
  if (A) A[p] = s;
  // 1/sqrt(w^2 * sum(x^2) - sum(x)^2) in double (matchers.cpp:127-128)
  if (C) C[p] = 1.0 / sqrt((double)((uint64_t)sq * q) - (double)s * (double)s);
  // double sum / w^2 (matchers.cpp:476-482); the sum of integers is exact
  if (M) M[p] = (double)s / (double)sq;
}

// ---------------------------------------------------------------- NCC
__global__ void k_ncc_cost(const uint8_t* __restrict__ L, const uint8_t* __restrict__ R,
                           const uint32_t* __restrict__ AL, const uint32_t* __restrict__ AR,
                           const double* __restrict__ CL, const double* __restrict__ CR, int H,
                           int W, int D, int w, double* __restrict__ out, VolStrides st,
                           double* __restrict__ mins) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  double* o = out + y * st.sy + x * st.sx;
  const bool cv = centre_valid(y, x, H, W, w);
  const int wc = w / 2;
  const int sq = w * w;
  double m = kRandMax;
  const int64_t p = (int64_t)y * W + x;
  const uint32_t al = cv ? AL[p] : 0;
  const double cl = cv ? CL[p] : 0.0;
  for (int d = 0; d < D; d++) {
    double v = kRandMax;
    if (cv && voxel_valid(y, x, d, H, W, w)) {
      const double cr = CR[p - d];
      if (isfinite(cl) && isfinite(cr)) {
        uint32_t lr = 0;
        for (int a = 0; a < w; a++) {
          const int64_t rowoff = (int64_t)(y - wc + a) * W + (x - wc);
          for (int b = 0; b < w; b++) lr += (uint32_t)L[rowoff + b] * (uint32_t)R[rowoff + b - d];
        }
        // -(double)(w^2 * sum(LR) - A_l * A_r) * C_l * C_r, left to right
        // (matchers.cpp:175-179); the difference is an exact integer.
        const double num = (double)sq * (double)lr - (double)((uint64_t)al * AR[p - d]);
        v = -num * cl * cr;
      } else {
        v = 1.0;
      }
    }
    o[d * st.sd] = v;
    m = fmin(m, v);
  }
  if (mins) mins[p] = m;
}

// ---------------------------------------------------------------- Sobel
template <typename OutT>
__global__ void k_sobel(const uint8_t* __restrict__ img, int H, int W, OutT* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  int v = 0;
  // written at (i+1, j+1) for i < H-3, j < W-3 (matchers.cpp:539-548)
  if (y >= 1 && y <= H - 3 && x >= 1 && x <= W - 3) {
    const uint8_t* r0 = img + (int64_t)(y - 1) * W + (x - 1);
    const uint8_t* r1 = r0 + W;
    const uint8_t* r2 = r1 + W;
    v = -(int)r0[0] + (int)r0[2] - 2 * (int)r1[0] + 2 * (int)r1[2] - (int)r2[0] + (int)r2[2];
  }
  out[(int64_t)y * W + x] = (OutT)v;
}

// ---------------------------------------------------------------- SAD of Sobel
template <typename InT, typename AccT, typename OutT>
__global__ void k_sadsob_cost(const InT* __restrict__ Ls, const InT* __restrict__ Rs, int H,
                              int W, int D, int w, OutT* __restrict__ out, VolStrides st,
                              OutT invalid, double* __restrict__ mins) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  OutT* o = out + y * st.sy + x * st.sx;
  const bool cv = centre_valid(y, x, H, W, w);
  const int wc = w / 2;
  OutT m = invalid;
  for (int d = 0; d < D; d++) {
    OutT v = invalid;
    if (cv && voxel_valid(y, x, d, H, W, w)) {
      AccT s = 0;
      for (int a = 0; a < w; a++) {
        const int64_t rowoff = (int64_t)(y - wc + a) * W + (x - wc);
        for (int b = 0; b < w; b++) {
          const AccT diff = (AccT)Ls[rowoff + b] - (AccT)Rs[rowoff + b - d];
          s += diff < 0 ? -diff : diff;
        }
      }
      v = (OutT)s;
    }
    o[d * st.sd] = v;
    m = v < m ? v : m;
  }
  if (mins) mins[(int64_t)y * W + x] = m == invalid ? kRandMax : (double)m;
}

// ---------------------------------------------------------------- zero-mean SAD
__global__ void k_zsad_cost(const uint8_t* __restrict__ L, const uint8_t* __restrict__ R,
                            const double* __restrict__ ML, const double* __restrict__ MR, int H,
                            int W, int D, int w, double* __restrict__ out, VolStrides st,
                            double* __restrict__ mins) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  double* o = out + y * st.sy + x * st.sx;
  const bool cv = centre_valid(y, x, H, W, w);
  const int wc = w / 2;
  const int64_t p = (int64_t)y * W + x;
  const double ml = cv ? ML[p] : 0.0;
  double m = kRandMax;
  for (int d = 0; d < D; d++) {
    double v = kRandMax;
    if (cv && voxel_valid(y, x, d, H, W, w)) {
      const double mr = MR[p - d];
      double s = 0.0;
      for (int a = 0; a < w; a++) {
        const int64_t rowoff = (int64_t)(y - wc + a) * W + (x - wc);
        for (int b = 0; b < w; b++)
          s += fabs((double)L[rowoff + b] - ml - (double)R[rowoff + b - d] + mr);
      }
      v = s;
    }
    o[d * st.sd] = v;
    m = fmin(m, v);
  }
  if (mins) mins[p] = m;
}

// ---------------------------------------------------------------- register windows
// The same three costs for a window size WS known at compile time. Moving
// from d to d+1 shifts the right-image window left by one column, so each
// thread keeps the WS x WS right window in registers as a ring of columns
// (column of window position b at disparity d in ring slot (b - d) mod WS;
// the d loop is unrolled WS times so that every slot index is a constant)
// and loads one new column per disparity instead of the whole window. The
// left window's terms do not depend on d and are formed once. The sums run
// in the same order (window row-major) with the same operations as the
// generic kernels above, so the volumes are identical.
//
// A pixel's valid disparities are d <= j = x - WS/2 (its centre being
// valid); their windows stay inside the image. The rest are written as the
// invalid value.

// zero-mean SAD: s += |((l - ml) - r) + mr| in double
template <int WS>
__global__ void __launch_bounds__(kBX * kBY) k_zsad_win(
    const uint8_t* __restrict__ L, const uint8_t* __restrict__ R, const double* __restrict__ ML,
    const double* __restrict__ MR, int H, int W, int D, double* __restrict__ out, VolStrides st,
    double* __restrict__ mins, unsigned long long* __restrict__ rm) {
  constexpr int wc = WS / 2;
  constexpr int NW = (WS + 3) / 4;  // dwords of a packed window column
  const int lane = threadIdx.x & 63;
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int x0 = x - lane;
  // every lane of a wavefront with a column in the image stays (values
  // move between its lanes, as in k_ncc_win)
  if (x0 >= W || y >= H) return;
  const bool inx = x < W;
  const int xc = min(x, W - 1);
  double* o = out + y * st.sy + xc * st.sx;
  // the minimum over d of the written values, in ascending d (the volume
  // holds no NaN)
  double m = kRandMax;
  auto put = [&](int d, double v) {
    if (inx) o[d * st.sd] = v;
    m = fmin(m, v);
  };
  if (y - wc < 0 || y - wc >= H - WS) {  // the whole row is invalid
    for (int d = 0; d < D; d++) put(d, kRandMax);
    if (mins && inx) mins[(int64_t)y * W + x] = m;
    if (rm && inx) rm[(int64_t)y * W + x] = dkey(kRandMax);
    return;
  }
  const int dmax = centre_valid(y, xc, H, W, WS) && inx ? min(D - 1, x - wc) : -1;
  const int64_t row0 = (int64_t)(y - wc) * W;
  const int64_t prow = (int64_t)y * W;
  const int64_t p = prow + xc;
  const double ml = ML[p];
  double t[WS][WS];   // [row][column] (l - ml)
  double rw[WS][WS];  // ring of right-window columns, [slot][row]
#pragma unroll
  for (int b = 0; b < WS; b++) {
    uint32_t lc[NW], rc[NW];
    window_column<WS>(L, row0, W, x - wc + b, lc);
    window_column<WS>(R, row0, W, x - wc + b, rc);
#pragma unroll
    for (int a = 0; a < WS; a++) {
      t[a][b] = (double)((lc[a / 4] >> (8 * (a % 4))) & 0xffu) - ml;
      rw[b][a] = (double)((rc[a / 4] >> (8 * (a % 4))) & 0xffu);
    }
  }
  // moving values: the new window column (NW dwords) and the mean of pixel x - d
  LaneFeed<NW + 2> mv;
  window_column<WS>(R, row0, W, x - wc, mv.cur);
  const uint64_t mr0 = (uint64_t)__double_as_longlong(MR[p]);
  mv.cur[NW] = (uint32_t)mr0;
  mv.cur[NW + 1] = (uint32_t)(mr0 >> 32);
  auto load = [&](int dn, uint32_t* f) {
    const int c = x0 - (dn + lane);
    window_column<WS>(R, row0, W, c - wc, f);
    const uint64_t mb = (uint64_t)__double_as_longlong(MR[prow + min(max(c, 0), W - 1)]);
    f[NW] = (uint32_t)mb;
    f[NW + 1] = (uint32_t)(mb >> 32);
  };
  RightMin<true> rmv;
  // One voxel: disparity d = d0 + dd takes right-window column x - wc - d
  // into ring slot (WS - dd) % WS (for d = 0 the initial column again).
  auto step = [&](int d, int dd) {
    if (d > 0) mv.advance(d, load);
#pragma unroll
    for (int a = 0; a < WS; a++)
      rw[(WS - dd) % WS][a] = (double)((mv.cur[a / 4] >> (8 * (a % 4))) & 0xffu);
    const double mr =
        __longlong_as_double((long long)(((uint64_t)mv.cur[NW + 1] << 32) | mv.cur[NW]));
    double sum = 0.0;
#pragma unroll
    for (int a = 0; a < WS; a++)
#pragma unroll
      for (int b = 0; b < WS; b++) sum += fabs(t[a][b] - rw[(b - dd + WS) % WS][a] + mr);
    const double v = d <= dmax ? sum : kRandMax;
    put(d, v);
    if (rm) rmv.step(v, lane);
  };
  // whole groups of WS disparities (the ring slots are compile-time), two
  // groups per iteration while two fit, then the last partial group
  int d0 = 0;
  for (; d0 + 2 * WS <= D; d0 += 2 * WS) {
#pragma unroll
    for (int dd = 0; dd < 2 * WS; dd++) step(d0 + dd, dd % WS);
    if (rm) rmv.sync(rm, prow, x0, lane, d0 + 2 * WS - 1, W, 2 * WS);
  }
  for (; d0 + WS <= D; d0 += WS) {
#pragma unroll
    for (int dd = 0; dd < WS; dd++) step(d0 + dd, dd);
    if (rm) rmv.sync(rm, prow, x0, lane, d0 + WS - 1, W, 2 * WS);
  }
#pragma unroll
  for (int dd = 0; dd < WS - 1; dd++)
    if (d0 + dd < D) step(d0 + dd, dd);
  if (rm) rmv.end(rm, prow, x0, x, lane, D, W);
  if (mins && inx) mins[(int64_t)y * W + x] = m;
}

// SAD of Sobel-x images (integers)
template <int WS>
__global__ void __launch_bounds__(kBX * kBY) k_sadsob_win(
    const int16_t* __restrict__ Ls, const int16_t* __restrict__ Rs, int H, int W, int D,
    uint16_t* __restrict__ out, VolStrides st,
    double* __restrict__ mins, unsigned long long* __restrict__ rm) {
  constexpr int wc = WS / 2;
  constexpr int NH = (WS + 1) / 2;  // dwords of a packed window column (two rows each)
  const int lane = threadIdx.x & 63;
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int x0 = x - lane;
  // every lane of a wavefront with a column in the image stays (values
  // move between its lanes, as in k_ncc_win)
  if (x0 >= W || y >= H) return;
  const bool inx = x < W;
  const int xc = min(x, W - 1);
  uint16_t* o = out + y * st.sy + xc * st.sx;
  // the minimum over d of the written codes (the invalid code is the largest)
  uint32_t m = kSobInvalid;
  auto put = [&](int d, uint32_t v) {
    if (inx) o[d * st.sd] = (uint16_t)v;
    m = min(m, v);
  };
  if (y - wc < 0 || y - wc >= H - WS) {  // the whole row is invalid
    for (int d = 0; d < D; d++) put(d, kSobInvalid);
    if (mins && inx) mins[(int64_t)y * W + x] = kRandMax;
    if (rm && inx) rm[(int64_t)y * W + x] = kSobInvalid;
    return;
  }
  const int dmax = inx && centre_valid(y, xc, H, W, WS) ? min(D - 1, x - wc) : -1;
  const int64_t row0 = (int64_t)(y - wc) * W;
  const int64_t prow = (int64_t)y * W;
  // |l - r| plus the running sum is one v_sad_u16. The Sobel values lie in
  // [-1020, 1020]; biased by 1024 they are in [4, 2044], where the low 16
  // bits read the same as signed or unsigned numbers. A window column
  // (clamped to the image) packs two rows' biased values per dword.
  auto column = [&](const int16_t* __restrict__ img, int c, uint32_t* pk) {
    c = min(max(c, 0), W - 1);
#pragma unroll
    for (int k = 0; k < NH; k++) pk[k] = 0;
#pragma unroll
    for (int a = 0; a < WS; a++)
      pk[a / 2] |= (uint32_t)(img[row0 + (int64_t)a * W + c] + 1024) << (16 * (a % 2));
  };
  auto unpack = [](const uint32_t* pk, int a) { return (pk[a / 2] >> (16 * (a % 2))) & 0xffffu; };
  uint32_t t[WS][WS], rw[WS][WS];  // t[row][column], rw[slot][row]
#pragma unroll
  for (int b = 0; b < WS; b++) {
    uint32_t lc[NH], rc[NH];
    column(Ls, x - wc + b, lc);
    column(Rs, x - wc + b, rc);
#pragma unroll
    for (int a = 0; a < WS; a++) {
      t[a][b] = unpack(lc, a);
      rw[b][a] = unpack(rc, a);
    }
  }
  // moving values: the new window column x - wc - d
  LaneFeed<NH> mv;
  column(Rs, x - wc, mv.cur);
  auto load = [&](int dn, uint32_t* f) { column(Rs, x0 - (dn + lane) - wc, f); };
  RightMin<false> rmv;
  // ring and groups as in k_zsad_win
  auto step = [&](int d, int dd) {
    if (d > 0) mv.advance(d, load);
#pragma unroll
    for (int a = 0; a < WS; a++) rw[(WS - dd) % WS][a] = unpack(mv.cur, a);
    uint32_t sum = 0;
#pragma unroll
    for (int a = 0; a < WS; a++)
#pragma unroll
      for (int b = 0; b < WS; b++) sum = __builtin_amdgcn_sad_u16(t[a][b], rw[(b - dd + WS) % WS][a], sum);
    const uint32_t v = d <= dmax ? sum : (uint32_t)kSobInvalid;
    put(d, v);
    if (rm) rmv.step(v, lane);
  };
  int d0 = 0;
  for (; d0 + 2 * WS <= D; d0 += 2 * WS) {
#pragma unroll
    for (int dd = 0; dd < 2 * WS; dd++) step(d0 + dd, dd % WS);
    if (rm) rmv.sync(rm, prow, x0, lane, d0 + 2 * WS - 1, W, 2 * WS);
  }
  for (; d0 + WS <= D; d0 += WS) {
#pragma unroll
    for (int dd = 0; dd < WS; dd++) step(d0 + dd, dd);
    if (rm) rmv.sync(rm, prow, x0, lane, d0 + WS - 1, W, 2 * WS);
  }
#pragma unroll
  for (int dd = 0; dd < WS - 1; dd++)
    if (d0 + dd < D) step(d0 + dd, dd);
  if (rm) rmv.end(rm, prow, x0, x, lane, D, W);
  if (mins && inx) mins[(int64_t)y * W + x] = m == kSobInvalid ? kRandMax : (double)m;
}

// NCC: integer window products, the reference's double finish
template <int WS>
__global__ void __launch_bounds__(kBX * kBY) k_ncc_win(
    const uint8_t* __restrict__ L, const uint8_t* __restrict__ R,
    const uint32_t* __restrict__ AL, const uint32_t* __restrict__ AR,
    const double* __restrict__ CL, const double* __restrict__ CR, int H, int W, int D,
    double* __restrict__ out, VolStrides st,
    double* __restrict__ mins, unsigned long long* __restrict__ rm) {
  constexpr int wc = WS / 2;
  constexpr int sq = WS * WS;
  constexpr int NW = (WS + 3) / 4;  // dwords of a packed window column
  const int lane = threadIdx.x & 63;
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int x0 = x - lane;  // the wavefront's first column
  // every lane of a wavefront with a column in the image stays (values
  // move between its lanes)
  if (x0 >= W || y >= H) return;
  const bool inx = x < W;
  const int xc = min(x, W - 1);
  double* o = out + y * st.sy + xc * st.sx;
  // the minimum over d of the written values, in ascending d (the volume
  // holds no NaN)
  double m = kRandMax;
  auto put = [&](int d, double v) {
    if (inx) o[d * st.sd] = v;
    m = fmin(m, v);
  };
  if (y - wc < 0 || y - wc >= H - WS) {  // the whole row is invalid
    for (int d = 0; d < D; d++) put(d, kRandMax);
    if (mins && inx) mins[(int64_t)y * W + x] = m;
    if (rm && inx) rm[(int64_t)y * W + x] = dkey(kRandMax);
    return;
  }
  const int dmax = centre_valid(y, xc, H, W, WS) && inx ? min(D - 1, x - wc) : -1;
  const int64_t row0 = (int64_t)(y - wc) * W;
  const int64_t prow = (int64_t)y * W;
  const int64_t p = prow + xc;
  const uint32_t al = AL[p];
  const double cl = CL[p];
  const bool clf = isfinite(cl);
  uint32_t tl[WS][NW], rw[WS][NW];
#pragma unroll
  for (int b = 0; b < WS; b++) {
    window_column<WS>(L, row0, W, x - wc + b, tl[b]);
    window_column<WS>(R, row0, W, x - wc + b, rw[b]);
  }
  // moving values: the new window column (NW dwords), A and C of pixel x - d
  LaneFeed<NW + 3> mv;
#pragma unroll
  for (int k = 0; k < NW; k++) mv.cur[k] = rw[0][k];
  const uint64_t cr0 = (uint64_t)__double_as_longlong(CR[p]);
  mv.cur[NW] = AR[p];
  mv.cur[NW + 1] = (uint32_t)cr0;
  mv.cur[NW + 2] = (uint32_t)(cr0 >> 32);
  auto load = [&](int dn, uint32_t* f) {
    const int c = x0 - (dn + lane);
    window_column<WS>(R, row0, W, c - wc, f);
    const int64_t q = prow + min(max(c, 0), W - 1);
    const uint64_t cb = (uint64_t)__double_as_longlong(CR[q]);
    f[NW] = AR[q];
    f[NW + 1] = (uint32_t)cb;
    f[NW + 2] = (uint32_t)(cb >> 32);
  };
  RightMin<true> rmv;
  auto step = [&](int d, int dd) {
    if (d > 0) mv.advance(d, load);
    // ring slot (WS - dd) % WS takes the new column; tap b reads slot
    // (b - dd) mod WS
#pragma unroll
    for (int k = 0; k < NW; k++) rw[(WS - dd) % WS][k] = mv.cur[k];
    uint32_t lr = 0;
#pragma unroll
    for (int b = 0; b < WS; b++)
#pragma unroll
      for (int k = 0; k < NW; k++)
        lr = __builtin_amdgcn_udot4(tl[b][k], rw[(b - dd + WS) % WS][k], lr, false);
    const double cr =
        __longlong_as_double((long long)(((uint64_t)mv.cur[NW + 2] << 32) | mv.cur[NW + 1]));
    const double num = (double)sq * (double)lr - (double)((uint64_t)al * mv.cur[NW]);
    const bool ok = clf & (bool)isfinite(cr);  // no short circuit: one select
    // a constant window (non-finite C) gives 1
    const double v = d <= dmax ? (ok ? -num * cl * cr : 1.0) : kRandMax;
    put(d, v);
    if (rm) rmv.step(v, lane);
  };
  // whole groups of WS disparities (the ring slots are compile-time), two
  // groups per iteration while two fit, then the last partial group
  int d0 = 0;
  for (; d0 + 2 * WS <= D; d0 += 2 * WS) {
#pragma unroll
    for (int dd = 0; dd < 2 * WS; dd++) step(d0 + dd, dd % WS);
    if (rm) rmv.sync(rm, prow, x0, lane, d0 + 2 * WS - 1, W, 2 * WS);
  }
  for (; d0 + WS <= D; d0 += WS) {
#pragma unroll
    for (int dd = 0; dd < WS; dd++) step(d0 + dd, dd);
    if (rm) rmv.sync(rm, prow, x0, lane, d0 + WS - 1, W, 2 * WS);
  }
#pragma unroll
  for (int dd = 0; dd < WS - 1; dd++)
    if (d0 + dd < D) step(d0 + dd, dd);
  if (rm) rmv.end(rm, prow, x0, x, lane, D, W);
  if (mins && inx) mins[(int64_t)y * W + x] = m;
}

inline dim3 pix_grid(int H, int W) { return dim3((W + kBX - 1) / kBX, (H + kBY - 1) / kBY); }
inline dim3 pix_block() { return dim3(kBX, kBY); }

}  // namespace

int census_words(int wsize) { return (wsize * wsize + 63) / 64; }

int launch_census_desc(const uint8_t* img, int H, int W, int wsize, uint64_t* desc, hipStream_t s) {
  switch (census_words(wsize)) {
    case 1: k_census_desc<1><<<pix_grid(H, W), pix_block(), 0, s>>>(img, H, W, wsize, desc); break;
    case 2: k_census_desc<2><<<pix_grid(H, W), pix_block(), 0, s>>>(img, H, W, wsize, desc); break;
    case 3: k_census_desc<3><<<pix_grid(H, W), pix_block(), 0, s>>>(img, H, W, wsize, desc); break;
    case 4: k_census_desc<4><<<pix_grid(H, W), pix_block(), 0, s>>>(img, H, W, wsize, desc); break;
    default: set_error("census window %d too large (max 16)", wsize); return CBMV_EINVAL;
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

template <typename OutT>
static int census_cost(const uint64_t* dl, const uint64_t* dr, int H, int W, int D, int wsize,
                       OutT* out, VolStrides st, OutT invalid, hipStream_t s,
                       double* mins = nullptr, unsigned long long* rm = nullptr) {
  switch (census_words(wsize)) {
    case 1: k_census_cost<1, OutT><<<pix_grid(H, W), pix_block(), 0, s>>>(dl, dr, H, W, D, wsize, out, st, invalid, mins, rm); break;
    case 2: k_census_cost<2, OutT><<<pix_grid(H, W), pix_block(), 0, s>>>(dl, dr, H, W, D, wsize, out, st, invalid, mins, rm); break;
    case 3: k_census_cost<3, OutT><<<pix_grid(H, W), pix_block(), 0, s>>>(dl, dr, H, W, D, wsize, out, st, invalid, mins, rm); break;
    case 4: k_census_cost<4, OutT><<<pix_grid(H, W), pix_block(), 0, s>>>(dl, dr, H, W, D, wsize, out, st, invalid, mins, rm); break;
    default: set_error("census window %d too large (max 16)", wsize); return CBMV_EINVAL;
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_census_cost_u8(const uint64_t* dl, const uint64_t* dr, int H, int W, int D, int wsize,
                          uint8_t* out, VolStrides st, hipStream_t s, double* mins,
                          unsigned long long* rm, bool* rm_done) {
  if (rm_done) *rm_done = rm != nullptr;
  return census_cost<uint8_t>(dl, dr, H, W, D, wsize, out, st, kCensusInvalid, s, mins, rm);
}

int launch_census_cost_f64(const uint64_t* dl, const uint64_t* dr, int H, int W, int D, int wsize,
                           double* out, VolStrides st, hipStream_t s) {
  return census_cost<double>(dl, dr, H, W, D, wsize, out, st, kRandMax, s);
}

int launch_window_stats(const uint8_t* img, int H, int W, int wsize, uint32_t* A, double* C,
                        double* M, hipStream_t s) {
  k_window_stats<<<pix_grid(H, W), pix_block(), 0, s>>>(img, H, W, wsize, A, C, M);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_ncc_cost(const uint8_t* L, const uint8_t* R, const uint32_t* AL, const uint32_t* AR,
                    const double* CL, const double* CR, int H, int W, int D, int wsize,
                    double* out, VolStrides st, hipStream_t s, double* mins,
                    unsigned long long* rm, bool* rm_done) {
  if (rm_done) *rm_done = rm != nullptr && (wsize == 3 || wsize == 5);
  switch (wsize) {
    case 3: k_ncc_win<3><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, AL, AR, CL, CR, H, W, D, out, st, mins, rm); break;
    case 5: k_ncc_win<5><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, AL, AR, CL, CR, H, W, D, out, st, mins, rm); break;
    default: k_ncc_cost<<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, AL, AR, CL, CR, H, W, D, wsize, out, st, mins);
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_sobel_f64(const uint8_t* img, int H, int W, double* out, hipStream_t s) {
  k_sobel<double><<<pix_grid(H, W), pix_block(), 0, s>>>(img, H, W, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_sobel_i16(const uint8_t* img, int H, int W, int16_t* out, hipStream_t s) {
  k_sobel<int16_t><<<pix_grid(H, W), pix_block(), 0, s>>>(img, H, W, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_sadsob_f64(const double* L, const double* R, int H, int W, int D, int wsize,
                      double* out, VolStrides st, hipStream_t s) {
  k_sadsob_cost<double, double, double><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, H, W, D, wsize, out, st, kRandMax, nullptr);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_sadsob_u16(const int16_t* L, const int16_t* R, int H, int W, int D, int wsize,
                      uint16_t* out, VolStrides st, hipStream_t s, double* mins,
                      unsigned long long* rm, bool* rm_done) {
  if (rm_done) *rm_done = rm != nullptr && (wsize == 3 || wsize == 5);
  switch (wsize) {
    case 3: k_sadsob_win<3><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, H, W, D, out, st, mins, rm); break;
    case 5: k_sadsob_win<5><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, H, W, D, out, st, mins, rm); break;
    default:
      k_sadsob_cost<int16_t, int, uint16_t><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, H, W, D, wsize, out, st, kSobInvalid, mins);
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int launch_zsad(const uint8_t* L, const uint8_t* R, const double* ML, const double* MR, int H,
                int W, int D, int wsize, double* out, VolStrides st, hipStream_t s, double* mins,
                unsigned long long* rm, bool* rm_done) {
  if (rm_done) *rm_done = rm != nullptr && (wsize == 3 || wsize == 5 || wsize == 7);
  switch (wsize) {
    case 3: k_zsad_win<3><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, ML, MR, H, W, D, out, st, mins, rm); break;
    case 5: k_zsad_win<5><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, ML, MR, H, W, D, out, st, mins, rm); break;
    case 7: k_zsad_win<7><<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, ML, MR, H, W, D, out, st, mins, rm); break;
    default: k_zsad_cost<<<pix_grid(H, W), pix_block(), 0, s>>>(L, R, ML, MR, H, W, D, wsize, out, st, mins);
  }
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

}  // namespace cbmv

// ===========================================================================
// Drop-in entry points (libmatchers): host buffers, synchronous.
// ===========================================================================
using namespace cbmv;

static int check_dims(int64_t H, int64_t W, int32_t D, int32_t w) {
  CBMV_REQUIRE(H > 0 && W > 0 && H < (1 << 24) && W < (1 << 24), "bad image size %lldx%lld",
               (long long)H, (long long)W);
  CBMV_REQUIRE(D > 0, "ndisp must be positive (got %d)", D);
  CBMV_REQUIRE(w > 0 && w <= 31, "window size must be in [1, 31] (got %d)", w);
  return CBMV_OK;
}

// ---------------------------------------------------------------- census without allocation
// For callers that cannot allocate device memory (code under graph capture,
// serving loops): census between device images into a caller's device
// volume, with both images' descriptors in a fixed device array instead of
// library allocations. kNoAllocScratchWords u64 = 48 MiB per device, held
// until the process ends: for an 11x11 window (two words per pixel) up to
// 1,572,864 pixels (1500x1000 fits). Every call on a device shares that
// device's array, so calls that overlap on one device race; nothing needs
// clearing between calls, since the cost kernel reads only descriptors of
// valid centres, which the descriptor kernel writes on every call.
constexpr int64_t kNoAllocScratchWords = (int64_t)3 << 21;
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchWords];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint64_t* g_scratch_ptr[kNoAllocMaxDevices];

// The device address of this device's copy of `scratch` (the array's name
// is only a host-side stand-in), looked up once per device.
static int noalloc_scratch(uint64_t** out) {
  int dev = 0;
  CBMV_HIP(hipGetDevice(&dev));
  CBMV_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "device %d beyond the %d supported", dev,
               kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    CBMV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CBMV_REQUIRE(p != nullptr, "no device address for the census scratch array");
    g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return CBMV_OK;
}

#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

extern "C" {

int cbmv_census(const uint8_t* left, const uint8_t* right, int64_t H, int64_t W, int32_t ndisp,
                int32_t wsize, double* out_hwd) {
  clear_error();
  RC(check_dims(H, W, ndisp, wsize));
  CBMV_REQUIRE(wsize <= 16, "census window must be <= 16 (got %d)", wsize);
  hipStream_t s = lib_stream();
  const size_t np = (size_t)H * W;
  const int words = census_words(wsize);
  DevBuf dL, dR, descL, descR, out;
  RC(to_device(dL, left, np));
  RC(to_device(dR, right, np));
  RC(descL.alloc(np * words * 8));
  RC(descR.alloc(np * words * 8));
  RC(out.alloc(np * ndisp * sizeof(double)));
  RC(launch_census_desc(dL.as<uint8_t>(), (int)H, (int)W, wsize, descL.as<uint64_t>(), s));
  RC(launch_census_desc(dR.as<uint8_t>(), (int)H, (int)W, wsize, descR.as<uint64_t>(), s));
  RC(launch_census_cost_f64(descL.as<uint64_t>(), descR.as<uint64_t>(), (int)H, (int)W, ndisp,
                            wsize, out.as<double>(), strides_hwd(H, W, ndisp), s));
  return to_host(out_hwd, out, np * ndisp * sizeof(double));
}

int cbmv_census_noalloc(const uint8_t* d_left, const uint8_t* d_right, int64_t H, int64_t W,
                        int32_t ndisp, int32_t wsize, double* d_out_hwd, void* stream) {
  clear_error();
  RC(check_dims(H, W, ndisp, wsize));
  CBMV_REQUIRE(wsize <= 16, "census window must be <= 16 (got %d)", wsize);
  CBMV_REQUIRE(d_left && d_right && d_out_hwd, "null device pointer");
  const int64_t np = H * W;
  const int words = census_words(wsize);
  CBMV_REQUIRE(2 * np * words <= kNoAllocScratchWords,
               "census_noalloc: %lldx%lld pixels with a %dx%d window need %lld descriptor words, "
               "more than the fixed %lld (%lld pixels at 2 words per pixel)",
               (long long)H, (long long)W, wsize, wsize, (long long)(2 * np * words),
               (long long)kNoAllocScratchWords, (long long)(kNoAllocScratchWords / 4));
  uint64_t* desc = nullptr;
  RC(noalloc_scratch(&desc));
  hipStream_t s = (hipStream_t)stream;
  RC(launch_census_desc(d_left, (int)H, (int)W, wsize, desc, s));
  RC(launch_census_desc(d_right, (int)H, (int)W, wsize, desc + np * words, s));
  return launch_census_cost_f64(desc, desc + np * words, (int)H, (int)W, ndisp, wsize, d_out_hwd,
                                strides_hwd(H, W, ndisp), s);
}

int cbmv_ncc_nister(const uint8_t* left, const uint8_t* right, int64_t H, int64_t W,
                    int32_t ndisp, int32_t wsize, double* out_dhw) {
  clear_error();
  RC(check_dims(H, W, ndisp, wsize));
  hipStream_t s = lib_stream();
  const size_t np = (size_t)H * W;
  DevBuf dL, dR, AL, AR, CL, CR, out;
  RC(to_device(dL, left, np));
  RC(to_device(dR, right, np));
  RC(AL.alloc(np * 4));
  RC(AR.alloc(np * 4));
  RC(CL.alloc(np * 8));
  RC(CR.alloc(np * 8));
  RC(out.alloc(np * ndisp * sizeof(double)));
  RC(launch_window_stats(dL.as<uint8_t>(), (int)H, (int)W, wsize, AL.as<uint32_t>(), CL.as<double>(), nullptr, s));
  RC(launch_window_stats(dR.as<uint8_t>(), (int)H, (int)W, wsize, AR.as<uint32_t>(), CR.as<double>(), nullptr, s));
  RC(launch_ncc_cost(dL.as<uint8_t>(), dR.as<uint8_t>(), AL.as<uint32_t>(), AR.as<uint32_t>(),
                     CL.as<double>(), CR.as<double>(), (int)H, (int)W, ndisp, wsize,
                     out.as<double>(), strides_dhw(H, W, ndisp), s));
  return to_host(out_dhw, out, np * ndisp * sizeof(double));
}

int cbmv_sobel(const uint8_t* img, int64_t H, int64_t W, double* out) {
  clear_error();
  RC(check_dims(H, W, 1, 1));
  hipStream_t s = lib_stream();
  const size_t np = (size_t)H * W;
  DevBuf dI, dO;
  RC(to_device(dI, img, np));
  RC(dO.alloc(np * 8));
  RC(launch_sobel_f64(dI.as<uint8_t>(), (int)H, (int)W, dO.as<double>(), s));
  return to_host(out, dO, np * 8);
}

int cbmv_sadsob(const double* left, const double* right, int64_t H, int64_t W, int32_t ndisp,
                int32_t wsize, double* out_dhw) {
  clear_error();
  RC(check_dims(H, W, ndisp, wsize));
  hipStream_t s = lib_stream();
  const size_t np = (size_t)H * W;
  DevBuf dL, dR, out;
  RC(to_device(dL, left, np * 8));
  RC(to_device(dR, right, np * 8));
  RC(out.alloc(np * ndisp * sizeof(double)));
  RC(launch_sadsob_f64(dL.as<double>(), dR.as<double>(), (int)H, (int)W, ndisp, wsize,
                       out.as<double>(), strides_dhw(H, W, ndisp), s));
  return to_host(out_dhw, out, np * ndisp * sizeof(double));
}

int cbmv_zsad(const uint8_t* left, const uint8_t* right, int64_t H, int64_t W, int32_t ndisp,
              int32_t wsize, double* out_dhw) {
  clear_error();
  RC(check_dims(H, W, ndisp, wsize));
  hipStream_t s = lib_stream();
  const size_t np = (size_t)H * W;
  DevBuf dL, dR, ML, MR, out;
  RC(to_device(dL, left, np));
  RC(to_device(dR, right, np));
  RC(ML.alloc(np * 8));
  RC(MR.alloc(np * 8));
  RC(out.alloc(np * ndisp * sizeof(double)));
  RC(launch_window_stats(dL.as<uint8_t>(), (int)H, (int)W, wsize, nullptr, nullptr, ML.as<double>(), s));
  RC(launch_window_stats(dR.as<uint8_t>(), (int)H, (int)W, wsize, nullptr, nullptr, MR.as<double>(), s));
  RC(launch_zsad(dL.as<uint8_t>(), dR.as<uint8_t>(), ML.as<double>(), MR.as<double>(), (int)H,
                 (int)W, ndisp, wsize, out.as<double>(), strides_dhw(H, W, ndisp), s));
  return to_host(out_dhw, out, np * ndisp * sizeof(double));
}

}  // extern "C"

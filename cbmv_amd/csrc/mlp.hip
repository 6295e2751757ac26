// MFMA MLP: pipeline inference (with the optional forest gate) and training
// (forward + backward + Adam). Training and inference layer 1 run on
// v_mfma_f32_16x16x32_bf16; inference layer 2 takes half-precision
// activations (evaluated two at a time in packed half arithmetic) on
// v_mfma_f32_16x16x32_f16, with the same operand layout.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 for lane l (i = l & 15,
// g = l >> 4): A fragment = A[i][8g .. 8g+7], B fragment = B[8g .. 8g+7][i],
// result = C[4g + r][i] for r = 0..3.
//
// Inference runs the layers transposed (H1^T = W1^T Z^T, H2^T = W2^T A1^T):
// the accumulator of layer 1 then already sits in the lanes that need it as
// the B operand of layer 2, once the K order of layer 2 is permuted the same
// way in its pre-arranged A fragments; no LDS round trip between layers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "internal.hpp"
#include "mlp.hpp"

namespace cbmv {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma(f16x8 a, f16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Hidden activations (kActGelu, kActRelu).
//
// GELU, tanh form: x * sigmoid(2 sqrt(2/pi) (x + 0.044715 x^3)). Training
// evaluates it as x * rcp(1 + exp2(x * (-k1 - k3 x^2))) with
// k1 = 2 sqrt(2/pi) log2(e) and k3 = 0.044715 k1.
constexpr float kG1 = 2.3022082f, kG3 = 0.10294324f;

__device__ __forceinline__ float gelu_sig(float x) {
  const float u = x * fmaf(x * x, -kG3, -kG1);
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(u));
}

__device__ __forceinline__ float gelu(float x) { return x * gelu_sig(x); }

__device__ __forceinline__ float gelu_grad(float x) {
  const float s = gelu_sig(x);
  return s + x * s * (1.f - s) * 1.5957691216f * (1.f + 0.134145f * x * x);
}

__device__ __forceinline__ float relu(float x) { return fmaxf(x, 0.f); }

__device__ __forceinline__ float relu_grad(float x) { return x > 0.f ? 1.f : 0.f; }

template <int ACT>
__device__ __forceinline__ float act_fn(float x) {
  return ACT == kActRelu ? relu(x) : gelu(x);
}

template <int ACT>
__device__ __forceinline__ float act_grad(float x) {
  return ACT == kActRelu ? relu_grad(x) : gelu_grad(x);
}

__device__ __forceinline__ float zclamp(float z) { return fminf(fmaxf(z, -8.f), 8.f); }

// ---------------------------------------------------------------- parameter block
// byte offsets of the packed inference parameters
constexpr int kOffMean = 0, kOffIstd = 128, kOffB1 = 256, kOffB2 = 768, kOffW3 = 1024,
              kOffB3 = 1280, kOffF1 = 1536, kOffF2 = kOffF1 + 8 * 64 * 16,
              kParamBytes = kOffF2 + 16 * 64 * 16;

struct MlpView {
  const float* mean;
  const float* istd;
  const float* b1;
  const float* b2;
  const float* w3;
  const float* b3;
  const bf16x8* f1;  // [8 tiles][64 lanes]
  const f16x8* f2;   // [4 tiles][4 k-steps][64 lanes]
};

MlpView view_of(const void* base) {
  const char* p = (const char*)base;
  return MlpView{(const float*)(p + kOffMean), (const float*)(p + kOffIstd),
                 (const float*)(p + kOffB1),   (const float*)(p + kOffB2),
                 (const float*)(p + kOffW3),   (const float*)(p + kOffB3),
                 (const bf16x8*)(p + kOffF1),  (const f16x8*)(p + kOffF2)};
}

uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

// IEEE half bits of a float, rounded to nearest even (NaN stays NaN, from
// 65520 up infinity, below 2^-14 a subnormal)
uint16_t f16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t sign = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (a < 0x38800000u) {
    float v;
    std::memcpy(&v, &a, 4);
    return (uint16_t)(sign | (uint32_t)std::nearbyint(v * 16777216.f));  // units of 2^-24
  }
  const uint32_t r = a - 0x38000000u;  // exponent bias 127 -> 15
  return (uint16_t)(sign | ((r + 0xfffu + ((r >> 13) & 1u)) >> 13));
}

// ---------------------------------------------------------------- inference
constexpr int kTX = 64, kTD = 64, kB = 256;

// Layer-1 input slots: lane group g of the B fragment holds the five
// features of matcher g (slot s < 5 = feature g + 4 s, the order of
// matcher_features_f), and slot 5 of group 0 is the constant 1 that carries
// the layer-1 bias; the packed W1^T fragments use the same slot order.
__device__ __forceinline__ int slot_feature(int g, int s) { return g + 4 * s; }

// Loads of one voxel, issued ahead of use: the four matcher values and the
// right view's statistics of matcher g.
struct VoxelLoads {
  uint32_t cen, sob;
  double ncc, sad;
  float mR, rsR;
};

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Two floats to packed bf16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

// ReLU of two packed 16-bit floats (bf16 or half): a negative one is a
// negative int16, so one v_pk_max_i16 against 0 clears it
// (relu(round(x)) == round(relu(x))).
__device__ __forceinline__ uint32_t relu_pk(uint32_t v) {
  const i16x2 s = __builtin_elementwise_max(__builtin_bit_cast(i16x2, v), i16x2{0, 0});
  return __builtin_bit_cast(uint32_t, s);
}

// Two floats to packed half (round to nearest even, v_cvt_pk_f16_f32).
__device__ __forceinline__ f16x2 to_f16x2(float a, float b) {
  return __builtin_convertvector(f32x2{a, b}, f16x2);
}

constexpr _Float16 hbits(uint16_t b) { return __builtin_bit_cast(_Float16, b); }

typedef float f32x8 __attribute__((ext_vector_type(8)));

// Inference GELU of eight values in packed half arithmetic, without
// transcendentals: x times the sigmoid factor q = clamp(0.5 + x P(x^2), 0, 1),
// P of degree 4. The coefficients are half values: the minimax fit of GELU
// (tanh form) on [0, 3.375] with q(3.375) >= 1, rounded to half from
// the x^8 coefficient down with the lower ones refitted after each rounding,
// the constant picked among its neighbours so that, evaluated in half, q is 1
// for every x above 3.375 and 0 for every x below -3.375 (checked over all
// 63488 finite half inputs: GELU is then exactly x, or exactly 0, there, and
// x needs no clamp). Absolute error against the tanh form, the rounding of x
// to half included: <= 5.6e-3 for |x| <= 12 (4.2e-3 on [-4, -2], 2.6e-3 on
// [-2, 2]), half an ulp of x beyond. The x^8 coefficient is a half subnormal
// (half denormals are kept, the default mode of gfx950 kernels). The clamp is
// the v_pk_fma_f16 clamp bit; eight instructions per pair: v_cvt_pk_f16_f32,
// v_pk_mul, five v_pk_fma, v_pk_mul. Each step is four independent packed
// instructions, so a wavefront issues them back to back instead of waiting on
// one pair's dependent chain.
__device__ __forceinline__ f16x8 gelu_h8(f32x4 lo, f32x4 hi) {
  const f16x8 x = __builtin_convertvector(
      __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7), f16x8);
  const f16x8 s = x * x;
  f16x8 p = hbits(0x00b8);
  p = __builtin_elementwise_fma(p, s, (f16x8)hbits(0x8ef6));
  p = __builtin_elementwise_fma(p, s, (f16x8)hbits(0x1ef2));
  p = __builtin_elementwise_fma(p, s, (f16x8)hbits(0xabac));
  p = __builtin_elementwise_fma(p, s, (f16x8)hbits(0x364f));
  f16x8 q = __builtin_elementwise_fma(x, p, (f16x8)(_Float16)0.5f);
  q = __builtin_elementwise_min(__builtin_elementwise_max(q, (f16x8)(_Float16)0.f),
                                (f16x8)(_Float16)1.f);
  return x * q;
}

// Hidden activation of eight layer outputs (lo, then hi) as half.
template <int ACT>
__device__ __forceinline__ f16x8 act_h8(f32x4 lo, f32x4 hi) {
  if constexpr (ACT == kActRelu) {
    auto r = [](float a, float b) { return relu_pk(__builtin_bit_cast(uint32_t, to_f16x2(a, b))); };
    return __builtin_bit_cast(f16x8, u32x4{r(lo[0], lo[1]), r(lo[2], lo[3]), r(hi[0], hi[1]),
                                           r(hi[2], hi[3])});
  } else {
    return gelu_h8(lo, hi);
  }
}

// Block = 64 x-values x 64 disparities of one row. Wave w takes the
// disparities d0 + w + 4 j; each step handles 16 voxels (16 consecutive x)
// with lane (i, g) computing matcher g's features of voxel i straight into
// its B-fragment slots, then the three layers on MFMA: no LDS staging of
// features and no barrier until the cost tile is written out transposed.
// Loads use the row's base pointers (wave-uniform) with 32-bit lane offsets.
template <int ACT>
__global__ void __launch_bounds__(kB) k_mlp_classify(MlpRun r, MlpView P) {
  __shared__ float tile[kTX][kTD + 1];
  // s_w3h[k]: output weights 2k and 2k + 1 as packed half
  __shared__ __attribute__((aligned(16))) float s_b2[kMlpH2];
  __shared__ __attribute__((aligned(16))) uint32_t s_w3h[kMlpH2 / 2];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int i = lane & 15, g = lane >> 4;
  for (int k = tid; k < kMlpH2; k += kB) s_b2[k] = P.b2[k];
  for (int k = tid; k < kMlpH2 / 2; k += kB)
    s_w3h[k] = __builtin_bit_cast(uint32_t, to_f16x2(P.w3[2 * k], P.w3[2 * k + 1]));
  bf16x8 f1[8];
  f16x8 f2[16];
#pragma unroll
  for (int t = 0; t < 8; t++) f1[t] = P.f1[t * 64 + lane];
#pragma unroll
  for (int q = 0; q < 16; q++) f2[q] = P.f2[q * 64 + lane];
  const float b3 = P.b3[0];
  // z = (f - mean) istd as one fma, f istd + (-mean istd): the rounding
  // differs from the two-step form by an ulp or so, far below the bf16
  // rounding that follows
  float nmi5[5], istd5[5];
#pragma unroll
  for (int s = 0; s < 5; s++) {
    istd5[s] = P.istd[slot_feature(g, s)];
    nmi5[s] = -P.mean[slot_feature(g, s)] * istd5[s];
  }
  const MatcherConst mc = matcher_const(r.fs, g);
  const int W = r.W, D = r.D;
  const int x0 = blockIdx.x * kTX, d0 = blockIdx.y * kTD, y = blockIdx.z;
  const int64_t HW = (int64_t)r.H * W;
  const int64_t vrow = (int64_t)y * D * W;
  const uint8_t* __restrict__ cen_row = r.vols.census + vrow;
  const uint16_t* __restrict__ sob_row = r.vols.sob + vrow;
  const double* __restrict__ ncc_row = r.vols.ncc + vrow;
  const double* __restrict__ sad_row = r.vols.sad + vrow;
  // this row's statistics: sf_row[k * HW + x] is statistic k of pixel (y, x)
  const float* __restrict__ sf_row = r.statsf.s + (int64_t)y * W;
  const int row0 = (y / r.rows_per_batch) * r.rows_per_batch;
  // offset of the batch-start pixel (row0, 0) from pixel (y, 0)
  const int fill_off = (row0 - y) * W;
  const uint32_t kofL = (uint32_t)(g * HW), kofR = (uint32_t)((8 + g) * HW),
                 kofLs = (uint32_t)((4 + g) * HW);
  // left statistics of the lane's four x positions; the batch-start cost
  int xq[4];
  float mL[4], rsL[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    xq[q] = min(x0 + 16 * q + i, W - 1);
    mL[q] = sf_row[kofL + xq[q]];
    rsL[q] = sf_row[kofLs + xq[q]];
  }
  const uint32_t ivalid_bad = g == 0 ? (uint32_t)kCensusInvalid : (uint32_t)kSobInvalid;
  float c0;
  {
    const int64_t b = (int64_t)row0 * D * W;
    const RawCost rc = RawCost{r.vols.census[b], r.vols.sob[b], r.vols.ncc[b], r.vols.sad[b]};
    c0 = raw_cost_f(rc, g);
  }

  // Buffer loads over this row's [D][W] slab of each volume and over the
  // statistics planes from pixel (y, 0) on (32-bit byte counts: one row of a
  // volume and the 16 statistics planes stay below 4 GiB, checked on the
  // host). The disparity's part of a volume address, d W times the element
  // size, is wave-uniform and goes in the scalar offset; the lane's part is
  // its x position, fixed per x group.
  const uint32_t slab = (uint32_t)D * (uint32_t)W;
  const __amdgpu_buffer_rsrc_t bcen = buffer_rsrc(cen_row, slab);
  const __amdgpu_buffer_rsrc_t bsob = buffer_rsrc(sob_row, 2u * slab);
  const __amdgpu_buffer_rsrc_t bncc = buffer_rsrc(ncc_row, 8u * slab);
  const __amdgpu_buffer_rsrc_t bsad = buffer_rsrc(sad_row, 8u * slab);
  const __amdgpu_buffer_rsrc_t bst = buffer_rsrc(sf_row, (uint32_t)((16 * HW - (int64_t)y * W) * 4));
  const uint32_t srs = (uint32_t)(4 * HW) * 4u;  // plane 12 + g from plane 8 + g, in bytes
  auto issue = [&](int j, int q) {
    const int d = min(d0 + wave + 4 * j, D - 1);   // wave-uniform
    const uint32_t sd = (uint32_t)(d * W);
    const uint32_t xo = (uint32_t)xq[q];
    // right-view pixel x - d of this row, or the batch start (row0, 0)
    const int rp = xq[q] >= d ? xq[q] - d : fill_off;
    const uint32_t ro = (uint32_t)((int)kofR + rp) * 4u;
    VoxelLoads v;
    v.cen = buffer_ldu8(bcen, xo, sd);
    v.sob = buffer_ldu16(bsob, xo * 2u, sd * 2u);
    v.ncc = buffer_ldd(bncc, xo * 8u, sd * 8u);
    v.sad = buffer_ldd(bsad, xo * 8u, sd * 8u);
    v.mR = buffer_ldf(bst, ro, 0u);
    v.rsR = buffer_ldf(bst, ro, srs);
    return v;
  };
  VoxelLoads ring[4];
#pragma unroll
  for (int q = 0; q < 4; q++) ring[q] = issue(0, q);
  __syncthreads();

#pragma unroll 2
  for (int j = 0; j < kTD / 4; j++) {
    const int dd = wave + 4 * j;
    const int d = d0 + dd;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const VoxelLoads cur = ring[q];
      ring[q] = issue(min(j + 1, kTD / 4 - 1), q);
      const int x = x0 + 16 * q + i;
      // matcher g's cost without branches: one integer and one double
      // candidate, each selected per lane, then one of the two
      const uint32_t iv = g == 0 ? cur.cen : cur.sob;
      const double dv = g == 1 ? cur.ncc : cur.sad;
      const float fi = iv == ivalid_bad ? (float)kRandMax : (float)iv;
      const float c = (g & 1) ? (float)dv : fi;
      const float cr = x >= d ? c : c0;
      float f[5];
      matcher_features_f(c, cr, mL[q], rsL[q], cur.mR, cur.rsR, mc, f);
      float z[5];
#pragma unroll
      for (int s = 0; s < 5; s++) z[s] = zclamp(fmaf(f[s], istd5[s], nmi5[s]));
      const bf16x8 zb = __builtin_bit_cast(
          bf16x8, u32x4{pack_bf16(z[0], z[1]), pack_bf16(z[2], z[3]),
                        pack_bf16(z[4], g == 0 ? 1.f : 0.f), 0u});
      // layer 2 accumulators start at the bias; layer 1 runs two 16-unit
      // tiles at a time and feeds each 32-unit k-step as soon as it is ready
      f32x4 a2[4];
#pragma unroll
      for (int t2 = 0; t2 < 4; t2++) a2[t2] = *(const f32x4*)&s_b2[16 * t2 + 4 * g];
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        const f32x4 lo = mfma(f1[2 * ks], zb, f32x4{0.f, 0.f, 0.f, 0.f});
        const f32x4 hi = mfma(f1[2 * ks + 1], zb, f32x4{0.f, 0.f, 0.f, 0.f});
        // the activations in the layer-2 K order: units 4g + r of the low
        // tile, then of the high tile
        const f16x8 hb = act_h8<ACT>(lo, hi);
#pragma unroll
        for (int t2 = 0; t2 < 4; t2++) a2[t2] = mfma(f2[t2 * 4 + ks], hb, a2[t2]);
      }
      // layer 3: unit 16 t2 + 4 g + u of this lane, two at a time through
      // v_dot2_f32_f16 (half products, float sum); the pairs are taken with
      // shuffles (element reads of a u32x4 bit cast of the activations were
      // compiled to one pair for all four)
      float part = 0.f;
#pragma unroll
      for (int t2 = 0; t2 < 4; t2 += 2) {
        const f16x8 a = act_h8<ACT>(a2[t2], a2[t2 + 1]);
        const uint2 w3a = *(const uint2*)&s_w3h[8 * t2 + 2 * g];
        const uint2 w3b = *(const uint2*)&s_w3h[8 * (t2 + 1) + 2 * g];
        part = __builtin_amdgcn_fdot2(__builtin_shufflevector(a, a, 0, 1),
                                      __builtin_bit_cast(f16x2, w3a.x), part, false);
        part = __builtin_amdgcn_fdot2(__builtin_shufflevector(a, a, 2, 3),
                                      __builtin_bit_cast(f16x2, w3a.y), part, false);
        part = __builtin_amdgcn_fdot2(__builtin_shufflevector(a, a, 4, 5),
                                      __builtin_bit_cast(f16x2, w3b.x), part, false);
        part = __builtin_amdgcn_fdot2(__builtin_shufflevector(a, a, 6, 7),
                                      __builtin_bit_cast(f16x2, w3b.y), part, false);
      }
      part += __shfl_xor(part, 16, 64);
      part += __shfl_xor(part, 32, 64);
      // the logit; the sigmoid is applied once per voxel when the tile is
      // written out (here all 64 lanes would evaluate it for 16 voxels)
      if (g == 0) tile[16 * q + i][dd] = part + b3;
    }
  }
  __syncthreads();
  auto sigmoid = [](float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); };
  for (int j = 0; j < kTX / 4; j++) {
    const int rr = wave + 4 * j;
    const int xx = x0 + rr, d = d0 + lane;
    if (xx < W && d < D) r.cost[((int64_t)y * W + xx) * D + d] = 1.f - sigmoid(tile[rr][lane]);
  }
  if (r.pmax != nullptr && tid < kTX && x0 + tid < W) {
    // best match probability of the pixel over this block's disparities,
    // the sigmoid of the largest logit (the sigmoid is monotonic); p >= 0,
    // so the float's bits order like unsigned integers
    float zm = -INFINITY;
    for (int dd = 0; dd < kTD && d0 + dd < D; dd++) zm = fmaxf(zm, tile[tid][dd]);
    atomicMax((unsigned int*)&r.pmax[(int64_t)y * W + x0 + tid], __float_as_uint(sigmoid(zm)));
  }
}

// ---------------------------------------------------------------- gate
// Clears the per-pixel maxima and the list counter (a fill kernel, so the
// sequence stays valid under stream capture).
__global__ void k_gate_clear(float* __restrict__ pmax, int32_t* __restrict__ flags, int64_t HW) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < HW;
       k += (int64_t)gridDim.x * blockDim.x)
    pmax[k] = 0.f;
  if (blockIdx.x == 0 && threadIdx.x == 0) flags[HW] = 0;
}

// Pixels whose best match probability is below the threshold go to the forest.
__global__ void k_gate_list(const float* __restrict__ pmax, int64_t HW, float thr,
                            int32_t* __restrict__ flags) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= HW) return;
  if (pmax[p] < thr) {
    const int k = atomicAdd(&flags[HW], 1);
    flags[k] = (int32_t)p;
  }
}

// Random forest over every disparity of the listed pixels (float64 features,
// as in the forest path). Fixed grid; every block walks the list.
__global__ void __launch_bounds__(kB) k_gate_forest(MlpRun r) {
  __shared__ float xs[20][kB];
  const int64_t HW = (int64_t)r.H * r.W;
  const int count = r.gate_flags[HW];
  const int tid = threadIdx.x;
  for (int k = blockIdx.x; k < count; k += gridDim.x) {
    const int64_t p = r.gate_flags[k];
    const int y = (int)(p / r.W), x = (int)(p % r.W);
    for (int d = tid; d < r.D; d += kB) {
      voxel_features(r.vols, r.stats, r.fs, r.H, r.W, r.D, r.rows_per_batch, y, x, d,
                     [&](int q, float v) { xs[q][tid] = v; });
      const double acc = forest_sum0<8>(r.forest, [&](uint32_t q) { return xs[q][tid]; });
      r.cost[p * r.D + d] = (float)(acc / r.forest.n_trees);
    }
  }
}

// ---------------------------------------------------------------- training
// Parameter vector order (float32): w1[20x128], b1[128], w2[128x64], b2[64], w3[64], b3.
constexpr int kPW1 = 0, kPB1 = kPW1 + kMlpIn * kMlpH1, kPW2 = kPB1 + kMlpH1,
              kPB2 = kPW2 + kMlpH1 * kMlpH2, kPW3 = kPB2 + kMlpH2, kPB3 = kPW3 + kMlpH2,
              kNParams = kPB3 + 1;
constexpr int kChunk = 32;  // samples per block iteration

struct TrainView {
  float* P;       // parameters
  float* G;       // gradient accumulators
  float* M;       // Adam first moments
  float* V;       // Adam second moments
  float* mean;    // [32]
  float* istd;    // [32]
  __bf16* w1t;    // [128][32]  W1^T, zero-padded inputs
  __bf16* w2;     // [128][64]  W2
  __bf16* w2t;    // [64][128]  W2^T
  float* loss;    // [1]
};

// acc += A[m0.., k] * B[k, n0..] over K (multiple of 32); A row-major with
// leading dimension lda, B given transposed (BT[n][k], leading dimension ldb).
__device__ __forceinline__ f32x4 tile_gemm(const __bf16* a, int lda, const __bf16* bt, int ldb,
                                           int K, f32x4 acc, int i, int g) {
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bf16x8 af = *(const bf16x8*)(a + i * lda + k0 + 8 * g);
    const bf16x8 bf = *(const bf16x8*)(bt + i * ldb + k0 + 8 * g);
    acc = mfma(af, bf, acc);
  }
  return acc;
}

// mode 0: forward + backward, gradients accumulated into t.G, loss into t.loss
// mode 1: forward only, p(match) written to out
template <int ACT>
__global__ void __launch_bounds__(kB) k_mlp_train(TrainView t, const float* __restrict__ X,
                                                  const float* __restrict__ Y, int64_t B,
                                                  float inv_B, int mode, float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) __bf16 z[kChunk][kMlpInPad];
  __shared__ __attribute__((aligned(16))) __bf16 zT[kMlpInPad][kChunk];
  __shared__ __attribute__((aligned(16))) __bf16 w1t[kMlpH1][kMlpInPad];
  __shared__ __attribute__((aligned(16))) __bf16 a1[kChunk][kMlpH1];
  __shared__ __attribute__((aligned(16))) __bf16 a1T[kMlpH1][kChunk];  // reused as dh1T
  __shared__ __attribute__((aligned(16))) __bf16 w2t[kMlpH2][kMlpH1];
  __shared__ __attribute__((aligned(16))) __bf16 w2[kMlpH1][kMlpH2];
  __shared__ __attribute__((aligned(16))) __bf16 dh2[kChunk][kMlpH2];
  __shared__ __attribute__((aligned(16))) __bf16 dh2T[kMlpH2][kChunk];
  __shared__ float h1[kChunk][kMlpH1];
  __shared__ float h2[kChunk][kMlpH2];
  __shared__ float s_b1[kMlpH1], s_b2[kMlpH2], s_w3[kMlpH2], s_mean[kMlpInPad], s_istd[kMlpInPad];
  __shared__ float dlog[kChunk];
  __shared__ float s_gw3[kMlpH2], s_gb2[kMlpH2], s_gb1[kMlpH1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  for (int k = tid; k < kMlpH1 * kMlpInPad; k += kB) (&w1t[0][0])[k] = t.w1t[k];
  for (int k = tid; k < kMlpH1 * kMlpH2; k += kB) {
    (&w2[0][0])[k] = t.w2[k];
    (&w2t[0][0])[k] = t.w2t[k];
  }
  for (int k = tid; k < kMlpH1; k += kB) {
    s_b1[k] = t.P[kPB1 + k];
    s_gb1[k] = 0.f;
  }
  for (int k = tid; k < kMlpH2; k += kB) {
    s_b2[k] = t.P[kPB2 + k];
    s_w3[k] = t.P[kPW3 + k];
    s_gw3[k] = 0.f;
    s_gb2[k] = 0.f;
  }
  for (int k = tid; k < kMlpInPad; k += kB) {
    s_mean[k] = t.mean[k];
    s_istd[k] = t.istd[k];
  }
  const float b3 = t.P[kPB3];
  // per-wave weight-gradient accumulators: gW2 tiles (8 of 32) and gW1 tiles (4 of 16)
  f32x4 gw2[8], gw1[4];
#pragma unroll
  for (int q = 0; q < 8; q++) gw2[q] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 4; q++) gw1[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  float gb3 = 0.f, lsum = 0.f;
  __syncthreads();
  const int64_t nchunks = (B + kChunk - 1) / kChunk;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int64_t s0 = c * kChunk;
    // standardised inputs
    for (int k = tid; k < kChunk * kMlpInPad; k += kB) {
      const int s = k / kMlpInPad, f = k % kMlpInPad;
      float v = 0.f;
      if (f < kMlpIn && s0 + s < B) v = zclamp((X[(s0 + s) * kMlpIn + f] - s_mean[f]) * s_istd[f]);
      z[s][f] = (__bf16)v;
      zT[f][s] = (__bf16)v;
    }
    __syncthreads();
    // H1 = Z W1: 2 x 8 tiles, 4 per wave
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tile = wave * 4 + q, tm = tile >> 3, tn = tile & 7;
      f32x4 acc = tile_gemm(&z[16 * tm][0], kMlpInPad, &w1t[16 * tn][0], kMlpInPad, kMlpInPad,
                            f32x4{0.f, 0.f, 0.f, 0.f}, i, g);
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int s = 16 * tm + 4 * g + rr, h = 16 * tn + i;
        const float pre = acc[rr] + s_b1[h];
        h1[s][h] = pre;
        const __bf16 av = (__bf16)act_fn<ACT>(pre);
        a1[s][h] = av;
        a1T[h][s] = av;
      }
    }
    __syncthreads();
    // H2 = A1 W2: 2 x 4 tiles, 2 per wave
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int tile = wave * 2 + q, tm = tile >> 2, tn = tile & 3;
      f32x4 acc = tile_gemm(&a1[16 * tm][0], kMlpH1, &w2t[16 * tn][0], kMlpH1, kMlpH1,
                            f32x4{0.f, 0.f, 0.f, 0.f}, i, g);
#pragma unroll
      for (int rr = 0; rr < 4; rr++) h2[16 * tm + 4 * g + rr][16 * tn + i] = acc[rr] + s_b2[16 * tn + i];
    }
    __syncthreads();
    // logits: thread s < 32 sums over the 64 hidden units
    if (tid < kChunk) {
      const int s = tid;
      float l = b3;
      for (int h = 0; h < kMlpH2; h++) l += act_fn<ACT>(h2[s][h]) * s_w3[h];
      const float p = 1.f / (1.f + __expf(-l));
      if (s0 + s < B) {
        if (mode == 1) {
          out[s0 + s] = p;
        } else {
          const float yv = Y[s0 + s];
          // BCE with logits: softplus(l) - y l; d/dl = sigmoid(l) - y
          lsum += fmaxf(l, 0.f) + log1pf(__expf(-fabsf(l))) - yv * l;
          dlog[s] = (p - yv) * inv_B;
        }
      } else {
        dlog[s] = 0.f;
      }
    }
    if (mode == 1) {
      __syncthreads();
      continue;
    }
    __syncthreads();
    // dH2 = dlogit w3 * act'(h2); gw3, gb2, gb3 partial sums
    for (int k = tid; k < kChunk * kMlpH2; k += kB) {
      const int s = k / kMlpH2, h = k % kMlpH2;
      const float pre = h2[s][h];
      const float dl = dlog[s];
      const float dv = dl * s_w3[h] * act_grad<ACT>(pre);
      const __bf16 db = (__bf16)dv;
      dh2[s][h] = db;
      dh2T[h][s] = db;
      atomicAdd(&s_gw3[h], dl * act_fn<ACT>(pre));
      atomicAdd(&s_gb2[h], dv);
    }
    if (tid < kChunk) gb3 += dlog[tid];
    __syncthreads();
    // gW2 += A1^T dH2 (128 x 64, K = chunk): 32 tiles, 8 per wave
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int tile = wave * 8 + q, tm = tile >> 2, tn = tile & 3;
      gw2[q] = tile_gemm(&a1T[16 * tm][0], kChunk, &dh2T[16 * tn][0], kChunk, kChunk, gw2[q], i, g);
    }
    // dA1 = dH2 W2^T (chunk x 128, K = 64): 16 tiles, 4 per wave; dH1 -> h1 (f32)
    f32x4 da1[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tile = wave * 4 + q, tm = tile >> 3, tn = tile & 7;
      da1[q] = tile_gemm(&dh2[16 * tm][0], kMlpH2, &w2[16 * tn][0], kMlpH2, kMlpH2,
                         f32x4{0.f, 0.f, 0.f, 0.f}, i, g);
    }
    __syncthreads();  // a1T no longer read: reuse it as dH1^T
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tile = wave * 4 + q, tm = tile >> 3, tn = tile & 7;
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int s = 16 * tm + 4 * g + rr, h = 16 * tn + i;
        const float dv = da1[q][rr] * act_grad<ACT>(h1[s][h]);
        a1T[h][s] = (__bf16)dv;
        atomicAdd(&s_gb1[h], dv);
      }
    }
    __syncthreads();
    // gW1 += Z^T dH1 (32 x 128, K = chunk): 16 tiles, 4 per wave
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tile = wave * 4 + q, tm = tile >> 3, tn = tile & 7;
      gw1[q] = tile_gemm(&zT[16 * tm][0], kChunk, &a1T[16 * tn][0], kChunk, kChunk, gw1[q], i, g);
    }
    __syncthreads();
  }
  if (mode == 1) return;
  // flush the block's gradient sums
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int tile = wave * 8 + q, tm = tile >> 2, tn = tile & 3;
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
      atomicAdd(&t.G[kPW2 + (16 * tm + 4 * g + rr) * kMlpH2 + 16 * tn + i], gw2[q][rr]);
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int tile = wave * 4 + q, tm = tile >> 3, tn = tile & 7;
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const int f = 16 * tm + 4 * g + rr;
      if (f < kMlpIn) atomicAdd(&t.G[kPW1 + f * kMlpH1 + 16 * tn + i], gw1[q][rr]);
    }
  }
  __syncthreads();
  for (int k = tid; k < kMlpH1; k += kB) atomicAdd(&t.G[kPB1 + k], s_gb1[k]);
  for (int k = tid; k < kMlpH2; k += kB) {
    atomicAdd(&t.G[kPB2 + k], s_gb2[k]);
    atomicAdd(&t.G[kPW3 + k], s_gw3[k]);
  }
  if (tid < kChunk) {
    atomicAdd(&t.G[kPB3], gb3);
    atomicAdd(t.loss, lsum);
  }
}

__global__ void k_fill(float* a, int n, float v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}

// Adam (float32 master weights); clears the gradients.
__global__ void k_adam(TrainView t, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNParams) return;
  const float gr = t.G[k];
  const float m = b1 * t.M[k] + (1.f - b1) * gr;
  const float v = b2 * t.V[k] + (1.f - b2) * gr * gr;
  t.M[k] = m;
  t.V[k] = v;
  t.P[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  t.G[k] = 0.f;
}

// bf16 copies of the weights for the MFMA kernels.
__global__ void k_refresh_bf16(TrainView t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kMlpH1 * kMlpInPad) {
    const int h = k / kMlpInPad, f = k % kMlpInPad;
    t.w1t[k] = (__bf16)(f < kMlpIn ? t.P[kPW1 + f * kMlpH1 + h] : 0.f);
  }
  if (k < kMlpH1 * kMlpH2) {
    const int a = k / kMlpH2, b = k % kMlpH2;   // w2[a][b]
    t.w2[k] = (__bf16)t.P[kPW2 + k];
    t.w2t[b * kMlpH1 + a] = (__bf16)t.P[kPW2 + k];
  }
}

__global__ void k_probe(const float* A, const float* Bm, float* C) {
  const int l = threadIdx.x, i = l & 15, g = l >> 4;
  bf16x8 a, b;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    a[j] = (__bf16)A[i * 32 + 8 * g + j];
    b[j] = (__bf16)Bm[(8 * g + j) * 16 + i];
  }
  const f32x4 c = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
  for (int r = 0; r < 4; r++) C[(4 * g + r) * 16 + i] = c[r];
}

// The inference activation of groups of eight floats (diagnostic).
__global__ void k_act_probe(const float* __restrict__ x, float* __restrict__ y, int64_t groups,
                            int act) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= groups) return;
  const f32x4 lo = *(const f32x4*)&x[8 * k], hi = *(const f32x4*)&x[8 * k + 4];
  const f16x8 h = act == kActRelu ? act_h8<kActRelu>(lo, hi) : act_h8<kActGelu>(lo, hi);
#pragma unroll
  for (int j = 0; j < 8; j++) y[8 * k + j] = (float)h[j];
}

}  // namespace

// ---------------------------------------------------------------- host side
int mlp_upload(MlpDevice& m, const float* mean, const float* inv_std, const float* w1,
               const float* b1, const float* w2, const float* b2, const float* w3,
               const float* b3, int act) {
  if (act != kActGelu && act != kActRelu) {
    set_error("unknown MLP activation %d", act);
    return CBMV_EINVAL;
  }
  // every parameter finite
  auto all_finite = [](const float* v, int n) {
    for (int i = 0; i < n; i++)
      if (!std::isfinite(v[i])) return false;
    return true;
  };
  if (!(all_finite(mean, kMlpIn) && all_finite(inv_std, kMlpIn) &&
        all_finite(w1, kMlpIn * kMlpH1) && all_finite(b1, kMlpH1) &&
        all_finite(w2, kMlpH1 * kMlpH2) && all_finite(b2, kMlpH2) && all_finite(w3, kMlpH2) &&
        all_finite(b3, 1))) {
    set_error("MLP parameters must be finite");
    return CBMV_EINVAL;
  }
  // the hidden activations enter layers 2 and 3 as half values: bound the
  // pre-activations below the half range (standardised inputs are clamped to
  // +-8; |GELU(x)| and ReLU(x) are at most max(|x|, 0.17)). Layer 1's
  // weights and bias are used as bf16 (relative rounding at most 2^-9),
  // layer 2's and w3 as half (2^-11): the bounds take the rounded values'
  // largest magnitudes, |v| (1 + 2^-8) and |v| (1 + 2^-10).
  const double rb = 1.0 + 1.0 / 256, rh = 1.0 + 1.0 / 1024;
  double lim1 = 0.0, lim2 = 0.0, lim3 = 0.0, limw2 = 0.0;
  for (int h = 0; h < kMlpH1; h++) {
    double v = rb * std::fabs((double)b1[h]);
    for (int f = 0; f < kMlpIn; f++) v += 8.0 * rb * std::fabs((double)w1[f * kMlpH1 + h]);
    lim1 = std::max(lim1, v);
  }
  for (int k = 0; k < kMlpH2; k++) {
    double v = std::fabs((double)b2[k]);
    for (int h = 0; h < kMlpH1; h++) {
      const double a = rh * std::fabs((double)w2[h * kMlpH2 + k]);
      limw2 = std::max(limw2, a);
      v += std::max(lim1, 0.17) * a;
    }
    lim2 = std::max(lim2, v);
    lim3 = std::max(lim3, rh * std::fabs((double)w3[k]));
  }
  if (!(lim1 < 65504.0 && lim2 < 65504.0 && limw2 < 65504.0 && lim3 < 65504.0)) {
    set_error("MLP weights too large for half-precision hidden activations "
              "(pre-activation bounds %g, %g; |w2|, |w3| up to %g, %g; the limit is 65504)",
              lim1, lim2, limw2, lim3);
    return CBMV_EINVAL;
  }
  std::vector<char> buf(kParamBytes, 0);
  float* fm = (float*)(buf.data() + kOffMean);
  float* fi = (float*)(buf.data() + kOffIstd);
  for (int k = 0; k < kMlpIn; k++) {
    fm[k] = mean[k];
    fi[k] = inv_std[k];
  }
  std::memcpy(buf.data() + kOffB1, b1, kMlpH1 * 4);
  std::memcpy(buf.data() + kOffB2, b2, kMlpH2 * 4);
  std::memcpy(buf.data() + kOffW3, w3, kMlpH2 * 4);
  std::memcpy(buf.data() + kOffB3, b3, 4);
  uint16_t* f1 = (uint16_t*)(buf.data() + kOffF1);
  uint16_t* f2 = (uint16_t*)(buf.data() + kOffF2);
  // layer 1 A fragments: W1^T[16t + i][slot 8g + j], slot j < 5 of group g
  // = feature g + 4j, slot 5 of group 0 = the bias (its input is 1)
  for (int t = 0; t < 8; t++)
    for (int l = 0; l < 64; l++)
      for (int j = 0; j < 8; j++) {
        const int i = l & 15, g = l >> 4, h = 16 * t + i;
        float w = 0.f;
        if (j < 5)
          w = w1[(g + 4 * j) * kMlpH1 + h];
        else if (g == 0 && j == 5)
          w = b1[h];
        f1[(t * 64 + l) * 8 + j] = bf16_bits(w);
      }
  // layer 2 A fragments (half) with K permuted to the layer-1 accumulator order:
  // slot j of k-step ks in lane group g is hidden unit 32ks + 4g + j (j < 4)
  // or 32ks + 16 + 4g + (j - 4)
  for (int t2 = 0; t2 < 4; t2++)
    for (int ks = 0; ks < 4; ks++)
      for (int l = 0; l < 64; l++)
        for (int j = 0; j < 8; j++) {
          const int i = l & 15, g = l >> 4;
          const int h1 = j < 4 ? 32 * ks + 4 * g + j : 32 * ks + 16 + 4 * g + (j - 4);
          const int h2 = 16 * t2 + i;
          f2[((t2 * 4 + ks) * 64 + l) * 8 + j] = f16_bits(w2[h1 * kMlpH2 + h2]);
        }
  int rc = to_device(m.params, buf.data(), buf.size());
  if (rc) return rc;
  rc = sync();
  if (rc) return rc;
  m.act = act;
  m.ready = true;
  return CBMV_OK;
}

int mlp_classify(const MlpDevice& m, const MlpRun& r, hipStream_t s) {
  if (!m.ready) {
    set_error("MLP weights not set");
    return CBMV_ESTATE;
  }
  const MlpView P = view_of(m.params.p);
  const int64_t HW = (int64_t)r.H * r.W;
  // the kernel addresses a volume row and the statistics planes with 32-bit
  // byte offsets
  if ((int64_t)r.D * r.W * 8 > (int64_t)UINT32_MAX || HW * 16 * 4 > (int64_t)INT32_MAX) {
    set_error("MLP classifier: image too large (W*D*8 bytes must stay below 4 GiB, H*W*64 below 2 GiB)");
    return CBMV_EINVAL;
  }
  MlpRun rr = r;
  if (!r.gate) rr.pmax = nullptr;
  if (r.gate) k_gate_clear<<<1024, kB, 0, s>>>(r.pmax, r.gate_flags, HW);
  dim3 grid((r.W + kTX - 1) / kTX, (r.D + kTD - 1) / kTD, r.H);
  if (m.act == kActRelu)
    k_mlp_classify<kActRelu><<<grid, kB, 0, s>>>(rr, P);
  else
    k_mlp_classify<kActGelu><<<grid, kB, 0, s>>>(rr, P);
  CBMV_LAUNCH_CHECK();
  if (r.gate) {
    k_gate_list<<<grid_1d(HW, kB), kB, 0, s>>>(r.pmax, HW, r.gate_threshold, r.gate_flags);
    k_gate_forest<<<2048, kB, 0, s>>>(r);
    CBMV_LAUNCH_CHECK();
  }
  return CBMV_OK;
}

}  // namespace cbmv

// ===========================================================================
// Trainer C ABI
// ===========================================================================
using namespace cbmv;

struct cbmv_mlp_trainer {
  DevBuf P, G, M, V, mean, istd, w1t, w2, w2t, loss;
  int step = 0;
  int act = kActGelu;
  TrainView view() {
    return TrainView{P.as<float>(),   G.as<float>(),    M.as<float>(),   V.as<float>(),
                     mean.as<float>(), istd.as<float>(), w1t.as<__bf16>(), w2.as<__bf16>(),
                     w2t.as<__bf16>(), loss.as<float>()};
  }
};

#define RC(x)              \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

static int refresh(cbmv_mlp_trainer* t, hipStream_t s) {
  k_refresh_bf16<<<(kMlpH1 * kMlpH2 + 255) / 256, 256, 0, s>>>(t->view());
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

static int launch_train(int act, int grid, hipStream_t s, TrainView v, const float* X,
                        const float* Y, int64_t B, float inv_B, int mode, float* out) {
  if (act == kActRelu)
    k_mlp_train<kActRelu><<<grid, kB, 0, s>>>(v, X, Y, B, inv_B, mode, out);
  else
    k_mlp_train<kActGelu><<<grid, kB, 0, s>>>(v, X, Y, B, inv_B, mode, out);
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

static int grid_for(int64_t B) {
  const int64_t nchunks = (B + kChunk - 1) / kChunk;
  return (int)std::min<int64_t>(nchunks, 1024);
}

extern "C" {

cbmv_mlp_trainer* cbmv_mlp_trainer_create(const float* mean, const float* inv_std, const float* w1,
                                          const float* b1, const float* w2, const float* b2,
                                          const float* w3, const float* b3, int32_t activation) {
  clear_error();
  if (activation != kActGelu && activation != kActRelu) {
    set_error("unknown MLP activation %d", activation);
    return nullptr;
  }
  auto* t = new cbmv_mlp_trainer();
  t->act = activation;
  std::vector<float> p(kNParams, 0.f), zeros(kNParams, 0.f), mm(32, 0.f), ii(32, 0.f);
  std::memcpy(&p[kPW1], w1, kMlpIn * kMlpH1 * 4);
  std::memcpy(&p[kPB1], b1, kMlpH1 * 4);
  std::memcpy(&p[kPW2], w2, kMlpH1 * kMlpH2 * 4);
  std::memcpy(&p[kPB2], b2, kMlpH2 * 4);
  std::memcpy(&p[kPW3], w3, kMlpH2 * 4);
  p[kPB3] = b3[0];
  for (int k = 0; k < kMlpIn; k++) {
    mm[k] = mean[k];
    ii[k] = inv_std[k];
  }
  int rc = 0;
  rc |= to_device(t->P, p.data(), kNParams * 4);
  rc |= to_device(t->G, zeros.data(), kNParams * 4);
  rc |= to_device(t->M, zeros.data(), kNParams * 4);
  rc |= to_device(t->V, zeros.data(), kNParams * 4);
  rc |= to_device(t->mean, mm.data(), 32 * 4);
  rc |= to_device(t->istd, ii.data(), 32 * 4);
  rc |= t->w1t.alloc(kMlpH1 * kMlpInPad * 2);
  rc |= t->w2.alloc(kMlpH1 * kMlpH2 * 2);
  rc |= t->w2t.alloc(kMlpH1 * kMlpH2 * 2);
  rc |= to_device(t->loss, zeros.data(), 4);
  if (rc || refresh(t, lib_stream()) || sync()) {
    delete t;
    return nullptr;
  }
  return t;
}

int cbmv_mlp_trainer_step(cbmv_mlp_trainer* t, const float* X, const float* Y, int64_t B, float lr,
                          float* loss, void* stream) {
  clear_error();
  CBMV_REQUIRE(t && X && Y && B > 0, "bad trainer arguments");
  hipStream_t s = (hipStream_t)stream;
  TrainView v = t->view();
  k_fill<<<1, 64, 0, s>>>(v.loss, 1, 0.f);
  RC(launch_train(t->act, grid_for(B), s, v, X, Y, B, 1.f / (float)B, 0, nullptr));
  CBMV_LAUNCH_CHECK();
  t->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)t->step), c2 = 1.f - std::pow(b2, (float)t->step);
  k_adam<<<(kNParams + 255) / 256, 256, 0, s>>>(v, lr, b1, b2, 1e-8f, c1, c2);
  CBMV_LAUNCH_CHECK();
  RC(refresh(t, s));
  if (loss) {
    CBMV_HIP(hipStreamSynchronize(s));
    float l = 0;
    CBMV_HIP(hipMemcpy(&l, v.loss, 4, hipMemcpyDeviceToHost));
    *loss = l / (float)B;
  }
  return CBMV_OK;
}

int cbmv_mlp_trainer_predict(cbmv_mlp_trainer* t, const float* X, int64_t N, float* P,
                             void* stream) {
  clear_error();
  CBMV_REQUIRE(t && X && P && N > 0, "bad trainer arguments");
  RC(launch_train(t->act, grid_for(N), (hipStream_t)stream, t->view(), X, nullptr, N, 1.f, 1, P));
  CBMV_LAUNCH_CHECK();
  return CBMV_OK;
}

int cbmv_mlp_trainer_gradients(cbmv_mlp_trainer* t, const float* X, const float* Y, int64_t B,
                               float* grad_host, float* loss) {
  clear_error();
  CBMV_REQUIRE(t && X && Y && B > 0 && grad_host, "bad trainer arguments");
  hipStream_t s = lib_stream();
  TrainView v = t->view();
  k_fill<<<(kNParams + 255) / 256, 256, 0, s>>>(v.G, kNParams, 0.f);
  k_fill<<<1, 64, 0, s>>>(v.loss, 1, 0.f);
  RC(launch_train(t->act, grid_for(B), s, v, X, Y, B, 1.f / (float)B, 0, nullptr));
  CBMV_LAUNCH_CHECK();
  CBMV_HIP(hipStreamSynchronize(s));
  CBMV_HIP(hipMemcpy(grad_host, v.G, kNParams * 4, hipMemcpyDeviceToHost));
  float l = 0;
  CBMV_HIP(hipMemcpy(&l, v.loss, 4, hipMemcpyDeviceToHost));
  if (loss) *loss = l / (float)B;
  k_fill<<<(kNParams + 255) / 256, 256, 0, s>>>(v.G, kNParams, 0.f);
  CBMV_LAUNCH_CHECK();
  return sync();
}

int cbmv_mlp_trainer_weights(cbmv_mlp_trainer* t, float* w1, float* b1, float* w2, float* b2,
                             float* w3, float* b3) {
  clear_error();
  CBMV_REQUIRE(t, "null trainer");
  std::vector<float> p(kNParams);
  CBMV_HIP(hipDeviceSynchronize());
  CBMV_HIP(hipMemcpy(p.data(), t->P.p, kNParams * 4, hipMemcpyDeviceToHost));
  std::memcpy(w1, &p[kPW1], kMlpIn * kMlpH1 * 4);
  std::memcpy(b1, &p[kPB1], kMlpH1 * 4);
  std::memcpy(w2, &p[kPW2], kMlpH1 * kMlpH2 * 4);
  std::memcpy(b2, &p[kPB2], kMlpH2 * 4);
  std::memcpy(w3, &p[kPW3], kMlpH2 * 4);
  b3[0] = p[kPB3];
  return CBMV_OK;
}

void cbmv_mlp_trainer_destroy(cbmv_mlp_trainer* t) {
  if (t) (void)hipDeviceSynchronize();
  delete t;
}

int cbmv_mfma_probe(const float* A, const float* Bm, float* C) {
  clear_error();
  DevBuf da, db, dc;
  RC(to_device(da, A, 16 * 32 * 4));
  RC(to_device(db, Bm, 32 * 16 * 4));
  RC(dc.alloc(16 * 16 * 4));
  k_probe<<<1, 64, 0, lib_stream()>>>(da.as<float>(), db.as<float>(), dc.as<float>());
  CBMV_LAUNCH_CHECK();
  return to_host(C, dc, 16 * 16 * 4);
}

int cbmv_act_probe(const float* x, float* y, int64_t n, int32_t activation) {
  clear_error();
  if (n < 0 || (n & 7) || n > (int64_t(1) << 28) || (n > 0 && (!x || !y)) ||
      (activation != kActGelu && activation != kActRelu)) {
    set_error("cbmv_act_probe: bad arguments (n = %lld, activation = %d)", (long long)n,
              (int)activation);
    return CBMV_EINVAL;
  }
  if (n == 0) return CBMV_OK;
  DevBuf dx, dy;
  RC(to_device(dx, x, n * 4));
  RC(dy.alloc(n * 4));
  const int64_t groups = n / 8;
  k_act_probe<<<(unsigned)((groups + 255) / 256), 256, 0, lib_stream()>>>(dx.as<float>(),
                                                                          dy.as<float>(), groups,
                                                                          activation);
  CBMV_LAUNCH_CHECK();
  return to_host(y, dy, n * 4);
}

}  // extern "C"

// Shared host/device helpers of libcbmv_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/cbmv_abi.h"

namespace cbmv {

// The reference's "no valid cost" value (glibc RAND_MAX), emitted exactly in
// float64 outputs.
constexpr double kRandMax = 2147483647.0;
// Sentinels of the compact per-matcher volumes of the fused pipeline.
constexpr uint8_t kCensusInvalid = 255;
constexpr uint16_t kSobInvalid = 65535;

void set_error(const char* fmt, ...);
void clear_error();

// Library stream used by the synchronous drop-in entry points.
hipStream_t lib_stream();

// Number of compute units of the current device.
int device_cus();

// Checks a HIP call inside a function returning int status.
#define CBMV_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::cbmv::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                      \
      return CBMV_EHIP;                                                                 \
    }                                                                                   \
  } while (0)

// Checks the launch status of the last kernel.
#define CBMV_LAUNCH_CHECK() CBMV_HIP(hipGetLastError())

#define CBMV_REQUIRE(cond, ...)          \
  do {                                   \
    if (!(cond)) {                       \
      ::cbmv::set_error(__VA_ARGS__);    \
      return CBMV_EINVAL;                \
    }                                    \
  } while (0)

// Owning device allocation (drop-in entry points allocate per call).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int alloc(size_t n);
  void release();
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Host->device, device->host helpers on the library stream (synchronous).
int to_device(DevBuf& b, const void* host, size_t bytes);
int to_host(void* host, const DevBuf& b, size_t bytes);
int sync();

inline unsigned grid_1d(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > 0x7fffffff) g = 0x7fffffff;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace cbmv

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
namespace cbmv {

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// DPP cross-lane moves (register-to-register, no LDS round trip). A lane
// whose source is outside the row/wave, or whose row is masked off, keeps
// its own value.
#define CBMV_DPP_F(v, ctrl, rmask)                                                   \
  __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), \
                                             (ctrl), (rmask), 0xf, false))

// Minimum over the 64 lanes of a wavefront, returned to every lane:
// pair and quad swaps, half-row and row mirrors give each row of 16 its
// minimum; row_bcast:15 and row_bcast:31 fold the rows into lane 63.
__device__ __forceinline__ float wave_min_dpp(float v) {
  v = fminf(v, CBMV_DPP_F(v, 0xB1, 0xf));   // quad_perm [1,0,3,2]
  v = fminf(v, CBMV_DPP_F(v, 0x4E, 0xf));   // quad_perm [2,3,0,1]
  v = fminf(v, CBMV_DPP_F(v, 0x141, 0xf));  // row_half_mirror
  v = fminf(v, CBMV_DPP_F(v, 0x140, 0xf));  // row_mirror
  v = fminf(v, CBMV_DPP_F(v, 0x142, 0xa));  // row_bcast:15 into rows 1, 3
  v = fminf(v, CBMV_DPP_F(v, 0x143, 0xc));  // row_bcast:31 into rows 2, 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// The same over unsigned integers (no NaN handling; lanes without a source
// see all ones, the identity of the minimum, so each step folds into one
// v_min_u32 with a DPP operand).
#define CBMV_DPP_U(v, ctrl, rmask) \
  __builtin_amdgcn_update_dpp(~0u, (v), (ctrl), (rmask), 0xf, false)
__device__ __forceinline__ uint32_t wave_min_u32_dpp(uint32_t v) {
  v = min(v, (uint32_t)CBMV_DPP_U(v, 0xB1, 0xf));
  v = min(v, (uint32_t)CBMV_DPP_U(v, 0x4E, 0xf));
  v = min(v, (uint32_t)CBMV_DPP_U(v, 0x141, 0xf));
  v = min(v, (uint32_t)CBMV_DPP_U(v, 0x140, 0xf));
  v = min(v, (uint32_t)CBMV_DPP_U(v, 0x142, 0xa));
  v = min(v, (uint32_t)CBMV_DPP_U(v, 0x143, 0xc));
  return __builtin_amdgcn_readlane(v, 63);
}

// Order-preserving key of a float64 (no NaN): unsigned comparison of keys
// orders the values (x >= 0: the bits with the sign bit set; x < 0: the
// complemented bits), and dkey_value inverts it.
__device__ __forceinline__ uint64_t dkey(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_value(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

// Raw buffer access: a descriptor (64-bit base and byte count, in SGPRs)
// plus a 32-bit per-lane offset and a 32-bit uniform offset (an SGPR); loads
// beyond the byte count return 0 and stores beyond it are dropped. Loads
// count on the vector-memory counter like global loads (a uniform address
// would otherwise become a scalar load, which shares its counter with LDS),
// and the uniform part of an address costs no vector instruction.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buffer_ldf(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                            uint32_t soff) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ double buffer_ldd(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                             uint32_t soff) {
  return __builtin_bit_cast(double,
                            __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ uint32_t buffer_ldu8(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ uint32_t buffer_ldu16(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                                 uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void buffer_stf(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                           uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff,
                                        0);
}

}  // namespace cbmv

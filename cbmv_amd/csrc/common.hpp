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
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    T o = __shfl_xor(v, m, 64);
    v = o < v ? o : v;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

}  // namespace cbmv

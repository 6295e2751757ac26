// Library runtime: error reporting, the library stream, device buffers.
#include <mutex>

#include "common.hpp"

namespace cbmv {

namespace {
thread_local std::string g_error;
std::once_flag g_stream_once;
hipStream_t g_stream = nullptr;
}  // namespace

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

void clear_error() { g_error.clear(); }

hipStream_t lib_stream() {
  std::call_once(g_stream_once, [] {
    // Default priority on purpose: streams of different priorities are never mixed.
    if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) g_stream = nullptr;
  });
  return g_stream;
}

int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return cus;
}

int DevBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    set_error("hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
    return CBMV_ENOMEM;
  }
  bytes = n;
  return CBMV_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

int to_device(DevBuf& b, const void* host, size_t bytes) {
  int rc = b.alloc(bytes);
  if (rc) return rc;
  if (bytes) CBMV_HIP(hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, lib_stream()));
  return CBMV_OK;
}

int to_host(void* host, const DevBuf& b, size_t bytes) {
  if (bytes) CBMV_HIP(hipMemcpyAsync(host, b.p, bytes, hipMemcpyDeviceToHost, lib_stream()));
  CBMV_HIP(hipStreamSynchronize(lib_stream()));
  return CBMV_OK;
}

int sync() {
  CBMV_HIP(hipStreamSynchronize(lib_stream()));
  return CBMV_OK;
}

}  // namespace cbmv

extern "C" {

const char* cbmv_last_error(void) { return cbmv::g_error.c_str(); }

int cbmv_abi_version(void) { return 1; }

void* cbmv_device_alloc(int64_t bytes) {
  cbmv::clear_error();
  void* p = nullptr;
  if (bytes < 0 || hipMalloc(&p, bytes ? (size_t)bytes : 1) != hipSuccess) {
    cbmv::set_error("hipMalloc(%lld) failed", (long long)bytes);
    return nullptr;
  }
  return p;
}

int cbmv_device_free(void* p) {
  cbmv::clear_error();
  if (p) CBMV_HIP(hipFree(p));
  return CBMV_OK;
}

int cbmv_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
  cbmv::clear_error();
  if (bytes > 0) CBMV_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return CBMV_OK;
}

int cbmv_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
  cbmv::clear_error();
  CBMV_HIP(hipDeviceSynchronize());
  if (bytes > 0) CBMV_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return CBMV_OK;
}

int cbmv_device_synchronize(void) {
  cbmv::clear_error();
  CBMV_HIP(hipDeviceSynchronize());
  return CBMV_OK;
}

void* cbmv_event_create(void) {
  cbmv::clear_error();
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    cbmv::set_error("hipEventCreate failed");
    return nullptr;
  }
  return (void*)e;
}

int cbmv_event_record(void* ev, void* stream) {
  cbmv::clear_error();
  CBMV_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return CBMV_OK;
}

int cbmv_event_elapsed_ms(void* start, void* stop, float* ms) {
  cbmv::clear_error();
  CBMV_HIP(hipEventSynchronize((hipEvent_t)stop));
  CBMV_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return CBMV_OK;
}

int cbmv_event_destroy(void* ev) {
  cbmv::clear_error();
  if (ev) CBMV_HIP(hipEventDestroy((hipEvent_t)ev));
  return CBMV_OK;
}

int cbmv_initthreads(void) {
  int n = cbmv::device_cus();
  if (n <= 0) {
    cbmv::set_error("no HIP device");
    return CBMV_EHIP;
  }
  return n;
}

int cbmv_init_device(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    cbmv::set_error("There is no device.");
    return 0;
  }
  if (hipSetDevice(0) != hipSuccess) {
    cbmv::set_error("hipSetDevice(0) failed");
    return 0;
  }
  if (!cbmv::lib_stream()) {
    cbmv::set_error("could not create the library stream");
    return 0;
  }
  return 1;
}

}  // extern "C"

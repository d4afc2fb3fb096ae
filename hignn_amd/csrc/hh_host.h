// Host-side helpers shared by the C-ABI sources of libhh: error status,
// HIP call checks, device selection and growing device / pinned buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/hh.h"

namespace hh {

extern thread_local std::string g_last_error;
extern int g_device;
extern int g_cus;

struct HHError : std::runtime_error {
  int code;
  HHError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      throw ::hh::HHError(HH_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Runs f and maps exceptions to a status code and the thread's last error.
template <class F>
int guarded(F &&f) {
  try {
    f();
    return HH_OK;
  } catch (const HHError &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return HH_ERR_ALLOC;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return HH_ERR_ARG;
  }
}

// Selects the device once (LOCAL_RANK modulo the device count).
void ensure_device();

// Device buffer that only grows.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void reserve(size_t b) {
    if (b <= bytes) return;
    if (p) HIPCHK(hipFree(p));
    p = nullptr;
    bytes = 0;
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw HHError(HH_ERR_ALLOC, "hipMalloc(" + std::to_string(b) + "): " + hipGetErrorString(e));
    }
    bytes = b;
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Pinned host staging buffer that only grows.
struct PBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void reserve(size_t b) {
    if (b <= bytes) return;
    if (p) HIPCHK(hipHostFree(p));
    p = nullptr;
    bytes = 0;
    HIPCHK(hipHostMalloc(&p, b, hipHostMallocDefault));
    bytes = b;
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

}  // namespace hh

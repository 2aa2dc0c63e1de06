// hip_util.h -- small HIP helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "bude_internal.h"

namespace bude {

inline int hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? BUDE_ERR_OUT_OF_MEMORY : BUDE_ERR_RUNTIME,
              std::string(what) + ": " + hipGetErrorName(e) + ": " + hipGetErrorString(e));
}

inline int hip_check(hipError_t e, const char *what) { return e == hipSuccess ? BUDE_OK : hip_fail(e, what); }

// Owning device allocation (hipMalloc / hipFree).
struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;

  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  ~DeviceBuffer() { reset(); }

  void reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  int alloc(size_t n) {
    reset();
    if (n == 0) n = 4;
    if (int st = hip_check(hipMalloc(&ptr, n), "hipMalloc")) {
      ptr = nullptr;
      return st;
    }
    bytes = n;
    return BUDE_OK;
  }
  // Keeps the current allocation when it already holds n bytes; a growing
  // buffer gets a quarter of headroom, so sizes that vary a little from call
  // to call settle on one allocation.
  int reserve(size_t n) {
    if (ptr && bytes >= n) return BUDE_OK;
    return alloc(ptr ? n + n / 4 : n);
  }
  // Copies n bytes in, reusing the allocation when it is large enough.
  int assign(const void *src, size_t n) {
    if (int st = reserve(n)) return st;
    return n ? hip_check(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice), "hipMemcpy H2D") : BUDE_OK;
  }
  int upload(const void *src, size_t n) {
    if (int st = alloc(n)) return st;
    return n ? hip_check(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice), "hipMemcpy H2D") : BUDE_OK;
  }
};

// Owning pinned host allocation (hipHostMalloc / hipHostFree): the source of
// copies that run asynchronously to the host.
struct PinnedBuffer {
  void *ptr = nullptr;
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer &) = delete;
  PinnedBuffer &operator=(const PinnedBuffer &) = delete;
  ~PinnedBuffer() {
    if (ptr) (void)hipHostFree(ptr);
  }
  int alloc(size_t n) {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    if (int st = hip_check(hipHostMalloc(&ptr, n ? n : 4, hipHostMallocDefault), "hipHostMalloc")) {
      ptr = nullptr;
      return st;
    }
    return BUDE_OK;
  }
};

}  // namespace bude

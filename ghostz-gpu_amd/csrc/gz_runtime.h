// gz_runtime.h -- the opaque stream handle of the C ABI, and the library's
// device memory calls.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

struct gz_stream_s {
  hipStream_t s;
};

namespace gz {

// Stream capture and memory management exclude each other.  Freeing device
// or pinned memory synchronizes the device, which a capture in flight on
// another thread's stream forbids (the free fails with "operation not
// permitted when stream is capturing"); the library's captures and every
// allocation or free it makes mid-run hold this lock.
std::recursive_mutex &CaptureLock();

template <typename T>
inline hipError_t Malloc(T **p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> g(CaptureLock());
  return hipMalloc((void **)p, bytes);
}

template <typename T>
inline hipError_t HostMalloc(T **p, size_t bytes, unsigned flags = hipHostMallocDefault) {
  std::lock_guard<std::recursive_mutex> g(CaptureLock());
  return hipHostMalloc((void **)p, bytes, flags);
}

inline hipError_t Free(void *p) {
  std::lock_guard<std::recursive_mutex> g(CaptureLock());
  return hipFree(p);
}

inline hipError_t HostFree(void *p) {
  std::lock_guard<std::recursive_mutex> g(CaptureLock());
  return hipHostFree(p);
}

}  // namespace gz

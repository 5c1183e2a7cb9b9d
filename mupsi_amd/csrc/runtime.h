// runtime.h — host-side objects shared by the translation units of the library:
// the context (one HIP stream on one device), error reporting, and the owner of
// device allocations.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mupsi.h"
#include "scene_host.h"

struct mupsi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string name;
  int cus = 0;
  int march_blocks = 0;         // persistent grid of k_march
  int march_cached_blocks = 0;  // persistent grid of k_march_window
  int query_blocks = 0;         // resident blocks of the LDS-staging mesh-query kernels
  std::vector<mupsi_scene*> scenes;  // live scenes, freed by mupsi_close
  std::vector<mupsi_guide*> guides;  // live guides, freed by mupsi_close
  std::vector<mupsi_batch*> batches; // live batches, freed by mupsi_close
  // device time of the last mesh-query kernel (mupsi_last_kernel_ms)
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  float last_kernel_ms = 0.0f;
  // device address of the library's no-allocation scratch array on this device
  void* scratch = nullptr;
};

namespace mupsi {

inline mupsi_status fail(mupsi_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return ::mupsi::fail(e_ == hipErrorOutOfMemory ? MUPSI_ERR_NOMEM : MUPSI_ERR_HIP,      \
                           std::string(#expr) + ": " + hipGetErrorString(e_));               \
  } while (0)

inline unsigned grid_for(int64_t n, int block = 256) {
  return (unsigned)std::max<int64_t>(1, (n + block - 1) / block);
}

// Owns device allocations; frees them on destruction.
struct DeviceBuffers {
  std::vector<void*> ptrs;
  DeviceBuffers() = default;
  DeviceBuffers(const DeviceBuffers&) = delete;
  DeviceBuffers& operator=(const DeviceBuffers&) = delete;
  ~DeviceBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t upload(const T* src, size_t count, const T** dst) {
    *dst = nullptr;
    if (count == 0 || !src) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) return e;
    ptrs.push_back(p);
    e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
    *dst = static_cast<const T*>(p);
    return e;
  }
  template <typename T>
  hipError_t alloc(size_t count, T** dst) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T));
    if (e != hipSuccess) return e;
    ptrs.push_back(p);
    *dst = static_cast<T*>(p);
    return hipSuccess;
  }
};

}  // namespace mupsi

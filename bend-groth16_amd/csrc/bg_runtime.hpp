// Host-side runtime helpers shared by the library's translation units: error reporting
// across the C ABI, the per-process device context (one stream, grow-only scratch
// buffers) and small RAII wrappers.  No exception crosses the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>
#include <vector>
#include <mutex>
#include <functional>
#include "../../include/bg16.h"

namespace bg {

void set_error(const char* fmt, ...);

struct Error {
  int code;
};

#define BG_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      ::bg::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));   \
      throw ::bg::Error{BG_ERR_DEVICE};                                                    \
    }                                                                                      \
  } while (0)

#define BG_REQUIRE(cond, code, ...)                                                        \
  do {                                                                                     \
    if (!(cond)) {                                                                         \
      ::bg::set_error(__VA_ARGS__);                                                        \
      throw ::bg::Error{code};                                                             \
    }                                                                                      \
  } while (0)

// Runs f, converting thrown bg::Error / std::exception into an int status.
template <class Fn>
int guard(Fn&& f) {
  try {
    f();
    return BG_OK;
  } catch (const Error& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return BG_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error("%s", e.what());
    return BG_ERR_INTERNAL;
  }
}

// Device buffer that grows but never shrinks (reused across calls).
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  void* get(size_t bytes) {
    if (bytes > cap) {
      if (ptr) BG_HIP(hipFree(ptr));
      ptr = nullptr;
      cap = 0;
      size_t want = bytes + bytes / 8;
      hipError_t e = hipMalloc(&ptr, want);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        throw Error{BG_ERR_NOMEM};
      }
      cap = want;
    }
    return ptr;
  }
  template <class T>
  T* as(size_t count) {
    return static_cast<T*>(get(count * sizeof(T)));
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// Owned device allocation (pk arrays etc.)
struct DevArray {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  void alloc(size_t b) {
    release();
    if (b == 0) return;
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ptr = nullptr;
      set_error("hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
      throw Error{BG_ERR_NOMEM};
    }
    bytes = b;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  ~DevArray() { release(); }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// Pinned host staging buffer (grow-only).  flags: hipHostMallocCoherent for memory the
// device writes while the host polls it.
struct HostBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  void* get(size_t bytes, unsigned flags = hipHostMallocDefault) {
    if (bytes > cap) {
      if (ptr) (void)hipHostFree(ptr);
      ptr = nullptr;
      cap = 0;
      hipError_t e = hipHostMalloc(&ptr, bytes, flags);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipHostMalloc(%zu) failed", bytes);
        throw Error{BG_ERR_NOMEM};
      }
      cap = bytes;
    }
    return ptr;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// Persistent host worker pool for the small host-side reductions that finish each MSM
// (thread start-up would otherwise dominate them).  run(n, f) calls f(0..n-1) on the
// workers and the caller and returns when all calls are done; calls are serialised.
void host_parallel(int n, const std::function<void(int)>& f);

// Per-process device context.  The library drives one device; every entry point runs
// on the context stream and synchronises before returning host-visible results.
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t aux[3] = {nullptr, nullptr, nullptr};
  int cu_count = 0;
  std::mutex mu;
  // scratch used by the MSM / NTT drivers
  DevBuf scratch[16];
  HostBuf host_stage[4];
  bool timing = false;
};

Ctx& ctx();                 // initialises lazily on first use (device from bg_init or 0)
void ensure_init();

// True while stream s is being captured into a HIP graph.
bool stream_capturing(hipStream_t s);

// Records event e on stream s unless s is being captured into a graph (then nothing is
// recorded and false is returned): recorded launch sequences carry no host-visible events
// (a 2^10 proof whose graph recorded its MSM results through hipEventRecordExternal nodes
// measured 1.72 ms against 1.03 ms eager); results leave a graph through a flag instead.
bool record_if_eager(hipEvent_t e, hipStream_t s);

// Profiling events around the dominant kernels (enabled by bg_set_timing).
struct KTimer {
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
};
void timing_record(const char* name, hipStream_t s, hipEvent_t start, hipEvent_t stop);

// Records HIP events around the kernels launched on stream s inside its scope when
// timing is enabled (bg_set_timing); the events are read back lazily.
struct ScopedKTime {
  const char* name;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  ScopedKTime(const char* n, hipStream_t st) : name(n), s(st) {
    if (ctx().timing) {
      BG_HIP(hipEventCreate(&a));
      BG_HIP(hipEventCreate(&b));
      BG_HIP(hipEventRecord(a, s));
    }
  }
  ~ScopedKTime() {
    if (a) {
      (void)hipEventRecord(b, s);
      timing_record(name, s, a, b);
    }
  }
};

}  // namespace bg

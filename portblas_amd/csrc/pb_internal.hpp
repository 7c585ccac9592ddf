// Host- and device-side internals shared by every translation unit of
// libportblas_amd.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/portblas_amd.h"

namespace pb {

// ---------------------------------------------------------------------------
// Error reporting: every C entry point returns a status and leaves a
// per-thread message behind.
// ---------------------------------------------------------------------------
pb_status_t set_error(pb_status_t status, const char* fmt, ...);
void clear_error();

#define PB_CHECK_HIP(expr)                                                      \
  do {                                                                          \
    hipError_t pb_e_ = (expr);                                                  \
    if (pb_e_ != hipSuccess)                                                    \
      return ::pb::set_error(PB_STATUS_DEVICE_ERROR, "%s failed: %s (%s:%d)",   \
                             #expr, hipGetErrorString(pb_e_), __FILE__,         \
                             __LINE__);                                         \
  } while (0)

#define PB_CHECK_LAUNCH()                                                       \
  do {                                                                          \
    hipError_t pb_e_ = hipGetLastError();                                       \
    if (pb_e_ != hipSuccess)                                                    \
      return ::pb::set_error(PB_STATUS_DEVICE_ERROR, "kernel launch failed: %s " \
                             "(%s:%d)", hipGetErrorString(pb_e_), __FILE__,     \
                             __LINE__);                                         \
  } while (0)

#define PB_TRY(expr)                                                            \
  do {                                                                          \
    pb_status_t pb_s_ = (expr);                                                 \
    if (pb_s_ != PB_STATUS_SUCCESS) return pb_s_;                               \
  } while (0)

#define PB_REQUIRE(cond, ...)                                                   \
  do {                                                                          \
    if (!(cond)) return ::pb::set_error(PB_STATUS_INVALID_VALUE, __VA_ARGS__);  \
  } while (0)

}  // namespace pb

// ---------------------------------------------------------------------------
// The handle: one stream, the device's properties, a growable workspace and a
// small always-zero counter arena used by single-pass kernels (each kernel
// returns the counters it used to zero before it exits, so no memset is ever
// recorded into a stream or graph).
// ---------------------------------------------------------------------------
struct pb_selector_state {
  int n_layers = 0;
  std::vector<int> sizes;
  std::vector<std::vector<float>> w, b;
  std::vector<float> mean, stdv;
  // Recent decisions, direct-mapped by the feature vector (a repeated shape
  // skips the network); installing a network replaces the whole state.
  struct Decision {
    bool valid = false;
    float f[8];
    int choice;
  };
  static constexpr int kCacheSlots = 64;
  mutable Decision cache[kCacheSlots];
};

// One workspace allocation.  The handle holds the current one; every graph
// captured on the handle holds the one its kernels were recorded with, so a
// later growth of the handle's workspace never frees memory that a live graph
// still reads and writes (the buffer is freed when its last holder lets go).
struct pb_device_buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  pb_device_buffer(void* p, size_t n) : ptr(p), bytes(n) {}
  ~pb_device_buffer();
  pb_device_buffer(const pb_device_buffer&) = delete;
  pb_device_buffer& operator=(const pb_device_buffer&) = delete;
};

struct pb_handle_s {
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cus = 0;
  int max_threads = 1024;
  void* workspace = nullptr;         // == ws->ptr (kept for the kernels' plain access)
  size_t workspace_bytes = 0;
  std::shared_ptr<pb_device_buffer> ws;
  unsigned int* counters = nullptr;  // kCounterSlots zero-initialised slots
  void* scalar_slot = nullptr;       // 256 bytes for the _sync BLAS-1 results
  int* scale_guess = nullptr;        // scale exponents of the last FP16x3 call (A, B) and FP8 call (A, B), device side
  hipEvent_t switch_event = nullptr; // orders a stream switch after the old stream's work
  size_t counter_cursor = 0;
  bool capturing = false;
  // Guards the workspace pool (growth and hand-out), as the reference's
  // Temp_Mem_Pool guards its pool (include/sb_handle/temp_memory_pool.h:98-108).
  // Calls that overlap on one handle are still unsupported: they share the
  // stream, the workspace contents and the counter arena.
  std::mutex mu;
  pb_selector_state selector;        // BF16-input engine (40 configurations)
  pb_selector_state plain_selector;  // FP32 / FP64 plain engine (kPlainConfigs)
};

namespace pb {

constexpr int kCounterSlots = 4096;

// Validates the handle and makes its device current.
pb_status_t enter(pb_handle_t h);
// Returns a workspace of at least `bytes`, growing it if allowed.
pb_status_t workspace(pb_handle_t h, size_t bytes, void** ptr);
// Hands out `n` zeroed counter slots (round-robin over the arena).
unsigned int* counters(pb_handle_t h, int n);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Lower-cased transpose flag; 'c' counts as 't' for real types
// (src/interface/gemm_interface.hpp:141-151).  Returns 0 if invalid.
inline char norm_trans(char t) {
  char l = (t >= 'A' && t <= 'Z') ? char(t - 'A' + 'a') : t;
  if (l == 'n') return 'n';
  if (l == 't' || l == 'c') return 't';
  return 0;
}

}  // namespace pb

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
#if defined(__HIPCC__)
namespace pbdev {

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Element i of a BLAS vector view (negative increments walk backwards from
// the end, src/views/view_sycl.hpp:95-100).
__device__ __forceinline__ int64_t vidx(int64_t i, int64_t n, int64_t inc) {
  return inc >= 0 ? i * inc : (n - 1 - i) * (-inc);
}

typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));

// Streaming (non-temporal) load: data read once does not allocate in the
// caches, so it neither evicts nor forces write-back of other lines.  On
// MI355X this is worth ~2x of read bandwidth when the caches hold dirty data
// from earlier kernels.
template <typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
  return __builtin_nontemporal_load(p);
}

// Pointers that a kernel reads out of a descriptor in memory are generic to
// the compiler, so their accesses become flat instructions.  A flat access
// counts against both the vector-memory and the LDS/scalar wait counters, and
// while one is outstanding every wait the compiler inserts drains ALL pending
// memory operations: loads issued several steps ahead are then waited for at
// once.  Such pointers are cast to the global address space before use.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}

}  // namespace pbdev
#endif

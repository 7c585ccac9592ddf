// BLAS-1 reductions: dot, asum, nrm2, iamax, iamin.
//
// Reference behaviour (portBLAS):
//   _dot   src/interface/blas1_interface.hpp:611-642 (n == 0 -> rs untouched)
//   _asum  src/interface/blas1_interface.hpp:222-246
//   _nrm2  src/interface/blas1_interface.hpp:552-581 (sum of squares, then a
//          separate sqrt of the accumulated value; no scaling)
//   kernel src/operations/blas1/WGAtomicReduction.hpp:61-138 — the result is
//          ADDED to *rs with device atomics.
//   _iamax/_iamin src/interface/blas1_interface.hpp:278-405,
//          src/operations/blas1/IndexMaxMin.hpp:74-171,
//          operators src/operations/blas_operators.hpp:276-316.
//
// MI355X design: each 256-thread workgroup streams its share of the vector
// with 16-byte non-temporal loads (4 in flight per thread), reduces in
// registers, across the wave with shuffles and across the 4 waves in LDS, and
// writes one partial.  A one-workgroup fold kernel then combines the partials
// in a fixed order, so the result is deterministic for a given n (portBLAS's
// atomic order is not), and applies the same update to *rs that portBLAS's
// atomics + epilogue produce.
#include "pb_internal.hpp"

#include <cfloat>
#include <climits>
#include <limits>

namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;

enum ReduceOp { kDot = 0, kAsum = 1, kNrm2 = 2 };

template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = pbdev::f32x4_t; static constexpr int n = 4; };
template <> struct Vec16<double> { using type = pbdev::f64x2_t; static constexpr int n = 2; };

template <typename T, int OP>
__device__ __forceinline__ T term(T a, T b) {
  if constexpr (OP == kDot) return a * b;
  else if constexpr (OP == kAsum) return fabs(a);
  else return a * a;
}

template <typename T, int OP>
__device__ __forceinline__ T vterm_sum(const typename Vec16<T>::type& a,
                                       const typename Vec16<T>::type& b, T acc) {
  if constexpr (Vec16<T>::n == 4) {
    acc += term<T, OP>(a.x, b.x);
    acc += term<T, OP>(a.y, b.y);
    acc += term<T, OP>(a.z, b.z);
    acc += term<T, OP>(a.w, b.w);
  } else {
    acc += term<T, OP>(a.x, b.x);
    acc += term<T, OP>(a.y, b.y);
  }
  return acc;
}

// Block-wide sum; result valid in thread 0.
template <typename T>
__device__ __forceinline__ T block_sum(T v) {
  __shared__ T red[kBlock / 64];
  v = pbdev::wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  T r = T(0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) r += red[w];
  }
  return r;
}

// VEC: both operands contiguous (inc == 1) and 16-byte aligned.
template <typename T, int OP, bool VEC>
__global__ __launch_bounds__(kBlock) void reduce_kernel(
    int64_t n, const T* __restrict__ x, int64_t incx, const T* __restrict__ y,
    int64_t incy, T* __restrict__ partials) {
  T acc = T(0);
  if constexpr (VEC) {
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    const int64_t nvec = n / VN;
    const V* xv = reinterpret_cast<const V*>(x);
    const V* yv = reinterpret_cast<const V*>(OP == kDot ? y : x);
    T acc_u[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc_u[u] = T(0);
    const int64_t step = int64_t(gridDim.x) * kBlock * kUnroll;
    int64_t base = int64_t(blockIdx.x) * kBlock * kUnroll + threadIdx.x;
    // Full iterations: no bounds checks inside.
    for (; base + (kUnroll - 1) * kBlock < nvec; base += step) {
      V a[kUnroll], b[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) a[u] = pbdev::ld_stream(xv + base + u * kBlock);
      if constexpr (OP == kDot) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) b[u] = pbdev::ld_stream(yv + base + u * kBlock);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
        acc_u[u] = vterm_sum<T, OP>(a[u], OP == kDot ? b[u] : a[u], acc_u[u]);
    }
    // Ragged last iteration.
    if (base < nvec) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t i = base + u * kBlock;
        if (i < nvec) {
          V a = pbdev::ld_stream(xv + i);
          V b = (OP == kDot) ? pbdev::ld_stream(yv + i) : a;
          acc_u[u] = vterm_sum<T, OP>(a, b, acc_u[u]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc += acc_u[u];
    // Scalar tail (n not a multiple of the vector width).
    if (blockIdx.x == 0 && threadIdx.x < n - nvec * VN) {
      const int64_t i = nvec * VN + threadIdx.x;
      acc += term<T, OP>(x[i], OP == kDot ? y[i] : x[i]);
    }
  } else {
    const int64_t step = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += step) {
      const T a = pbdev::ld_stream(x + pbdev::vidx(i, n, incx));
      const T b = (OP == kDot) ? pbdev::ld_stream(y + pbdev::vidx(i, n, incy)) : a;
      acc += term<T, OP>(a, b);
    }
  }
  T bsum = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = bsum;
}

// Fixed-order fold of the per-workgroup partials (a second, one-workgroup
// launch: stream order makes the partials visible without device-scope
// fences, which on MI355X would write back the L2 once per workgroup) and the
// update of *rs that portBLAS's atomics + sqrt epilogue produce.
// add_sb: sdsdot's final *rs <- sb + *rs (blas1_interface.hpp:148-177).
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void reduce_fold_kernel(const T* __restrict__ partials,
                                                             int nparts, T* rs, T sb, int add_sb) {
  // All of a thread's loads are issued before the sums (partials come from
  // other XCDs' L2 write-back: a dependent chain would pay the latency per
  // element).
  constexpr int U = 8;
  T s = T(0);
  for (int base = threadIdx.x; base < nparts; base += kBlock * U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kBlock;
      v[u] = i < nparts ? partials[i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s += v[u];
  }
  const T total = block_sum(s);
  if (threadIdx.x == 0) {
    if constexpr (OP == kNrm2) {
      rs[0] = sqrt(rs[0] + total);
    } else {
      const T acc = rs[0] + total;
      rs[0] = add_sb ? sb + acc : acc;
    }
  }
}

template <typename T>
__global__ void sqrt_inplace_kernel(T* rs) { rs[0] = sqrt(rs[0]); }

// sdsdot with an empty or negative size: *rs <- sb (n == 0, the reference
// copies sb over the result) or *rs <- sb + *rs (n < 0: the dot adds nothing).
template <typename T>
__global__ void sdsdot_degenerate_kernel(T* rs, T sb, int overwrite) {
  rs[0] = overwrite ? sb : sb + rs[0];
}

// ---------------------------------------------------------------------------
// iamax / iamin.  A candidate (v, i) beats (bv, bi) iff v is strictly better,
// or equal with a lower index: exactly IMaxOperator / IMinOperator of
// blas_operators.hpp:276-316 (comparisons with NaN are false, so NaN never
// wins).  The running best starts from portBLAS's initial tuple
// (blas_constants.h:110-136): (sentinel, 0) for max and (sentinel, T_MAX) for
// min; the sentinel index is larger than any real index.
// ---------------------------------------------------------------------------
constexpr int64_t kSentinel = std::numeric_limits<int64_t>::max();

template <bool IS_MAX, typename T>
__device__ __forceinline__ bool better(T v, int64_t i, T bv, int64_t bi) {
  if constexpr (IS_MAX) return (bv < v) || (bv == v && bi > i);
  else return (bv > v) || (bv == v && bi > i);
}

template <bool IS_MAX, typename T>
__device__ __forceinline__ T init_val() {
  return IS_MAX ? T(0) : std::numeric_limits<T>::max();
}

template <bool IS_MAX, typename T>
__device__ __forceinline__ void block_best(T& v, int64_t& i) {
  __shared__ T sv[kBlock / 64];
  __shared__ int64_t si[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T ov = __shfl_xor(v, off, 64);
    int64_t oi = __shfl_xor(i, off, 64);
    if (better<IS_MAX>(ov, oi, v, i)) { v = ov; i = oi; }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { sv[wave] = v; si[wave] = i; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w)
      if (better<IS_MAX>(sv[w], si[w], v, i)) { v = sv[w]; i = si[w]; }
  }
}

// Per-workgroup best (value, index).  A thread visits its elements in
// increasing index order, so inside one thread "strictly better" already
// implements the lowest-index tie rule, and a running value below every
// candidate (-1 for max; +inf for min, which like the reference's FLT_MAX
// initial value lets FLT_MAX win and never inf) stands for the reference's
// initial tuple.  IDX is the in-thread index type: 32-bit whenever n fits.
// Threads and workgroups are then combined with the full better() rule.
template <bool IS_MAX, typename T>
__device__ __forceinline__ bool thread_better(T v, T bv) {
  return IS_MAX ? (v > bv) : (v < bv);
}

template <bool IS_MAX, typename T, typename IDX, bool VEC>
__global__ __launch_bounds__(kBlock) void iamax_kernel(
    int64_t n, const T* __restrict__ x, int64_t incx, T* __restrict__ pv,
    int64_t* __restrict__ pi) {
  T bv = IS_MAX ? T(-1) : std::numeric_limits<T>::infinity();
  IDX bi = IDX(-1);
  if constexpr (VEC) {
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    const int64_t nvec = n / VN;
    const V* xv = reinterpret_cast<const V*>(x);
    const int64_t step = int64_t(gridDim.x) * kBlock * kUnroll;
    int64_t base = int64_t(blockIdx.x) * kBlock * kUnroll + threadIdx.x;
    // Full iterations: no bounds checks inside.  One running best per load
    // slot u (independent compare chains); each chain still visits its
    // elements in increasing index order, and the chains are merged with the
    // full better() rule below.
    T cv[kUnroll];
    IDX ci[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) { cv[u] = bv; ci[u] = bi; }
    for (; base + (kUnroll - 1) * kBlock < nvec; base += step) {
      V a[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) a[u] = pbdev::ld_stream(xv + base + u * kBlock);
      const IDX b0 = IDX(base * VN);
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
#pragma unroll
        for (int j = 0; j < VN; ++j) {
          const T v = fabs(a[u][j]);
          if (thread_better<IS_MAX>(v, cv[u])) { cv[u] = v; ci[u] = b0 + IDX(u * kBlock * VN + j); }
        }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (ci[u] != IDX(-1) &&
          (bi == IDX(-1) || better<IS_MAX>(cv[u], int64_t(ci[u]), bv, int64_t(bi)))) {
        bv = cv[u];
        bi = ci[u];
      }
    if (base < nvec) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t iv = base + u * kBlock;
        if (iv < nvec) {
          const V a = pbdev::ld_stream(xv + iv);
#pragma unroll
          for (int j = 0; j < VN; ++j) {
            const T v = fabs(a[j]);
            if (thread_better<IS_MAX>(v, bv)) { bv = v; bi = IDX(iv * VN + j); }
          }
        }
      }
    }
    // Scalar tail: indices above every vector index, visited last.
    if (blockIdx.x == 0 && threadIdx.x < n - nvec * VN) {
      const int64_t idx = nvec * VN + threadIdx.x;
      const T v = fabs(x[idx]);
      if (thread_better<IS_MAX>(v, bv)) { bv = v; bi = IDX(idx); }
    }
  } else {
    const int64_t step = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += step) {
      const T v = fabs(pbdev::ld_stream(x + i * incx));
      if (thread_better<IS_MAX>(v, bv)) { bv = v; bi = IDX(i); }
    }
  }
  T v = init_val<IS_MAX, T>();
  int64_t i = kSentinel;
  if (bi != IDX(-1)) { v = bv; i = int64_t(bi); }
  block_best<IS_MAX>(v, i);
  if (threadIdx.x == 0) { pv[blockIdx.x] = v; pi[blockIdx.x] = i; }
}

template <bool IS_MAX, typename T, typename IndexT>
__global__ __launch_bounds__(kBlock) void iamax_fold_kernel(const T* __restrict__ pv,
                                                            const int64_t* __restrict__ pi,
                                                            int nparts, IndexT* rs) {
  T bv = init_val<IS_MAX, T>();
  int64_t bi = kSentinel;
  constexpr int U = 8;
  for (int base = threadIdx.x; base < nparts; base += kBlock * U) {
    T v[U];
    int64_t ix[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int w = base + u * kBlock;
      v[u] = w < nparts ? pv[w] : init_val<IS_MAX, T>();
      ix[u] = w < nparts ? pi[w] : kSentinel;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (better<IS_MAX>(v[u], ix[u], bv, bi)) { bv = v[u]; bi = ix[u]; }
  }
  block_best<IS_MAX>(bv, bi);
  if (threadIdx.x == 0)
    rs[0] = (bi == kSentinel) ? std::numeric_limits<IndexT>::max() : IndexT(bi);
}

template <typename IndexT>
__global__ void fill_index_kernel(IndexT* rs, IndexT v) { rs[0] = v; }

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

int grid_for(pb_handle_t h, int64_t n, int elems_per_thread_iter) {
  const int64_t per_block = int64_t(kBlock) * elems_per_thread_iter;
  int64_t g = pb::ceil_div(n, per_block);
  const int64_t cap = int64_t(h->num_cus) * 8;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return int(g);
}

template <typename T, int OP>
pb_status_t reduce_launch(pb_handle_t h, int64_t n, const T* x, int64_t incx,
                          const T* y, int64_t incy, T* rs, const T* sb = nullptr) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(rs != nullptr, "null result pointer");
  if (n <= 0) {
    // dot: rs untouched (blas1_interface.hpp:621); asum adds nothing;
    // nrm2 still takes the square root of *rs (.hpp:577-580); sdsdot
    // writes sb (n == 0) or adds it (.hpp:153-166).
    if (sb) {
      hipLaunchKernelGGL(sdsdot_degenerate_kernel<T>, dim3(1), dim3(1), 0, h->stream, rs, *sb,
                         int(n == 0));
      PB_CHECK_LAUNCH();
      return PB_STATUS_SUCCESS;
    }
    if constexpr (OP == kNrm2) {
      hipLaunchKernelGGL(sqrt_inplace_kernel<T>, dim3(1), dim3(1), 0, h->stream, rs);
      PB_CHECK_LAUNCH();
    }
    return PB_STATUS_SUCCESS;
  }
  PB_REQUIRE(x != nullptr && (OP != kDot || y != nullptr), "null vector pointer");
  const bool vec = incx == 1 && aligned16(x) &&
                   (OP != kDot || (incy == 1 && aligned16(y)));
  constexpr int VN = Vec16<T>::n;
  const int grid = vec ? grid_for(h, n, VN * kUnroll) : grid_for(h, n, 4);
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, sizeof(T) * grid, &ws));
  T* partials = static_cast<T*>(ws);
  if (vec)
    hipLaunchKernelGGL((reduce_kernel<T, OP, true>), dim3(grid), dim3(kBlock), 0,
                       h->stream, n, x, incx, y, incy, partials);
  else
    hipLaunchKernelGGL((reduce_kernel<T, OP, false>), dim3(grid), dim3(kBlock), 0,
                       h->stream, n, x, incx, y, incy, partials);
  hipLaunchKernelGGL((reduce_fold_kernel<T, OP>), dim3(1), dim3(kBlock), 0, h->stream, partials,
                     grid, rs, sb ? *sb : T(0), int(sb != nullptr));
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

template <bool IS_MAX, typename T, typename IndexT>
pb_status_t iamax_launch(pb_handle_t h, int64_t n, const T* x, int64_t incx,
                         IndexT* rs) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(rs != nullptr, "null result pointer");
  if (incx < 0 || n < 0) {
    // blas1_interface.hpp:375-379: negative increment or size -> index 0.
    hipLaunchKernelGGL(fill_index_kernel<IndexT>, dim3(1), dim3(1), 0, h->stream, rs,
                       IndexT(0));
    PB_CHECK_LAUNCH();
    return PB_STATUS_SUCCESS;
  }
  if (n == 0) {
    hipLaunchKernelGGL(fill_index_kernel<IndexT>, dim3(1), dim3(1), 0, h->stream, rs,
                       std::numeric_limits<IndexT>::max());
    PB_CHECK_LAUNCH();
    return PB_STATUS_SUCCESS;
  }
  PB_REQUIRE(x != nullptr, "null vector pointer");
  PB_REQUIRE(uint64_t(n) <= uint64_t(std::numeric_limits<IndexT>::max()),
             "n = %lld does not fit the %d-byte index type", (long long)n,
             int(sizeof(IndexT)));
  const bool vec = incx == 1 && aligned16(x);
  constexpr int VN = Vec16<T>::n;
  const int grid = vec ? grid_for(h, n, VN * kUnroll) : grid_for(h, n, 4);
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, (sizeof(T) + sizeof(int64_t)) * grid + 256, &ws));
  int64_t* pi = static_cast<int64_t*>(ws);
  T* pv = reinterpret_cast<T*>(pi + grid);
  const bool idx32 = n <= int64_t(std::numeric_limits<int32_t>::max());
#define PB_IAMAX(IDX, V)                                                                   \
  hipLaunchKernelGGL((iamax_kernel<IS_MAX, T, IDX, V>), dim3(grid), dim3(kBlock), 0, h->stream, \
                     n, x, incx, pv, pi)
  if (vec && idx32) PB_IAMAX(int32_t, true);
  else if (vec) PB_IAMAX(int64_t, true);
  else if (idx32) PB_IAMAX(int32_t, false);
  else PB_IAMAX(int64_t, false);
#undef PB_IAMAX
  hipLaunchKernelGGL((iamax_fold_kernel<IS_MAX, T, IndexT>), dim3(1), dim3(kBlock), 0, h->stream,
                     pv, pi, grid, rs);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

// Synchronous forms: start from the initial value, run the asynchronous
// kernel into the handle's device scalar slot, wait and copy back
// (blas1_interface.hpp:958-1178 does the same with a per-call allocation).
template <typename T, typename F>
pb_status_t sync_scalar(pb_handle_t h, T init, T* host_out, F&& launch) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(host_out != nullptr, "null host result pointer");
  T* d = static_cast<T*>(h->scalar_slot);
  PB_CHECK_HIP(hipMemcpyAsync(d, &init, sizeof(T), hipMemcpyHostToDevice, h->stream));
  PB_TRY(launch(d));
  PB_CHECK_HIP(hipMemcpyAsync(host_out, d, sizeof(T), hipMemcpyDeviceToHost, h->stream));
  PB_CHECK_HIP(hipStreamSynchronize(h->stream));
  return PB_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

pb_status_t pb_sdot(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                    const float* y, int64_t incy, float* rs) {
  return reduce_launch<float, kDot>(h, n, x, incx, y, incy, rs);
}
pb_status_t pb_ddot(pb_handle_t h, int64_t n, const double* x, int64_t incx,
                    const double* y, int64_t incy, double* rs) {
  return reduce_launch<double, kDot>(h, n, x, incx, y, incy, rs);
}
pb_status_t pb_sdsdot(pb_handle_t h, int64_t n, float sb, const float* x, int64_t incx,
                      const float* y, int64_t incy, float* rs) {
  return reduce_launch<float, kDot>(h, n, x, incx, y, incy, rs, &sb);
}
pb_status_t pb_sasum(pb_handle_t h, int64_t n, const float* x, int64_t incx, float* rs) {
  return reduce_launch<float, kAsum>(h, n, x, incx, nullptr, 0, rs);
}
pb_status_t pb_dasum(pb_handle_t h, int64_t n, const double* x, int64_t incx, double* rs) {
  return reduce_launch<double, kAsum>(h, n, x, incx, nullptr, 0, rs);
}
pb_status_t pb_snrm2(pb_handle_t h, int64_t n, const float* x, int64_t incx, float* rs) {
  return reduce_launch<float, kNrm2>(h, n, x, incx, nullptr, 0, rs);
}
pb_status_t pb_dnrm2(pb_handle_t h, int64_t n, const double* x, int64_t incx, double* rs) {
  return reduce_launch<double, kNrm2>(h, n, x, incx, nullptr, 0, rs);
}

pb_status_t pb_isamax(pb_handle_t h, int64_t n, const float* x, int64_t incx, int32_t* rs) {
  return iamax_launch<true, float, int32_t>(h, n, x, incx, rs);
}
pb_status_t pb_isamin(pb_handle_t h, int64_t n, const float* x, int64_t incx, int32_t* rs) {
  return iamax_launch<false, float, int32_t>(h, n, x, incx, rs);
}
pb_status_t pb_idamax(pb_handle_t h, int64_t n, const double* x, int64_t incx, int32_t* rs) {
  return iamax_launch<true, double, int32_t>(h, n, x, incx, rs);
}
pb_status_t pb_idamin(pb_handle_t h, int64_t n, const double* x, int64_t incx, int32_t* rs) {
  return iamax_launch<false, double, int32_t>(h, n, x, incx, rs);
}
pb_status_t pb_isamax_64(pb_handle_t h, int64_t n, const float* x, int64_t incx, int64_t* rs) {
  return iamax_launch<true, float, int64_t>(h, n, x, incx, rs);
}
pb_status_t pb_isamin_64(pb_handle_t h, int64_t n, const float* x, int64_t incx, int64_t* rs) {
  return iamax_launch<false, float, int64_t>(h, n, x, incx, rs);
}

pb_status_t pb_sdot_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                         const float* y, int64_t incy, float* out) {
  return sync_scalar<float>(h, 0.f, out, [&](float* d) {
    return reduce_launch<float, kDot>(h, n, x, incx, y, incy, d);
  });
}
pb_status_t pb_sdsdot_sync(pb_handle_t h, int64_t n, float sb, const float* x, int64_t incx,
                           const float* y, int64_t incy, float* out) {
  return sync_scalar<float>(h, 0.f, out, [&](float* d) {
    return reduce_launch<float, kDot>(h, n, x, incx, y, incy, d, &sb);
  });
}
pb_status_t pb_sasum_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                          float* out) {
  return sync_scalar<float>(h, 0.f, out, [&](float* d) {
    return reduce_launch<float, kAsum>(h, n, x, incx, nullptr, 0, d);
  });
}
pb_status_t pb_snrm2_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                          float* out) {
  return sync_scalar<float>(h, 0.f, out, [&](float* d) {
    return reduce_launch<float, kNrm2>(h, n, x, incx, nullptr, 0, d);
  });
}
pb_status_t pb_isamax_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                           int32_t* out) {
  return sync_scalar<int32_t>(h, -1, out, [&](int32_t* d) {
    return iamax_launch<true, float, int32_t>(h, n, x, incx, d);
  });
}
pb_status_t pb_isamin_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                           int32_t* out) {
  return sync_scalar<int32_t>(h, -1, out, [&](int32_t* d) {
    return iamax_launch<false, float, int32_t>(h, n, x, incx, d);
  });
}

}  // extern "C"

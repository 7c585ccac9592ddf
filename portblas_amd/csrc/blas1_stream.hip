// Streaming BLAS-1: axpy, scal, copy, swap.
//
// Reference behaviour (portBLAS, src/interface/blas1_interface.hpp):
//   _axpy  :46-60    y <- alpha * x + y      (ScalarOp product, then Add)
//   _copy  :62-74    y <- x
//   _swap  :419-430  x <-> y                 (DoubleAssign)
//   _scal  :440-456  x <- alpha * x; alpha == 0 assigns zero (so NaN and inf
//                    become 0, not NaN)
// Vectors are BLAS views: element i at x[i * inc], negative increments walk
// from the end (src/views/view_sycl.hpp:95-100).
//
// MI355X design: HBM-bound grid-stride loops, 256 threads per workgroup, at
// most 8 workgroups per CU.  With unit increments and 16-byte aligned
// vectors each lane moves 4 x 16 bytes per iteration; data read once uses
// non-temporal loads.  axpy is one fused multiply-add per element (one
// rounding), the evaluation of OpenBLAS (the reference's test oracle) and of
// a contracting device compiler.
#include "pb_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;

template <typename T> struct V16;
template <> struct V16<float> { using type = pbdev::f32x4_t; static constexpr int n = 4; };
template <> struct V16<double> { using type = pbdev::f64x2_t; static constexpr int n = 2; };

enum StreamOp { kAxpy = 0, kScal = 1, kCopy = 2, kSwap = 3 };

__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float fma_rn(float a, float b, float c) { return __fmaf_rn(a, b, c); }
__device__ __forceinline__ double fma_rn(double a, double b, double c) { return __fma_rn(a, b, c); }

template <typename T, int OP>
__device__ __forceinline__ void apply(T alpha, T& x, T& y) {
  if constexpr (OP == kAxpy) {
    y = fma_rn(alpha, x, y);
  } else if constexpr (OP == kScal) {
    x = (alpha == T(0)) ? T(0) : mul_rn(alpha, x);
  } else if constexpr (OP == kCopy) {
    y = x;
  } else {
    const T t = x;
    x = y;
    y = t;
  }
}

// VEC: unit increments, both vectors 16-byte aligned.
template <typename T, int OP, bool VEC>
__global__ __launch_bounds__(kBlock) void stream_kernel(int64_t n, T alpha, T* __restrict__ x,
                                                        int64_t incx, T* __restrict__ y, int64_t incy) {
  if constexpr (VEC) {
    using V = typename V16<T>::type;
    constexpr int VN = V16<T>::n;
    const int64_t nvec = n / VN;
    V* xv = reinterpret_cast<V*>(x);
    V* yv = reinterpret_cast<V*>(y);
    const int64_t step = int64_t(gridDim.x) * kBlock * kUnroll;
    for (int64_t base = int64_t(blockIdx.x) * kBlock * kUnroll + threadIdx.x; base < nvec; base += step) {
      V a[kUnroll], b[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t i = base + u * kBlock;
        if (i < nvec) {
          if constexpr (OP == kCopy) a[u] = pbdev::ld_stream(xv + i);
          else if constexpr (OP == kScal) a[u] = xv[i];
          else a[u] = pbdev::ld_stream(xv + i);
          if constexpr (OP == kAxpy || OP == kSwap) b[u] = yv[i];
        }
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int64_t i = base + u * kBlock;
        if (i >= nvec) continue;
#pragma unroll
        for (int e = 0; e < VN; ++e) {
          T xs = a[u][e];
          T ys = (OP == kAxpy || OP == kSwap) ? b[u][e] : T(0);
          apply<T, OP>(alpha, xs, ys);
          a[u][e] = xs;
          b[u][e] = ys;
        }
        if constexpr (OP == kScal || OP == kSwap) xv[i] = a[u];
        if constexpr (OP != kScal) yv[i] = b[u];
      }
    }
    // Scalar tail (n not a multiple of the vector width).
    if (blockIdx.x == 0 && threadIdx.x < n - nvec * VN) {
      const int64_t i = nvec * VN + threadIdx.x;
      T xs = x[i];
      T ys = (OP == kAxpy || OP == kSwap) ? y[i] : T(0);
      apply<T, OP>(alpha, xs, ys);
      if constexpr (OP == kScal || OP == kSwap) x[i] = xs;
      if constexpr (OP != kScal) y[i] = ys;
    }
  } else {
    const int64_t step = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += step) {
      const int64_t ix = pbdev::vidx(i, n, incx);
      const int64_t iy = (OP == kScal) ? 0 : pbdev::vidx(i, n, incy);
      T xs = x[ix];
      T ys = (OP == kAxpy || OP == kSwap) ? y[iy] : T(0);
      apply<T, OP>(alpha, xs, ys);
      if constexpr (OP == kScal || OP == kSwap) x[ix] = xs;
      if constexpr (OP != kScal) y[iy] = ys;
    }
  }
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <typename T, int OP>
pb_status_t stream_launch(pb_handle_t h, int64_t n, T alpha, T* x, int64_t incx, T* y, int64_t incy) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  if (n <= 0) return PB_STATUS_SUCCESS;  // an empty view: nothing to assign
  PB_REQUIRE(x != nullptr && (OP == kScal || y != nullptr), "null vector pointer");
  const bool vec = incx == 1 && aligned16(x) && (OP == kScal || (incy == 1 && aligned16(y)));
  constexpr int VN = V16<T>::n;
  const int64_t per_block = int64_t(kBlock) * (vec ? VN * kUnroll : 1);
  int64_t grid = pb::ceil_div(n, per_block);
  grid = std::max<int64_t>(1, std::min<int64_t>(grid, int64_t(h->num_cus) * 8));
  if (vec)
    hipLaunchKernelGGL((stream_kernel<T, OP, true>), dim3(unsigned(grid)), dim3(kBlock), 0, h->stream, n,
                       alpha, x, incx, y, incy);
  else
    hipLaunchKernelGGL((stream_kernel<T, OP, false>), dim3(unsigned(grid)), dim3(kBlock), 0, h->stream, n,
                       alpha, x, incx, y, incy);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

pb_status_t pb_saxpy(pb_handle_t h, int64_t n, float alpha, const float* x, int64_t incx, float* y,
                     int64_t incy) {
  return stream_launch<float, kAxpy>(h, n, alpha, const_cast<float*>(x), incx, y, incy);
}
pb_status_t pb_daxpy(pb_handle_t h, int64_t n, double alpha, const double* x, int64_t incx, double* y,
                     int64_t incy) {
  return stream_launch<double, kAxpy>(h, n, alpha, const_cast<double*>(x), incx, y, incy);
}
pb_status_t pb_sscal(pb_handle_t h, int64_t n, float alpha, float* x, int64_t incx) {
  return stream_launch<float, kScal>(h, n, alpha, x, incx, nullptr, 0);
}
pb_status_t pb_dscal(pb_handle_t h, int64_t n, double alpha, double* x, int64_t incx) {
  return stream_launch<double, kScal>(h, n, alpha, x, incx, nullptr, 0);
}
pb_status_t pb_scopy(pb_handle_t h, int64_t n, const float* x, int64_t incx, float* y, int64_t incy) {
  return stream_launch<float, kCopy>(h, n, 0.f, const_cast<float*>(x), incx, y, incy);
}
pb_status_t pb_dcopy(pb_handle_t h, int64_t n, const double* x, int64_t incx, double* y, int64_t incy) {
  return stream_launch<double, kCopy>(h, n, 0.0, const_cast<double*>(x), incx, y, incy);
}
pb_status_t pb_sswap(pb_handle_t h, int64_t n, float* x, int64_t incx, float* y, int64_t incy) {
  return stream_launch<float, kSwap>(h, n, 0.f, x, incx, y, incy);
}
pb_status_t pb_dswap(pb_handle_t h, int64_t n, double* x, int64_t incx, double* y, int64_t incy) {
  return stream_launch<double, kSwap>(h, n, 0.0, x, incx, y, incy);
}

}  // extern "C"

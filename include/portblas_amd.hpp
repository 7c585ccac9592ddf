// portblas_amd.hpp — header-only C++ drop-in for portBLAS's hot-path API,
// layered on the C ABI in portblas_amd.h.
//
// Same names, argument order and meaning as the reference:
//   blas::SB_Handle                       include/sb_handle/portblas_handle.h:46-200
//   blas::_gemm / _gemm_batched /
//         _gemm_strided_batched           include/interface/blas3_interface.h:86-123
//   blas::_gemv                           include/interface/blas2_interface.h:632-654
//   blas::_dot / _asum / _nrm2 /
//         _iamax / _iamin (async + sync)  include/interface/blas1_interface.h:582-718,905-1000
//   blas::extension::_reduction           include/interface/extension_interface.h:453-462
// Containers are device pointers (the USM form of the reference); events are
// HIP events wrapped in shared pointers.  Errors throw the reference's
// exception types: std::invalid_argument for bad arguments,
// blas::unsupported_exception, std::runtime_error for device failures.
#ifndef PORTBLAS_AMD_HPP
#define PORTBLAS_AMD_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "portblas_amd.h"

namespace blas {

// include/blas_meta.h:225-235
class unsupported_exception : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// include/operations/blas3_trees.h:59
enum class gemm_batch_type_t : int { strided = 0, interleaved = 1 };
// include/operations/extension/reduction.h:40
enum class reduction_dim_t : int { inner = 0, outer = 1 };

namespace detail {
inline void check(pb_status_t s) {
  if (s == PB_STATUS_SUCCESS) return;
  const std::string msg = std::string(pb_status_string(s)) + ": " + pb_last_error();
  if (s == PB_STATUS_INVALID_VALUE) throw std::invalid_argument(msg);
  if (s == PB_STATUS_NOT_SUPPORTED) throw unsupported_exception(msg);
  throw std::runtime_error(msg);
}
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace detail

class SB_Handle {
 public:
  using event_t = std::vector<std::shared_ptr<ihipEvent_t>>;

  explicit SB_Handle(hipStream_t stream = nullptr) {
    detail::check(pb_handle_create(stream, &h_));
  }
  SB_Handle(const SB_Handle&) = delete;
  SB_Handle& operator=(const SB_Handle&) = delete;
  ~SB_Handle() {
    if (scalar_) (void)hipFree(scalar_);
    if (h_) (void)pb_handle_destroy(h_);
  }

  hipStream_t get_queue() const { return static_cast<hipStream_t>(pb_handle_get_stream(h_)); }
  int get_num_compute_units() const { return pb_handle_get_num_compute_units(h_); }
  void wait() { detail::check(pb_handle_wait(h_)); }
  void wait(const event_t& events) {
    for (const auto& e : events) detail::hip_check(hipEventSynchronize(e.get()), "hipEventSynchronize");
  }
  pb_handle_t native() const { return h_; }

  // Stream-ordered dependency on `deps` (the reference's dependency lists).
  void depend(const event_t& deps) {
    for (const auto& e : deps) detail::check(pb_handle_wait_event(h_, e.get()));
  }
  // An event at the current end of the handle's stream.
  event_t record() {
    hipEvent_t e = nullptr;
    detail::hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    std::shared_ptr<ihipEvent_t> sp(e, [](hipEvent_t p) { (void)hipEventDestroy(p); });
    detail::check(pb_handle_record_event(h_, e));
    return {sp};
  }
  // 256 bytes of device memory for the synchronous scalar forms.
  void* scalar_slot() {
    if (!scalar_) detail::hip_check(hipMalloc(&scalar_, 256), "hipMalloc");
    return scalar_;
  }

 private:
  pb_handle_t h_ = nullptr;
  void* scalar_ = nullptr;
};

using event_t = SB_Handle::event_t;

namespace detail {
template <typename F>
event_t run(SB_Handle& h, const event_t& deps, F&& f) {
  h.depend(deps);
  check(f(h.native()));
  return h.record();
}
inline char trans_char(char t) { return t; }
}  // namespace detail

// ---------------------------------------------------------------------------
// BLAS-3
// ---------------------------------------------------------------------------
template <typename T>
event_t _gemm(SB_Handle& h, char transa, char transb, int64_t M, int64_t N, int64_t K, T alpha,
              const T* a, int64_t lda, const T* b, int64_t ldb, T beta, T* c, int64_t ldc,
              const event_t& deps = {}) {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value ||
                    std::is_same<T, std::complex<float>>::value ||
                    std::is_same<T, std::complex<double>>::value,
                "float, double, std::complex<float> or std::complex<double>");
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, std::complex<float>>::value)
      return pb_cgemm(p, transa, transb, M, N, K, reinterpret_cast<const float*>(&alpha),
                      reinterpret_cast<const float*>(a), lda, reinterpret_cast<const float*>(b), ldb,
                      reinterpret_cast<const float*>(&beta), reinterpret_cast<float*>(c), ldc);
    else if constexpr (std::is_same<T, std::complex<double>>::value)
      return pb_zgemm(p, transa, transb, M, N, K, reinterpret_cast<const double*>(&alpha),
                      reinterpret_cast<const double*>(a), lda, reinterpret_cast<const double*>(b), ldb,
                      reinterpret_cast<const double*>(&beta), reinterpret_cast<double*>(c), ldc);
    else if constexpr (std::is_same<T, float>::value)
      return pb_sgemm(p, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc);
    else
      return pb_dgemm(p, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc);
  });
}

template <typename T>
event_t _gemm_batched(SB_Handle& h, char transa, char transb, int64_t M, int64_t N, int64_t K,
                      T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta, T* c,
                      int64_t ldc, int64_t batch_size,
                      gemm_batch_type_t batch_type = gemm_batch_type_t::strided,
                      const event_t& deps = {}) {
  const auto bt = static_cast<pb_batch_type_t>(batch_type);
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value)
      return pb_sgemm_batched(p, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc,
                              batch_size, bt);
    else
      return pb_dgemm_batched(p, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc,
                              batch_size, bt);
  });
}

template <typename T>
event_t _gemm_strided_batched(SB_Handle& h, char transa, char transb, int64_t M, int64_t N,
                              int64_t K, T alpha, const T* a, int64_t lda, int64_t stridea,
                              const T* b, int64_t ldb, int64_t strideb, T beta, T* c, int64_t ldc,
                              int64_t stridec, int64_t batch_size, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value)
      return pb_sgemm_strided_batched(p, transa, transb, M, N, K, alpha, a, lda, stridea, b, ldb,
                                      strideb, beta, c, ldc, stridec, batch_size);
    else
      return pb_dgemm_strided_batched(p, transa, transb, M, N, K, alpha, a, lda, stridea, b, ldb,
                                      strideb, beta, c, ldc, stridec, batch_size);
  });
}

template <typename T>
event_t _symm(SB_Handle& h, char side, char uplo, int64_t M, int64_t N, T alpha, const T* a, int64_t lda,
              const T* b, int64_t ldb, T beta, T* c, int64_t ldc, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value)
      return pb_ssymm(p, side, uplo, M, N, alpha, a, lda, b, ldb, beta, c, ldc);
    else
      return pb_dsymm(p, side, uplo, M, N, alpha, a, lda, b, ldb, beta, c, ldc);
  });
}

template <typename T>
event_t _trsm(SB_Handle& h, char side, char uplo, char trans, char diag, int64_t M, int64_t N, T alpha,
              const T* a, int64_t lda, T* b, int64_t ldb, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value)
      return pb_strsm(p, side, uplo, trans, diag, M, N, alpha, a, lda, b, ldb);
    else
      return pb_dtrsm(p, side, uplo, trans, diag, M, N, alpha, a, lda, b, ldb);
  });
}

// ---------------------------------------------------------------------------
// BLAS-2
// ---------------------------------------------------------------------------
template <typename T>
event_t _gemv(SB_Handle& h, char trans, int64_t M, int64_t N, T alpha, const T* a, int64_t lda,
              const T* x, int64_t incx, T beta, T* y, int64_t incy, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value)
      return pb_sgemv(p, trans, M, N, alpha, a, lda, x, incx, beta, y, incy);
    else
      return pb_dgemv(p, trans, M, N, alpha, a, lda, x, incx, beta, y, incy);
  });
}

// ---------------------------------------------------------------------------
// BLAS-1.  Asynchronous forms add the result to *rs (device memory), as the
// reference's atomics do; the synchronous forms return the value.
// ---------------------------------------------------------------------------
template <typename T>
event_t _dot(SB_Handle& h, int64_t N, const T* x, int64_t incx, const T* y, int64_t incy, T* rs,
             const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value) return pb_sdot(p, N, x, incx, y, incy, rs);
    else return pb_ddot(p, N, x, incx, y, incy, rs);
  });
}
// sdsdot: *rs <- sb + (*rs + x . y) in FP32 (blas1_interface.h:116-125).
inline event_t _sdsdot(SB_Handle& h, int64_t N, float sb, const float* x, int64_t incx, const float* y,
                       int64_t incy, float* rs, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) { return pb_sdsdot(p, N, sb, x, incx, y, incy, rs); });
}
template <typename T>
event_t _axpy(SB_Handle& h, int64_t N, T alpha, const T* x, int64_t incx, T* y, int64_t incy,
              const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value) return pb_saxpy(p, N, alpha, x, incx, y, incy);
    else return pb_daxpy(p, N, alpha, x, incx, y, incy);
  });
}
template <typename T>
event_t _scal(SB_Handle& h, int64_t N, T alpha, T* x, int64_t incx, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value) return pb_sscal(p, N, alpha, x, incx);
    else return pb_dscal(p, N, alpha, x, incx);
  });
}
template <typename T>
event_t _copy(SB_Handle& h, int64_t N, const T* x, int64_t incx, T* y, int64_t incy,
              const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value) return pb_scopy(p, N, x, incx, y, incy);
    else return pb_dcopy(p, N, x, incx, y, incy);
  });
}
template <typename T>
event_t _swap(SB_Handle& h, int64_t N, T* x, int64_t incx, T* y, int64_t incy, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value) return pb_sswap(p, N, x, incx, y, incy);
    else return pb_dswap(p, N, x, incx, y, incy);
  });
}
template <typename T>
event_t _asum(SB_Handle& h, int64_t N, const T* x, int64_t incx, T* rs, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value) return pb_sasum(p, N, x, incx, rs);
    else return pb_dasum(p, N, x, incx, rs);
  });
}
template <typename T>
event_t _nrm2(SB_Handle& h, int64_t N, const T* x, int64_t incx, T* rs, const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value) return pb_snrm2(p, N, x, incx, rs);
    else return pb_dnrm2(p, N, x, incx, rs);
  });
}
// Index results: int32_t (portBLAS's default index_t) or int64_t.
template <typename T, typename I>
event_t _iamax(SB_Handle& h, int64_t N, const T* x, int64_t incx, I* rs, const event_t& deps = {}) {
  static_assert(std::is_same<I, int32_t>::value || std::is_same<I, int64_t>::value, "index type");
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<I, int64_t>::value) {
      static_assert(std::is_same<T, float>::value, "64-bit indices: float only");
      return pb_isamax_64(p, N, x, incx, rs);
    } else if constexpr (std::is_same<T, float>::value) {
      return pb_isamax(p, N, x, incx, rs);
    } else {
      return pb_idamax(p, N, x, incx, rs);
    }
  });
}
template <typename T, typename I>
event_t _iamin(SB_Handle& h, int64_t N, const T* x, int64_t incx, I* rs, const event_t& deps = {}) {
  static_assert(std::is_same<I, int32_t>::value || std::is_same<I, int64_t>::value, "index type");
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<I, int64_t>::value) {
      static_assert(std::is_same<T, float>::value, "64-bit indices: float only");
      return pb_isamin_64(p, N, x, incx, rs);
    } else if constexpr (std::is_same<T, float>::value) {
      return pb_isamin(p, N, x, incx, rs);
    } else {
      return pb_idamin(p, N, x, incx, rs);
    }
  });
}

namespace detail {
// Synchronous scalar: rs <- init on the device, run, copy back, wait
// (src/interface/blas1_interface.hpp:958-1178).
template <typename R, typename F>
R sync_scalar(SB_Handle& h, R init, const event_t& deps, F&& launch) {
  R* d = static_cast<R*>(h.scalar_slot());
  h.depend(deps);
  hip_check(hipMemcpyAsync(d, &init, sizeof(R), hipMemcpyHostToDevice, h.get_queue()), "hipMemcpyAsync");
  check(launch(d));
  R out{};
  hip_check(hipMemcpyAsync(&out, d, sizeof(R), hipMemcpyDeviceToHost, h.get_queue()), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(h.get_queue()), "hipStreamSynchronize");
  return out;
}
}  // namespace detail

template <typename T>
T _dot(SB_Handle& h, int64_t N, const T* x, int64_t incx, const T* y, int64_t incy,
       const event_t& deps = {}) {
  return detail::sync_scalar<T>(h, T(0), deps, [&](T* d) {
    if constexpr (std::is_same<T, float>::value) return pb_sdot(h.native(), N, x, incx, y, incy, d);
    else return pb_ddot(h.native(), N, x, incx, y, incy, d);
  });
}
inline float _sdsdot(SB_Handle& h, int64_t N, float sb, const float* x, int64_t incx, const float* y,
                     int64_t incy, const event_t& deps = {}) {
  return detail::sync_scalar<float>(h, 0.f, deps, [&](float* d) {
    return pb_sdsdot(h.native(), N, sb, x, incx, y, incy, d);
  });
}
template <typename T>
T _asum(SB_Handle& h, int64_t N, const T* x, int64_t incx, const event_t& deps = {}) {
  return detail::sync_scalar<T>(h, T(0), deps, [&](T* d) {
    if constexpr (std::is_same<T, float>::value) return pb_sasum(h.native(), N, x, incx, d);
    else return pb_dasum(h.native(), N, x, incx, d);
  });
}
template <typename T>
T _nrm2(SB_Handle& h, int64_t N, const T* x, int64_t incx, const event_t& deps = {}) {
  return detail::sync_scalar<T>(h, T(0), deps, [&](T* d) {
    if constexpr (std::is_same<T, float>::value) return pb_snrm2(h.native(), N, x, incx, d);
    else return pb_dnrm2(h.native(), N, x, incx, d);
  });
}
template <typename T>
int32_t _iamax(SB_Handle& h, int64_t N, const T* x, int64_t incx, const event_t& deps = {}) {
  return detail::sync_scalar<int32_t>(h, -1, deps, [&](int32_t* d) {
    if constexpr (std::is_same<T, float>::value) return pb_isamax(h.native(), N, x, incx, d);
    else return pb_idamax(h.native(), N, x, incx, d);
  });
}
template <typename T>
int32_t _iamin(SB_Handle& h, int64_t N, const T* x, int64_t incx, const event_t& deps = {}) {
  return detail::sync_scalar<int32_t>(h, -1, deps, [&](int32_t* d) {
    if constexpr (std::is_same<T, float>::value) return pb_isamin(h.native(), N, x, incx, d);
    else return pb_idamin(h.native(), N, x, incx, d);
  });
}

// ---------------------------------------------------------------------------
// extension::_reduction<operator_t, element_t> (operators of
// src/operations/blas_operators.hpp:149-274)
// ---------------------------------------------------------------------------
struct AddOperator { static constexpr pb_reduction_op_t op = PB_OP_ADD; };
struct ProductOperator { static constexpr pb_reduction_op_t op = PB_OP_PRODUCT; };
struct MaxOperator { static constexpr pb_reduction_op_t op = PB_OP_MAX; };
struct MinOperator { static constexpr pb_reduction_op_t op = PB_OP_MIN; };
struct AbsoluteAddOperator { static constexpr pb_reduction_op_t op = PB_OP_ABS_ADD; };
struct MeanOperator { static constexpr pb_reduction_op_t op = PB_OP_MEAN; };

namespace extension {
template <typename operator_t, typename T>
event_t _reduction(SB_Handle& h, const T* in, int64_t ld, T* out, int64_t rows, int64_t cols,
                   reduction_dim_t dim, const event_t& deps = {}) {
  const auto d = static_cast<pb_reduction_dim_t>(dim);
  return detail::run(h, deps, [&](pb_handle_t p) {
    if constexpr (std::is_same<T, float>::value)
      return pb_sreduction(p, operator_t::op, d, in, ld, out, rows, cols);
    else
      return pb_dreduction(p, operator_t::op, d, in, ld, out, rows, cols);
  });
}
}  // namespace extension

// ---------------------------------------------------------------------------
// MI355X extensions without a reference counterpart: precision-selectable
// FP32 GEMM (BF16x3 / BF16x4 / BF16 / FP8-e4m3 / FP32 SIMT / FP32 MFMA).
// ---------------------------------------------------------------------------
enum class gemm_precision_t : int {
  fp32_mfma = PB_GEMM_FP32_MFMA,
  fp32_simt = PB_GEMM_FP32_SIMT,
  bf16x3 = PB_GEMM_BF16X3,
  bf16 = PB_GEMM_BF16,
  fp8_e4m3 = PB_GEMM_FP8_E4M3,
  bf16x4 = PB_GEMM_BF16X4,
  fp16x3 = PB_GEMM_FP16X3
};

inline event_t _gemm_ex(SB_Handle& h, gemm_precision_t precision, char transa, char transb,
                        int64_t M, int64_t N, int64_t K, float alpha, const float* a, int64_t lda,
                        int64_t stridea, const float* b, int64_t ldb, int64_t strideb, float beta,
                        float* c, int64_t ldc, int64_t stridec, int64_t batch_size = 1,
                        const event_t& deps = {}) {
  return detail::run(h, deps, [&](pb_handle_t p) {
    return pb_gemm_ex(p, static_cast<pb_gemm_precision_t>(precision), transa, transb, M, N, K,
                      alpha, a, lda, stridea, b, ldb, strideb, beta, c, ldc, stridec, batch_size);
  });
}

}  // namespace blas

#endif  // PORTBLAS_AMD_HPP

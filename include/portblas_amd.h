/*
 * portblas_amd.h — C ABI of the MI355X-native portBLAS hot path.
 *
 * Every entry point takes plain device pointers, 64-bit sizes and a handle that
 * owns one HIP stream.  Nothing here throws; every call returns a pb_status_t
 * and, on failure, pb_last_error() holds a message for the calling thread.
 * Matrices are column-major exactly as in portBLAS (element (r, c) of a matrix
 * with leading dimension ld is at ptr[r + c * ld]); vectors follow the BLAS
 * stride rule portBLAS uses (include/views/view.h:47-131,
 * src/views/view_sycl.hpp:95-157): element i of a vector with increment
 * inc < 0 is ptr[(n - 1 - i) * |inc|].
 *
 * Calls are asynchronous with respect to the host (stream-ordered on the
 * handle's stream) unless the name ends in _sync.  Scalar BLAS-1 results of
 * the asynchronous calls are written to caller-owned device memory.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the portBLAS source tree).
 */
#ifndef PORTBLAS_AMD_H
#define PORTBLAS_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PB_ABI_VERSION 1

typedef enum pb_status_t {
  PB_STATUS_SUCCESS = 0,
  PB_STATUS_INVALID_VALUE = 1,   /* std::invalid_argument in portBLAS */
  PB_STATUS_NOT_SUPPORTED = 2,   /* blas::unsupported_exception */
  PB_STATUS_DEVICE_ERROR = 3,    /* a HIP call failed */
  PB_STATUS_ALLOC_FAILED = 4,    /* workspace could not be grown */
  PB_STATUS_NOT_INITIALIZED = 5  /* null handle */
} pb_status_t;

/* gemm_batch_type_t, include/operations/blas3_trees.h:59 */
typedef enum pb_batch_type_t {
  PB_BATCH_STRIDED = 0,
  PB_BATCH_INTERLEAVED = 1
} pb_batch_type_t;

/* reduction_dim_t, include/operations/extension/reduction.h:40 */
typedef enum pb_reduction_dim_t {
  PB_REDUCE_INNER = 0, /* out[c] = op over rows r of A[r, c] (length cols) */
  PB_REDUCE_OUTER = 1  /* out[r] = op over cols c of A[r, c] (length rows) */
} pb_reduction_dim_t;

/* Operators of blas::extension::_reduction, src/operations/blas_operators.hpp:149-274 */
typedef enum pb_reduction_op_t {
  PB_OP_ADD = 0,
  PB_OP_PRODUCT = 1,
  PB_OP_MAX = 2,
  PB_OP_MIN = 3,
  PB_OP_ABS_ADD = 4,
  PB_OP_MEAN = 5
} pb_reduction_op_t;

/* Arithmetic used by pb_gemm_ex for an FP32-in / FP32-out GEMM. */
typedef enum pb_gemm_precision_t {
  PB_GEMM_FP32_MFMA = 0, /* plain FP32 on v_mfma_f32_*_f32 (the default of pb_sgemm) */
  PB_GEMM_FP32_SIMT = 1, /* plain FP32 on vector FMA, LDS-tiled */
  PB_GEMM_BF16X3 = 2,    /* split precision: hi*hi + hi*lo + lo*hi on BF16 MFMA, FP32 accumulate */
  PB_GEMM_BF16 = 3,      /* inputs rounded to BF16, BF16 MFMA, FP32 accumulate */
  PB_GEMM_FP8_E4M3 = 4,  /* inputs scaled by per-tensor powers of two (amax to at most 448) to OCP e4m3,
                          * FP8 MFMA, FP32 accumulate */
  PB_GEMM_BF16X4 = 5,    /* BF16X3 plus the lo*lo term: smaller representation error */
  PB_GEMM_FP16X3 = 6     /* split precision in IEEE half: each operand scaled by a power of two
                          * (amax over its finite elements to [2^14, 2^15)), then hi = f16(x),
                          * lo = f16(x - hi) (22 significand bits); hi*hi + hi*lo + lo*hi on
                          * FP16 MFMA, FP32 accumulate, unscaled in the epilogue.  Elements
                          * below 2^-28 of their operand's amax lose significand bits (FP16
                          * subnormals); amax must lie in [2^-59, 2^78] */
} pb_gemm_precision_t;

typedef struct pb_handle_s* pb_handle_t;
typedef struct pb_graph_s* pb_graph_t;
typedef struct pb_grouped_plan_s* pb_grouped_plan_t;

/* ------------------------------------------------------------------------ */
/* Handle: replaces blas::SB_Handle(sycl::queue)                             */
/* include/sb_handle/portblas_handle.h:46-200                                */
/* ------------------------------------------------------------------------ */

/* stream: a hipStream_t (NULL = the null stream).  The handle keeps a
 * workspace pool (replaces Temp_Mem_Pool, include/sb_handle/temp_memory_pool.h:33-114). */
pb_status_t pb_handle_create(void* stream, pb_handle_t* handle);
pb_status_t pb_handle_destroy(pb_handle_t handle);
pb_status_t pb_handle_set_stream(pb_handle_t handle, void* stream);
void* pb_handle_get_stream(pb_handle_t handle);
/* SB_Handle::get_num_compute_units, portblas_handle.h:178-180 */
int pb_handle_get_num_compute_units(pb_handle_t handle);
/* SB_Handle::wait(), portblas_handle.h:160-170: blocks until the stream drains. */
pb_status_t pb_handle_wait(pb_handle_t handle);
/* Grow the workspace up front so later calls (e.g. under graph capture) never allocate. */
pb_status_t pb_handle_reserve_workspace(pb_handle_t handle, size_t bytes);
size_t pb_handle_workspace_bytes(pb_handle_t handle);
/* Events: portBLAS returns event_t = std::vector<sycl::event> and takes
 * dependency lists; here a dependency is a hipEvent_t waited on the stream. */
pb_status_t pb_handle_wait_event(pb_handle_t handle, void* hip_event);
pb_status_t pb_handle_record_event(pb_handle_t handle, void* hip_event);

const char* pb_status_string(pb_status_t status);
/* Message of the last failure on the calling thread ("" if none). */
const char* pb_last_error(void);
int pb_abi_version(void);

/* ------------------------------------------------------------------------ */
/* GEMM: C <- alpha * op(A) * op(B) + beta * C                               */
/* blas::_gemm, include/interface/blas3_interface.h:86-95;                  */
/* semantics src/interface/gemm_interface.hpp:105-198                        */
/* transa/transb: 'n','t','c' in either case ('c' == 't' for real types);   */
/* anything else -> PB_STATUS_INVALID_VALUE.  alpha == 0 scales C by beta   */
/* as gemm_interface.hpp:112-139 does; beta == 0 never reads C.             */
/* ------------------------------------------------------------------------ */
pb_status_t pb_sgemm(pb_handle_t h, char transa, char transb, int64_t m, int64_t n,
                     int64_t k, float alpha, const float* a, int64_t lda,
                     const float* b, int64_t ldb, float beta, float* c, int64_t ldc);
pb_status_t pb_dgemm(pb_handle_t h, char transa, char transb, int64_t m, int64_t n,
                     int64_t k, double alpha, const double* a, int64_t lda,
                     const double* b, int64_t ldb, double beta, double* c, int64_t ldc);

/* Allocation-free pb_sgemm.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * validation, plan, kernels and results as pb_sgemm (except where an installed
 * plain selector picks a split whose partials exceed the array: then the
 * built-in rule's plan), but the split-K partial
 * tiles live in a 16 MiB array the library holds on each GPU from load to
 * process exit, and the call never grows the handle workspace and never
 * allocates.  A plan needing more partials than the array holds returns
 * PB_STATUS_ALLOC_FAILED (a plain FP32 plan needs at most 8 MiB on 256 CUs).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two streams, two threads) race; serialize them. */
pb_status_t pb_sgemm_noalloc(pb_handle_t h, char transa, char transb, int64_t m,
                             int64_t n, int64_t k, float alpha, const float* a,
                             int64_t lda, const float* b, int64_t ldb, float beta,
                             float* c, int64_t ldc);

/* blas::_gemm_batched, blas3_interface.h:97-109; default strides
 * gemm_interface.hpp:202-226 (strided: A stride = (ta=='n' ? K : M) * lda,
 * B stride = (tb=='n' ? N : K) * ldb, C stride = N * ldc).  Interleaved:
 * element (r, c) of batch b is at ptr[b + batch * (r + ld * c)]. */
pb_status_t pb_sgemm_batched(pb_handle_t h, char transa, char transb, int64_t m,
                             int64_t n, int64_t k, float alpha, const float* a,
                             int64_t lda, const float* b, int64_t ldb, float beta,
                             float* c, int64_t ldc, int64_t batch_size,
                             pb_batch_type_t batch_type);
pb_status_t pb_dgemm_batched(pb_handle_t h, char transa, char transb, int64_t m,
                             int64_t n, int64_t k, double alpha, const double* a,
                             int64_t lda, const double* b, int64_t ldb, double beta,
                             double* c, int64_t ldc, int64_t batch_size,
                             pb_batch_type_t batch_type);

/* blas::_gemm_strided_batched, blas3_interface.h:111-123; stride checks
 * gemm_interface.hpp:153-166 (stridec < ldc*N, or any negative stride, with
 * batch > 1 -> PB_STATUS_INVALID_VALUE).  A/B stride 0 broadcasts. */
pb_status_t pb_sgemm_strided_batched(pb_handle_t h, char transa, char transb,
                                     int64_t m, int64_t n, int64_t k, float alpha,
                                     const float* a, int64_t lda, int64_t stridea,
                                     const float* b, int64_t ldb, int64_t strideb,
                                     float beta, float* c, int64_t ldc,
                                     int64_t stridec, int64_t batch_size);
pb_status_t pb_dgemm_strided_batched(pb_handle_t h, char transa, char transb,
                                     int64_t m, int64_t n, int64_t k, double alpha,
                                     const double* a, int64_t lda, int64_t stridea,
                                     const double* b, int64_t ldb, int64_t strideb,
                                     double beta, double* c, int64_t ldc,
                                     int64_t stridec, int64_t batch_size);

/* FP32-in / FP32-out GEMM with a selectable arithmetic (no reference
 * counterpart beyond blas::_gemm itself; the reduced-precision precedent is
 * src/operations/blas3/gemm_local_joint_matrix.hpp).  Same semantics as
 * pb_sgemm_strided_batched.  PB_GEMM_FP8_E4M3 derives one power-of-two scale
 * per operand from amax over all of that operand's batches. */
pb_status_t pb_gemm_ex(pb_handle_t h, pb_gemm_precision_t precision, char transa,
                       char transb, int64_t m, int64_t n, int64_t k, float alpha,
                       const float* a, int64_t lda, int64_t stridea, const float* b,
                       int64_t ldb, int64_t strideb, float beta, float* c,
                       int64_t ldc, int64_t stridec, int64_t batch_size);

/* Split once, multiply many times (no reference counterpart): writes op(X)
 * as BF16 hi = bf16(x) and lo = bf16(x - hi) in the packed format of the
 * BF16x3 kernel: K-slab-major, element (r, k) of the rows x ldp operand at
 * p[(k / 16) * rows * 16 + r * 16 + k % 16], zeros for k in [k, ldp); ldp a
 * multiple of 64.  operand 0: X is A and op(A) is rows x k; operand 1: X is
 * B and op(B) is k x rows. */
pb_status_t pb_pack_bf16x3(pb_handle_t h, int operand, char trans, int64_t rows, int64_t k,
                           const float* x, int64_t ldx, uint16_t* hi, uint16_t* lo,
                           int64_t ldp);
/* C = alpha * (A_hi B_lo + A_lo B_hi + A_hi B_hi) + beta * C on packed
 * operands from pb_pack_bf16x3 (the arithmetic of PB_GEMM_BF16X3); both
 * operands share one packed pitch (ldpa == ldpb). */
pb_status_t pb_gemm_bf16x3_packed(pb_handle_t h, int64_t m, int64_t n, int64_t k, float alpha,
                                  const uint16_t* a_hi, const uint16_t* a_lo, int64_t ldpa,
                                  const uint16_t* b_hi, const uint16_t* b_lo, int64_t ldpb,
                                  float beta, float* c, int64_t ldc);

/* The same pair for PB_GEMM_FP16X3 (no reference counterpart): op(X) is
 * scaled by 2^s (amax of its finite elements to [2^14, 2^15), s clamped to
 * [-63, 74]) and written as IEEE half hi = f16(x 2^s), lo = f16(x 2^s - hi)
 * in the packed layout above; *factor (device memory) receives 2^-s. */
pb_status_t pb_pack_fp16x3(pb_handle_t h, int operand, char trans, int64_t rows, int64_t k,
                           const float* x, int64_t ldx, uint16_t* hi, uint16_t* lo,
                           int64_t ldp, float* factor);
/* C = alpha * fa * fb * (A_hi B_lo + A_lo B_hi + A_hi B_hi) + beta * C on
 * operands from pb_pack_fp16x3, fa and fb read from the device factors it
 * wrote. */
pb_status_t pb_gemm_fp16x3_packed(pb_handle_t h, int64_t m, int64_t n, int64_t k, float alpha,
                                  const uint16_t* a_hi, const uint16_t* a_lo, const float* a_factor,
                                  int64_t ldpa, const uint16_t* b_hi, const uint16_t* b_lo,
                                  const float* b_factor, int64_t ldpb, float beta, float* c,
                                  int64_t ldc);

/* BF16-in (raw uint16 bit patterns), FP32-accumulate, FP32-out GEMM. */
pb_status_t pb_gemm_bf16(pb_handle_t h, char transa, char transb, int64_t m,
                         int64_t n, int64_t k, float alpha, const uint16_t* a,
                         int64_t lda, int64_t stridea, const uint16_t* b,
                         int64_t ldb, int64_t strideb, float beta, float* c,
                         int64_t ldc, int64_t stridec, int64_t batch_size);

/* Grouped (uneven) GEMM: no reference counterpart (portBLAS batches share
 * one shape, gemm_interface.hpp:202-240).  All problems run in ONE persistent
 * launch.  Inputs are BF16 (uint16 bit patterns) or FP32 per `precision`
 * (PB_GEMM_BF16 or PB_GEMM_FP32_MFMA); C is always FP32. */
typedef struct pb_gemm_problem_t {
  char transa, transb;
  int64_t m, n, k;
  float alpha, beta;
  const void* a; int64_t lda;
  const void* b; int64_t ldb;
  float* c; int64_t ldc;
} pb_gemm_problem_t;

/* Upload the problem list once (allocates); the plan can then be run any
 * number of times, including inside a captured graph. */
pb_status_t pb_gemm_grouped_plan_create(pb_handle_t h, pb_gemm_precision_t precision,
                                        const pb_gemm_problem_t* problems,
                                        int64_t count, pb_grouped_plan_t* plan);
pb_status_t pb_gemm_grouped_plan_run(pb_handle_t h, pb_grouped_plan_t plan);
pb_status_t pb_gemm_grouped_plan_destroy(pb_grouped_plan_t plan);
/* One-shot convenience: create, run, destroy (synchronises before freeing). */
pb_status_t pb_gemm_grouped(pb_handle_t h, pb_gemm_precision_t precision,
                            const pb_gemm_problem_t* problems, int64_t count);
/* The schedule choices of a BF16 grouped plan, for the autotuner
 * (portblas_amd.tuning.GroupedAutotuner times them on the caller's problems).
 * No reference counterpart.  Zero in every field means the built-in choice,
 * which pb_gemm_grouped_plan_create uses:
 *   ring:         1 = the ring kernel for the problems it takes (dimensions
 *                 multiples of 8, aligned), 2 = the lean kernel for all;
 *   rank_weights: 1 = weigh the workgroups sharing a CU by their measured
 *                 speeds, 2 = balance them as equals;
 *   spread_div:   problems worth more than 1/spread_div of the launch are
 *                 spread over all XCDs (built-in 16; 1 = never spread);
 *   per_cu:       workgroups per CU of the grid (built-in: as many as are
 *                 resident; larger values are clamped to that). */
typedef struct pb_grouped_options_t {
  int ring, rank_weights, spread_div, per_cu;
} pb_grouped_options_t;
pb_status_t pb_gemm_grouped_plan_create_ex(pb_handle_t h, pb_gemm_precision_t precision,
                                           const pb_gemm_problem_t* problems, int64_t count,
                                           const pb_grouped_options_t* options,
                                           pb_grouped_plan_t* plan);

/* ------------------------------------------------------------------------ */
/* GEMV: y <- alpha * op(A) * x + beta * y                                   */
/* blas::_gemv, include/interface/blas2_interface.h:629-654; semantics      */
/* src/interface/blas2_interface.hpp:68-217,1163-1192 (trans 'n' vs any     */
/* other character; beta == 0 never reads y).                               */
/* ------------------------------------------------------------------------ */
pb_status_t pb_sgemv(pb_handle_t h, char trans, int64_t m, int64_t n, float alpha,
                     const float* a, int64_t lda, const float* x, int64_t incx,
                     float beta, float* y, int64_t incy);
pb_status_t pb_dgemv(pb_handle_t h, char trans, int64_t m, int64_t n, double alpha,
                     const double* a, int64_t lda, const double* x, int64_t incx,
                     double beta, double* y, int64_t incy);

/* ------------------------------------------------------------------------ */
/* BLAS-1 reductions (asynchronous; result in device memory)                */
/* _dot: *rs <- *rs + sum x_i y_i   (blas1_interface.h:582-590;              */
/*       blas1_interface.hpp:611-642; n == 0 leaves *rs untouched)           */
/* _asum: *rs <- *rs + sum |x_i|    (blas1_interface.h:628-634; .hpp:222-246)*/
/* _nrm2: *rs <- sqrt(*rs + sum x_i^2) (blas1_interface.h:709-715;           */
/*       .hpp:552-581, no scaling)                                           */
/* These accumulate into *rs exactly as portBLAS's atomics do; the _sync     */
/* variants start from zero like blas1_interface.hpp:958-1178.              */
/* ------------------------------------------------------------------------ */
pb_status_t pb_sdot(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                    const float* y, int64_t incy, float* rs);
pb_status_t pb_ddot(pb_handle_t h, int64_t n, const double* x, int64_t incx,
                    const double* y, int64_t incy, double* rs);
pb_status_t pb_sasum(pb_handle_t h, int64_t n, const float* x, int64_t incx, float* rs);
pb_status_t pb_dasum(pb_handle_t h, int64_t n, const double* x, int64_t incx, double* rs);
pb_status_t pb_snrm2(pb_handle_t h, int64_t n, const float* x, int64_t incx, float* rs);
pb_status_t pb_dnrm2(pb_handle_t h, int64_t n, const double* x, int64_t incx, double* rs);

/* _sdsdot: *rs <- sb + (*rs + sum x_i y_i) in FP32 (blas1_interface.h:116-125;*/
/*       blas1_interface.hpp:148-177: the dot accumulates into *rs, then sb   */
/*       is added); n == 0 writes sb.  The _sync form starts from zero       */
/*       (blas1_interface.hpp:1013-1041).                                    */
pb_status_t pb_sdsdot(pb_handle_t h, int64_t n, float sb, const float* x, int64_t incx,
                      const float* y, int64_t incy, float* rs);
pb_status_t pb_sdsdot_sync(pb_handle_t h, int64_t n, float sb, const float* x, int64_t incx,
                           const float* y, int64_t incy, float* result_host);

/* Streaming BLAS-1 (blas1_interface.h:45-70,193-211; blas1_interface.hpp:   */
/* 46-74,419-456).  n <= 0 does nothing.                                     */
/* _axpy: y <- alpha x + y   (one fused multiply-add per element)          */
/* _scal: x <- alpha x; alpha == 0 assigns zero (NaN and inf included)      */
/* _copy: y <- x              _swap: x <-> y                                 */
pb_status_t pb_saxpy(pb_handle_t h, int64_t n, float alpha, const float* x, int64_t incx, float* y,
                     int64_t incy);
pb_status_t pb_daxpy(pb_handle_t h, int64_t n, double alpha, const double* x, int64_t incx,
                     double* y, int64_t incy);
pb_status_t pb_sscal(pb_handle_t h, int64_t n, float alpha, float* x, int64_t incx);
pb_status_t pb_dscal(pb_handle_t h, int64_t n, double alpha, double* x, int64_t incx);
pb_status_t pb_scopy(pb_handle_t h, int64_t n, const float* x, int64_t incx, float* y,
                     int64_t incy);
pb_status_t pb_dcopy(pb_handle_t h, int64_t n, const double* x, int64_t incx, double* y,
                     int64_t incy);
pb_status_t pb_sswap(pb_handle_t h, int64_t n, float* x, int64_t incx, float* y, int64_t incy);
pb_status_t pb_dswap(pb_handle_t h, int64_t n, double* x, int64_t incx, double* y, int64_t incy);

/* _iamax/_iamin (blas1_interface.h:643-667; blas1_interface.hpp:278-405;   */
/* operators blas_operators.hpp:276-316): 0-based index of the first max   */
/* (min) of |x_i|; the lowest index wins ties; NaN never wins; incx < 0 or  */
/* n < 0 writes 0; n == 0 (or every element NaN for iamax) writes the       */
/* index type's maximum, the index of portBLAS's initial tuple              */
/* (blas_constants.h:110-136).  `rs` is int32 (portBLAS's default int        */
/* index_t) or int64 for the _64 variants.                                   */
pb_status_t pb_isamax(pb_handle_t h, int64_t n, const float* x, int64_t incx, int32_t* rs);
pb_status_t pb_isamin(pb_handle_t h, int64_t n, const float* x, int64_t incx, int32_t* rs);
pb_status_t pb_idamax(pb_handle_t h, int64_t n, const double* x, int64_t incx, int32_t* rs);
pb_status_t pb_idamin(pb_handle_t h, int64_t n, const double* x, int64_t incx, int32_t* rs);
pb_status_t pb_isamax_64(pb_handle_t h, int64_t n, const float* x, int64_t incx, int64_t* rs);
pb_status_t pb_isamin_64(pb_handle_t h, int64_t n, const float* x, int64_t incx, int64_t* rs);

/* Synchronous scalar-returning forms (blas1_interface.h:905-1000). */
pb_status_t pb_sdot_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                         const float* y, int64_t incy, float* result_host);
pb_status_t pb_sasum_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                          float* result_host);
pb_status_t pb_snrm2_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                          float* result_host);
pb_status_t pb_isamax_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                           int32_t* result_host);
pb_status_t pb_isamin_sync(pb_handle_t h, int64_t n, const float* x, int64_t incx,
                           int32_t* result_host);

/* ------------------------------------------------------------------------ */
/* FP16 GEMMs: the reference's _gemm on sycl::half (BLAS_ENABLE_HALF;       */
/* src/interface/blas3/backend/default.hpp:150-246, amd_gpu.hpp:292-359).    */
/* Operands are IEEE binary16 bit patterns, accumulation is FP32 (MFMA).     */
/*   pb_hgemm:    A, B, C half; alpha and beta passed as float; the result   */
/*                is rounded to half once (beta == 0 never reads C).         */
/*   pb_gemm_f16: A, B half, C FP32 (strided batches), like pb_gemm_bf16.    */
/* ------------------------------------------------------------------------ */
pb_status_t pb_hgemm(pb_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                     float alpha, const uint16_t* a, int64_t lda, const uint16_t* b, int64_t ldb, float beta,
                     uint16_t* c, int64_t ldc);
pb_status_t pb_gemm_f16(pb_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                        float alpha, const uint16_t* a, int64_t lda, int64_t stridea, const uint16_t* b,
                        int64_t ldb, int64_t strideb, float beta, float* c, int64_t ldc, int64_t stridec,
                        int64_t batch);

/* ------------------------------------------------------------------------ */
/* Complex GEMM: blas::_gemm on std::complex<float / double> (built with     */
/* BLAS_ENABLE_COMPLEX; src/interface/gemm_interface.hpp:58-64,140-198,     */
/* test/unittest/blas3/blas3_gemm_common.hpp:480-580).  Matrices hold        */
/* interleaved (re, im) pairs, leading dimensions count complex elements,     */
/* alpha and beta point to (re, im) pairs in host memory.  'c' is taken as   */
/* 't' without conjugation, as the reference does; beta == 0 never reads C; */
/* alpha == 0 (or K == 0) gives C <- beta C.                                */
/* ------------------------------------------------------------------------ */
pb_status_t pb_cgemm(pb_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                     const float* alpha, const float* a, int64_t lda, const float* b, int64_t ldb,
                     const float* beta, float* c, int64_t ldc);
pb_status_t pb_zgemm(pb_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                     const double* alpha, const double* a, int64_t lda, const double* b, int64_t ldb,
                     const double* beta, double* c, int64_t ldc);

/* ------------------------------------------------------------------------ */
/* blas::_symm (include/interface/blas3_interface.h:78-84,138-145;          */
/* src/interface/symm_interface.hpp:36-80): C <- alpha A B + beta C (side    */
/* 'l', A M x M) or alpha B A + beta C (side 'r', A N x N); A symmetric,     */
/* only its uplo ('u' / 'l') triangle is read; side and uplo are case-       */
/* insensitive, anything else is PB_STATUS_INVALID_VALUE ("invalid _side" /  */
/* "invalid _uplo").  alpha == 0 gives C <- beta C.                          */
/* ------------------------------------------------------------------------ */
pb_status_t pb_ssymm(pb_handle_t h, char side, char uplo, int64_t m, int64_t n, float alpha,
                     const float* a, int64_t lda, const float* b, int64_t ldb, float beta, float* c,
                     int64_t ldc);
pb_status_t pb_dsymm(pb_handle_t h, char side, char uplo, int64_t m, int64_t n, double alpha,
                     const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                     double* c, int64_t ldc);

/* ------------------------------------------------------------------------ */
/* blas::_trsm (include/interface/blas3_interface.h:70-75,127-134;         */
/* src/interface/trsm_interface.hpp:36-420): solves op(A) X = alpha B (side  */
/* 'l', A M x M) or X op(A) = alpha B (side 'r', A N x N) and overwrites B   */
/* with X.  A is triangular (uplo 'u' / 'l'; only that triangle is read),    */
/* op(A) = A or A^T (trans 'n' / 't'), diag 'u' treats the diagonal as ones  */
/* without reading it.  Flags are case-insensitive.  M, N, lda or ldb equal  */
/* to zero, or an invalid flag, is PB_STATUS_INVALID_VALUE with the          */
/* reference's messages ("invalid matrix size argument", "invalid Side      */
/* argument", ...).                                                         */
/* ------------------------------------------------------------------------ */
pb_status_t pb_strsm(pb_handle_t h, char side, char uplo, char trans, char diag, int64_t m, int64_t n,
                     float alpha, const float* a, int64_t lda, float* b, int64_t ldb);
pb_status_t pb_dtrsm(pb_handle_t h, char side, char uplo, char trans, char diag, int64_t m, int64_t n,
                     double alpha, const double* a, int64_t lda, double* b, int64_t ldb);

/* ------------------------------------------------------------------------ */
/* blas::extension::_reduction<Op, T>                                       */
/* include/interface/extension_interface.h:453-462;                         */
/* src/interface/extension_interface.hpp:344-410,579-594;                   */
/* src/operations/extension/reduction.hpp:35-230.  Column-major rows x cols */
/* input with leading dimension ld; output (length cols for INNER, rows for */
/* OUTER) is overwritten.  MEAN divides by the reduced length.             */
/* ------------------------------------------------------------------------ */
pb_status_t pb_sreduction(pb_handle_t h, pb_reduction_op_t op, pb_reduction_dim_t dim,
                          const float* in, int64_t ld, float* out, int64_t rows,
                          int64_t cols);
pb_status_t pb_dreduction(pb_handle_t h, pb_reduction_op_t op, pb_reduction_dim_t dim,
                          const double* in, int64_t ld, double* out, int64_t rows,
                          int64_t cols);

/* ------------------------------------------------------------------------ */
/* Graph capture of the handle's stream (replaces nothing in portBLAS; SYCL */
/* has no graphs).  Calls between begin and end are recorded, not run.      */
/* ------------------------------------------------------------------------ */
pb_status_t pb_graph_capture_begin(pb_handle_t h);
pb_status_t pb_graph_capture_end(pb_handle_t h, pb_graph_t* graph);
pb_status_t pb_graph_launch(pb_handle_t h, pb_graph_t graph);
pb_status_t pb_graph_destroy(pb_graph_t graph);

/* ------------------------------------------------------------------------ */
/* Tile selection: autotuner and learned selector                           */
/* (replaces the hand heuristic of src/interface/blas3/backend/amd_gpu.hpp: */
/* 52-288 and the exhaustive search of tools/auto_tuner/include/           */
/* gemm_tuner.hpp:99-180).                                                  */
/* ------------------------------------------------------------------------ */
/* Number of tile/schedule configurations of the BF16-MFMA GEMM family. */
int pb_tune_num_configs(void);
/* Human-readable name of config `id` (static storage). */
const char* pb_tune_config_name(int id);
/* Run GEMM config `id` (PB_GEMM_BF16 or PB_GEMM_FP32_MFMA arithmetic on
 * FP32/BF16 data as for pb_gemm_ex / pb_gemm_bf16) and return the mean
 * device time per call in microseconds over `reps` timed launches after
 * `warmup` untimed ones.  Inputs are BF16 bit patterns. */
pb_status_t pb_tune_time_config(pb_handle_t h, int id, char transa, char transb,
                                int64_t m, int64_t n, int64_t k, const uint16_t* a,
                                int64_t lda, const uint16_t* b, int64_t ldb, float* c,
                                int64_t ldc, int64_t batch, int warmup, int reps,
                                float* usec_out);
/* Run one BF16 GEMM with an explicit config (for parity tests of every config). */
pb_status_t pb_gemm_bf16_config(pb_handle_t h, int id, char transa, char transb,
                                int64_t m, int64_t n, int64_t k, float alpha,
                                const uint16_t* a, int64_t lda, int64_t stridea,
                                const uint16_t* b, int64_t ldb, int64_t strideb,
                                float beta, float* c, int64_t ldc, int64_t stridec,
                                int64_t batch_size);
/* Install a selector: the library then picks the config for BF16 GEMMs by
 * evaluating this MLP (weights are host FP32 arrays, copied).  layer_sizes
 * has n_layers + 1 entries (input width first).  Pass n_layers = 0 to
 * return to the built-in rule. */
pb_status_t pb_tune_set_selector(pb_handle_t h, int n_layers, const int* layer_sizes,
                                 const float* const* weights, const float* const* biases,
                                 const float* feat_mean, const float* feat_std);
/* The config the handle would use for this shape (-1 on error). */
int pb_tune_select(pb_handle_t h, char transa, char transb, int64_t m, int64_t n,
                   int64_t k, int64_t batch);

/* The same for the plain FP32 / FP64 engine behind pb_sgemm / pb_dgemm and
 * their batched forms (src/interface/blas3/backend/amd_gpu.hpp:52-288 picks
 * these for the reference's own types).  Configurations: 128 x 128 tiles
 * with split-K 1, 2, ..., 128 (ids 0-7), 64 x 64 tiles with the same splits
 * (8-15), and the small-problem batch kernel (16, only for M, N <= 64 with
 * at least one problem per CU).  dtype: 1 = FP32, 2 = FP64. */
int pb_tune_num_plain_configs(void);
const char* pb_tune_plain_config_name(int id);
/* One strided-batched GEMM (default strides) with config `id` (-1: the
 * built-in rule); alpha and beta are converted to the element type.
 * PB_STATUS_NOT_SUPPORTED for the small-batch config on a shape it cannot run. */
pb_status_t pb_gemm_plain_config(pb_handle_t h, int dtype, int id, char transa, char transb,
                                 int64_t m, int64_t n, int64_t k, double alpha, const void* a,
                                 int64_t lda, const void* b, int64_t ldb, double beta, void* c,
                                 int64_t ldc, int64_t batch);
/* Mean device time per call (microseconds) of config `id` (-1: the built-in
 * rule), alpha 1, beta 0, over `reps` launches after `warmup`. */
pb_status_t pb_tune_time_plain(pb_handle_t h, int dtype, int id, char transa, char transb,
                               int64_t m, int64_t n, int64_t k, const void* a, int64_t lda,
                               const void* b, int64_t ldb, void* c, int64_t ldc, int64_t batch,
                               int warmup, int reps, float* usec_out);
/* Install a network with pb_tune_num_plain_configs() outputs for the plain
 * engine (features: log2 m, n, k, transa, transb, dtype, log2 batch, log2 of
 * lda over the rows of A); n_layers = 0 returns to the built-in rule. */
pb_status_t pb_tune_set_plain_selector(pb_handle_t h, int n_layers, const int* layer_sizes,
                                       const float* const* weights, const float* const* biases,
                                       const float* feat_mean, const float* feat_std);
/* The config the installed plain network picks; -1 when none is installed or
 * the shape lies outside the network's training domain (work above 2^35
 * flop, operands above 2^26 elements or batches above 256), where the
 * library uses its built-in rule. */
int pb_tune_select_plain(pb_handle_t h, int dtype, char transa, char transb, int64_t m, int64_t n,
                         int64_t k, int64_t batch, int64_t lda);

/* ------------------------------------------------------------------------ */
/* Learned-selector kernels: MLP forward/backward/update on the GPU.        */
/* Layer l maps width d_l to d_{l+1}: Z = X W^T + b, ReLU on hidden layers, */
/* linear output; loss = mean squared error over all outputs.  All arrays   */
/* are device FP32, row-major [batch x width], W_l row-major [d_{l+1} x d_l]*/
/* ------------------------------------------------------------------------ */
/* One forward pass; activations (n_layers + 1 arrays) are kept for backward. */
pb_status_t pb_mlp_forward(pb_handle_t h, int n_layers, const int* layer_sizes,
                           const float* const* w, const float* const* b,
                           float* const* acts, int64_t batch);
/* BF16 inference: weights/activations rounded to BF16, FP32 accumulate. */
pb_status_t pb_mlp_forward_bf16(pb_handle_t h, int n_layers, const int* layer_sizes,
                                const uint16_t* const* w, const float* const* b,
                                const float* x, float* y, int64_t batch);
/* Backward of mean-squared-error loss against `target`; writes dW, db and
 * the loss (device scalar). `grad_scratch` holds 2 * batch * max_width floats. */
pb_status_t pb_mlp_backward(pb_handle_t h, int n_layers, const int* layer_sizes,
                            const float* const* w, const float* const* acts,
                            const float* target, const float* mask,
                            float* const* dw, float* const* db, float* loss,
                            float* grad_scratch, int64_t batch);
/* Adam update of one parameter array. */
pb_status_t pb_mlp_adam(pb_handle_t h, float* param, const float* grad, float* m,
                        float* v, int64_t count, float lr, float beta1, float beta2,
                        float eps, int step);

#ifdef __cplusplus
}
#endif

#endif /* PORTBLAS_AMD_H */

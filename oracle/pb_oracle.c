/*
 * pb_oracle.c — CPU restatement of portBLAS's hot path, used ONLY as test
 * infrastructure: the parity checker of the tests, the `smoke()` check and
 * the CPU baseline leg of bench.py.  Nothing in the shipped library links,
 * loads or calls this file.
 *
 * Two families of functions:
 *   orc_*_f64      exact-as-practical references: FP64 accumulation of the
 *                  textbook definitions (the oracle the GPU results are
 *                  compared with, within stated FP32/FP64 tolerances);
 *   orc_pb_*       restatements of portBLAS's own CPU ("DEFAULT" tuning
 *                  target) schedules in the element type, timed as the CPU
 *                  baseline; each cites the reference lines it follows.
 * iamax/iamin are integer results and are restated exactly.
 *
 * Paths in citations are relative to the portBLAS source tree.
 * Build: oracle/Makefile (gcc -O3 -fopenmp -shared).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static int lower(int c) { return (c >= 'A' && c <= 'Z') ? c - 'A' + 'a' : c; }

/* Element (r, k) of op(A) for a column-major A with leading dimension lda. */
#define OPA(ta, A, lda, r, k) ((ta) == 'n' ? (A)[(r) + (int64_t)(k) * (lda)] : (A)[(k) + (int64_t)(r) * (lda)])
/* Element (k, c) of op(B). */
#define OPB(tb, B, ldb, k, c) ((tb) == 'n' ? (B)[(k) + (int64_t)(c) * (ldb)] : (B)[(c) + (int64_t)(k) * (ldb)])

/* -------------------------------------------------------------------------
 * GEMM reference in FP64: C = alpha * op(A) op(B) + beta * C; beta == 0 does
 * not read C (src/operations/blas3/gemm_ref.hpp:234-251).  'c' == 't'.
 * ------------------------------------------------------------------------- */
void orc_gemm_f64(char ta_, char tb_, int64_t M, int64_t N, int64_t K, double alpha,
                  const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                  double* C, int64_t ldc) {
  const int ta = lower(ta_) == 'n' ? 'n' : 't';
  const int tb = lower(tb_) == 'n' ? 'n' : 't';
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j) {
    double* acc = (double*)calloc((size_t)(M > 0 ? M : 1), sizeof(double));
    for (int64_t k = 0; k < K; ++k) {
      const double b = OPB(tb, B, ldb, k, j);
      if (ta == 'n') {
        const double* a = A + k * lda;
        for (int64_t i = 0; i < M; ++i) acc[i] += a[i] * b;
      } else {
        for (int64_t i = 0; i < M; ++i) acc[i] += A[k + i * lda] * b;
      }
    }
    for (int64_t i = 0; i < M; ++i) {
      double* c = C + i + j * ldc;
      *c = (beta != 0.0) ? alpha * acc[i] + beta * *c : alpha * acc[i];
    }
    free(acc);
  }
}

/* |op(A)| |op(B)| in FP64: the scale of the rounding error of every entry. */
void orc_gemm_absprod_f64(char ta_, char tb_, int64_t M, int64_t N, int64_t K,
                          const double* A, int64_t lda, const double* B, int64_t ldb,
                          double* C, int64_t ldc) {
  const int ta = lower(ta_) == 'n' ? 'n' : 't';
  const int tb = lower(tb_) == 'n' ? 'n' : 't';
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j)
    for (int64_t i = 0; i < M; ++i) {
      double s = 0.0;
      for (int64_t k = 0; k < K; ++k) s += fabs(OPA(ta, A, lda, i, k)) * fabs(OPB(tb, B, ldb, k, j));
      C[i + j * ldc] = s;
    }
}

/* -------------------------------------------------------------------------
 * portBLAS CPU path restated: Gemm<no_local, standard, partial> with
 * Tile<4,4,4,4> (src/interface/blas3/backend/default.hpp:98-112;
 * src/operations/blas3/gemm_no_local_partial_vec.hpp:122-416).  Each work item
 * owns a 4 x 4 register tile; the accumulator starts at (beta/alpha) * C
 * (:129, :307-347; zero when beta == 0, C not read) and takes one
 * multiply-add per k (sycl::mad, src/operations/blas3/gemm_common.hpp:45-59);
 * the result stored is alpha * acc.  FP32 throughout.
 * ------------------------------------------------------------------------- */
void orc_pb_sgemm(char ta_, char tb_, int64_t M, int64_t N, int64_t K, float alpha,
                  const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                  float* C, int64_t ldc) {
  const int ta = lower(ta_) == 'n' ? 'n' : 't';
  const int tb = lower(tb_) == 'n' ? 'n' : 't';
  const int64_t tiles_m = (M + 3) / 4, tiles_n = (N + 3) / 4;
#pragma omp parallel for schedule(static) collapse(2)
  for (int64_t tj = 0; tj < tiles_n; ++tj)
    for (int64_t ti = 0; ti < tiles_m; ++ti) {
      float acc[4][4];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
          const int64_t i = ti * 4 + r, j = tj * 4 + c;
          acc[r][c] = (beta != 0.f && i < M && j < N) ? (beta / alpha) * C[i + j * ldc] : 0.f;
        }
      for (int64_t k = 0; k < K; ++k) {
        float a[4], b[4];
        for (int r = 0; r < 4; ++r) {
          const int64_t i = ti * 4 + r;
          a[r] = i < M ? OPA(ta, A, lda, i, k) : 0.f;
        }
        for (int c = 0; c < 4; ++c) {
          const int64_t j = tj * 4 + c;
          b[c] = j < N ? OPB(tb, B, ldb, k, j) : 0.f;
        }
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) acc[r][c] = fmaf(a[r], b[c], acc[r][c]);
      }
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
          const int64_t i = ti * 4 + r, j = tj * 4 + c;
          if (i < M && j < N) C[i + j * ldc] = alpha * acc[r][c];
        }
    }
}

/* -------------------------------------------------------------------------
 * GEMV reference in FP64: y = alpha * op(A) x + beta * y, trans 'n' vs any
 * other character (src/interface/blas2_interface.hpp:1184-1191); beta == 0
 * does not read y (:206-212).  Negative increments walk backwards
 * (src/views/view_sycl.hpp:95-100).
 * ------------------------------------------------------------------------- */
static int64_t vidx(int64_t i, int64_t n, int64_t inc) {
  return inc >= 0 ? i * inc : (n - 1 - i) * (-inc);
}

void orc_gemv_f64(char trans, int64_t M, int64_t N, double alpha, const double* A, int64_t lda,
                  const double* x, int64_t incx, double beta, double* y, int64_t incy) {
  const int tr = lower(trans) != 'n';
  const int64_t ylen = tr ? N : M, xlen = tr ? M : N;
#pragma omp parallel for schedule(static)
  for (int64_t o = 0; o < ylen; ++o) {
    double s = 0.0;
    for (int64_t q = 0; q < xlen; ++q) {
      const double a = tr ? A[q + o * lda] : A[o + q * lda];
      s += a * x[vidx(q, xlen, incx)];
    }
    double* yo = y + vidx(o, ylen, incy);
    *yo = (beta != 0.0) ? alpha * s + beta * *yo : alpha * s;
  }
}

/* portBLAS CPU GEMV restated in FP32: partial dot products per work item then
 * the column sum and the epilogue beta*y + alpha*sum
 * (src/interface/blas2_interface.hpp:68-217, DEFAULT config
 * src/interface/blas2/backend/default.hpp:32-49: one row per work item, no
 * column chunking, so the sum runs over the whole contracted dimension). */
void orc_pb_sgemv(char trans, int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                  const float* x, int64_t incx, float beta, float* y, int64_t incy) {
  const int tr = lower(trans) != 'n';
  const int64_t ylen = tr ? N : M, xlen = tr ? M : N;
  if (!tr) {
    /* Column sweep (contiguous in A) with one accumulator per row. */
    float* acc = (float*)calloc((size_t)(M > 0 ? M : 1), sizeof(float));
    for (int64_t q = 0; q < xlen; ++q) {
      const float xq = x[vidx(q, xlen, incx)];
      const float* a = A + q * lda;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < M; ++i) acc[i] = fmaf(a[i], xq, acc[i]);
    }
    for (int64_t o = 0; o < ylen; ++o) {
      float* yo = y + vidx(o, ylen, incy);
      *yo = (beta != 0.f) ? beta * *yo + alpha * acc[o] : alpha * acc[o];
    }
    free(acc);
    return;
  }
#pragma omp parallel for schedule(static)
  for (int64_t o = 0; o < ylen; ++o) {
    float s = 0.f;
    const float* a = A + o * lda;
    for (int64_t q = 0; q < xlen; ++q) s = fmaf(a[q], x[vidx(q, xlen, incx)], s);
    float* yo = y + vidx(o, ylen, incy);
    *yo = (beta != 0.f) ? beta * *yo + alpha * s : alpha * s;
  }
}

/* -------------------------------------------------------------------------
 * BLAS-1 references (FP64 accumulation).  Results are what the reference's
 * asynchronous API leaves in *rs starting from `rs0`:
 *   dot:  rs0 + sum x_i y_i   (n == 0: rs0 untouched, blas1_interface.hpp:621)
 *   asum: rs0 + sum |x_i|
 *   nrm2: sqrt(rs0 + sum x_i^2) (blas1_interface.hpp:552-581, no scaling)
 * ------------------------------------------------------------------------- */
double orc_dot_f64(int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
                   double rs0) {
  if (n <= 0) return rs0;
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += x[vidx(i, n, incx)] * y[vidx(i, n, incy)];
  return rs0 + s;
}

double orc_asum_f64(int64_t n, const double* x, int64_t incx, double rs0) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += fabs(x[vidx(i, n, incx)]);
  return rs0 + s;
}

double orc_nrm2_f64(int64_t n, const double* x, int64_t incx, double rs0) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double v = x[vidx(i, n, incx)];
    s += v * v;
  }
  return sqrt(rs0 + s);
}

/* portBLAS CPU reduction schedule restated in FP32: 8 work groups x 16 work
 * items stride over the vector (src/interface/blas1/backend/default.hpp:
 * 32-147), each work item accumulates in the element type, the work group
 * sums its items and adds to *rs (WGAtomicReduction.hpp:61-90).  op: 0 dot,
 * 1 asum, 2 nrm2 (sqrt of the accumulated value at the end). */
float orc_pb_sreduce(int op, int64_t n, const float* x, int64_t incx, const float* y,
                     int64_t incy, float rs0) {
  if (op == 0 && n <= 0) return rs0;
  const int groups = 8, items = 16, total = groups * items;
  float part[128];
#pragma omp parallel for schedule(static)
  for (int t = 0; t < total; ++t) {
    float acc = 0.f;
    for (int64_t i = t; i < n; i += total) {
      const float a = x[vidx(i, n, incx)];
      if (op == 0) acc = fmaf(a, y[vidx(i, n, incy)], acc);
      else if (op == 1) acc += fabsf(a);
      else acc = fmaf(a, a, acc);
    }
    part[t] = acc;
  }
  float rs = rs0;
  for (int g = 0; g < groups; ++g) {
    float w = 0.f;
    for (int t = 0; t < items; ++t) w += part[g * items + t];
    rs += w;
  }
  return op == 2 ? sqrtf(rs) : rs;
}

/* -------------------------------------------------------------------------
 * iamax / iamin, restated exactly from IMaxOperator / IMinOperator
 * (src/operations/blas_operators.hpp:276-316) folded left to right from the
 * initial tuple (blas_constants.h:110-136, constant_pair<max, zero> for max,
 * <max, abs_max> for min).  r replaces the running l iff |l| < |r| (max) or
 * |l| > |r| (min), or |l| == |r| and l.index > r.index.  Comparisons with NaN
 * are false, so a NaN never replaces the running value.  incx < 0 or n < 0
 * returns 0 (blas1_interface.hpp:375-379); sentinel = index type's maximum.
 * ------------------------------------------------------------------------- */
int64_t orc_iamax_f64(int is_max, int64_t n, const double* x, int64_t incx, int64_t sentinel,
                      int elem_is_float) {
  if (incx < 0 || n < 0) return 0;
  int64_t li = sentinel;
  double lv = is_max ? 0.0 : (elem_is_float ? (double)FLT_MAX : DBL_MAX);
  for (int64_t i = 0; i < n; ++i) {
    const double rv = fabs(x[i * incx]);
    const int take = is_max ? (lv < rv || (lv == rv && li > i)) : (lv > rv || (lv == rv && li > i));
    if (take) { lv = rv; li = i; }
  }
  return li;
}

/* -------------------------------------------------------------------------
 * extension::_reduction reference (FP64 accumulation), column-major rows x
 * cols with leading dimension ld; dim 0 = inner (one value per column),
 * 1 = outer (one value per row).  op: 0 add, 1 product, 2 max, 3 min,
 * 4 abs-add, 5 mean (divide by the reduced length), following
 * src/operations/blas_operators.hpp:149-274 and the hand loop of
 * test/unittest/extension/reduction_test.cpp:155-180.
 * ------------------------------------------------------------------------- */
void orc_reduction_f64(int op, int dim, const double* in, int64_t ld, double* out, int64_t rows,
                       int64_t cols) {
  const int64_t outer_len = dim == 0 ? cols : rows;
  const int64_t red_len = dim == 0 ? rows : cols;
#pragma omp parallel for schedule(static)
  for (int64_t o = 0; o < outer_len; ++o) {
    double acc;
    switch (op) {
      case 1: acc = 1.0; break;
      case 2: acc = -DBL_MAX; break;
      case 3: acc = DBL_MAX; break;
      default: acc = 0.0;
    }
    for (int64_t q = 0; q < red_len; ++q) {
      const double v = dim == 0 ? in[q + o * ld] : in[o + q * ld];
      switch (op) {
        case 0: case 5: acc += v; break;
        case 1: acc *= v; break;
        case 2: acc = acc > v ? acc : v; break;
        case 3: acc = acc < v ? acc : v; break;
        case 4: acc += fabs(v); break;
      }
    }
    if (op == 5) acc /= (double)red_len;
    out[o] = acc;
  }
}

/* -------------------------------------------------------------------------
 * Seeded inputs shared by tests and benchmarks: splitmix64 (Steele, Lea,
 * Flood 2014) -> the top 24 bits -> uniform in [lo, hi).  Element i of the
 * stream with seed s is f(splitmix64 state s + (i+1) * golden gamma).
 * ------------------------------------------------------------------------- */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

void orc_uniform_f32(uint64_t seed, int64_t n, float lo, float hi, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t z = mix64(seed + (uint64_t)(i + 1) * 0x9e3779b97f4a7c15ULL);
    const float u = (float)(z >> 40) * (1.0f / 16777216.0f);
    out[i] = lo + (hi - lo) * u;
  }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* _axpy (src/interface/blas1_interface.hpp:46-60): y <- alpha * x + y for   */
/* each element of the views (view_sycl.hpp:95-100), evaluated with one      */
/* rounding (a fused multiply-add, fmaf / fma), as OpenBLAS's kernels and    */
/* contracting device compilers evaluate the expression.                    */
/* ------------------------------------------------------------------------ */
static int64_t orc_vidx(int64_t i, int64_t n, int64_t inc) {
  return inc >= 0 ? i * inc : (n - 1 - i) * (-inc);
}

void orc_saxpy(int64_t n, float alpha, const float* x, int64_t incx, float* y, int64_t incy) {
  for (int64_t i = 0; i < n; ++i) {
    const int64_t iy = orc_vidx(i, n, incy);
    y[iy] = fmaf(alpha, x[orc_vidx(i, n, incx)], y[iy]);
  }
}

void orc_daxpy(int64_t n, double alpha, const double* x, int64_t incx, double* y, int64_t incy) {
  for (int64_t i = 0; i < n; ++i) {
    const int64_t iy = orc_vidx(i, n, incy);
    y[iy] = fma(alpha, x[orc_vidx(i, n, incx)], y[iy]);
  }
}

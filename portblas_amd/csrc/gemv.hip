// GEMV: y <- alpha * op(A) * x + beta * y (column-major A, M x N, lda).
//
// Reference behaviour (portBLAS): blas::_gemv dispatches on trans == 'n' vs
// anything else (src/interface/blas2_interface.hpp:1163-1192); _gemv_impl
// (:68-217) writes partial dot products of each column chunk into a temporary
// ld x WGs_per_C matrix, then SumMatrixColumns + the epilogue
// y = beta*y + alpha*sum (beta == 0 -> y not read, :185-212).
//
// MI355X design (HBM-bound; A is read once with non-temporal 16-byte loads):
//  'n'  a workgroup of 16 waves owns 256 consecutive rows; every lane holds
//       4 consecutive rows (one 16-byte load per column for float) and the
//       16 waves split the workgroup's column chunk.  Partial sums are
//       folded across waves in LDS; with several column chunks, each writes
//       its partials to the workspace and a fold kernel sums them in a fixed
//       order and applies the epilogue.
//  't'  a workgroup owns a group of columns; each lane keeps its 32 elements
//       of a 2048-row chunk of x in registers, each wave walks its columns,
//       one column chunk (8 KB, 8 16-byte loads per lane) at a time, and
//       reduces across the wave with shuffles.  With enough columns to fill
//       the GPU, a workgroup walks every row chunk itself and adds each
//       column's chunk sums in chunk order (one launch); otherwise the row
//       chunks are spread over the grid and folded like the 'n' column
//       chunks, in the same order.
#include "pb_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kTQ = 8;              // 't': 16-byte row groups per lane
constexpr int kTRows = 64 * 4 * kTQ;  // 't': rows per chunk (2048)

template <typename T> struct Four { T v[4]; };

// Four consecutive elements; STREAM selects a non-temporal load (A is read
// once; x is shared by every workgroup and stays cacheable).
template <typename T, bool STREAM = true>
__device__ __forceinline__ Four<T> load4(const T* p) {
  if constexpr (sizeof(T) == 4) {
    const auto* q = reinterpret_cast<const pbdev::f32x4_t*>(p);
    const pbdev::f32x4_t a = STREAM ? pbdev::ld_stream(q) : *q;
    return {{a.x, a.y, a.z, a.w}};
  } else {
    const auto* q = reinterpret_cast<const pbdev::f64x2_t*>(p);
    const pbdev::f64x2_t a = STREAM ? pbdev::ld_stream(q) : q[0];
    const pbdev::f64x2_t b = STREAM ? pbdev::ld_stream(q + 1) : q[1];
    return {{a.x, a.y, b.x, b.y}};
  }
}

template <typename T>
__device__ __forceinline__ void store_y(T* y, int64_t idx, T alpha, T beta, T sum) {
  y[idx] = (beta != T(0)) ? beta * y[idx] + alpha * sum : alpha * sum;
}

// ---------------------------------------------------------------------------
// 'n': grid (row_tiles, col_chunks), kNBlock threads (kNWaves waves).  A
// workgroup owns 256 Q consecutive rows; every lane holds 4 of every 256
// (rows r0 + 256 q + 4 lane + e: one 16-byte load per column for float, VEC
// = A 16-byte aligned with lda a multiple of 4 elements; scalar: r0 + 256 q
// + 64 e + lane).  The waves interleave the chunk's columns, kNU columns per step, so
// each wave keeps kNU 1 KB column segments in flight.  The waves' sums are
// folded in LDS in a fixed order.  Sixteen waves per workgroup keep the
// number of column chunks (and so the fold traffic) small.
// ---------------------------------------------------------------------------
constexpr int kNWaves = 16;
constexpr int kNBlock = 64 * kNWaves;
constexpr int kNU = 8;

// Q: 256-row groups per workgroup (each lane holds 4 Q rows; with Q > 1 a
// wave reads Q KB of every column).
template <typename T, bool VEC, int Q>
__global__ __launch_bounds__(kNBlock) void gemv_n_kernel(
    int64_t M, int64_t N, T alpha, const T* __restrict__ A, int64_t lda,
    const T* __restrict__ x, int64_t incx, T beta, T* y, int64_t incy,
    int64_t cols_per_chunk, T* __restrict__ partials) {
  constexpr int ROWS = 256 * Q;
  constexpr int U = kNU / Q;  // columns per step: U x Q 16-byte loads in flight
  __shared__ T red[kNWaves][ROWS];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t r0 = int64_t(blockIdx.x) * ROWS;
  const int64_t c0 = int64_t(blockIdx.y) * cols_per_chunk;
  int64_t c1 = c0 + cols_per_chunk;
  if (c1 > N) c1 = N;
  T acc[Q][4];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[q][e] = T(0);
  if (VEC && r0 + ROWS <= M) {
    const T* Ap = A + r0 + 4 * lane;
    int64_t j = c0 + wave;
    for (; j + (U - 1) * kNWaves < c1; j += U * kNWaves) {
      Four<T> a[U][Q];
      T xs[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < Q; ++q) a[u][q] = load4(Ap + (j + u * kNWaves) * lda + 256 * q);
#pragma unroll
      for (int u = 0; u < U; ++u) xs[u] = x[pbdev::vidx(j + u * kNWaves, N, incx)];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
          for (int e = 0; e < 4; ++e) acc[q][e] += a[u][q].v[e] * xs[u];
    }
    for (; j < c1; j += kNWaves) {
      const T xj = x[pbdev::vidx(j, N, incx)];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const Four<T> a = load4(Ap + j * lda + 256 * q);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[q][e] += a.v[e] * xj;
      }
    }
  } else {
    for (int64_t j = c0 + wave; j < c1; j += kNWaves) {
      const T xj = x[pbdev::vidx(j, N, incx)];
      const T* col = A + j * lda;
#pragma unroll
      for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t r = VEC ? r0 + 256 * q + 4 * lane + e : r0 + 256 * q + 64 * e + lane;
          if (r < M) acc[q][e] += pbdev::ld_stream(col + r) * xj;
        }
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) red[wave][256 * q + (VEC ? 4 * lane + e : 64 * e + lane)] = acc[q][e];
  __syncthreads();
  for (int rr = threadIdx.x; rr < ROWS; rr += kNBlock) {
    T s = red[0][rr];
#pragma unroll
    for (int w = 1; w < kNWaves; ++w) s += red[w][rr];
    const int64_t r = r0 + rr;
    if (r >= M) continue;
    if (gridDim.y == 1) store_y(y, pbdev::vidx(r, M, incy), alpha, beta, s);
    else partials[int64_t(blockIdx.y) * M + r] = s;
  }
}

// Folds the per-chunk partial sums in a fixed order and applies the epilogue
// (a separate launch: stream order publishes the partials without
// device-scope fences).
template <typename T>
__global__ __launch_bounds__(kBlock) void gemv_fold_kernel(int64_t len, int chunks,
                                                           const T* __restrict__ partials,
                                                           T alpha, T beta, T* y, int64_t incy) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= len) return;
  constexpr int U = 8;  // loads in flight before the (fixed-order) sums
  T t = T(0);
  for (int c0 = 0; c0 < chunks; c0 += U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = c0 + u < chunks ? partials[int64_t(c0 + u) * len + i] : T(0);
#pragma unroll
    for (int u = 0; u < U; ++u) t += v[u];
  }
  store_y(y, pbdev::vidx(i, len, incy), alpha, beta, t);
}

// ---------------------------------------------------------------------------
// 't': grid (col_groups, row_chunks).  y has length N, x has length M.
// VEC: A aligned as above and x contiguous + 16-byte aligned.
// ---------------------------------------------------------------------------
template <typename T, bool VEC>
__global__ __launch_bounds__(kBlock) void gemv_t_kernel(
    int64_t M, int64_t N, T alpha, const T* __restrict__ A, int64_t lda,
    const T* __restrict__ x, int64_t incx, T beta, T* y, int64_t incy,
    int64_t cols_per_group, T* __restrict__ partials) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t r0 = int64_t(blockIdx.y) * kTRows;
  const int64_t j0 = int64_t(blockIdx.x) * cols_per_group;
  int64_t j1 = j0 + cols_per_group;
  if (j1 > N) j1 = N;
  const bool full = r0 + kTRows <= M;
  // x slice in registers: element (q, e) is row r0 + 256*q + 4*lane + e (VEC)
  // or r0 + 64*(4*q + e) + lane (scalar).
  T xr[kTQ][4];
#pragma unroll
  for (int q = 0; q < kTQ; ++q) {
    if (VEC && full) {
      Four<T> v = load4<T, false>(x + r0 + 256 * q + 4 * lane);
#pragma unroll
      for (int e = 0; e < 4; ++e) xr[q][e] = v.v[e];
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = VEC ? r0 + 256 * q + 4 * lane + e : r0 + 64 * (4 * q + e) + lane;
        xr[q][e] = (r < M) ? x[pbdev::vidx(r, M, incx)] : T(0);
      }
    }
  }
  auto emit = [&](int64_t j, T s) {
    s = pbdev::wave_sum(s);
    if (lane == 0) {
      if (gridDim.y == 1) store_y(y, pbdev::vidx(j, N, incy), alpha, beta, s);
      else partials[int64_t(blockIdx.y) * N + j] = s;
    }
  };
  int64_t j = j0 + wave;
  for (; j < j1; j += kWaves) {
    const T* col = A + j * lda;
    T acc = T(0);
    if (VEC && full) {
      Four<T> a[kTQ];
#pragma unroll
      for (int q = 0; q < kTQ; ++q) a[q] = load4(col + r0 + 256 * q + 4 * lane);
#pragma unroll
      for (int q = 0; q < kTQ; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc += a[q].v[e] * xr[q][e];
    } else {
#pragma unroll
      for (int q = 0; q < kTQ; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t r = VEC ? r0 + 256 * q + 4 * lane + e : r0 + 64 * (4 * q + e) + lane;
          if (r < M) acc += pbdev::ld_stream(col + r) * xr[q][e];
        }
    }
    emit(j, acc);
  }
}

// 't' without row chunks in the grid: grid (col_groups), each workgroup walks
// every 2048-row chunk of its columns (at most kTCols per wave), reloading
// the chunk's x slice (cached) and adding each column's chunk sum to a
// running total in chunk order: the same sums, in the same order, as the
// partials and the fold launch of gemv_t_kernel, without the partials and the
// second launch.
constexpr int kTCols = 4;  // columns a wave keeps a running sum for

template <typename T, bool VEC>
__global__ __launch_bounds__(kBlock) void gemv_t_full_kernel(
    int64_t M, int64_t N, T alpha, const T* __restrict__ A, int64_t lda,
    const T* __restrict__ x, int64_t incx, T beta, T* y, int64_t incy,
    int64_t cols_per_group) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t j0 = int64_t(blockIdx.x) * cols_per_group;
  int64_t j1 = j0 + cols_per_group;
  if (j1 > N) j1 = N;
  T colsum[kTCols];
#pragma unroll
  for (int c = 0; c < kTCols; ++c) colsum[c] = T(0);
  for (int64_t r0 = 0; r0 < M; r0 += kTRows) {
    const bool full = r0 + kTRows <= M;
    T xr[kTQ][4];
#pragma unroll
    for (int q = 0; q < kTQ; ++q) {
      if (VEC && full) {
        Four<T> v = load4<T, false>(x + r0 + 256 * q + 4 * lane);
#pragma unroll
        for (int e = 0; e < 4; ++e) xr[q][e] = v.v[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t r = VEC ? r0 + 256 * q + 4 * lane + e : r0 + 64 * (4 * q + e) + lane;
          xr[q][e] = (r < M) ? x[pbdev::vidx(r, M, incx)] : T(0);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kTCols; ++c) {
      const int64_t j = j0 + wave + int64_t(kWaves) * c;
      if (j >= j1) break;
      const T* col = A + j * lda;
      T acc = T(0);
      if (VEC && full) {
        Four<T> a[kTQ];
#pragma unroll
        for (int q = 0; q < kTQ; ++q) a[q] = load4(col + r0 + 256 * q + 4 * lane);
#pragma unroll
        for (int q = 0; q < kTQ; ++q)
#pragma unroll
          for (int e = 0; e < 4; ++e) acc += a[q].v[e] * xr[q][e];
      } else {
#pragma unroll
        for (int q = 0; q < kTQ; ++q)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t r = VEC ? r0 + 256 * q + 4 * lane + e : r0 + 64 * (4 * q + e) + lane;
            if (r < M) acc += pbdev::ld_stream(col + r) * xr[q][e];
          }
      }
      colsum[c] += pbdev::wave_sum(acc);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < kTCols; ++c) {
      const int64_t j = j0 + wave + int64_t(kWaves) * c;
      if (j < j1) store_y(y, pbdev::vidx(j, N, incy), alpha, beta, colsum[c]);
    }
  }
}

// y <- beta * y for the degenerate shapes (an empty contraction).
template <typename T>
__global__ void scale_y_kernel(int64_t n, T beta, T* y, int64_t incy) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    const int64_t k = pbdev::vidx(i, n, incy);
    y[k] = (beta != T(0)) ? beta * y[k] : T(0);
  }
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <typename T>
pb_status_t gemv_launch(pb_handle_t h, char trans, int64_t M, int64_t N, T alpha,
                        const T* A, int64_t lda, const T* x, int64_t incx, T beta,
                        T* y, int64_t incy) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  const char l = (trans >= 'A' && trans <= 'Z') ? char(trans - 'A' + 'a') : trans;
  // portBLAS: 'n' -> Normal, any other character -> Transposed
  // (blas2_interface.hpp:1184-1191).
  const bool tr = l != 'n';
  PB_REQUIRE(M >= 0 && N >= 0, "negative dimension (M=%lld, N=%lld)", (long long)M,
             (long long)N);
  const int64_t ylen = tr ? N : M;
  const int64_t xlen = tr ? M : N;
  if (ylen == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(y != nullptr, "null y");
  if (xlen == 0) {
    const int g = int(pb::ceil_div(ylen, 256));
    hipLaunchKernelGGL(scale_y_kernel<T>, dim3(g), dim3(256), 0, h->stream, ylen, beta,
                       y, incy);
    PB_CHECK_LAUNCH();
    return PB_STATUS_SUCCESS;
  }
  PB_REQUIRE(A != nullptr && x != nullptr, "null A or x");
  PB_REQUIRE(lda >= (M > 1 ? M : 1), "lda (%lld) < max(1, M)", (long long)lda);
  const bool avec = aligned16(A) && (lda * int64_t(sizeof(T))) % 16 == 0;
  const int target = 4 * h->num_cus;
  if (!tr) {
    // One 16-wave workgroup per CU; every chunk keeps >= 8 columns per wave.
    // 256-row tiles (1 KB of every column per wave) measured faster than
    // 512- or 1024-row ones at 8192 x 8192, and 16-wave workgroups faster
    // than twice as many 8-wave ones (half the partials to fold).
    const int q = 1;
    const int64_t rows = 256 * q;
    const int64_t tiles = pb::ceil_div(M, rows);
    PB_REQUIRE(tiles <= 0x7fffffffLL, "M too large");
    int64_t chunks = pb::ceil_div(h->num_cus, tiles);
    const int64_t max_chunks = pb::ceil_div(N, 8 * kNWaves);
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    if (chunks > 65535) chunks = 65535;
    const int64_t cpc = pb::ceil_div(N, chunks);
    chunks = pb::ceil_div(N, cpc);
    void* ws = nullptr;
    if (chunks > 1) PB_TRY(pb::workspace(h, sizeof(T) * size_t(chunks) * size_t(M), &ws));
    T* part = static_cast<T*>(ws);
    dim3 grid{unsigned(tiles), unsigned(chunks)};
#define PB_GEMV_N(V, Q)                                                                      \
  hipLaunchKernelGGL((gemv_n_kernel<T, V, Q>), grid, dim3(kNBlock), 0, h->stream, M, N, alpha, A, \
                     lda, x, incx, beta, y, incy, cpc, part)
    if (avec && q == 2) PB_GEMV_N(true, 2);
    else if (avec) PB_GEMV_N(true, 1);
    else if (q == 2) PB_GEMV_N(false, 2);
    else PB_GEMV_N(false, 1);
#undef PB_GEMV_N
    if (chunks > 1)
      hipLaunchKernelGGL(gemv_fold_kernel<T>, dim3(unsigned(pb::ceil_div(M, kBlock))),
                         dim3(kBlock), 0, h->stream, M, int(chunks), part, alpha, beta, y, incy);
  } else {
    const bool vec = avec && incx == 1 && aligned16(x);
    const int64_t chunks = pb::ceil_div(M, kTRows);
    // Enough columns for `target` workgroups of at most kTCols columns per
    // wave: no row chunks in the grid, no partials, no fold launch.  (A
    // workgroup then walks every row chunk; beyond 32 chunks, or with too few
    // columns to fill the GPU, the chunked grid below spreads the rows.)
    // Twice `target` workgroups (8 per CU; one column per wave at 8192 x
    // 8192): 44.5-45.9 us for the whole call, against 48.0-49.8 us with 4 per
    // CU and 47.7-48.9 us for the chunked grid with its fold launch.
    if (chunks > 1 && chunks <= 32 && N >= int64_t(target) * kWaves) {
      int64_t cpg = pb::ceil_div(N, 2 * int64_t(target));
      cpg = pb::ceil_div(cpg, kWaves) * kWaves;
      if (cpg > int64_t(kTCols) * kWaves) cpg = int64_t(kTCols) * kWaves;
      const int64_t groups = pb::ceil_div(N, cpg);
      PB_REQUIRE(groups <= 0x7fffffffLL, "N too large");
      if (vec)
        hipLaunchKernelGGL((gemv_t_full_kernel<T, true>), dim3(unsigned(groups)), dim3(kBlock), 0, h->stream,
                           M, N, alpha, A, lda, x, incx, beta, y, incy, cpg);
      else
        hipLaunchKernelGGL((gemv_t_full_kernel<T, false>), dim3(unsigned(groups)), dim3(kBlock), 0, h->stream,
                           M, N, alpha, A, lda, x, incx, beta, y, incy, cpg);
      PB_CHECK_LAUNCH();
      return PB_STATUS_SUCCESS;
    }
    PB_REQUIRE(chunks <= 65535, "M too large for the transposed GEMV (max %d rows)",
               65535 * kTRows);
    int64_t groups = pb::ceil_div(target, chunks);
    const int64_t max_groups = pb::ceil_div(N, kWaves);
    if (groups > max_groups) groups = max_groups;
    if (groups < 1) groups = 1;
    const int64_t cpg = pb::ceil_div(N, groups);
    groups = pb::ceil_div(N, cpg);
    void* ws = nullptr;
    if (chunks > 1) PB_TRY(pb::workspace(h, sizeof(T) * size_t(chunks) * size_t(N), &ws));
    T* part = static_cast<T*>(ws);
    dim3 grid{unsigned(groups), unsigned(chunks)};
    if (vec)
      hipLaunchKernelGGL((gemv_t_kernel<T, true>), grid, dim3(kBlock), 0, h->stream, M, N,
                         alpha, A, lda, x, incx, beta, y, incy, cpg, part);
    else
      hipLaunchKernelGGL((gemv_t_kernel<T, false>), grid, dim3(kBlock), 0, h->stream, M, N,
                         alpha, A, lda, x, incx, beta, y, incy, cpg, part);
    if (chunks > 1)
      hipLaunchKernelGGL(gemv_fold_kernel<T>, dim3(unsigned(pb::ceil_div(N, kBlock))),
                         dim3(kBlock), 0, h->stream, N, int(chunks), part, alpha, beta, y, incy);
  }
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

pb_status_t pb_sgemv(pb_handle_t h, char trans, int64_t m, int64_t n, float alpha,
                     const float* a, int64_t lda, const float* x, int64_t incx,
                     float beta, float* y, int64_t incy) {
  return gemv_launch<float>(h, trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

pb_status_t pb_dgemv(pb_handle_t h, char trans, int64_t m, int64_t n, double alpha,
                     const double* a, int64_t lda, const double* x, int64_t incx,
                     double beta, double* y, int64_t incy) {
  return gemv_launch<double>(h, trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

}  // extern "C"

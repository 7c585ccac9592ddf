// GEMV: y <- alpha * op(A) * x + beta * y (column-major A, M x N, lda).
//
// Reference behaviour (portBLAS): blas::_gemv dispatches on trans == 'n' vs
// anything else (src/interface/blas2_interface.hpp:1163-1192); _gemv_impl
// (:68-217) writes partial dot products of each column chunk into a temporary
// ld x WGs_per_C matrix, then SumMatrixColumns + the epilogue
// y = beta*y + alpha*sum (beta == 0 -> y not read, :185-212).
//
// MI355X design (one launch, HBM-bound):
//  'n'  a workgroup of 4 waves owns 256 consecutive rows; every lane holds 4
//       consecutive rows (one 16-byte load per column for float) and the 4
//       waves split the workgroup's column chunk.  Partial sums are folded
//       across waves in LDS; column chunks write partials to the workspace and
//       the last chunk to finish for a row tile (device-scope counter) folds
//       them in a fixed order and applies the epilogue.
//  't'  a workgroup owns a 2048-row chunk and a group of columns; each lane
//       keeps its 32 elements of x in registers for the whole workgroup, each
//       wave walks its columns with 16-byte loads and reduces across the wave
//       with shuffles.  Row chunks are folded like the 'n' column chunks.
#include "pb_internal.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kRowsPerTile = 256;   // 'n': rows per workgroup (64 lanes x 4)
constexpr int kTQ = 8;              // 't': 16-byte row groups per lane
constexpr int kTRows = 64 * 4 * kTQ;  // 't': rows per chunk (2048)

template <typename T> struct Four { T v[4]; };

template <typename T>
__device__ __forceinline__ Four<T> load4(const T* p);
template <>
__device__ __forceinline__ Four<float> load4<float>(const float* p) {
  float4 a = *reinterpret_cast<const float4*>(p);
  return {{a.x, a.y, a.z, a.w}};
}
template <>
__device__ __forceinline__ Four<double> load4<double>(const double* p) {
  double2 a = reinterpret_cast<const double2*>(p)[0];
  double2 b = reinterpret_cast<const double2*>(p)[1];
  return {{a.x, a.y, b.x, b.y}};
}

__device__ __forceinline__ bool arrive_last(unsigned int* counter, unsigned int total) {
  __shared__ bool last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned int prev = atomicAdd(counter, 1u);
    last = (prev == total - 1);
    if (last) *counter = 0u;  // every other arrival for this tile is done
  }
  __syncthreads();
  if (last) __threadfence();
  return last;
}

template <typename T>
__device__ __forceinline__ void store_y(T* y, int64_t idx, T alpha, T beta, T sum) {
  y[idx] = (beta != T(0)) ? beta * y[idx] + alpha * sum : alpha * sum;
}

// ---------------------------------------------------------------------------
// 'n': grid (row_tiles, col_chunks).  VEC: A 16-byte aligned with lda a
// multiple of 4 elements, so lane rows [4l, 4l+4) load as one vector.
// ---------------------------------------------------------------------------
template <typename T, bool VEC>
__global__ __launch_bounds__(kBlock) void gemv_n_kernel(
    int64_t M, int64_t N, T alpha, const T* __restrict__ A, int64_t lda,
    const T* __restrict__ x, int64_t incx, T beta, T* y, int64_t incy,
    int64_t cols_per_chunk, T* __restrict__ partials, unsigned int* counters) {
  __shared__ T red[kWaves][kRowsPerTile + 4];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t row0 = int64_t(blockIdx.x) * kRowsPerTile;
  const int64_t c0 = int64_t(blockIdx.y) * cols_per_chunk;
  int64_t c1 = c0 + cols_per_chunk;
  if (c1 > N) c1 = N;
  T acc[4] = {T(0), T(0), T(0), T(0)};
  // Row r of this lane: VEC -> row0 + 4*lane + q; scalar -> row0 + lane + 64*q.
  const bool full = row0 + kRowsPerTile <= M;
  if (VEC && full) {
    const T* Ap = A + row0 + 4 * lane;
    int64_t j = c0 + wave;
    // 4 columns per step, each wave strides by kWaves columns.
    for (; j + 3 * kWaves < c1; j += 4 * kWaves) {
      Four<T> a0 = load4(Ap + (j + 0 * kWaves) * lda);
      Four<T> a1 = load4(Ap + (j + 1 * kWaves) * lda);
      Four<T> a2 = load4(Ap + (j + 2 * kWaves) * lda);
      Four<T> a3 = load4(Ap + (j + 3 * kWaves) * lda);
      const T x0 = x[pbdev::vidx(j + 0 * kWaves, N, incx)];
      const T x1 = x[pbdev::vidx(j + 1 * kWaves, N, incx)];
      const T x2 = x[pbdev::vidx(j + 2 * kWaves, N, incx)];
      const T x3 = x[pbdev::vidx(j + 3 * kWaves, N, incx)];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[q] += a0.v[q] * x0;
        acc[q] += a1.v[q] * x1;
        acc[q] += a2.v[q] * x2;
        acc[q] += a3.v[q] * x3;
      }
    }
    for (; j < c1; j += kWaves) {
      Four<T> a0 = load4(Ap + j * lda);
      const T x0 = x[pbdev::vidx(j, N, incx)];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] += a0.v[q] * x0;
    }
  } else {
    for (int64_t j = c0 + wave; j < c1; j += kWaves) {
      const T xj = x[pbdev::vidx(j, N, incx)];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t r = VEC ? row0 + 4 * lane + q : row0 + lane + 64 * q;
        if (r < M) acc[q] += A[r + j * lda] * xj;
      }
    }
  }
  // Fold the 4 waves in LDS (fixed order).
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int rr = VEC ? 4 * lane + q : lane + 64 * q;
    red[wave][rr] = acc[q];
  }
  __syncthreads();
  const int rr = threadIdx.x;  // one row of the tile per thread
  T s = red[0][rr];
#pragma unroll
  for (int w = 1; w < kWaves; ++w) s += red[w][rr];
  const int64_t r = row0 + rr;
  if (gridDim.y == 1) {
    if (r < M) store_y(y, pbdev::vidx(r, M, incy), alpha, beta, s);
    return;
  }
  if (r < M) partials[int64_t(blockIdx.y) * M + r] = s;
  if (!arrive_last(counters + blockIdx.x, gridDim.y)) return;
  if (r < M) {
    T t = T(0);
    for (unsigned c = 0; c < gridDim.y; ++c) t += partials[int64_t(c) * M + r];
    store_y(y, pbdev::vidx(r, M, incy), alpha, beta, t);
  }
}

// ---------------------------------------------------------------------------
// 't': grid (col_groups, row_chunks).  y has length N, x has length M.
// VEC: A aligned as above and x contiguous + 16-byte aligned.
// ---------------------------------------------------------------------------
template <typename T, bool VEC>
__global__ __launch_bounds__(kBlock) void gemv_t_kernel(
    int64_t M, int64_t N, T alpha, const T* __restrict__ A, int64_t lda,
    const T* __restrict__ x, int64_t incx, T beta, T* y, int64_t incy,
    int64_t cols_per_group, T* __restrict__ partials, unsigned int* counters) {
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
      Four<T> v = load4(x + r0 + 256 * q + 4 * lane);
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
  for (int64_t j = j0 + wave; j < j1; j += kWaves) {
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
          if (r < M) acc += col[r] * xr[q][e];
        }
    }
    acc = pbdev::wave_sum(acc);
    if (lane == 0) {
      if (gridDim.y == 1) store_y(y, pbdev::vidx(j, N, incy), alpha, beta, acc);
      else partials[int64_t(blockIdx.y) * N + j] = acc;
    }
  }
  if (gridDim.y == 1) return;
  if (!arrive_last(counters + blockIdx.x, gridDim.y)) return;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += kBlock) {
    T t = T(0);
    for (unsigned c = 0; c < gridDim.y; ++c) t += partials[int64_t(c) * N + j];
    store_y(y, pbdev::vidx(j, N, incy), alpha, beta, t);
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
  unsigned int* ctr = pb::counters(h, 0);
  if (!tr) {
    const int64_t tiles = pb::ceil_div(M, kRowsPerTile);
    PB_REQUIRE(tiles <= 65535 * 4096LL, "M too large");
    int64_t chunks = pb::ceil_div(target, tiles);
    const int64_t max_chunks = pb::ceil_div(N, 4 * kWaves);  // >= 16 columns per chunk
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    if (tiles > pb::kCounterSlots) chunks = 1;  // one pass, no cross-chunk fold
    if (chunks > 65535) chunks = 65535;
    const int64_t cpc = pb::ceil_div(N, chunks);
    chunks = pb::ceil_div(N, cpc);
    void* ws = nullptr;
    if (chunks > 1) PB_TRY(pb::workspace(h, sizeof(T) * size_t(chunks) * size_t(M), &ws));
    dim3 grid{unsigned(tiles), unsigned(chunks)};
    if (avec)
      hipLaunchKernelGGL((gemv_n_kernel<T, true>), grid, dim3(kBlock), 0, h->stream, M, N,
                         alpha, A, lda, x, incx, beta, y, incy, cpc, static_cast<T*>(ws), ctr);
    else
      hipLaunchKernelGGL((gemv_n_kernel<T, false>), grid, dim3(kBlock), 0, h->stream, M, N,
                         alpha, A, lda, x, incx, beta, y, incy, cpc, static_cast<T*>(ws), ctr);
  } else {
    const bool vec = avec && incx == 1 && aligned16(x);
    const int64_t chunks = pb::ceil_div(M, kTRows);
    PB_REQUIRE(chunks <= 65535, "M too large for the transposed GEMV (max %d rows)",
               65535 * kTRows);
    int64_t groups = pb::ceil_div(target, chunks);
    const int64_t max_groups = pb::ceil_div(N, kWaves);
    if (groups > max_groups) groups = max_groups;
    if (groups < 1) groups = 1;
    if (chunks > 1 && groups > pb::kCounterSlots) groups = pb::kCounterSlots;
    const int64_t cpg = pb::ceil_div(N, groups);
    groups = pb::ceil_div(N, cpg);
    void* ws = nullptr;
    if (chunks > 1) PB_TRY(pb::workspace(h, sizeof(T) * size_t(chunks) * size_t(N), &ws));
    dim3 grid{unsigned(groups), unsigned(chunks)};
    if (vec)
      hipLaunchKernelGGL((gemv_t_kernel<T, true>), grid, dim3(kBlock), 0, h->stream, M, N,
                         alpha, A, lda, x, incx, beta, y, incy, cpg, static_cast<T*>(ws), ctr);
    else
      hipLaunchKernelGGL((gemv_t_kernel<T, false>), grid, dim3(kBlock), 0, h->stream, M, N,
                         alpha, A, lda, x, incx, beta, y, incy, cpg, static_cast<T*>(ws), ctr);
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

// blas::extension::_reduction<Op, T>: partial reduction of a column-major
// rows x cols matrix (leading dimension ld) along one dimension.
//
// Reference behaviour (portBLAS): include/interface/extension_interface.h:
// 453-462; src/interface/extension_interface.hpp:344-410 (one or two steps;
// Mean's second step is Add, :46-53); src/operations/extension/
// reduction.hpp:35-230 (inner -> one value per column, outer -> one value per
// row; Mean divides by the reduced length, :188); operators
// src/operations/blas_operators.hpp:149-274; output is overwritten.
//
// Note on ld: the reference addresses element (r, c) as r + (ld / rows) * rows * c
// (integer ld_mul_, reduction.hpp:41,135-140), which equals the BLAS address
// r + ld * c whenever rows divides ld (every case its tests use).  This
// implementation uses the BLAS address for every ld >= rows.
//
// MI355X design (HBM-bound, non-temporal loads):
//   INNER: each wave reduces whole columns (lanes along the contiguous rows,
//          4 loads in flight per lane) and folds across the wave with shuffles;
//          very tall matrices split the rows into chunks whose partials a
//          fold kernel combines in a fixed order.
//   OUTER: each workgroup owns 256 rows (one per lane, 4 waves split the
//          columns), folds the waves in LDS; column chunks are combined by
//          the fold kernel.
#include "pb_internal.hpp"

#include <limits>

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

template <typename T, int OP>
struct Op;

template <typename T> struct Op<T, PB_OP_ADD> {
  static __device__ T init() { return T(0); }
  static __device__ T apply(T a, T x) { return a + x; }
  static __device__ T combine(T a, T b) { return a + b; }
  static __device__ T finish(T a, int64_t) { return a; }
};
template <typename T> struct Op<T, PB_OP_MEAN> {
  static __device__ T init() { return T(0); }
  static __device__ T apply(T a, T x) { return a + x; }
  static __device__ T combine(T a, T b) { return a + b; }
  static __device__ T finish(T a, int64_t n) { return a / T(n); }
};
template <typename T> struct Op<T, PB_OP_PRODUCT> {
  static __device__ T init() { return T(1); }
  static __device__ T apply(T a, T x) { return a * x; }
  static __device__ T combine(T a, T b) { return a * b; }
  static __device__ T finish(T a, int64_t) { return a; }
};
// MaxOperator/MinOperator: (l > r) ? l : r  /  (l < r) ? l : r.
template <typename T> struct Op<T, PB_OP_MAX> {
  static __device__ T init() { return std::numeric_limits<T>::lowest(); }
  static __device__ T apply(T a, T x) { return (a > x) ? a : x; }
  static __device__ T combine(T a, T b) { return (a > b) ? a : b; }
  static __device__ T finish(T a, int64_t) { return a; }
};
template <typename T> struct Op<T, PB_OP_MIN> {
  static __device__ T init() { return std::numeric_limits<T>::max(); }
  static __device__ T apply(T a, T x) { return (a < x) ? a : x; }
  static __device__ T combine(T a, T b) { return (a < b) ? a : b; }
  static __device__ T finish(T a, int64_t) { return a; }
};
// AbsoluteAddOperator: |l| + |r|.
template <typename T> struct Op<T, PB_OP_ABS_ADD> {
  static __device__ T init() { return T(0); }
  static __device__ T apply(T a, T x) { return fabs(a) + fabs(x); }
  static __device__ T combine(T a, T b) { return fabs(a) + fabs(b); }
  static __device__ T finish(T a, int64_t) { return a; }
};

// INNER: grid (col_groups, row_chunks); out[c] for c in [0, cols).
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void reduce_inner_kernel(
    int64_t rows, int64_t cols, const T* __restrict__ in, int64_t ld, T* out,
    int64_t cols_per_group, int64_t rows_per_chunk, T* __restrict__ partials) {
  using O = Op<T, OP>;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t c0 = int64_t(blockIdx.x) * cols_per_group;
  int64_t c1 = c0 + cols_per_group;
  if (c1 > cols) c1 = cols;
  const int64_t r0 = int64_t(blockIdx.y) * rows_per_chunk;
  int64_t r1 = r0 + rows_per_chunk;
  if (r1 > rows) r1 = rows;
  for (int64_t c = c0 + wave; c < c1; c += kWaves) {
    const T* col = in + c * ld;
    T a0 = O::init(), a1 = O::init(), a2 = O::init(), a3 = O::init();
    int64_t r = r0 + lane;
    for (; r + 192 < r1; r += 256) {
      const T x0 = pbdev::ld_stream(col + r), x1 = pbdev::ld_stream(col + r + 64);
      const T x2 = pbdev::ld_stream(col + r + 128), x3 = pbdev::ld_stream(col + r + 192);
      a0 = O::apply(a0, x0);
      a1 = O::apply(a1, x1);
      a2 = O::apply(a2, x2);
      a3 = O::apply(a3, x3);
    }
    for (; r < r1; r += 64) a0 = O::apply(a0, pbdev::ld_stream(col + r));
    T a = O::combine(O::combine(a0, a1), O::combine(a2, a3));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = O::combine(a, __shfl_xor(a, off, 64));
    if (lane == 0) {
      if (gridDim.y == 1) out[c] = O::finish(a, rows);
      else partials[int64_t(blockIdx.y) * cols + c] = a;
    }
  }
}

// Folds the per-chunk partials in a fixed order (a separate launch: stream
// order publishes the partials without device-scope fences).
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void reduce_fold_kernel(int64_t len, int chunks,
                                                             const T* __restrict__ partials,
                                                             T* out, int64_t count) {
  using O = Op<T, OP>;
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= len) return;
  constexpr int U = 8;  // loads in flight before the (fixed-order) combines
  T t = partials[i];
  for (int k0 = 1; k0 < chunks; k0 += U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u < chunks) v[u] = partials[int64_t(k0 + u) * len + i];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (k0 + u < chunks) t = O::combine(t, v[u]);
  }
  out[i] = O::finish(t, count);
}

// OUTER: grid (row_tiles, col_chunks); out[r] for r in [0, rows).
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void reduce_outer_kernel(
    int64_t rows, int64_t cols, const T* __restrict__ in, int64_t ld, T* out,
    int64_t cols_per_chunk, T* __restrict__ partials) {
  using O = Op<T, OP>;
  __shared__ T red[kWaves][64 * 4 + 1];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t row0 = int64_t(blockIdx.x) * 256;
  const int64_t c0 = int64_t(blockIdx.y) * cols_per_chunk;
  int64_t c1 = c0 + cols_per_chunk;
  if (c1 > cols) c1 = cols;
  T acc[4] = {O::init(), O::init(), O::init(), O::init()};
  for (int64_t c = c0 + wave; c < c1; c += kWaves) {
    const T* col = in + c * ld;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r = row0 + lane + 64 * q;
      if (r < rows) acc[q] = O::apply(acc[q], pbdev::ld_stream(col + r));
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[wave][lane + 64 * q] = acc[q];
  __syncthreads();
  const int rr = threadIdx.x;
  T s = red[0][rr];
#pragma unroll
  for (int w = 1; w < kWaves; ++w) s = O::combine(s, red[w][rr]);
  const int64_t r = row0 + rr;
  if (gridDim.y == 1) {
    if (r < rows) out[r] = O::finish(s, cols);
    return;
  }
  if (r < rows) partials[int64_t(blockIdx.y) * rows + r] = s;
}

template <typename T, int OP>
pb_status_t launch_op(pb_handle_t h, pb_reduction_dim_t dim, const T* in, int64_t ld,
                      T* out, int64_t rows, int64_t cols) {
  const int target = 8 * h->num_cus;
  if (dim == PB_REDUCE_INNER) {
    // Chunks of at least 4096 rows; only split rows when columns are few.
    int64_t groups = pb::ceil_div(cols, kWaves);
    int64_t chunks = 1;
    if (groups < target) {
      chunks = pb::ceil_div(target, groups);
      const int64_t max_chunks = pb::ceil_div(rows, 4096);
      if (chunks > max_chunks) chunks = max_chunks;
      if (chunks > 65535) chunks = 65535;
    } else {
      groups = target * 4 < groups ? target * 4 : groups;
    }
    const int64_t cpg = pb::ceil_div(cols, groups);
    groups = pb::ceil_div(cols, cpg);
    const int64_t rpc = pb::ceil_div(rows, chunks);
    chunks = pb::ceil_div(rows, rpc);
    void* ws = nullptr;
    if (chunks > 1) PB_TRY(pb::workspace(h, sizeof(T) * size_t(chunks * cols), &ws));
    T* part = static_cast<T*>(ws);
    hipLaunchKernelGGL((reduce_inner_kernel<T, OP>), dim3(unsigned(groups), unsigned(chunks)),
                       dim3(kBlock), 0, h->stream, rows, cols, in, ld, out, cpg, rpc, part);
    if (chunks > 1)
      hipLaunchKernelGGL((reduce_fold_kernel<T, OP>), dim3(unsigned(pb::ceil_div(cols, kBlock))),
                         dim3(kBlock), 0, h->stream, cols, int(chunks), part, out, rows);
  } else {
    const int64_t tiles = pb::ceil_div(rows, 256);
    PB_REQUIRE(tiles <= 2147483647LL, "rows too large");
    int64_t chunks = pb::ceil_div(target, tiles);
    const int64_t max_chunks = pb::ceil_div(cols, 16);
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    if (chunks > 65535) chunks = 65535;
    const int64_t cpc = pb::ceil_div(cols, chunks);
    chunks = pb::ceil_div(cols, cpc);
    void* ws = nullptr;
    if (chunks > 1) PB_TRY(pb::workspace(h, sizeof(T) * size_t(chunks * rows), &ws));
    T* part = static_cast<T*>(ws);
    hipLaunchKernelGGL((reduce_outer_kernel<T, OP>), dim3(unsigned(tiles), unsigned(chunks)),
                       dim3(kBlock), 0, h->stream, rows, cols, in, ld, out, cpc, part);
    if (chunks > 1)
      hipLaunchKernelGGL((reduce_fold_kernel<T, OP>), dim3(unsigned(pb::ceil_div(rows, kBlock))),
                         dim3(kBlock), 0, h->stream, rows, int(chunks), part, out, cols);
  }
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

template <typename T>
pb_status_t reduction(pb_handle_t h, pb_reduction_op_t op, pb_reduction_dim_t dim,
                      const T* in, int64_t ld, T* out, int64_t rows, int64_t cols) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(dim == PB_REDUCE_INNER || dim == PB_REDUCE_OUTER, "invalid reduction dim %d",
             int(dim));
  PB_REQUIRE(rows >= 0 && cols >= 0, "negative dimension");
  const int64_t out_len = dim == PB_REDUCE_INNER ? cols : rows;
  const int64_t red_len = dim == PB_REDUCE_INNER ? rows : cols;
  if (out_len == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(red_len > 0, "empty reduced dimension");
  PB_REQUIRE(in != nullptr && out != nullptr, "null pointer");
  PB_REQUIRE(ld >= rows, "ld (%lld) < rows (%lld)", (long long)ld, (long long)rows);
  switch (op) {
    case PB_OP_ADD: return launch_op<T, PB_OP_ADD>(h, dim, in, ld, out, rows, cols);
    case PB_OP_PRODUCT: return launch_op<T, PB_OP_PRODUCT>(h, dim, in, ld, out, rows, cols);
    case PB_OP_MAX: return launch_op<T, PB_OP_MAX>(h, dim, in, ld, out, rows, cols);
    case PB_OP_MIN: return launch_op<T, PB_OP_MIN>(h, dim, in, ld, out, rows, cols);
    case PB_OP_ABS_ADD: return launch_op<T, PB_OP_ABS_ADD>(h, dim, in, ld, out, rows, cols);
    case PB_OP_MEAN: return launch_op<T, PB_OP_MEAN>(h, dim, in, ld, out, rows, cols);
  }
  return pb::set_error(PB_STATUS_INVALID_VALUE, "invalid reduction operator %d", int(op));
}

}  // namespace

extern "C" {

pb_status_t pb_sreduction(pb_handle_t h, pb_reduction_op_t op, pb_reduction_dim_t dim,
                          const float* in, int64_t ld, float* out, int64_t rows,
                          int64_t cols) {
  return reduction<float>(h, op, dim, in, ld, out, rows, cols);
}

pb_status_t pb_dreduction(pb_handle_t h, pb_reduction_op_t op, pb_reduction_dim_t dim,
                          const double* in, int64_t ld, double* out, int64_t rows,
                          int64_t cols) {
  return reduction<double>(h, op, dim, in, ld, out, rows, cols);
}

}  // extern "C"

// FP32 / FP64 GEMM entry points: _gemm, _gemm_batched, _gemm_strided_batched.
//
// Front end (validation, alpha == 0, default strides) restates
// src/interface/gemm_interface.hpp:105-240.  The tile kernels are the MFMA
// engine of gemm_core.hpp (the "plain" FP32/FP64 method) and an FP32 vector
// (SIMT) kernel used as the FP32-SIMT variant of pb_gemm_ex.
#include "gemm_launch.hpp"
#include "gemm_front.hpp"
#include "gemm_small.hpp"

#include <algorithm>

namespace pbgemm {

using F128 = TileShape<128, 128, 32, 64, 64>;
using F64 = TileShape<64, 64, 32, 32, 32>;
using D128 = TileShape<128, 128, 16, 64, 64>;
using D64 = TileShape<64, 64, 16, 32, 32>;

// ---------------------------------------------------------------------------
// alpha == 0: C <- beta * C exactly as gemm_interface.hpp:112-139 does it.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void scal_vector_kernel(int64_t n, T beta, T* c) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    c[i] = (beta == T(0)) ? T(0) : beta * c[i];  // _scal: alpha == 0 assigns zero
}

template <typename T>
__global__ void scal_matrix_kernel(int64_t rows, int64_t cols, T beta, T* c, int64_t ldc) {
  const int64_t total = rows * cols;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = i % rows, col = i / rows;
    c[r + col * ldc] *= beta;  // _scal_matrix: multiply (NaN * 0 stays NaN)
  }
}

template <typename T>
pb_status_t scal_matrix(pb_handle_t h, int64_t rows, int64_t cols, T beta, T* c, int64_t ldc) {
  if (beta == T(1) || rows <= 0 || cols <= 0) return PB_STATUS_SUCCESS;  // :505-506
  const int64_t total = rows * cols;
  int64_t g = pb::ceil_div(total, 256);
  if (g > 8 * h->num_cus) g = 8 * h->num_cus;
  hipLaunchKernelGGL(scal_matrix_kernel<T>, dim3(unsigned(g)), dim3(256), 0, h->stream, rows,
                     cols, beta, c, ldc);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

template <typename T>
pb_status_t alpha_zero(pb_handle_t h, int64_t M, int64_t N, T beta, T* C, int64_t ldc,
                       int64_t stridec, int64_t batch, bool interleaved) {
  if (M <= 0 || N <= 0 || batch <= 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(C != nullptr, "null C");
  if (interleaved) {
    // Rows r*batch + b are contiguous: one (M*batch) x N matrix of pitch ldc*batch.
    return scal_matrix<T>(h, M * batch, N, beta, C, ldc * batch);
  }
  const int64_t size_c = ldc * N;
  if (size_c == stridec) {
    if (ldc == M) {
      const int64_t n = size_c * batch;
      int64_t g = pb::ceil_div(n, 256);
      if (g > 8 * h->num_cus) g = 8 * h->num_cus;
      hipLaunchKernelGGL(scal_vector_kernel<T>, dim3(unsigned(g)), dim3(256), 0, h->stream, n,
                         beta, C);
      PB_CHECK_LAUNCH();
      return PB_STATUS_SUCCESS;
    }
    return scal_matrix<T>(h, M, N * batch, beta, C, ldc);
  }
  for (int64_t b = 0; b < batch; ++b) PB_TRY(scal_matrix<T>(h, M, N, beta, C + b * stridec, ldc));
  return PB_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// FP32 SIMT kernel: 128 x 128 tile, 16 x 16 threads each owning an 8 x 8
// register tile (rows ty + 16 i, cols tx + 16 j), sequential FMA over k.
// ---------------------------------------------------------------------------
template <bool KCA, bool KCB, bool EDGE>
__global__ __launch_bounds__(256) void gemm_simt_kernel(GemmProblem<float> p) {
  using TL = Tile<float, F128, KCA, KCB>;
  __shared__ __attribute__((aligned(16))) float lds[2 * TL::STAGE_ELEMS];
  int64_t tm, tn;
  tile_of(p, blockIdx.x, tm, tn);
  const int64_t m0 = tm * 128, n0 = tn * 128;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  for (int64_t b = blockIdx.y; b < p.batch; b += gridDim.y) {
    float acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = 0.f;
    Operand<float> A{p.a[0] + b * p.a_stride, p.a_rs, p.a_ks, p.M};
    Operand<float> B{p.b[0] + b * p.b_stride, p.b_rs, p.b_ks, p.N};
    TL::template mainloop_fn<EDGE>(A, B, m0, n0, 0, p.K, lds, [&](const float* st) {
      const float* As = st;
      const float* Bs = st + TL::A_ELEMS;
#pragma unroll 2
      for (int kk = 0; kk < 32; kk += 4) {
        float4 a[8], bb[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
          a[i] = *reinterpret_cast<const float4*>(As + (ty + 16 * i) * TL::LDK + kk);
#pragma unroll
        for (int j = 0; j < 8; ++j)
          bb[j] = *reinterpret_cast<const float4*>(Bs + (tx + 16 * j) * TL::LDK + kk);
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            acc[i][j] = fmaf(a[i].x, bb[j].x, acc[i][j]);
            acc[i][j] = fmaf(a[i].y, bb[j].y, acc[i][j]);
            acc[i][j] = fmaf(a[i].z, bb[j].z, acc[i][j]);
            acc[i][j] = fmaf(a[i].w, bb[j].w, acc[i][j]);
          }
      }
    });
    float* c = p.c + b * p.c_stride;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t m = m0 + ty + 16 * i;
      if (m >= p.M) continue;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t n = n0 + tx + 16 * j;
        if (n >= p.N) continue;
        float* cp = c + m * p.c_rs + n * p.ldc;
        const float v = p.alpha * acc[i][j];
        *cp = (p.beta != 0.f) ? v + p.beta * *cp : v;
      }
    }
  }
}

template <bool EDGE>
void dispatch_simt(pb_handle_t h, const GemmProblem<float>& p, dim3 grid, bool kca, bool kcb) {
#define PB_LAUNCH(KA, KB)                                                                 \
  hipLaunchKernelGGL((gemm_simt_kernel<KA, KB, EDGE>), grid, dim3(256), 0, h->stream, p)
  if (kca && kcb) PB_LAUNCH(true, true);
  else if (kca) PB_LAUNCH(true, false);
  else if (kcb) PB_LAUNCH(false, true);
  else PB_LAUNCH(false, false);
#undef PB_LAUNCH
}

pb_status_t launch_simt(pb_handle_t h, GemmProblem<float> p, bool kca, bool kcb, bool vec) {
  p.tiles_m = pb::ceil_div(p.M, 128);
  p.tiles_n = pb::ceil_div(p.N, 128);
  p.full_m = p.M / 128;
  p.full_n = p.N / 128;
  p.splits = 1;
  const int64_t tiles = p.tiles_m * p.tiles_n;
  PB_REQUIRE(tiles <= 0x7fffffffLL, "problem too large");
  const unsigned gy = unsigned(p.batch < 65535 ? p.batch : 65535);
  const int64_t interior = vec ? p.full_m * p.full_n : 0;
  if (interior > 0) {
    p.tile_set = kInteriorTiles;
    dispatch_simt<false>(h, p, dim3{unsigned(interior), gy}, kca, kcb);
    PB_CHECK_LAUNCH();
  }
  if (tiles - interior > 0) {
    p.tile_set = vec ? kEdgeTiles : kAllTiles;
    dispatch_simt<true>(h, p, dim3{unsigned(tiles - interior), gy}, kca, kcb);
    PB_CHECK_LAUNCH();
  }
  return PB_STATUS_SUCCESS;
}

// Tile shape and split count for the plain MFMA path: big tiles when there
// are enough of them to fill the chip, else 64 x 64 tiles, else split K.
// small: the small-problem batch kernel (one problem per workgroup).
struct Plan { bool big; int splits; bool small = false; };

inline Plan plan_plain(pb_handle_t h, int64_t M, int64_t N, int64_t K, int64_t batch, int bk) {
  const int64_t cus = h->num_cus;
  const int64_t t128 = pb::ceil_div(M, 128) * pb::ceil_div(N, 128) * batch;
  const int64_t t64 = pb::ceil_div(M, 64) * pb::ceil_div(N, 64) * batch;
  // Padded area of the 128 x 128 tiling relative to the problem: large tiles
  // only when they are nearly full (otherwise every tile runs guarded).
  const double pad128 = double(pb::ceil_div(M, 128) * 128) * double(pb::ceil_div(N, 128) * 128) /
                        (double(M) * double(N));
  if (t128 >= 2 * cus && pad128 < 1.15) return {true, 1};
  if (t64 >= cus) return {false, 1};
  int64_t s = pb::ceil_div(cus, t64);
  const int64_t smax = K / (4 * bk);
  if (s > smax) s = smax;
  if (s < 1) s = 1;
  return {false, int(s)};
}

}  // namespace pbgemm

// ---------------------------------------------------------------------------
// The plain engine's configurations, the autotuner's search space for FP32
// and FP64 _gemm: 2 tile shapes (128 x 128, 64 x 64) x 8 split-K factors,
// then the small-problem batch kernel (only for M, N <= 64 with at least one
// problem per CU).  The reference's counterpart is the hand-written choice
// of src/interface/blas3/backend/amd_gpu.hpp:52-288, tuned offline by
// tools/auto_tuner/include/gemm_tuner.hpp:99-180.
// ---------------------------------------------------------------------------
namespace pbgemm {

constexpr int kPlainSplits[] = {1, 2, 4, 8, 16, 32, 64, 128};
constexpr int kPlainNumSplits = 8;
constexpr int kPlainSmall = 2 * kPlainNumSplits;
constexpr int kPlainConfigs = kPlainSmall + 1;

int selector_eval(const pb_selector_state& s, const float f[8]);

// Features of an FP32 (dtype 1) or FP64 (dtype 2) plain GEMM, in the
// selector's order: log2 M, N, K, transa, transb, element type, log2 batch,
// log2 of A's leading-dimension padding.
inline void plain_features(int dtype, char ta, char tb, int64_t M, int64_t N, int64_t K, int64_t batch,
                           int64_t lda, float f[8]) {
  auto lg = [](double v) { return float(std::log2(v > 1.0 ? v : 1.0)); };
  f[0] = lg(double(M));
  f[1] = lg(double(N));
  f[2] = lg(double(K));
  f[3] = ta == 't' ? 1.f : 0.f;
  f[4] = tb == 't' ? 1.f : 0.f;
  f[5] = float(dtype);
  f[6] = lg(double(batch));
  const int64_t arows = ta == 'n' ? M : K;
  f[7] = lg(double(lda) / double(arows > 0 ? arows : 1));
}

// Tiles x splits of plain config `id` (< kPlainSmall) for an M x N x batch problem.
inline int64_t plain_split_tiles(pb_handle_t, int64_t M, int64_t N, int64_t batch, int id) {
  const int64_t bm = id < kPlainNumSplits ? 128 : 64;
  const int64_t tiles = pb::ceil_div(M, bm) * pb::ceil_div(N, bm) * batch;
  const int splits = kPlainSplits[id % kPlainNumSplits];
  return splits > 1 ? tiles * splits : 0;
}

// The shapes the plain network was trained and validated on: work up to
// 2^35 flop, operands up to 2^26 elements, batches up to 256.  Outside them
// the built-in rule decides (the network would be extrapolating).
inline bool plain_selector_domain(int64_t M, int64_t N, int64_t K, int64_t batch) {
  const double flop = 2.0 * double(M) * double(N) * double(K) * double(batch);
  const double elems = double(std::max(std::max(M * K, K * N), M * N)) * double(batch);
  return batch <= 256 && flop <= 34359738368.0 && elems <= 67108864.0;
}

inline Plan plain_config_plan(int id) {
  if (id == kPlainSmall) return {false, 1, true};
  return {id < kPlainNumSplits, kPlainSplits[id % kPlainNumSplits], false};
}

// The plan run_plain uses: `forced` (a configuration id) when given and
// valid for the shape, else the handle's plain selector when one is
// installed (forced == kRuleOnly skips it), else the built-in rule.
constexpr int kRuleOnly = -2;

template <typename T>
Plan choose_plain(pb_handle_t h, const GemmCall<T>& g, int forced = -1) {
  const bool small_ok = small::eligible(h, g.M, g.N, g.batch, g.interleaved);
  int id = forced;
  if (id == -1 && h->plain_selector.n_layers > 0 && plain_selector_domain(g.M, g.N, g.K, g.batch)) {
    float f[8];
    plain_features(sizeof(T) == 4 ? 1 : 2, g.ta, g.tb, g.M, g.N, g.K, g.batch, g.lda, f);
    id = selector_eval(h->plain_selector, f);
    // The tuner never times splits beyond 16 split tiles per CU (their
    // partials only cost memory traffic); such a pick gives way to the rule.
    if (id >= 0 && id < kPlainSmall && plain_split_tiles(h, g.M, g.N, g.batch, id) > 16 * int64_t(h->num_cus))
      id = -1;
  }
  if (id >= 0 && id < kPlainConfigs && (id != kPlainSmall || small_ok)) return plain_config_plan(id);
  if (small_ok) return {false, 1, true};
  return plan_plain(h, g.M, g.N, g.K, g.batch, sizeof(T) == 4 ? 32 : 16);
}

// Split-K partial bytes run_plain(h, g, false) needs, so that a caller can put
// its own data after them in one workspace allocation (0 without a split).
template <typename T>
size_t plain_partial_bytes(pb_handle_t h, const GemmCall<T>& g, int forced = -1) {
  const Plan pl = choose_plain<T>(h, g, forced);
  if (pl.small) return 0;
  if constexpr (sizeof(T) == 4) {
    return pl.big ? splitk_bytes<T, F128>(g.M, g.N, g.K, g.batch, pl.splits)
                  : splitk_bytes<T, F64>(g.M, g.N, g.K, g.batch, pl.splits);
  } else {
    return pl.big ? splitk_bytes<T, D128>(g.M, g.N, g.K, g.batch, pl.splits)
                  : splitk_bytes<T, D64>(g.M, g.N, g.K, g.batch, pl.splits);
  }
}

// `partial`: caller-provided split-K partial buffer (plain_partial_bytes), or
// null for the handle workspace.  `forced`: a plain configuration id, or -1.
template <typename T>
pb_status_t run_plain(pb_handle_t h, const GemmCall<T>& g, bool simt, T* partial = nullptr, int forced = -1) {
  GemmProblem<T> p{};
  p.M = g.M; p.N = g.N; p.K = g.K;
  p.alpha = g.alpha; p.beta = g.beta;
  p.nterms = 1;
  p.a[0] = g.a;
  p.b[0] = g.b;
  p.c = g.c;
  p.batch = g.batch;
  p.partial = partial;
  const bool vec = set_operands(p, g.ta, g.tb, g.lda, g.sa, g.ldb, g.sb, g.interleaved);
  p.c_rs = g.interleaved ? g.batch : 1;
  p.ldc = g.interleaved ? g.ldc * g.batch : g.ldc;
  p.c_stride = g.interleaved ? 1 : g.sc;
  p.c_vec = !g.interleaved && aligned16(g.c) && g.ldc % 4 == 0 &&
            (g.batch == 1 || g.sc % 4 == 0);
  const bool kca = g.ta == 't', kcb = g.tb == 'n';
  if constexpr (sizeof(T) == 4) {
    if (simt) return launch_simt(h, p, kca, kcb, vec);
  } else {
    (void)simt;
  }
  const Plan pl = choose_plain<T>(h, g, forced);
  if (pl.small) return small::launch<T>(h, p, vec);
  if constexpr (sizeof(T) == 4) {
    if (pl.big) return launch_gemm<T, F128>(h, p, kca, kcb, vec, pl.splits);
    return launch_gemm<T, F64>(h, p, kca, kcb, vec, pl.splits);
  } else {
    if (pl.big) return launch_gemm<T, D128>(h, p, kca, kcb, vec, pl.splits);
    return launch_gemm<T, D64>(h, p, kca, kcb, vec, pl.splits);
  }
}

template <typename T>
pb_status_t gemm_entry(pb_handle_t h, char ta, char tb, int64_t M, int64_t N, int64_t K, T alpha,
                       const T* a, int64_t lda, int64_t sa, const T* b, int64_t ldb, int64_t sb,
                       T beta, T* c, int64_t ldc, int64_t sc, int64_t batch,
                       pb_batch_type_t bt, bool simt) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  GemmCall<T> g;
  bool done = false;
  PB_TRY(gemm_front<T>(h, ta, tb, M, N, K, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                       batch, bt, g, done));
  if (done) return PB_STATUS_SUCCESS;
  if (g.alpha == T(0))
    return alpha_zero<T>(h, g.M, g.N, g.beta, g.c, g.ldc, g.sc, g.batch, g.interleaved);
  return run_plain<T>(h, g, simt);
}

}  // namespace pbgemm

// Used by pb_gemm_ex (gemm_lowp.hip) for the FP32 arithmetics.
namespace pbgemm {
pb_status_t sgemm_plain(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                        float alpha, const float* a, int64_t lda, int64_t sa, const float* b,
                        int64_t ldb, int64_t sb, float beta, float* c, int64_t ldc, int64_t sc,
                        int64_t batch, bool simt) {
  return gemm_entry<float>(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                           batch, PB_BATCH_STRIDED, simt);
}
pb_status_t sgemm_alpha_zero(pb_handle_t h, int64_t M, int64_t N, float beta, float* C,
                             int64_t ldc, int64_t stridec, int64_t batch) {
  return alpha_zero<float>(h, M, N, beta, C, ldc, stridec, batch, false);
}
// ---------------------------------------------------------------------------
// SYMM (src/interface/symm_interface.hpp:36-80): C <- alpha A B + beta C
// (side 'l', A M x M) or alpha B A + beta C (side 'r', A N x N), A symmetric
// with only its `uplo` triangle referenced.  The reference runs its GEMM with
// a symmetric-operand loader; here the stored triangle is mirrored into a
// full matrix in the workspace (a 64 x 64 tiled pass through LDS: k^2 reads
// and writes, against k^2 n flops of the product) and the MFMA GEMM engine
// runs on it.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void symmetrize_kernel(int64_t k, const T* __restrict__ a,
                                                         int64_t lda, bool upper, T* __restrict__ full) {
  __shared__ T mirror[64][65];  // the transposed tile (tj, ti): mirror[r][c] = A(tj*64+c, ti*64+r)
  const int64_t i0 = int64_t(blockIdx.x) * 64, j0 = int64_t(blockIdx.y) * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // Element (i, j) comes from A(i, j) when it is in the stored triangle
  // (upper: i <= j), else from A(j, i).  The mirrored tile is needed when any
  // element of the output tile lies outside the stored triangle.
  const bool need_mirror = upper ? (i0 + 63 > j0) : (j0 + 63 > i0);
  if (need_mirror) {
    for (int r = ty; r < 64; r += 4) {  // A(j0 + tx, i0 + r): coalesced along tx
      const int64_t si = j0 + tx, sj = i0 + r;
      mirror[r][tx] = (si < k && sj < k) ? a[si + sj * lda] : T(0);
    }
  }
  __syncthreads();
  for (int c = ty; c < 64; c += 4) {
    const int64_t i = i0 + tx, j = j0 + c;
    if (i >= k || j >= k) continue;
    const bool stored = upper ? (i <= j) : (i >= j);
    full[i + j * k] = stored ? a[i + j * lda] : mirror[tx][c];
  }
}

template <typename T>
pb_status_t symm_entry(pb_handle_t h, char side, char uplo, int64_t M, int64_t N, T alpha,
                       const T* a, int64_t lda, const T* b, int64_t ldb, T beta, T* c, int64_t ldc) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  const char s = char(tolower(side)), u = char(tolower(uplo));
  PB_REQUIRE(u == 'u' || u == 'l', "invalid _uplo");
  PB_REQUIRE(s == 'l' || s == 'r', "invalid _side");
  PB_REQUIRE(M >= 0 && N >= 0, "negative dimension (M=%lld N=%lld)", (long long)M, (long long)N);
  const int64_t k = s == 'l' ? M : N;
  PB_REQUIRE(lda >= std::max<int64_t>(1, k), "lda (%lld) < max(1, %lld)", (long long)lda, (long long)k);
  PB_REQUIRE(ldb >= std::max<int64_t>(1, M), "ldb (%lld) < max(1, M)", (long long)ldb);
  PB_REQUIRE(ldc >= std::max<int64_t>(1, M), "ldc (%lld) < max(1, M)", (long long)ldc);
  if (M == 0 || N == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(c != nullptr && a != nullptr && b != nullptr, "null matrix pointer");
  // alpha == 0: C <- beta C (the product is not formed).
  if (alpha == T(0)) return alpha_zero<T>(h, M, N, beta, c, ldc, ldc * N, 1, false);
  GemmCall<T> g;
  g.ta = 'n'; g.tb = 'n';
  g.M = M; g.N = N; g.K = k;
  g.alpha = alpha; g.beta = beta;
  g.c = c; g.ldc = ldc;
  // Leading dimensions before the plan is sized: an installed selector reads
  // them, and the run below must plan exactly the split sized here.
  g.lda = s == 'l' ? k : ldb;
  g.ldb = s == 'l' ? ldb : k;
  const size_t part = pb::align_up(plain_partial_bytes<T>(h, g), 256);
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, part + size_t(k) * size_t(k) * sizeof(T), &ws));
  T* full = reinterpret_cast<T*>(static_cast<char*>(ws) + part);
  const dim3 grid(unsigned(pb::ceil_div(k, 64)), unsigned(pb::ceil_div(k, 64)));
  PB_REQUIRE(grid.y <= 65535u, "matrix too large");
  hipLaunchKernelGGL(symmetrize_kernel<T>, grid, dim3(256), 0, h->stream, k, a, lda, u == 'u', full);
  PB_CHECK_LAUNCH();
  if (s == 'l') {
    g.a = full; g.lda = k;
    g.b = b; g.ldb = ldb;
  } else {
    g.a = b; g.lda = ldb;
    g.b = full; g.ldb = k;
  }
  return run_plain<T>(h, g, false, part ? static_cast<T*>(ws) : nullptr);
}

// ---------------------------------------------------------------------------
// TRSM (src/interface/trsm_interface.hpp:36-420): solves op(A) X = alpha B
// (side 'l', A M x M) or X op(A) = alpha B (side 'r', A N x N) for X, A
// triangular (uplo), unit or not (diag: a unit diagonal is never read),
// and overwrites B with X.  The reference inverts 16 x 16 diagonal blocks
// and walks them with two GEMMs per block.  Here the diagonal blocks are
// 64 x 64 (all inverted by one launch, one workgroup per block, in LDS), so
// a solve of order k takes ceil(k / 64) steps, each an MFMA GEMM
//   X_i <- op(inv(A_ii)) (alpha_i W_i)
// where W is a workspace copy of B, alpha_i is alpha at the first step and 1
// after, and the steps run forward or backward as op(A) is lower or upper.
// The steps are grouped in panels of kTrsmPanel blocks: after block i the
// panel's later blocks get W_p <- alpha_i W_p - op(A)_(p, i) X_i, and after the
// panel the rows (columns) beyond it get one update of rank up to 512 from
// the whole panel, so W is read and written once per panel instead of once
// per block.  Below order kTrsmPanelMinK the panel is one block: the wider
// updates do not pay for the extra launches there.
// ---------------------------------------------------------------------------
constexpr int kTrsmNb = 64;
constexpr int kTrsmPanel = 8;
constexpr int64_t kTrsmPanelMinK = 2048;

// Inverts the triangular diagonal block `blk` of A (order min(64, k - 64 blk),
// padded to 64 with the identity) into inv + blk * 64 * 64 (ld 64, same
// triangle).  Thread c computes column c of the inverse by substitution,
// right-looking: its 64 running sums stay in registers, and each solved
// entry x_i updates the sums of the rows still to come with independent
// FMAs (a left-looking dot product per entry would chain one LDS read per
// term).  Only the stored triangle (and the diagonal unless unit) of A is read.
template <typename T>
__global__ __launch_bounds__(64) void trsm_invert_kernel(int64_t k, const T* __restrict__ a, int64_t lda,
                                                         bool upper, bool unit, T* __restrict__ inv) {
  constexpr int NB = kTrsmNb;
  // LT[j][i] = A(i0 + i, i0 + j) in the stored triangle, 0 elsewhere (column j
  // of the block contiguous, read as a broadcast); X[i][c]: the inverse.
  __shared__ T LT[NB][NB];
  __shared__ T X[NB][NB];
  const int64_t i0 = int64_t(blockIdx.x) * NB;
  const int bs = int(k - i0 < NB ? k - i0 : NB);
  const int c = threadIdx.x;
  for (int j = 0; j < NB; ++j) {  // thread c loads row c of the block, column by column (coalesced per j)
    const bool in_tri = upper ? (c <= j) : (c >= j);
    T v = T(0);
    if (c < bs && j < bs && in_tri && !(unit && c == j)) v = a[(i0 + c) + (i0 + j) * lda];
    if (c == j && (unit || c >= bs)) v = T(1);
    LT[j][c] = v;
  }
  __syncthreads();
  T acc[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) acc[r] = (r == c) ? T(1) : T(0);
  if (!upper) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const T xi = (i < c) ? T(0) : acc[i] / LT[i][i];
      X[i][c] = xi;
      if (i >= c) {
#pragma unroll
        for (int r = i + 1; r < NB; ++r) acc[r] -= LT[i][r] * xi;
      }
    }
  } else {
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      const T xi = (i > c) ? T(0) : acc[i] / LT[i][i];
      X[i][c] = xi;
      if (i <= c) {
#pragma unroll
        for (int r = 0; r < i; ++r) acc[r] -= LT[i][r] * xi;
      }
    }
  }
  __syncthreads();
  T* out = inv + int64_t(blockIdx.x) * NB * NB;
  for (int j = 0; j < NB; ++j) out[c + j * NB] = X[c][j];  // column-major, coalesced per j
}

template <typename T>
__global__ void copy_matrix_kernel(int64_t rows, int64_t cols, const T* __restrict__ src, int64_t lds,
                                   T* __restrict__ dst, int64_t ldd) {
  const int64_t total = rows * cols;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t r = e % rows, col = e / rows;
    dst[r + col * ldd] = src[r + col * lds];
  }
}

template <typename T>
pb_status_t trsm_entry(pb_handle_t h, char side, char uplo, char trans, char diag, int64_t M, int64_t N,
                       T alpha, const T* a, int64_t lda, T* b, int64_t ldb) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  // trsm_interface.hpp:44-62
  PB_REQUIRE(M != 0 && N != 0 && lda != 0 && ldb != 0, "invalid matrix size argument");
  const char s = char(tolower(side)), u = char(tolower(uplo)), t = char(tolower(trans)),
             d = char(tolower(diag));
  PB_REQUIRE(s == 'l' || s == 'r', "invalid Side argument");
  PB_REQUIRE(u == 'u' || u == 'l', "invalid Triangle argument");
  PB_REQUIRE(t == 'n' || t == 't', "invalid Transpose argument");
  PB_REQUIRE(d == 'u' || d == 'n', "invalid Diagonal argument");
  PB_REQUIRE(M > 0 && N > 0, "negative dimension (M=%lld N=%lld)", (long long)M, (long long)N);
  const int64_t k = s == 'l' ? M : N;
  PB_REQUIRE(lda >= k && ldb >= M, "leading dimension too small (lda %lld, ldb %lld)", (long long)lda,
             (long long)ldb);
  PB_REQUIRE(a != nullptr && b != nullptr, "null matrix pointer");
  const bool lower_eff = (u == 'l') == (t == 'n');  // op(A) is lower triangular
  const int64_t nblk = pb::ceil_div(k, kTrsmNb);
  PB_REQUIRE(nblk <= 0x7fffffffLL, "matrix too large");
  const bool forward = (s == 'l') == lower_eff;  // left+lower or right+upper walk from block 0
  // Block q of the walk (0 first): rows (or columns) [start, start + size).
  auto blk_start = [&](int64_t q) { return (forward ? q : nblk - 1 - q) * kTrsmNb; };
  auto blk_size = [&](int64_t q) { return std::min<int64_t>(kTrsmNb, k - blk_start(q)); };
  const int64_t panel = k >= kTrsmPanelMinK ? kTrsmPanel : 1;
  // The walk, in launch order: solve(i, bs, al) solves block [i, i + bs);
  // update(r0, r1, c0, c1, be) takes the solved blocks [c0, c1) out of the
  // rows (columns) [r0, r1) of W.  Inside a panel each solved block updates
  // only the panel's later blocks; the rest of W gets one update per panel.
  // W is scaled by alpha where it is first touched.
  auto walk = [&](auto&& solve, auto&& update) -> pb_status_t {
    for (int64_t q0 = 0; q0 < nblk; q0 += panel) {
      const int64_t q1 = std::min<int64_t>(q0 + panel, nblk);
      const int64_t plo = forward ? blk_start(q0) : blk_start(q1 - 1);
      const int64_t phi = forward ? blk_start(q1 - 1) + blk_size(q1 - 1) : blk_start(q0) + blk_size(q0);
      for (int64_t q = q0; q < q1; ++q) {
        const int64_t i = blk_start(q), bs = blk_size(q);
        const T al = q == 0 ? alpha : T(1);
        PB_TRY(solve(i, bs, al));
        if (forward) PB_TRY(update(i + bs, phi, i, i + bs, al));
        else PB_TRY(update(plo, i, i, i + bs, al));
      }
      const T be = q0 == 0 ? alpha : T(1);
      if (forward) PB_TRY(update(phi, k, plo, phi, be));
      else PB_TRY(update(0, plo, plo, phi, be));
    }
    return PB_STATUS_SUCCESS;
  };
  // Workspace: [split-K partials][inverted blocks][W].  The partial region
  // holds the largest split of any GEMM of the walk: a narrow trailing update
  // can split K where the full-width one does not, and must not fall back to
  // the handle workspace, which holds the blocks and W.  The walk's GEMMs are
  // sized here from their shapes alone, so they plan with the built-in rule
  // (which reads nothing else) rather than an installed selector (whose
  // features include the transposes and leading dimensions).
  size_t part = 0;
  {
    auto need = [&](int64_t m, int64_t n, int64_t kk) {
      GemmCall<T> g;
      g.M = m; g.N = n; g.K = kk;
      if (m > 0 && n > 0) part = std::max(part, plain_partial_bytes<T>(h, g, kRuleOnly));
    };
    PB_TRY(walk(
        [&](int64_t, int64_t bs, T) -> pb_status_t {
          if (s == 'l') need(bs, N, bs); else need(M, bs, bs);
          return PB_STATUS_SUCCESS;
        },
        [&](int64_t r0, int64_t r1, int64_t c0, int64_t c1, T) -> pb_status_t {
          if (r1 > r0 && c1 > c0) { if (s == 'l') need(r1 - r0, N, c1 - c0); else need(M, r1 - r0, c1 - c0); }
          return PB_STATUS_SUCCESS;
        }));
  }
  part = pb::align_up(part, 256);
  const size_t inv_bytes = pb::align_up(size_t(nblk) * kTrsmNb * kTrsmNb * sizeof(T), 256);
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, part + inv_bytes + size_t(M) * size_t(N) * sizeof(T), &ws));
  T* partial = part ? static_cast<T*>(ws) : nullptr;
  T* inv = reinterpret_cast<T*>(static_cast<char*>(ws) + part);
  T* W = reinterpret_cast<T*>(static_cast<char*>(ws) + part + inv_bytes);
  const int64_t ldw = M;
  hipLaunchKernelGGL(trsm_invert_kernel<T>, dim3(unsigned(nblk)), dim3(kTrsmNb), 0, h->stream, k, a, lda,
                     u == 'u', d == 'u', inv);
  PB_CHECK_LAUNCH();
  {
    int64_t g = pb::ceil_div(M * N, 256);
    if (g > 8 * h->num_cus) g = 8 * h->num_cus;
    hipLaunchKernelGGL(copy_matrix_kernel<T>, dim3(unsigned(g)), dim3(256), 0, h->stream, M, N, b, ldb, W, ldw);
    PB_CHECK_LAUNCH();
  }
  auto gemm = [&](char ta, char tb, int64_t m, int64_t n, int64_t kk, T al, const T* A_, int64_t la,
                  const T* B_, int64_t lb, T be, T* C_, int64_t lc) -> pb_status_t {
    if (m <= 0 || n <= 0) return PB_STATUS_SUCCESS;
    GemmCall<T> g;
    g.ta = ta; g.tb = tb; g.M = m; g.N = n; g.K = kk;
    g.alpha = al; g.beta = be;
    g.a = A_; g.lda = la; g.b = B_; g.ldb = lb; g.c = C_; g.ldc = lc;
    return run_plain<T>(h, g, false, partial, kRuleOnly);
  };
  // A(r, c) of op(A): A + r + c lda ('n') or, transposed, A + c + r lda.
  auto opA = [&](int64_t r, int64_t c) { return t == 'n' ? a + r + c * lda : a + c + r * lda; };
  // X_i <- op(inv_ii) (al W_i)  (left)  or  (al W_i) op(inv_ii)  (right).
  auto solve = [&](int64_t i, int64_t bs, T al) -> pb_status_t {
    const T* invb = inv + (i / kTrsmNb) * kTrsmNb * kTrsmNb;
    if (s == 'l') return gemm(t, 'n', bs, N, bs, al, invb, kTrsmNb, W + i, ldw, T(0), b + i, ldb);
    return gemm('n', t, M, bs, bs, al, W + i * ldw, ldw, invb, kTrsmNb, T(0), b + i * ldb, ldb);
  };
  // W_[r0, r1) <- be W_[r0, r1) - op(A)[r0, r1) x [c0, c1) X_[c0, c1)  (left; the
  // right side is the transposed statement on columns).
  auto update = [&](int64_t r0, int64_t r1, int64_t c0, int64_t c1, T be) -> pb_status_t {
    if (r1 <= r0 || c1 <= c0) return PB_STATUS_SUCCESS;
    if (s == 'l')
      return gemm(t, 'n', r1 - r0, N, c1 - c0, T(-1), opA(r0, c0), lda, b + c0, ldb, be, W + r0, ldw);
    return gemm('n', t, M, r1 - r0, c1 - c0, T(-1), b + c0 * ldb, ldb, opA(c0, r0), lda, be, W + r0 * ldw,
                ldw);
  };
  return walk(solve, update);
}

// ---------------------------------------------------------------------------
// Complex GEMM (the reference's _gemm on complex types, BLAS_ENABLE_COMPLEX:
// src/interface/gemm_interface.hpp:58-64,140-198; 'c' is taken as 't', as
// the reference does, without conjugation).  Interleaved (re, im) operands
// are packed into one real problem of twice the size,
//   [T_re]   [op(A)_re  -op(A)_im] [op(B)_re]
//   [T_im] = [op(A)_im   op(A)_re] [op(B)_im]      (2M x 2K times 2K x N),
// which one MFMA GEMM solves with K-contiguous-friendly unit strides; an
// epilogue then forms C <- alpha T + beta C in complex arithmetic (beta == 0
// never reads C).  The 4 real products of the textbook method are all
// there, summed over 2K terms in FP32 / FP64.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void cplx_pack_a_kernel(int64_t M, int64_t K, const T* __restrict__ a, int64_t rs, int64_t ks,
                                   T* __restrict__ ahat) {
  const int64_t total = M * K, ld = 2 * M;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e % M, k = e / M;
    const int64_t src = 2 * (i * rs + k * ks);
    const T re = a[src], im = a[src + 1];
    ahat[i + k * ld] = re;
    ahat[i + (K + k) * ld] = -im;
    ahat[(M + i) + k * ld] = im;
    ahat[(M + i) + (K + k) * ld] = re;
  }
}

template <typename T>
__global__ void cplx_pack_b_kernel(int64_t K, int64_t N, const T* __restrict__ b, int64_t ks, int64_t ns,
                                   T* __restrict__ bhat) {
  const int64_t total = K * N, ld = 2 * K;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t k = e % K, j = e / K;
    const int64_t src = 2 * (k * ks + j * ns);
    bhat[k + j * ld] = b[src];
    bhat[(K + k) + j * ld] = b[src + 1];
  }
}

// C <- alpha T + beta C, T = [T_re; T_im] (2M x N, ld 2M); beta == 0: C not
// read (beta_zero), alpha == 0: T not read (C <- beta C).
template <typename T>
__global__ void cplx_epilogue_kernel(int64_t M, int64_t N, T ar, T ai, T br, T bi, const T* __restrict__ t,
                                     T* __restrict__ c, int64_t ldc, bool beta_zero, bool alpha_zero) {
  const int64_t total = M * N;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e % M, j = e / M;
    T re = T(0), im = T(0);
    if (!alpha_zero) {
      const T tr = t[i + j * 2 * M], ti = t[M + i + j * 2 * M];
      re = ar * tr - ai * ti;
      im = ar * ti + ai * tr;
    }
    T* cp = c + 2 * (i + j * ldc);
    if (!beta_zero) {
      const T cr = cp[0], ci = cp[1];
      re += br * cr - bi * ci;
      im += br * ci + bi * cr;
    }
    cp[0] = re;
    cp[1] = im;
  }
}

template <typename T>
pb_status_t cgemm_entry(pb_handle_t h, char ta, char tb, int64_t M, int64_t N, int64_t K, const T* alpha,
                        const T* a, int64_t lda, const T* b, int64_t ldb, const T* beta, T* c, int64_t ldc) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(alpha != nullptr && beta != nullptr, "null scalar pointer");
  const char nta = pb::norm_trans(ta), ntb = pb::norm_trans(tb);
  PB_REQUIRE(nta, "invalid _TransA");
  PB_REQUIRE(ntb, "invalid _TransB");
  PB_REQUIRE(M >= 0 && N >= 0 && K >= 0, "negative dimension");
  if (M == 0 || N == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(ldc >= M && lda >= (nta == 'n' ? M : K) && ldb >= (ntb == 'n' ? K : N) && lda >= 1 && ldb >= 1,
             "leading dimension too small");
  PB_REQUIRE(c != nullptr, "null C");
  const bool alpha_zero = alpha[0] == T(0) && alpha[1] == T(0);
  const bool beta_zero = beta[0] == T(0) && beta[1] == T(0);
  int64_t g = pb::ceil_div(M * N, 256);
  if (g > 8 * h->num_cus) g = 8 * h->num_cus;
  if (alpha_zero || K == 0) {
    hipLaunchKernelGGL(cplx_epilogue_kernel<T>, dim3(unsigned(g)), dim3(256), 0, h->stream, M, N, T(0), T(0),
                       beta[0], beta[1], static_cast<const T*>(nullptr), c, ldc, beta_zero, true);
    PB_CHECK_LAUNCH();
    return PB_STATUS_SUCCESS;
  }
  PB_REQUIRE(a != nullptr && b != nullptr, "null matrix pointer");
  GemmCall<T> gc;
  gc.ta = 'n'; gc.tb = 'n';
  gc.M = 2 * M; gc.N = N; gc.K = 2 * K;
  gc.alpha = T(1); gc.beta = T(0);
  gc.lda = 2 * M; gc.ldb = 2 * K; gc.ldc = 2 * M;  // the packed problem's (read by a selector)
  const size_t part = pb::align_up(plain_partial_bytes<T>(h, gc), 256);
  const size_t ahat_b = pb::align_up(size_t(4) * M * K * sizeof(T), 256);
  const size_t bhat_b = pb::align_up(size_t(2) * K * N * sizeof(T), 256);
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, part + ahat_b + bhat_b + size_t(2) * M * N * sizeof(T), &ws));
  char* base = static_cast<char*>(ws);
  T* ahat = reinterpret_cast<T*>(base + part);
  T* bhat = reinterpret_cast<T*>(base + part + ahat_b);
  T* that = reinterpret_cast<T*>(base + part + ahat_b + bhat_b);
  // op(A)(i, k) at a[2 (i rs + k ks)], op(B)(k, j) at b[2 (k ks + j ns)] (complex elements).
  const int64_t a_rs = nta == 'n' ? 1 : lda, a_ks = nta == 'n' ? lda : 1;
  const int64_t b_ks = ntb == 'n' ? 1 : ldb, b_ns = ntb == 'n' ? ldb : 1;
  int64_t ga = std::min<int64_t>(pb::ceil_div(M * K, 256), 8 * h->num_cus);
  int64_t gb = std::min<int64_t>(pb::ceil_div(K * N, 256), 8 * h->num_cus);
  hipLaunchKernelGGL(cplx_pack_a_kernel<T>, dim3(unsigned(ga)), dim3(256), 0, h->stream, M, K, a, a_rs, a_ks, ahat);
  hipLaunchKernelGGL(cplx_pack_b_kernel<T>, dim3(unsigned(gb)), dim3(256), 0, h->stream, K, N, b, b_ks, b_ns, bhat);
  PB_CHECK_LAUNCH();
  gc.a = ahat; gc.lda = 2 * M;
  gc.b = bhat; gc.ldb = 2 * K;
  gc.c = that; gc.ldc = 2 * M;
  PB_TRY(run_plain<T>(h, gc, false, part ? static_cast<T*>(ws) : nullptr));
  hipLaunchKernelGGL(cplx_epilogue_kernel<T>, dim3(unsigned(g)), dim3(256), 0, h->stream, M, N, alpha[0],
                     alpha[1], beta[0], beta[1], static_cast<const T*>(that), c, ldc, beta_zero, false);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // namespace pbgemm

// ---------------------------------------------------------------------------
// Scratch of the allocation-free SGEMM (pb_sgemm_noalloc): the split-K partial
// tiles, in device memory the library owns from load to process exit (one
// copy per GPU).  A plain FP32 GEMM splits K only while its 64 x 64 tiles are
// fewer than the compute units, into ceil(CUs / tiles) splits, so it never
// writes more than 2 * CUs - 1 partial tiles of 16 KiB: 8 MiB on 256 CUs.
// Every split writes its whole tile before the fold reads it, so nothing here
// needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace pbgemm {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device (which must be `device`),
// looked up on the first call per device and kept.  The array's name in host
// code is only a host-side stand-in, never a device address.
pb_status_t scratch_address(int device, void** out) {
  PB_REQUIRE(device >= 0 && device < kMaxDevices, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    PB_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return PB_STATUS_SUCCESS;
}
}  // namespace
}  // namespace pbgemm

using namespace pbgemm;

extern "C" {

int pb_tune_num_plain_configs(void) { return kPlainConfigs; }

const char* pb_tune_plain_config_name(int id) {
  static const char* names[kPlainConfigs] = {
      "128x128/splitk1", "128x128/splitk2", "128x128/splitk4", "128x128/splitk8", "128x128/splitk16",
      "128x128/splitk32", "128x128/splitk64", "128x128/splitk128", "64x64/splitk1", "64x64/splitk2",
      "64x64/splitk4", "64x64/splitk8", "64x64/splitk16", "64x64/splitk32", "64x64/splitk64",
      "64x64/splitk128", "small-batch"};
  return (id >= 0 && id < kPlainConfigs) ? names[id] : "";
}

pb_status_t pb_gemm_plain_config(pb_handle_t h, int dtype, int id, char ta, char tb, int64_t m, int64_t n,
                                 int64_t k, double alpha, const void* a, int64_t lda, const void* b,
                                 int64_t ldb, double beta, void* c, int64_t ldc, int64_t batch) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(dtype == 1 || dtype == 2, "dtype must be 1 (FP32) or 2 (FP64)");
  PB_REQUIRE(id >= -1 && id < kPlainConfigs, "config id %d out of range", id);
  auto run = [&](auto zero) -> pb_status_t {
    using T = decltype(zero);
    int64_t sa = 0, sb = 0, sc = 0;
    default_strides(ta, tb, m, n, k, lda, ldb, ldc, PB_BATCH_STRIDED, sa, sb, sc);
    GemmCall<T> g;
    bool done = false;
    PB_TRY(gemm_front<T>(h, ta, tb, m, n, k, T(alpha), static_cast<const T*>(a), lda, sa,
                         static_cast<const T*>(b), ldb, sb, T(beta), static_cast<T*>(c), ldc, sc, batch,
                         PB_BATCH_STRIDED, g, done));
    if (done) return PB_STATUS_SUCCESS;
    if (g.alpha == T(0)) return alpha_zero<T>(h, g.M, g.N, g.beta, g.c, g.ldc, g.sc, g.batch, false);
    if (id == kPlainSmall && !small::eligible(h, g.M, g.N, g.batch, g.interleaved))
      return pb::set_error(PB_STATUS_NOT_SUPPORTED, "the small-batch kernel needs M, N <= 64 and a batch "
                           "of at least one problem per CU");
    return run_plain<T>(h, g, false, nullptr, id < 0 ? kRuleOnly : id);
  };
  return dtype == 1 ? run(0.f) : run(0.0);
}

pb_status_t pb_tune_time_plain(pb_handle_t h, int dtype, int id, char ta, char tb, int64_t m, int64_t n,
                               int64_t k, const void* a, int64_t lda, const void* b, int64_t ldb, void* c,
                               int64_t ldc, int64_t batch, int warmup, int reps, float* usec_out) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(usec_out != nullptr && reps > 0 && warmup >= 0, "bad timing arguments");
  for (int i = 0; i < warmup; ++i)
    PB_TRY(pb_gemm_plain_config(h, dtype, id, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc, batch));
  hipEvent_t e0, e1;
  PB_CHECK_HIP(hipEventCreate(&e0));
  PB_CHECK_HIP(hipEventCreate(&e1));
  PB_CHECK_HIP(hipEventRecord(e0, h->stream));
  pb_status_t st = PB_STATUS_SUCCESS;
  for (int i = 0; i < reps && st == PB_STATUS_SUCCESS; ++i)
    st = pb_gemm_plain_config(h, dtype, id, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc, batch);
  hipError_t e = hipEventRecord(e1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != PB_STATUS_SUCCESS) return st;
  if (e != hipSuccess)
    return pb::set_error(PB_STATUS_DEVICE_ERROR, "timing failed: %s", hipGetErrorString(e));
  *usec_out = 1000.f * ms / float(reps);
  return PB_STATUS_SUCCESS;
}

int pb_tune_select_plain(pb_handle_t h, int dtype, char ta, char tb, int64_t m, int64_t n, int64_t k,
                         int64_t batch, int64_t lda) {
  if (!h || h->plain_selector.n_layers == 0 || !plain_selector_domain(m, n, k, batch)) return -1;
  float f[8];
  plain_features(dtype, pb::norm_trans(ta), pb::norm_trans(tb), m, n, k, batch, lda, f);
  return selector_eval(h->plain_selector, f);
}

pb_status_t pb_sgemm_noalloc(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                             float alpha, const float* a, int64_t lda, const float* b,
                             int64_t ldb, float beta, float* c, int64_t ldc) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  GemmCall<float> g;
  bool done = false;
  PB_TRY(gemm_front<float>(h, ta, tb, m, n, k, alpha, a, lda, 0, b, ldb, 0, beta, c, ldc, 0, 1,
                           PB_BATCH_STRIDED, g, done));
  if (done) return PB_STATUS_SUCCESS;
  if (g.alpha == 0.f)
    return alpha_zero<float>(h, g.M, g.N, g.beta, g.c, g.ldc, g.sc, g.batch, g.interleaved);
  // Same plan and kernels as pb_sgemm; only the partials' home differs.  A
  // plan from an installed selector that needs more partials than the
  // scratch holds gives way to the built-in rule (at most 2 CUs - 1 tiles).
  int forced = -1;
  size_t need = plain_partial_bytes<float>(h, g);
  if (need > kNoAllocScratchBytes) {
    forced = kRuleOnly;
    need = plain_partial_bytes<float>(h, g, forced);
  }
  if (need > kNoAllocScratchBytes)
    return pb::set_error(PB_STATUS_ALLOC_FAILED,
                         "pb_sgemm_noalloc: %zu bytes of split-K partials exceed the %zu-byte "
                         "scratch",
                         need, kNoAllocScratchBytes);
  void* part = nullptr;
  PB_TRY(scratch_address(h->device, &part));
  return run_plain<float>(h, g, false, static_cast<float*>(part), forced);
}

pb_status_t pb_cgemm(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, const float* alpha,
                     const float* a, int64_t lda, const float* b, int64_t ldb, const float* beta, float* c,
                     int64_t ldc) {
  return cgemm_entry<float>(h, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

pb_status_t pb_zgemm(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, const double* alpha,
                     const double* a, int64_t lda, const double* b, int64_t ldb, const double* beta, double* c,
                     int64_t ldc) {
  return cgemm_entry<double>(h, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

pb_status_t pb_strsm(pb_handle_t h, char side, char uplo, char trans, char diag, int64_t m, int64_t n,
                     float alpha, const float* a, int64_t lda, float* b, int64_t ldb) {
  return trsm_entry<float>(h, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb);
}

pb_status_t pb_dtrsm(pb_handle_t h, char side, char uplo, char trans, char diag, int64_t m, int64_t n,
                     double alpha, const double* a, int64_t lda, double* b, int64_t ldb) {
  return trsm_entry<double>(h, side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb);
}

pb_status_t pb_ssymm(pb_handle_t h, char side, char uplo, int64_t m, int64_t n, float alpha,
                     const float* a, int64_t lda, const float* b, int64_t ldb, float beta, float* c,
                     int64_t ldc) {
  return symm_entry<float>(h, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}

pb_status_t pb_dsymm(pb_handle_t h, char side, char uplo, int64_t m, int64_t n, double alpha,
                     const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                     double* c, int64_t ldc) {
  return symm_entry<double>(h, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}

pb_status_t pb_sgemm(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                     float alpha, const float* a, int64_t lda, const float* b, int64_t ldb,
                     float beta, float* c, int64_t ldc) {
  // internal::_gemm: batch 1, strides 0 (gemm_interface.hpp:189-198).
  return gemm_entry<float>(h, ta, tb, m, n, k, alpha, a, lda, 0, b, ldb, 0, beta, c, ldc, 0, 1,
                           PB_BATCH_STRIDED, false);
}

pb_status_t pb_dgemm(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                     double alpha, const double* a, int64_t lda, const double* b, int64_t ldb,
                     double beta, double* c, int64_t ldc) {
  return gemm_entry<double>(h, ta, tb, m, n, k, alpha, a, lda, 0, b, ldb, 0, beta, c, ldc, 0, 1,
                            PB_BATCH_STRIDED, false);
}

pb_status_t pb_sgemm_batched(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                             float alpha, const float* a, int64_t lda, const float* b,
                             int64_t ldb, float beta, float* c, int64_t ldc, int64_t batch,
                             pb_batch_type_t bt) {
  int64_t sa = 0, sb = 0, sc = 0;
  default_strides(ta, tb, m, n, k, lda, ldb, ldc, bt, sa, sb, sc);
  return gemm_entry<float>(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                           batch, bt, false);
}

pb_status_t pb_dgemm_batched(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                             double alpha, const double* a, int64_t lda, const double* b,
                             int64_t ldb, double beta, double* c, int64_t ldc, int64_t batch,
                             pb_batch_type_t bt) {
  int64_t sa = 0, sb = 0, sc = 0;
  default_strides(ta, tb, m, n, k, lda, ldb, ldc, bt, sa, sb, sc);
  return gemm_entry<double>(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                            batch, bt, false);
}

pb_status_t pb_sgemm_strided_batched(pb_handle_t h, char ta, char tb, int64_t m, int64_t n,
                                     int64_t k, float alpha, const float* a, int64_t lda,
                                     int64_t sa, const float* b, int64_t ldb, int64_t sb,
                                     float beta, float* c, int64_t ldc, int64_t sc,
                                     int64_t batch) {
  return gemm_entry<float>(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                           batch, PB_BATCH_STRIDED, false);
}

pb_status_t pb_dgemm_strided_batched(pb_handle_t h, char ta, char tb, int64_t m, int64_t n,
                                     int64_t k, double alpha, const double* a, int64_t lda,
                                     int64_t sa, const double* b, int64_t ldb, int64_t sb,
                                     double beta, double* c, int64_t ldc, int64_t sc,
                                     int64_t batch) {
  return gemm_entry<double>(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                            batch, PB_BATCH_STRIDED, false);
}

}  // extern "C"

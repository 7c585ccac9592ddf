// Grouped (uneven) GEMM: many independent problems of different shapes in
// ONE persistent launch per operand layout.  portBLAS has no counterpart (its
// batches share one shape: src/interface/gemm_interface.hpp:202-240, with the
// in-kernel batch loop of src/operations/blas3/gemm_local.hpp:273-279,511-516).
//
// The plan uploads one descriptor per problem and the prefix sum of their
// tile counts.  A persistent grid (a few workgroups per CU) pulls tile indices
// from a device-scope work queue, finds the owning problem by binary search
// over the prefix sums, and runs the MFMA tile engine on that tile (the fast
// path for interior tiles of vector-aligned problems, the guarded one
// otherwise).  The last workgroup to leave resets the queue, so the launch
// needs no memset and can be captured in a graph and replayed.
#include "gemm_front.hpp"
#include "gemm_launch.hpp"

#include <algorithm>
#include <vector>

namespace pbgemm {

struct DevProblem {
  int64_t M, N, K;
  float alpha, beta;
  const void* a;
  int64_t a_rs, a_ks;
  const void* b;
  int64_t b_rs, b_ks;
  float* c;
  int64_t ldc;
  int64_t tile_begin;
  int32_t tiles_m, tiles_n;
  int32_t vec, c_vec;
};

template <typename T, typename Shape, bool KCA, bool KCB>
__global__ __launch_bounds__(256) void grouped_kernel(const DevProblem* __restrict__ probs,
                                                      int count, int64_t total_tiles,
                                                      unsigned int* queue) {
  using TL = Tile<T, Shape, KCA, KCB>;
  using Raw = typename Traits<T>::Raw;
  __shared__ __attribute__((aligned(16))) Raw lds[2 * TL::STAGE_ELEMS];
  __shared__ int64_t s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = int64_t(atomicAdd(&queue[0], 1u));
    __syncthreads();
    const int64_t t = s_tile;
    __syncthreads();  // everyone has read s_tile before it is rewritten
    if (t >= total_tiles) break;
    // Owning problem: last p with tile_begin <= t.
    int lo = 0, hi = count - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    const DevProblem& P = probs[lo];
    const int64_t local = t - P.tile_begin;
    int64_t tm, tn;
    tile_coords(local, P.tiles_m, P.tiles_n, tm, tn);
    const int64_t m0 = tm * TL::BM, n0 = tn * TL::BN;
    Operand<T> A{static_cast<const Raw*>(P.a), P.a_rs, P.a_ks, P.M};
    Operand<T> B{static_cast<const Raw*>(P.b), P.b_rs, P.b_ks, P.N};
    typename TL::AccV acc[TL::FM][TL::FN];
    TL::zero(acc);
    const bool interior = P.vec && m0 + TL::BM <= P.M && n0 + TL::BN <= P.N;
    if (interior) TL::template mainloop<false>(A, B, m0, n0, 0, P.K, lds, acc);
    else TL::template mainloop<true>(A, B, m0, n0, 0, P.K, lds, acc);
    TL::epilogue(acc, P.alpha, P.beta, P.c, 1, P.ldc, P.M, P.N, m0, n0, P.c_vec != 0);
  }
  // Last workgroup out resets the queue and the exit counter.
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(&queue[1], 1u);
    if (prev == gridDim.x - 1) {
      queue[0] = 0u;
      queue[1] = 0u;
      __threadfence();
    }
  }
}

// BF16 problems: a lean 64 x 64 tile.  18 KB of LDS and few registers let
// about 8 workgroups share a CU, so the latency of one tile's loads is
// hidden by the other resident tiles rather than by deep pipelining (most
// grouped problems are a handful of tiles and a few K chunks).  Per K chunk
// of 64 every thread issues its 2 + 2 16-byte loads, then writes them to LDS
// K-contiguous ([row][k], pitch 72): a K-contiguous operand with one
// 16-byte store, an M/N-contiguous one transposed with eight 2-byte stores.
// Each wave computes a 32 x 32 quarter with v_mfma_f32_32x32x16_bf16.
constexpr int kLT = 64, kLP = 72;

template <bool KC>
__device__ __forceinline__ void lean_load(const uint16_t* x, int64_t rs, int64_t ks, int64_t rows,
                                          int64_t K, int64_t r0, int64_t k0, bool vec, int e,
                                          u32x4& out) {
  // e in [0, 512): KC -> row e / 8, k group (e % 8) * 8; else k e / 8, row group (e % 8) * 8.
  const int64_t r = KC ? r0 + e / 8 : r0 + (e % 8) * 8;
  const int64_t k = KC ? k0 + (e % 8) * 8 : k0 + e / 8;
  if (vec && (KC ? (r < rows && k + 8 <= K) : (k < K && r + 8 <= rows))) {
    out = *reinterpret_cast<const u32x4*>(x + r * rs + k * ks);
    return;
  }
  // Edge or unaligned: element by element (kept rolled: it is rare, and
  // unrolled it would cost the registers that keep 3+ workgroups per CU).
  uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll 1
  for (int i = 0; i < 8; ++i) {
    const int64_t ri = KC ? r : r + i, ki = KC ? k + i : k;
    const uint32_t v = (ri < rows && ki < K) ? x[ri * rs + ki * ks] : 0u;
    w[i / 2] |= v << (16 * (i % 2));
  }
  out = u32x4{w[0], w[1], w[2], w[3]};
}

template <bool KC>
__device__ __forceinline__ void lean_store(uint16_t (*dst)[kLP], int e, const u32x4& v) {
  if constexpr (KC) {
    *reinterpret_cast<u32x4*>(&dst[e / 8][(e % 8) * 8]) = v;
  } else {
    uint16_t h[8];
    __builtin_memcpy(h, &v, 16);
#pragma unroll
    for (int i = 0; i < 8; ++i) dst[(e % 8) * 8 + i][e / 8] = h[i];
  }
}

template <bool KCA, bool KCB>
__global__ __launch_bounds__(256) void grouped_lean_kernel(const DevProblem* __restrict__ probs,
                                                           int count, int64_t total_tiles,
                                                           unsigned int* queue) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kLT][kLP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kLT][kLP];
  __shared__ int64_t s_tile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  for (;;) {
    if (threadIdx.x == 0) s_tile = int64_t(atomicAdd(&queue[0], 1u));
    __syncthreads();
    const int64_t t = s_tile;
    __syncthreads();  // everyone has read s_tile before it is rewritten
    if (t >= total_tiles) break;
    int lo = 0, hi = count - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    const DevProblem& P = probs[lo];
    int64_t tm, tn;
    tile_coords(t - P.tile_begin, P.tiles_m, P.tiles_n, tm, tn);
    const int64_t m0 = tm * kLT, n0 = tn * kLT;
    const auto* A = static_cast<const uint16_t*>(P.a);
    const auto* B = static_cast<const uint16_t*>(P.b);
    const bool vec = P.vec != 0;
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.f;
    for (int64_t k0 = 0; k0 < P.K; k0 += kLT) {
      u32x4 va[2], vb[2];
#pragma unroll
      for (int it = 0; it < 2; ++it) {
        const int e = threadIdx.x + 256 * it;
        lean_load<KCA>(A, P.a_rs, P.a_ks, P.M, P.K, m0, k0, vec, e, va[it]);
        lean_load<KCB>(B, P.b_rs, P.b_ks, P.N, P.K, n0, k0, vec, e, vb[it]);
      }
      if (k0 > 0) __syncthreads();  // the previous chunk is consumed
#pragma unroll
      for (int it = 0; it < 2; ++it) {
        const int e = threadIdx.x + 256 * it;
        lean_store<KCA>(As, e, va[it]);
        lean_store<KCB>(Bs, e, vb[it]);
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < kLT; kk += 16) {
        bf16x8 a, b;
        __builtin_memcpy(&a, &As[wm + lane % 32][kk + 8 * (lane / 32)], 16);
        __builtin_memcpy(&b, &Bs[wn + lane % 32][kk + 8 * (lane / 32)], 16);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
    }
    __syncthreads();  // LDS free for the next tile's first chunk
    // Epilogue: lane holds column wn + lane % 32, rows wm + 8 (e / 4) +
    // 4 (lane / 32) + e % 4.
    const int64_t n = n0 + wn + lane % 32;
    if (n < P.N) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t m = m0 + wm + 8 * q + 4 * (lane / 32);
        float* cp = P.c + m + n * P.ldc;
        if (P.c_vec && m + 3 < P.M) {
          f32x4 o;
          if (P.beta != 0.f) {
            const f32x4 old = *reinterpret_cast<const f32x4*>(cp);
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = P.alpha * acc[4 * q + e] + P.beta * old[e];
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = P.alpha * acc[4 * q + e];
          }
          *reinterpret_cast<f32x4*>(cp) = o;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (m + e < P.M) {
              const float v = P.alpha * acc[4 * q + e];
              cp[e] = (P.beta != 0.f) ? v + P.beta * cp[e] : v;
            }
        }
      }
    }
  }
  // Last workgroup out resets the queue and the exit counter.
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(&queue[1], 1u);
    if (prev == gridDim.x - 1) {
      queue[0] = 0u;
      queue[1] = 0u;
      __threadfence();
    }
  }
}

using G64 = TileShape<64, 64, 64, 32, 32>;
using GF64 = TileShape<64, 64, 32, 32, 32>;

struct Launch {
  int layout;   // 2 * kca + kcb
  int count;
  int64_t tiles;
  DevProblem* dev;
};

}  // namespace pbgemm

struct pb_grouped_plan_s {
  pb_gemm_precision_t precision;
  bool big_tiles = false;
  std::vector<pbgemm::Launch> launches;
  void* dev_mem = nullptr;
  int device = 0;
};

namespace pbgemm {

void launch_lean(pb_handle_t h, const Launch& L, unsigned grid, unsigned int* queue) {
  switch (L.layout) {
    case 3: hipLaunchKernelGGL((grouped_lean_kernel<true, true>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
    case 2: hipLaunchKernelGGL((grouped_lean_kernel<true, false>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
    case 1: hipLaunchKernelGGL((grouped_lean_kernel<false, true>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
    default: hipLaunchKernelGGL((grouped_lean_kernel<false, false>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
  }
}

template <typename T, typename Shape>
void launch_grouped(pb_handle_t h, const Launch& L, unsigned grid, unsigned int* queue) {
  switch (L.layout) {
    case 3: hipLaunchKernelGGL((grouped_kernel<T, Shape, true, true>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
    case 2: hipLaunchKernelGGL((grouped_kernel<T, Shape, true, false>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
    case 1: hipLaunchKernelGGL((grouped_kernel<T, Shape, false, true>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
    default: hipLaunchKernelGGL((grouped_kernel<T, Shape, false, false>), dim3(grid), dim3(256), 0, h->stream, L.dev, L.count, L.tiles, queue); break;
  }
}

}  // namespace pbgemm

using namespace pbgemm;

extern "C" {

pb_status_t pb_gemm_grouped_plan_create(pb_handle_t h, pb_gemm_precision_t prec,
                                        const pb_gemm_problem_t* problems, int64_t count,
                                        pb_grouped_plan_t* out) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(out != nullptr, "null plan pointer");
  *out = nullptr;
  PB_REQUIRE(prec == PB_GEMM_BF16 || prec == PB_GEMM_FP32_MFMA,
             "grouped GEMM supports PB_GEMM_BF16 and PB_GEMM_FP32_MFMA inputs");
  PB_REQUIRE(count >= 0 && count <= (1 << 24), "bad problem count");
  PB_REQUIRE(count == 0 || problems != nullptr, "null problem array");
  const bool bf16 = prec == PB_GEMM_BF16;
  const int vw = bf16 ? 8 : 4;
  const size_t esz = bf16 ? 2 : 4;
  // Tile size: 128 x 128 only when most problems fill such tiles.
  int64_t big = 0;
  for (int64_t i = 0; i < count; ++i)
    if (problems[i].m >= 512 && problems[i].n >= 512) ++big;
  auto* plan = new pb_grouped_plan_s();
  plan->precision = prec;
  plan->device = h->device;
  plan->big_tiles = false;  // BF16: the lean 64 x 64 kernel; FP32: 64 x 64 tile engine
  (void)big;
  const int bm = plan->big_tiles ? 128 : 64, bn = bm;
  std::vector<DevProblem> by_layout[4];
  for (int64_t i = 0; i < count; ++i) {
    const pb_gemm_problem_t& q = problems[i];
    const char ta = pb::norm_trans(q.transa), tb = pb::norm_trans(q.transb);
    auto fail = [&](const char* what) {
      delete plan;
      return pb::set_error(PB_STATUS_INVALID_VALUE, "problem %lld: %s", (long long)i, what);
    };
    if (!ta || !tb) return fail("invalid transpose flag");
    if (q.m < 0 || q.n < 0 || q.k < 0) return fail("negative dimension");
    if (q.m == 0 || q.n == 0) continue;
    if (q.ldc < q.m || q.lda < (ta == 'n' ? q.m : q.k) || q.ldb < (tb == 'n' ? q.k : q.n))
      return fail("leading dimension too small");
    if (!q.c || (q.k > 0 && (!q.a || !q.b))) return fail("null pointer");
    DevProblem d{};
    d.M = q.m; d.N = q.n; d.K = q.k;
    d.alpha = q.alpha; d.beta = q.beta;
    d.a = q.a; d.b = q.b; d.c = q.c; d.ldc = q.ldc;
    d.a_rs = ta == 'n' ? 1 : q.lda; d.a_ks = ta == 'n' ? q.lda : 1;
    d.b_rs = tb == 'n' ? q.ldb : 1; d.b_ks = tb == 'n' ? 1 : q.ldb;
    d.tiles_m = int32_t(pb::ceil_div(q.m, bm));
    d.tiles_n = int32_t(pb::ceil_div(q.n, bn));
    d.vec = aligned16(q.a) && aligned16(q.b) && q.lda % vw == 0 && q.ldb % vw == 0;
    d.c_vec = aligned16(q.c) && q.ldc % 4 == 0;
    (void)esz;
    by_layout[2 * (ta == 't') + (tb == 'n')].push_back(d);
  }
  size_t total = 0;
  for (auto& v : by_layout) total += v.size();
  if (total > 0) {
    hipError_t e = hipMalloc(&plan->dev_mem, total * sizeof(DevProblem));
    if (e != hipSuccess) {
      delete plan;
      return pb::set_error(PB_STATUS_ALLOC_FAILED, "plan allocation failed: %s", hipGetErrorString(e));
    }
  }
  DevProblem* dst = static_cast<DevProblem*>(plan->dev_mem);
  for (int lay = 3; lay >= 0; --lay) {
    auto& v = by_layout[lay];
    if (v.empty()) continue;
    // Largest problems first: their tiles are handed out first, so the tail
    // of the queue holds the small ones (better balance at the end).
    std::stable_sort(v.begin(), v.end(), [](const DevProblem& x, const DevProblem& y) {
      return x.M * x.N * x.K > y.M * y.N * y.K;
    });
    int64_t t = 0;
    for (auto& d : v) {
      d.tile_begin = t;
      t += int64_t(d.tiles_m) * d.tiles_n;
    }
    hipError_t e = hipMemcpy(dst, v.data(), v.size() * sizeof(DevProblem), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(plan->dev_mem);
      delete plan;
      return pb::set_error(PB_STATUS_DEVICE_ERROR, "plan upload failed: %s", hipGetErrorString(e));
    }
    plan->launches.push_back({lay, int(v.size()), t, dst});
    dst += v.size();
  }
  *out = plan;
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_gemm_grouped_plan_run(pb_handle_t h, pb_grouped_plan_t plan) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(plan != nullptr, "null plan");
  PB_REQUIRE(plan->device == h->device, "plan belongs to another device");
  unsigned int* queue = pb::counters(h, 2);
  for (const Launch& L : plan->launches) {
    const int64_t per_cu = plan->precision == PB_GEMM_BF16 ? 6 : 4;  // resident workgroups per CU
    int64_t grid = std::min<int64_t>(L.tiles, per_cu * h->num_cus);
    if (grid < 1) grid = 1;
    if (plan->precision == PB_GEMM_BF16) {
      launch_lean(h, L, unsigned(grid), queue);
    } else {
      launch_grouped<float, GF64>(h, L, unsigned(grid), queue);
    }
    PB_CHECK_LAUNCH();
  }
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_gemm_grouped_plan_destroy(pb_grouped_plan_t plan) {
  pb::clear_error();
  if (!plan) return PB_STATUS_SUCCESS;
  int cur = 0;
  if (hipGetDevice(&cur) == hipSuccess && cur != plan->device) (void)hipSetDevice(plan->device);
  if (plan->dev_mem) (void)hipFree(plan->dev_mem);
  if (cur != plan->device) (void)hipSetDevice(cur);
  delete plan;
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_gemm_grouped(pb_handle_t h, pb_gemm_precision_t prec,
                            const pb_gemm_problem_t* problems, int64_t count) {
  pb_grouped_plan_t plan = nullptr;
  PB_TRY(pb_gemm_grouped_plan_create(h, prec, problems, count, &plan));
  pb_status_t st = pb_gemm_grouped_plan_run(h, plan);
  hipError_t e = hipStreamSynchronize(h->stream);
  pb_gemm_grouped_plan_destroy(plan);
  if (st != PB_STATUS_SUCCESS) return st;
  if (e != hipSuccess)
    return pb::set_error(PB_STATUS_DEVICE_ERROR, "grouped GEMM failed: %s", hipGetErrorString(e));
  return PB_STATUS_SUCCESS;
}

}  // extern "C"

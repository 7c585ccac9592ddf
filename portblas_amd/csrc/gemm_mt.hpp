// Large-tile BF16 MFMA kernel for packed operands (the hot GEMM of the
// BF16x3 / BF16 paths at large M, N).
//
// Operands are the pre-pass output: op(A) as [M][kp] and op(B) as [N][kp]
// BF16, K-contiguous, kp a multiple of 64 with zeros in [K, kp).  Up to two
// parts per operand (hi, lo); TERMS products are accumulated per K step:
//   TERMS 1: Ah Bh
//   TERMS 3: Ah Bl + Al Bh + Ah Bh           (BF16x3)
//   TERMS 4: Ah Bl + Al Bh + Al Bl + Ah Bh   (BF16x4: adds the lo*lo term)
//
// Workgroup: 256 threads = 4 waves (2 x 2), C tile 256 x 256, wave tile
// 128 x 128 = 4 x 4 v_mfma_f32_32x32x16_bf16 tiles (256 FP32 accumulators per
// lane).  Per K step of BK:
//   * global -> VGPR: every thread loads 16-byte chunks with buffer loads
//     (the buffer range is the operand, so rows past M / N read as zero and
//     no edge code is needed in the main loop);
//   * VGPR -> LDS: double-buffered stages, rows of BK elements without
//     padding; the 16-byte chunk c of row r is stored at slot c ^ swz(r), so
//     both the stores (8 lanes fill one 128-byte row segment) and the MFMA
//     fragment reads (8 lanes read 8 different rows) hit 8 distinct 16-byte
//     bank groups;
//   * LDS -> MFMA: fragments of the next 16-k group are read while the MFMAs
//     of the current group run; the stage barrier sits before the last group
//     so the first fragments of the next stage are in flight during it.
// One barrier per K step; the global loads for step k+2 are issued at the
// start of step k.
#pragma once

#include "gemm_core.hpp"

namespace pbgemm {
namespace mt {

typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 256, kBN = 256;

struct Problem {
  int M, N, kp;          // kp: packed row pitch (elements), multiple of BK
  float alpha, beta;
  const float* alpha_dev;  // optional device factor multiplied into alpha
  const uint16_t* a[2];    // hi, lo parts of op(A): [batch][M][kp]
  const uint16_t* b[2];    // hi, lo parts of op(B): [batch][N][kp]
  int64_t a_stride, b_stride;  // batch strides (elements)
  float* c;
  int64_t ldc, c_stride;
  int batch;
  int c_vec;             // C 16-byte aligned with ldc % 4 == 0
  int tiles_m, tiles_n;
};

template <int BK, int TERMS>
struct Cfg {
  static constexpr int NPARTS = TERMS == 1 ? 1 : 2;
  static constexpr int CH = BK / 8;                // 16-byte chunks per row
  static constexpr int ROW_BYTES = BK * 2;
  static constexpr int TILE_BYTES = kBM * ROW_BYTES;  // one part of one operand
  static constexpr int LOADS = TILE_BYTES / (16 * 256);  // chunks per thread per part
  static constexpr int ROWS_PER_LOAD = 256 / CH;   // rows covered by one chunk index
  static constexpr int STAGE_BYTES = 2 * NPARTS * TILE_BYTES;
  static constexpr int LDS_BYTES = 2 * STAGE_BYTES;
  static constexpr int GROUPS = BK / 16;
  static_assert(BK == 32 || BK == 64, "BK");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");
};

// Slot permutation: rows that share a 128-byte bank line set get different
// chunk slots.
template <int CH>
__device__ __forceinline__ int swz(int row) {
  return (row / (8 / CH)) % CH;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  // Raw buffer (stride 0); third dword = default data format for gfx9 parts.
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(bytes),
                                           0x00020000);
}

// LDS-only workgroup barrier: waits for this wave's LDS operations, not for
// its outstanding global loads (those land in registers two steps ahead).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Scheduling: the next N instructions of class MASK (0x020 global load,
// 0x100 LDS read, 0x200 LDS write) go PER at a time between consecutive
// MFMAs.
template <int N, int MASK, int PER>
__device__ __forceinline__ void interleave() {
  static_assert(N % PER == 0, "uneven interleave");
#pragma unroll
  for (int i = 0; i < N / PER; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(MASK, PER, 0);
  }
}

template <int TERMS> constexpr int kFragReads = TERMS == 1 ? 8 : 16;
template <int BK, int TERMS>
constexpr int kStoresPerStep = 2 * (TERMS == 1 ? 1 : 2) * (256 * BK * 2 / (16 * 256));

template <int BK, int TERMS>
__global__ __launch_bounds__(256, 1) void mt_kernel(Problem p) {
  using C = Cfg<BK, TERMS>;
  constexpr int NP = C::NPARTS;
  constexpr int G = C::GROUPS;
  __shared__ __attribute__((aligned(16))) uint8_t lds[C::LDS_BYTES];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave & 1) * 128, wn = (wave >> 1) * 128;

  int64_t tm64, tn64;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tm64, tn64);
  const int m0 = int(tm64) * kBM, n0 = int(tn64) * kBN;
  const int bidx = blockIdx.y;

  // Buffer resources: one per part per operand; the range is the operand.
  __amdgpu_buffer_rsrc_t ra[NP], rb[NP];
  const uint32_t a_bytes = uint32_t(p.M) * uint32_t(p.kp) * 2u;
  const uint32_t b_bytes = uint32_t(p.N) * uint32_t(p.kp) * 2u;
#pragma unroll
  for (int t = 0; t < NP; ++t) {
    ra[t] = make_rsrc(p.a[t] + int64_t(bidx) * p.a_stride, a_bytes);
    rb[t] = make_rsrc(p.b[t] + int64_t(bidx) * p.b_stride, b_bytes);
  }
  // This thread's chunks: row tid / CH + ROWS_PER_LOAD * i, chunk tid % CH.
  const int lrow = tid / C::CH, lkc = tid % C::CH;
  const uint32_t pitch = uint32_t(p.kp) * 2u;
  const uint32_t voff_a = uint32_t(m0 + lrow) * pitch + uint32_t(lkc) * 16u;
  const uint32_t voff_b = uint32_t(n0 + lrow) * pitch + uint32_t(lkc) * 16u;
  const uint32_t vstep = uint32_t(C::ROWS_PER_LOAD) * pitch;
  // LDS byte offset of this thread's chunk i within a part tile.
  const int st_off = lrow * C::ROW_BYTES + ((lkc ^ swz<C::CH>(lrow)) * 16);

  // Staging registers: [operand part][chunk]; parts 0..NP-1 = A, NP.. = B.
  i32x4 stg[2 * NP][C::LOADS];

  auto gload = [&](int k0) {
    const int soff = k0 * 2;
#pragma unroll
    for (int t = 0; t < NP; ++t)
#pragma unroll
      for (int i = 0; i < C::LOADS; ++i) {
        stg[t][i] = __builtin_amdgcn_raw_buffer_load_b128(ra[t], voff_a + uint32_t(i) * vstep,
                                                          soff, 0);
        stg[NP + t][i] = __builtin_amdgcn_raw_buffer_load_b128(
            rb[t], voff_b + uint32_t(i) * vstep, soff, 0);
      }
  };
  // Part tile order inside a stage: A parts, then B parts.
  auto lstore = [&](int stage) {
    uint8_t* base = lds + stage * C::STAGE_BYTES + st_off;
#pragma unroll
    for (int t = 0; t < 2 * NP; ++t)
#pragma unroll
      for (int i = 0; i < C::LOADS; ++i)
        *reinterpret_cast<i32x4*>(base + t * C::TILE_BYTES +
                                  i * C::ROWS_PER_LOAD * C::ROW_BYTES) = stg[t][i];
  };

  // Fragment read offsets (bytes within a part tile) for group g: row
  // w + 32 i + lane % 32, chunk 2 g + lane / 32.
  const int frow = lane & 31;
  const int fh = lane >> 5;
  const int fswz = swz<C::CH>(frow);  // (w + 32 i) is a multiple of 8
  int foff[G];
#pragma unroll
  for (int g = 0; g < G; ++g) foff[g] = frow * C::ROW_BYTES + (((2 * g + fh) ^ fswz) * 16);

  // Fragments: [buffer][part][i] for A and B.
  bf16x8 fa[2][NP][4], fb[2][NP][4];
  auto fread = [&](int stage, int g, int buf) {
    const uint8_t* base = lds + stage * C::STAGE_BYTES + foff[g];
#pragma unroll
    for (int t = 0; t < NP; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        fa[buf][t][i] = *reinterpret_cast<const bf16x8*>(
            base + t * C::TILE_BYTES + (wm + 32 * i) * C::ROW_BYTES);
        fb[buf][t][i] = *reinterpret_cast<const bf16x8*>(
            base + (NP + t) * C::TILE_BYTES + (wn + 32 * i) * C::ROW_BYTES);
      }
  };

  f32x16 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;

  auto mma = [&](int buf) {
    // Small terms first, the leading Ah Bh term last; term-outer so that
    // consecutive MFMAs write different accumulators.
    if constexpr (TERMS >= 3) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[buf][0][i], fb[buf][1][j],
                                                              acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[buf][1][i], fb[buf][0][j],
                                                              acc[i][j], 0, 0, 0);
    }
    if constexpr (TERMS == 4) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[buf][1][i], fb[buf][1][j],
                                                              acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[buf][0][i], fb[buf][0][j],
                                                            acc[i][j], 0, 0, 0);
  };

  const int nk = p.kp / BK;
  // The loop body is branch-free (so the interleaving below applies to all of
  // it): the loads past the last step re-read the last K tile and the stores
  // and fragment reads they feed are never consumed.
  auto kofs = [&](int kt) { return (kt < nk ? kt : nk - 1) * BK; };
  gload(0);
  lstore(0);
  gload(kofs(1));
  lds_barrier();
  fread(0, 0, 0);

  constexpr int NST = kStoresPerStep<BK, TERMS>;  // LDS stores = global loads per step
  constexpr int NFR = kFragReads<TERMS>;
  for (int kt = 0; kt < nk; ++kt) {
    const int s = kt & 1;
    // Groups before the last: the next stage's LDS stores, the global loads
    // two steps ahead and the fragment reads of the following group are
    // spread between the MFMAs.
    if constexpr (G == 2) {
      lstore(s ^ 1);
      gload(kofs(kt + 2));
      fread(s, 1, 1);
      mma(0);
      interleave<NST, 0x200, 1>();
      interleave<NST, 0x020, 1>();
      interleave<NFR, 0x100, 1>();
      __builtin_amdgcn_sched_barrier(0);
    } else {
      lstore(s ^ 1);
      fread(s, 1, 1);
      mma(0);
      interleave<NST, 0x200, 2>();
      interleave<NFR, 0x100, 1>();
      __builtin_amdgcn_sched_barrier(0);
      gload(kofs(kt + 2));
      fread(s, 2, 0);
      mma(1);
      interleave<NST, 0x020, 2>();
      interleave<NFR, 0x100, 1>();
      __builtin_amdgcn_sched_barrier(0);
      fread(s, 3, 1);
      mma(0);
      interleave<NFR, 0x100, 1>();
      __builtin_amdgcn_sched_barrier(0);
    }
    // Every wave has stored the next stage and read its last fragments of
    // this one: swap stages, then fetch the next step's first group while
    // the last group's MFMAs run.
    lds_barrier();
    fread(s ^ 1, 0, 0);
    mma(1);
    interleave<NFR, 0x100, 1>();
    __builtin_amdgcn_sched_barrier(0);
  }

  // Epilogue: C = alpha * acc + beta * C.  Lane holds column wn + 32 j +
  // lane % 32 and rows wm + 32 i + 8 (e / 4) + 4 (lane / 32) + e % 4.
  float* cb = p.c + int64_t(bidx) * p.c_stride;
  const float alpha = p.alpha_dev ? p.alpha * *p.alpha_dev : p.alpha;
  const float beta = p.beta;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + wn + 32 * j + frow;
    if (n >= p.N) continue;
    float* ccol = cb + int64_t(n) * p.ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + 32 * i + 8 * q + 4 * fh;
        if (p.c_vec && m + 3 < p.M) {
          f32x4 o;
          if (beta != 0.f) {
            const f32x4 old = *reinterpret_cast<const f32x4*>(ccol + m);
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = alpha * acc[i][j][4 * q + e] + beta * old[e];
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = alpha * acc[i][j][4 * q + e];
          }
          *reinterpret_cast<f32x4*>(ccol + m) = o;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if (m + e < p.M) {
              const float v = alpha * acc[i][j][4 * q + e];
              ccol[m + e] = (beta != 0.f) ? v + beta * ccol[m + e] : v;
            }
          }
        }
      }
  }
}

// True when the packed operands fit the kernel's 32-bit buffer offsets.
inline bool fits(int64_t M, int64_t N, int64_t kp) {
  const int64_t lim = int64_t(1) << 31;
  return (M + kBM) * kp * 2 < lim && (N + kBN) * kp * 2 < lim;
}

template <int BK, int TERMS>
pb_status_t launch(pb_handle_t h, Problem p) {
  p.tiles_m = int(pb::ceil_div(p.M, kBM));
  p.tiles_n = int(pb::ceil_div(p.N, kBN));
  PB_REQUIRE(p.kp % BK == 0, "packed pitch must be a multiple of %d", BK);
  PB_REQUIRE(p.batch >= 1 && p.batch <= 65535, "batch out of range");
  const dim3 grid{unsigned(p.tiles_m) * unsigned(p.tiles_n), unsigned(p.batch)};
  hipLaunchKernelGGL((mt_kernel<BK, TERMS>), grid, dim3(256), 0, h->stream, p);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // namespace mt
}  // namespace pbgemm

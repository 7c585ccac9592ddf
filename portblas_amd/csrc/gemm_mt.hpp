// Large-tile MFMA kernel for packed operands (the hot GEMM of the BF16x3 /
// BF16x4 / BF16 / FP8 paths at large M, N).
//
// Operands are the pre-pass output: op(A) (M x kp) and op(B) (N x kp) in
// BF16 or FP8 e4m3, K-slab-major ([kp / BK][rows][BK], BK the kernel's K
// step), kp a multiple of 4 K steps (64 for the split formats, 128 for BF16,
// 256 for FP8) with zeros in [K, kp).  Up to two parts per operand (hi, lo);
// TERMS products are accumulated per K step:
//   TERMS 1: Ah Bh                           (BF16, or FP8 with ESZ 1)
//   TERMS 3: Ah Bl + Al Bh + Ah Bh           (BF16x3, or FP16x3 with F16)
//   TERMS 4: Ah Bl + Al Bh + Al Bl + Ah Bh   (BF16x4: adds the lo*lo term)
// F16 selects IEEE binary16 parts and v_mfma_f32_32x32x16_f16 (same fragment
// shape and rate as the BF16 instruction).
//
// Workgroup: 256 threads = 4 waves (2 x 2), C tile 256 x 256, wave tile
// 128 x 128 = 4 x 4 v_mfma_f32_32x32x16_bf16 tiles (256 FP32 accumulators per
// lane).  K advances in steps of BK (16 for two-part operands, 32 for BF16,
// 64 for FP8 with v_mfma_scale_f32_32x32x64_f8f6f4), so one stage (every
// part of both operand tiles for one step) is 32 KB.
//   * global -> LDS: buffer loads straight into LDS (no VGPR staging) fill a
//     ring of 4 stages, three steps ahead of the MFMAs (about 2 us of memory
//     latency hidden).  The buffer range is the operand, so rows past M / N
//     read as zero and no edge code is needed in the main loop.  Each wave
//     fills 8 x 1 KB of a stage per step (one part of one operand for the
//     split formats).
//   * layout: rows of BK elements without padding; the 16-byte chunk c of
//     row r sits at slot c ^ swz(r).  A DMA instruction writes 1 KB of
//     consecutive slots, so the swizzle is applied by choosing the global
//     chunk each lane fetches; MFMA fragment reads (16 lanes on 16 rows at
//     one logical chunk) then hit 16 different bank groups.
//   * LDS -> MFMA: the fragments of the next 16-k group are read while the
//     MFMAs of the current group run.
// One barrier per K step; it waits only for this wave's oldest in-flight
// stage, never for the loads still in flight behind it.
#pragma once

#include <type_traits>

#include "gemm_core.hpp"

namespace pbgemm {
namespace mt {

typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 256, kBN = 256;

struct Problem {
  int M, N, kp;          // kp: padded K (elements), a multiple of 4 K steps
  float alpha, beta;
  const float* alpha_dev;  // optional device factor multiplied into alpha
  const float* alpha_dev2; // optional second device factor (one per operand)
  const void* a[2];        // hi, lo parts of op(A): [batch][kp / BK][M][BK]
  const void* b[2];        // hi, lo parts of op(B): [batch][kp / BK][N][BK]
  int64_t a_stride, b_stride;  // batch strides (elements)
  float* c;
  int64_t ldc, c_stride;
  int batch;
  int c_vec;             // C 16-byte aligned with ldc % 4 == 0
  int tiles_m, tiles_n;
};

// ESZ: element bytes, 2 (BF16) or 1 (FP8 e4m3).
template <int BK, int TERMS, int ESZ = 2>
struct Cfg {
  static constexpr int NPARTS = TERMS == 1 ? 1 : 2;
  static constexpr int PARTS = 2 * NPARTS;            // A parts, then B parts
  static constexpr int ROW_BYTES = BK * ESZ;
  static constexpr int CH = ROW_BYTES / 16;           // 16-byte chunks per row
  static constexpr int TILE_BYTES = kBM * ROW_BYTES;  // one part of one operand
  static constexpr int STAGE_BYTES = PARTS * TILE_BYTES;
  static constexpr int NSTAGE = 4;
  // MFMA K groups per step: 16 (BF16, 32x32x16) or 64 (FP8, 32x32x64).
  static constexpr int GROUPS = ESZ == 2 ? BK / 16 : BK / 64;
  // 16-byte LDS reads per fragment (a lane's share of one MFMA operand).
  static constexpr int FRAG_CHUNKS = ESZ == 2 ? 1 : 2;
  static constexpr int ROWS_PER_DMA = 1024 / ROW_BYTES;  // rows one wave DMA fills
  static constexpr int DMA_PER_PART = kBM / ROWS_PER_DMA;
  static constexpr int DMA_PER_WAVE = PARTS * DMA_PER_PART / 4;
  static_assert((ESZ == 2 && ((BK == 16 && TERMS > 1) || (BK == 32 && TERMS == 1))) ||
                    (ESZ == 1 && BK == 64 && TERMS == 1),
                "BK");
  static_assert(STAGE_BYTES == 32 * 1024 && DMA_PER_WAVE == 8, "stage shape");
  static_assert(NSTAGE * STAGE_BYTES <= 160 * 1024, "LDS");
};

// K steps per ring revolution; the packed pitch must be a multiple of this
// many elements.
template <int BK> constexpr int kKAlign = 4 * BK;

// Slot permutation of a row's 16-byte chunks.  A fragment read has 16
// consecutive lanes on 16 consecutive rows at one logical chunk, a stage
// store has consecutive lanes along consecutive chunks; with the slot
// chunk ^ swz(row) both land on 16 different 16-byte bank groups of every
// 256-byte window (and 8 of every 128-byte one).  Depends on row bits 0-4
// only, so it is a per-lane constant for fragment rows w + 32 i + lane % 32.
template <int CH>
__device__ __forceinline__ int swz(int row) {
  if constexpr (CH == 2) return ((row >> 2) ^ (row >> 3)) & 1;
  else if constexpr (CH == 4) return ((row >> 1) ^ (row >> 3)) & 3;
  else return (row ^ ((row >> 3) & 3)) & 7;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  // Raw buffer (stride 0); third dword = default data format for gfx9 parts.
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(bytes),
                                           0x00020000);
}

// Stage barrier: this wave's DMA into the next stage to be read has landed
// (at most 2 x 8 of its LDS loads, the two younger stages, still in flight)
// and its LDS reads are done; then the workgroup barrier.
__device__ __forceinline__ void stage_barrier() {
  asm volatile("s_waitcnt vmcnt(16) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Scheduling: the next N instructions of class MASK (0x020 global load,
// 0x100 LDS read) go PER at a time between consecutive MFMAs.
template <int N, int MASK, int PER>
__device__ __forceinline__ void interleave() {
  static_assert(N % PER == 0, "uneven interleave");
#pragma unroll
  for (int i = 0; i < N / PER; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(MASK, PER, 0);
  }
}

template <int TERMS, int ESZ> constexpr int kFragReads = (TERMS == 1 ? 8 : 16) * (ESZ == 2 ? 1 : 2);

typedef __attribute__((address_space(3))) void* lds_ptr_t;

template <int BK, int TERMS, int ESZ, bool F16>
__global__ __launch_bounds__(256, 1) void mt_kernel(Problem p) {
  // The body is device code only (address-space casts and LDS loads have no
  // host meaning); the host pass needs just the launch stub.
#if defined(__HIP_DEVICE_COMPILE__)
  using C = Cfg<BK, TERMS, ESZ>;
  constexpr int NP = C::NPARTS;
  constexpr int G = C::GROUPS;
  // One array per ring stage: the compiler then tells the LDS loads of one
  // stage from the reads of another.
  __shared__ __attribute__((aligned(1024))) uint8_t st0[C::STAGE_BYTES];
  __shared__ __attribute__((aligned(1024))) uint8_t st1[C::STAGE_BYTES];
  __shared__ __attribute__((aligned(1024))) uint8_t st2[C::STAGE_BYTES];
  __shared__ __attribute__((aligned(1024))) uint8_t st3[C::STAGE_BYTES];
  auto stage = [&](auto r) -> uint8_t* {
    if constexpr (decltype(r)::value == 0) return st0;
    else if constexpr (decltype(r)::value == 1) return st1;
    else if constexpr (decltype(r)::value == 2) return st2;
    else return st3;
  };

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // Wave-uniform values go through readfirstlane so that the buffer
  // descriptor and the LDS destination stay in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wave & 1) * 128, wn = (wave >> 1) * 128;

  int64_t tm64, tn64;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tm64, tn64);
  const int m0 = int(tm64) * kBM, n0 = int(tn64) * kBN;
  const int bidx = blockIdx.y;

  // This wave's DMA work per step: 8 x 1 KB of part `part` (A parts first),
  // row blocks rb0 .. rb0 + 7.
  const int q0 = wave * C::DMA_PER_WAVE;
  const int part = q0 / C::DMA_PER_PART;
  const int rb0 = q0 % C::DMA_PER_PART;
  const bool is_a = part < NP;
  const int pidx = is_a ? part : part - NP;
  const uint8_t* base = is_a ? static_cast<const uint8_t*>(p.a[pidx]) + int64_t(bidx) * p.a_stride * ESZ
                             : static_cast<const uint8_t*>(p.b[pidx]) + int64_t(bidx) * p.b_stride * ESZ;
  const uint32_t rows = uint32_t(is_a ? p.M : p.N);
  // Packed operands are K-slab-major, [kp / BK][rows][BK]: step kt of rows
  // r .. r + ROWS_PER_DMA - 1 is one contiguous kilobyte.  Each step gets its
  // own buffer descriptor whose range is that slab, so rows past the
  // operand's end read as zero.
  const uint32_t slab_bytes = rows * uint32_t(C::ROW_BYTES);
  const int row0 = is_a ? m0 : n0;
  // Lane l of DMA j writes slot l % CH of tile row (rb0 + j) * ROWS_PER_DMA +
  // l / CH, so it fetches chunk (l % CH) ^ swz(row) of that row.
  uint32_t voff[C::DMA_PER_WAVE];
#pragma unroll
  for (int j = 0; j < C::DMA_PER_WAVE; ++j) {
    const int r = (rb0 + j) * C::ROWS_PER_DMA + lane / C::CH;
    voff[j] = uint32_t(row0 + r) * uint32_t(C::ROW_BYTES) +
              uint32_t((lane % C::CH) ^ swz<C::CH>(r)) * 16u;
  }
  const int dma_lds = part * C::TILE_BYTES + rb0 * 1024;

  const int nk = p.kp / BK;
  auto dma = [&](auto r, int kt) {
    const int ks = kt < nk ? kt : nk - 1;  // past the end: re-read, never consumed
    const __amdgpu_buffer_rsrc_t rsrc =
        make_rsrc(base + int64_t(ks) * slab_bytes, slab_bytes);
    uint8_t* dst = stage(r) + dma_lds;
#pragma unroll
    for (int j = 0; j < C::DMA_PER_WAVE; ++j)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr_t)(dst + j * 1024), 16, voff[j], 0,
                                               0, 0);
  };

  // Fragment reads: row w + 32 i + lane % 32 of a part tile, chunk 2 g +
  // lane / 32, at slot chunk ^ swz(row) (swz depends on lane % 32 only).
  const int frow = lane & 31;
  const int fh = lane >> 5;
  const int fswz = swz<C::CH>(frow);
  // BF16: group g is chunk 2 g + fh (8 k).  FP8: the lane's 32 k of the
  // step's 64 are chunks 2 fh and 2 fh + 1.
  constexpr int FC = C::FRAG_CHUNKS;
  int foff[G][FC];
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int h = 0; h < FC; ++h) {
      const int chunk = ESZ == 2 ? 2 * g + fh : 2 * fh + h;
      foff[g][h] = frow * C::ROW_BYTES + ((chunk ^ fswz) * 16);
    }

  using Frag2 = typename std::conditional<F16, f16x8, bf16x8>::type;
  using FragT = typename std::conditional<ESZ == 2, Frag2, i32x8>::type;
  FragT fa[2][NP][4], fb[2][NP][4];
  auto ldfrag = [&](const uint8_t* b, int g, FragT& f) {
    u32x4 v[FC];
#pragma unroll
    for (int h = 0; h < FC; ++h) v[h] = *reinterpret_cast<const u32x4*>(b + foff[g][h]);
    __builtin_memcpy(&f, v, sizeof(FragT));
  };
  auto fread = [&](auto r, int g, int buf) {
    const uint8_t* b = stage(r);
#pragma unroll
    for (int t = 0; t < NP; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ldfrag(b + t * C::TILE_BYTES + (wm + 32 * i) * C::ROW_BYTES, g, fa[buf][t][i]);
        ldfrag(b + (NP + t) * C::TILE_BYTES + (wn + 32 * i) * C::ROW_BYTES, g, fb[buf][t][i]);
      }
  };
  // One MFMA: BF16 or FP16 32x32x16, or FP8 e4m3 32x32x64 with unit scales
  // (E8M0 127 = 2^0; the per-tensor scales are folded into alpha).
  auto mfma = [&](const FragT& x, const FragT& y, f32x16 c) -> f32x16 {
    if constexpr (ESZ == 2 && F16) return __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, c, 0, 0, 0);
    else if constexpr (ESZ == 2) return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x, y, c, 0, 0, 0, 127, 0, 127);
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
          acc[i][j] = mfma(fa[buf][0][i], fb[buf][1][j], acc[i][j]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = mfma(fa[buf][1][i], fb[buf][0][j], acc[i][j]);
    }
    if constexpr (TERMS == 4) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = mfma(fa[buf][1][i], fb[buf][1][j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = mfma(fa[buf][0][i], fb[buf][0][j], acc[i][j]);
  };

  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  constexpr int NFR = kFragReads<TERMS, ESZ>;

  // Prologue: steps 0..3 in flight; wait for step 0 and read its group 0.
  dma(I0{}, 0);
  dma(I1{}, 1);
  dma(I2{}, 2);
  dma(I3{}, 3);
  asm volatile("s_waitcnt vmcnt(24)\n\ts_barrier" ::: "memory");
  fread(I0{}, 0, 0);

  // Step kt lives in stage kt % 4.  Per step: wait for step kt + 1, refill
  // the stage step kt's fragments came from with step kt + 4 (G == 1), and
  // run step kt's MFMAs while the next group's fragments are read.
  auto step = [&](auto r, int kt) {
    constexpr int R = decltype(r)::value;
    using RN = std::integral_constant<int, (R + 1) % 4>;
    if constexpr (G == 1) {
      stage_barrier();
      dma(r, kt + 4);
      fread(RN{}, 0, (R + 1) & 1);
      mma(R & 1);
      interleave<C::DMA_PER_WAVE, 0x020, 1>();
      interleave<NFR, 0x100, 1>();
      __builtin_amdgcn_sched_barrier(0);
    } else {
      fread(r, 1, 1);
      mma(0);
      interleave<NFR, 0x100, 1>();
      __builtin_amdgcn_sched_barrier(0);
      stage_barrier();
      dma(r, kt + 4);
      fread(RN{}, 0, 0);
      mma(1);
      interleave<C::DMA_PER_WAVE, 0x020, 1>();
      interleave<NFR, 0x100, 1>();
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  for (int kt = 0; kt < nk; kt += 4) {
    step(I0{}, kt);
    step(I1{}, kt + 1);
    step(I2{}, kt + 2);
    step(I3{}, kt + 3);
  }
  // The ring's last loads (re-reads past the end) must land before the
  // workgroup's LDS can be handed to another workgroup.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

  // Epilogue: C = alpha * acc + beta * C.  Lane holds column wn + 32 j +
  // lane % 32 and rows wm + 32 i + 8 (e / 4) + 4 (lane / 32) + e % 4.
  float* cb = p.c + int64_t(bidx) * p.c_stride;
  float alpha = p.alpha;
  if (p.alpha_dev) alpha *= *p.alpha_dev;
  if (p.alpha_dev2) alpha *= *p.alpha_dev2;
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
#endif
}

// True when the kernel's 32-bit buffer ranges and offsets hold (one slab
// of up to 64 bytes per row, plus the tile overhang).
inline bool fits(int64_t M, int64_t N, int64_t kp) {
  (void)kp;
  const int64_t lim = int64_t(1) << 31;
  return (M + kBM) * 64 < lim && (N + kBN) * 64 < lim;
}

template <int BK, int TERMS, int ESZ = 2, bool F16 = false>
pb_status_t launch(pb_handle_t h, Problem p) {
  p.tiles_m = int(pb::ceil_div(p.M, kBM));
  p.tiles_n = int(pb::ceil_div(p.N, kBN));
  PB_REQUIRE(p.kp % kKAlign<BK> == 0, "packed pitch must be a multiple of %d", kKAlign<BK>);
  PB_REQUIRE(p.batch >= 1 && p.batch <= 65535, "batch out of range");
  const dim3 grid{unsigned(p.tiles_m) * unsigned(p.tiles_n), unsigned(p.batch)};
  hipLaunchKernelGGL((mt_kernel<BK, TERMS, ESZ, F16>), grid, dim3(256), 0, h->stream, p);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // namespace mt
}  // namespace pbgemm

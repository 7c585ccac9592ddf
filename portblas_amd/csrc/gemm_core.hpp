// The MFMA tile engine shared by every GEMM entry point.
//
// One workgroup (4 waves, 2 x 2) computes a BM x BN tile of C over a K range:
//   * global -> registers: the tile of op(A) (BM x BK) and op(B) (BN x BK) is
//     read in VW x VW blocks of 16-byte vectors (VW = 16 / sizeof(element)).
//     Operands stored K-contiguous in memory are read row by row; operands
//     stored M/N-contiguous are read column by column and transposed in
//     registers, so LDS always holds K-contiguous rows ([rows][BK + pad]).
//   * registers -> LDS: double-buffered; the next K tile's global loads are in
//     flight while the MFMAs run on the current one (one barrier per K tile).
//   * LDS -> MFMA: every lane fetches 16 (or 32) contiguous bytes of K for its
//     fragment row.  Within a fragment group of KG k-values the k order of the
//     hardware steps is permuted identically for A and B, which leaves the
//     product unchanged and lets one ds_read_b128 feed several MFMA steps.
//   * epilogue: C = alpha * acc + beta * C (beta == 0 never reads C), with
//     16-byte stores along the contiguous rows of C, or a split-K partial that
//     the last split to finish folds in a fixed order.
//
// Element types and their MFMA (gfx950):
//   float  : v_mfma_f32_16x16x4_f32   (FP32 in, FP32 accumulate)
//   double : v_mfma_f64_16x16x4_f64   (FP64 in, FP64 accumulate)
//   bf16   : v_mfma_f32_32x32x16_bf16 (BF16 in, FP32 accumulate)
//   fp16   : v_mfma_f32_32x32x16_f16  (FP16 in, FP32 accumulate)
//   fp8    : v_mfma_scale_f32_16x16x128_f8f6f4, e4m3 operands, unit scales
//
// The reference kernel this replaces is Gemm<local, standard, full>
// (src/operations/blas3/gemm_local.hpp:262-581): LDS-staged tiles, register
// tile FMAs, accumulator pre-scaled by beta/alpha.  Here the accumulator
// starts at zero and the epilogue applies alpha and beta.
#pragma once

#include "pb_internal.hpp"

namespace pbgemm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16-byte staging unit

struct bf16_t { uint16_t bits; };
struct fp16_t { uint16_t bits; };  // IEEE binary16
struct fp8_t { uint8_t bits; };

// ---------------------------------------------------------------------------
// Per-element-type traits.
// ---------------------------------------------------------------------------
template <typename T> struct Traits;

template <> struct Traits<float> {
  using Raw = float;          // storage type
  using Acc = float;          // accumulator / C type
  using AccV = f32x4;         // per-lane accumulator of one MFMA tile
  static constexpr int VW = 4;    // elements per 16-byte vector
  static constexpr int MR = 16;   // MFMA tile edge
  static constexpr int KG = 16;   // k values per fragment group
  static constexpr int KPL = 4;   // k values one lane holds per group
  static constexpr int NACC = 4;  // accumulator registers per MFMA tile
};
template <> struct Traits<double> {
  using Raw = double;
  using Acc = double;
  using AccV = f64x4;
  static constexpr int VW = 2;
  static constexpr int MR = 16;
  static constexpr int KG = 16;
  static constexpr int KPL = 4;
  static constexpr int NACC = 4;
};
template <> struct Traits<bf16_t> {
  using Raw = uint16_t;
  using Acc = float;
  using AccV = f32x16;
  static constexpr int VW = 8;
  static constexpr int MR = 32;
  static constexpr int KG = 16;
  static constexpr int KPL = 8;
  static constexpr int NACC = 16;
};
template <> struct Traits<fp16_t> {  // same fragment shape as BF16 (32x32x16)
  using Raw = uint16_t;
  using Acc = float;
  using AccV = f32x16;
  static constexpr int VW = 8;
  static constexpr int MR = 32;
  static constexpr int KG = 16;
  static constexpr int KPL = 8;
  static constexpr int NACC = 16;
};
template <> struct Traits<fp8_t> {
  using Raw = uint8_t;
  using Acc = float;
  using AccV = f32x4;
  static constexpr int VW = 16;
  static constexpr int MR = 16;
  static constexpr int KG = 128;
  static constexpr int KPL = 32;
  static constexpr int NACC = 4;
};

// Fragment: the KPL raw elements a lane contributes to one group.
template <typename T> struct Frag { typename Traits<T>::Raw v[Traits<T>::KPL]; };

template <typename T>
__device__ __forceinline__ void lds_read_frag(const typename Traits<T>::Raw* p, Frag<T>& f) {
  constexpr int bytes = Traits<T>::KPL * sizeof(typename Traits<T>::Raw);
  static_assert(bytes == 16 || bytes == 32, "fragment size");
  const u32x4* q = reinterpret_cast<const u32x4*>(p);
  u32x4* d = reinterpret_cast<u32x4*>(f.v);
  d[0] = q[0];
  if constexpr (bytes == 32) d[1] = q[1];
}

template <typename T>
__device__ __forceinline__ void mma_group(typename Traits<T>::AccV& acc, const Frag<T>& a,
                                          const Frag<T>& b);

template <>
__device__ __forceinline__ void mma_group<float>(f32x4& acc, const Frag<float>& a,
                                                 const Frag<float>& b) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[s], b.v[s], acc, 0, 0, 0);
}
template <>
__device__ __forceinline__ void mma_group<double>(f64x4& acc, const Frag<double>& a,
                                                  const Frag<double>& b) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v[s], b.v[s], acc, 0, 0, 0);
}
template <>
__device__ __forceinline__ void mma_group<bf16_t>(f32x16& acc, const Frag<bf16_t>& a,
                                                  const Frag<bf16_t>& b) {
  bf16x8 av, bv;
  __builtin_memcpy(&av, a.v, 16);
  __builtin_memcpy(&bv, b.v, 16);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
}
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
template <>
__device__ __forceinline__ void mma_group<fp16_t>(f32x16& acc, const Frag<fp16_t>& a,
                                                  const Frag<fp16_t>& b) {
  f16x8 av, bv;
  __builtin_memcpy(&av, a.v, 16);
  __builtin_memcpy(&bv, b.v, 16);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc, 0, 0, 0);
}
template <>
__device__ __forceinline__ void mma_group<fp8_t>(f32x4& acc, const Frag<fp8_t>& a,
                                                 const Frag<fp8_t>& b) {
  i32x8 av, bv;
  __builtin_memcpy(&av, a.v, 32);
  __builtin_memcpy(&bv, b.v, 32);
  // Formats 0/0 = FP8 e4m3 for both operands; E8M0 scales 127 = 2^0.
  acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av, bv, acc, 0, 0, 0, 127, 0, 127);
}

// Row of the C tile that accumulator register e of `lane` holds (column is
// lane % MR for every type).  The FP64 16x16x4 MFMA interleaves its rows
// across lane groups (row = group + 4 e); the FP32/BF16/FP8 forms give each
// lane group 4 consecutive rows.
template <typename T>
__device__ __forceinline__ int acc_row(int lane, int e) {
  if constexpr (sizeof(typename Traits<T>::Acc) == 8) return (lane >> 4) + 4 * e;
  else if constexpr (Traits<T>::MR == 16) return 4 * (lane >> 4) + e;
  else return 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
}

// True when accumulator registers e0..e0+3 hold 4 consecutive rows.
template <typename T>
constexpr bool acc_rows_contiguous() { return sizeof(typename Traits<T>::Acc) == 4; }

// ---------------------------------------------------------------------------
// Operand description.  Element (r, k) of op(X) (r indexes M for A, N for B)
// lives at base[r * rs + k * ks]:  K-contiguous ("kc") operands have ks == 1,
// M/N-contiguous ones have rs == 1 (or rs == batch for interleaved batches).
// ---------------------------------------------------------------------------
template <typename T>
struct Operand {
  const typename Traits<T>::Raw* base;
  int64_t rs, ks;
  int64_t rows;    // M (A) or N (B)
};

template <int BM, int BN, int BK, int WM, int WN>
struct TileShape {
  static constexpr int kBM = BM, kBN = BN, kBK = BK, kWM = WM, kWN = WN;
  static constexpr int kThreads = 256;
  static_assert((BM / WM) * (BN / WN) == 4, "4 waves per workgroup");
};

// KCA / KCB: op(A) / op(B) stored K-contiguous.
template <typename T, typename Shape, bool KCA, bool KCB>
struct Tile {
  using Raw = typename Traits<T>::Raw;
  using Acc = typename Traits<T>::Acc;
  using AccV = typename Traits<T>::AccV;
  static constexpr int VW = Traits<T>::VW;
  static constexpr int MR = Traits<T>::MR;
  static constexpr int KG = Traits<T>::KG;
  static constexpr int KPL = Traits<T>::KPL;
  static constexpr int BM = Shape::kBM, BN = Shape::kBN, BK = Shape::kBK;
  static constexpr int WM = Shape::kWM, WN = Shape::kWN;
  static constexpr int FM = WM / MR, FN = WN / MR;
  static constexpr int LDK = BK + VW;  // LDS row pitch (elements), 16-byte pad
  // FP64 64 x 64 tiles keep two K tiles in flight (a second register stage):
  // at two workgroups per CU their MFMA work per K tile is too short to
  // cover one load latency.  (For FP32 the extra registers cost a third
  // resident workgroup and measured slower.)
  static constexpr bool kDeep = BM * BN <= 64 * 64 && sizeof(Raw) == 8;
  static constexpr int A_ELEMS = BM * LDK, B_ELEMS = BN * LDK;
  static constexpr int STAGE_ELEMS = A_ELEMS + B_ELEMS;
  static constexpr int LDS_BYTES = 2 * STAGE_ELEMS * int(sizeof(Raw));
  // VW x VW load blocks per operand and per thread.
  static constexpr int A_BLOCKS = (BM / VW) * (BK / VW);
  static constexpr int B_BLOCKS = (BN / VW) * (BK / VW);
  static constexpr int APT = (A_BLOCKS + 255) / 256;  // A blocks per thread
  static constexpr int BPTB = (B_BLOCKS + 255) / 256; // B blocks per thread
  static_assert(BK % KG == 0, "BK must be a multiple of the fragment group");
  static_assert(BM % VW == 0 && BN % VW == 0 && BK % VW == 0, "vector blocking");
  static_assert(WM % MR == 0 && WN % MR == 0, "wave tile");

  // Registers of the in-flight global loads: VW vectors per block.
  struct Stage { u32x4 a[APT][VW]; u32x4 b[BPTB][VW]; };

  // Block b of the combined (A then B) index space -> row block, k block.
  // K-contiguous operands walk K fastest (coalesced rows), the others walk
  // rows fastest (coalesced columns).
  template <bool KC, int NROWS>
  static __device__ __forceinline__ void block_coords(int b, int& rb, int& kb) {
    constexpr int KB = BK / VW;
    constexpr int RB = NROWS / VW;
    if constexpr (KC) { rb = b / KB; kb = b % KB; }
    else { rb = b % RB; kb = b / RB; }
  }


  // 16-byte load through a global-address-space pointer: a global_load
  // counts against vmcnt only, so the LDS-read waits before the MFMAs do not
  // also wait for the next K tile's loads (a flat load would count on both).
  static __device__ __forceinline__ u32x4 ld_global16(const Raw* p) {
    typedef const __attribute__((address_space(1))) u32x4 gvec;
    return *(gvec*)(p);
  }

  // Fast path: 16-byte loads from a block base pointer (interior tile, K
  // tile complete, operands vector-aligned).  A K-contiguous block arrives
  // as VW rows; an M/N-contiguous one as VW columns (out[j] = elements
  // (r0 .. r0+VW-1, k0 + j)), transposed only when it is written to LDS so
  // that nothing waits for the loads before the current tile's MFMAs.
  template <bool KC>
  static __device__ __forceinline__ void load_fast(const Raw* p, int64_t rs, int64_t ks,
                                                   u32x4 out[VW]) {
    if constexpr (KC) {
#pragma unroll
      for (int i = 0; i < VW; ++i) {
        out[i] = ld_global16(p);
        p += rs;
      }
    } else {
#pragma unroll
      for (int j = 0; j < VW; ++j) out[j] = ld_global16(p + j * ks);
    }
  }

  // Row i of a block: as loaded (K-contiguous) or gathered from the VW
  // loaded columns.
  template <bool KC>
  static __device__ __forceinline__ u32x4 block_row(const u32x4 (&v)[VW], int i) {
    if constexpr (KC) {
      return v[i];
    } else {
      Raw col[VW][VW];  // col[j][i] = element (r0 + i, k0 + j)
#pragma unroll
      for (int j = 0; j < VW; ++j) __builtin_memcpy(col[j], &v[j], 16);
      Raw row[VW];
#pragma unroll
      for (int j = 0; j < VW; ++j) row[j] = col[j][i];
      u32x4 out;
      __builtin_memcpy(&out, row, 16);
      return out;
    }
  }

  // Guarded path: element loads written straight into one LDS stage, zero
  // outside [0, rows) x [0, K).  Consecutive threads walk the operand's
  // contiguous dimension.  Kept as a plain loop so that it costs few
  // registers: it only runs on edge tiles and on the last partial K tile.
  template <bool KC, int NROWS>
  static __device__ __forceinline__ void fill_guarded_op(const Operand<T>& op, int64_t r0,
                                                         int64_t k0, int64_t K, Raw* dst) {
    constexpr int TOTAL = NROWS * BK;
#pragma unroll 4
    for (int e = threadIdx.x; e < TOTAL; e += 256) {
      const int r = KC ? e / BK : e % NROWS;
      const int k = KC ? e % BK : e / NROWS;
      const int64_t gr = r0 + r, gk = k0 + k;
      dst[r * LDK + k] = (gr < op.rows && gk < K) ? op.base[gr * op.rs + gk * op.ks] : Raw(0);
    }
  }

  static __device__ __forceinline__ void fill_guarded(const Operand<T>& A, const Operand<T>& B,
                                                      int64_t m0, int64_t n0, int64_t k0,
                                                      int64_t K, Raw* stage) {
    fill_guarded_op<KCA, BM>(A, m0, k0, K, stage);
    fill_guarded_op<KCB, BN>(B, n0, k0, K, stage + A_ELEMS);
  }

  static __device__ __forceinline__ void load_all(const Raw* const (&pa)[APT],
                                                  const Raw* const (&pb)[BPTB], int64_t ars,
                                                  int64_t aks, int64_t brs, int64_t bks,
                                                  Stage& st) {
#pragma unroll
    for (int t = 0; t < APT; ++t)
      if (A_BLOCKS % 256 == 0 || threadIdx.x + 256 * t < A_BLOCKS)
        load_fast<KCA>(pa[t], ars, aks, st.a[t]);
#pragma unroll
    for (int t = 0; t < BPTB; ++t)
      if (B_BLOCKS % 256 == 0 || threadIdx.x + 256 * t < B_BLOCKS)
        load_fast<KCB>(pb[t], brs, bks, st.b[t]);
  }

  static __device__ __forceinline__ void lds_store(const Stage& st, Raw* lds) {
#pragma unroll
    for (int t = 0; t < APT; ++t) {
      const int blk = threadIdx.x + 256 * t;
      if (blk < A_BLOCKS) {
        int rb, kb;
        block_coords<KCA, BM>(blk, rb, kb);
#pragma unroll
        for (int i = 0; i < VW; ++i)
          *reinterpret_cast<u32x4*>(lds + (rb * VW + i) * LDK + kb * VW) = block_row<KCA>(st.a[t], i);
      }
    }
#pragma unroll
    for (int t = 0; t < BPTB; ++t) {
      const int blk = threadIdx.x + 256 * t;
      if (blk < B_BLOCKS) {
        int rb, kb;
        block_coords<KCB, BN>(blk, rb, kb);
#pragma unroll
        for (int i = 0; i < VW; ++i)
          *reinterpret_cast<u32x4*>(lds + A_ELEMS + (rb * VW + i) * LDK + kb * VW) =
              block_row<KCB>(st.b[t], i);
      }
    }
  }

  static __device__ __forceinline__ void compute(const Raw* lds, AccV acc[FM][FN]) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = (wave % (BM / WM)) * WM;
    const int wn = (wave / (BM / WM)) * WN;
    const int r = lane % MR;
    const int g = lane / MR;
    const Raw* As = lds;
    const Raw* Bs = lds + A_ELEMS;
    if constexpr (sizeof(Acc) == 8) {
      // FP64: a lane's 4 k values per group are read as two 16-byte halves
      // (2 MFMA steps each), which halves the fragment registers; each
      // accumulator still sees its k steps in the order 0, 1, 2, 3.
#pragma unroll
      for (int kg = 0; kg < BK; kg += KG) {
#pragma unroll
        for (int h = 0; h < KPL; h += 2) {
          u32x4 fa[FM], fb[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i)
            fa[i] = *reinterpret_cast<const u32x4*>(As + (wm + i * MR + r) * LDK + kg + g * KPL + h);
#pragma unroll
          for (int j = 0; j < FN; ++j)
            fb[j] = *reinterpret_cast<const u32x4*>(Bs + (wn + j * MR + r) * LDK + kg + g * KPL + h);
#pragma unroll
          for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
              for (int j = 0; j < FN; ++j) {
                double av, bv;
                __builtin_memcpy(&av, reinterpret_cast<const char*>(&fa[i]) + 8 * s, 8);
                __builtin_memcpy(&bv, reinterpret_cast<const char*>(&fb[j]) + 8 * s, 8);
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[i][j], 0, 0, 0);
              }
        }
      }
      return;
    }
#pragma unroll
    for (int kg = 0; kg < BK; kg += KG) {
      Frag<T> fa[FM], fb[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        lds_read_frag<T>(As + (wm + i * MR + r) * LDK + kg + g * KPL, fa[i]);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        lds_read_frag<T>(Bs + (wn + j * MR + r) * LDK + kg + g * KPL, fb[j]);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) mma_group<T>(acc[i][j], fa[i], fb[j]);
    }
  }

  // Accumulates op(A)[m0:m0+BM, k_begin:k_end] * op(B)[n0:n0+BN, ...]^T.
  // EDGE == false requires the whole BM x BN tile inside M x N and both
  // operands vector-aligned; the K tail (k_end - k_begin not a multiple of
  // BK) is handled by one guarded step after the pipelined loop.  `consume`
  // is called once per LDS stage.  Ends with a barrier (LDS reusable).
  template <bool EDGE, typename F>
  static __device__ __forceinline__ void mainloop_fn(const Operand<T>& A, const Operand<T>& B,
                                                     int64_t m0, int64_t n0, int64_t k_begin,
                                                     int64_t k_end, Raw* lds, F&& consume) {
    if (k_end <= k_begin) return;
    Stage st;
    if constexpr (EDGE) {
      // Two stages alternate so one barrier per K tile suffices: the fill of
      // stage s+1 cannot race the reads of stage s.
      const int64_t nk = (k_end - k_begin + BK - 1) / BK;
      for (int64_t kt = 0; kt < nk; ++kt) {
        Raw* cur = lds + (kt & 1) * STAGE_ELEMS;
        fill_guarded(A, B, m0, n0, k_begin + kt * BK, k_end, cur);
        __syncthreads();
        consume(cur);
      }
      __syncthreads();
    } else {
      const int64_t nk = (k_end - k_begin) / BK;
      if (nk > 0) {
        // Per-slot block pointers, advanced by one K tile per step.
        const Raw* pa[APT];
        const Raw* pb[BPTB];
#pragma unroll
        for (int t = 0; t < APT; ++t) {
          int rb, kb;
          block_coords<KCA, BM>(threadIdx.x + 256 * t, rb, kb);
          pa[t] = A.base + (m0 + rb * VW) * A.rs + (k_begin + kb * VW) * A.ks;
        }
#pragma unroll
        for (int t = 0; t < BPTB; ++t) {
          int rb, kb;
          block_coords<KCB, BN>(threadIdx.x + 256 * t, rb, kb);
          pb[t] = B.base + (n0 + rb * VW) * B.rs + (k_begin + kb * VW) * B.ks;
        }
        const int64_t step_a = int64_t(BK) * A.ks, step_b = int64_t(BK) * B.ks;
        auto advance = [&]() {
#pragma unroll
          for (int t = 0; t < APT; ++t) pa[t] += step_a;
#pragma unroll
          for (int t = 0; t < BPTB; ++t) pb[t] += step_b;
        };
        if constexpr (kDeep) {
          // Two K tiles in flight: tile kt + 2 is loaded while tile kt is
          // consumed and tile kt + 1 waits in the other register stage.
          Stage st2;
          load_all(pa, pb, A.rs, A.ks, B.rs, B.ks, st);
          if (nk > 1) {
            advance();
            load_all(pa, pb, A.rs, A.ks, B.rs, B.ks, st2);
          }
          lds_store(st, lds);
          __syncthreads();
          // hold: tile kt + 1 (in flight); spare: free for tile kt + 2.
          auto step = [&](int64_t kt, Raw* cur, Raw* nxt, Stage& hold, Stage& spare) {
            if (kt + 2 < nk) {
              advance();
              load_all(pa, pb, A.rs, A.ks, B.rs, B.ks, spare);
            }
            consume(cur);
            if (kt + 1 < nk) lds_store(hold, nxt);
            __syncthreads();
          };
          Raw* lds1 = lds + STAGE_ELEMS;
          for (int64_t kt = 0; kt < nk; kt += 2) {
            step(kt, lds, lds1, st2, st);
            if (kt + 1 < nk) step(kt + 1, lds1, lds, st, st2);
          }
        } else {
          load_all(pa, pb, A.rs, A.ks, B.rs, B.ks, st);
          lds_store(st, lds);
          __syncthreads();
          for (int64_t kt = 0; kt < nk; ++kt) {
            Raw* cur = lds + (kt & 1) * STAGE_ELEMS;
            Raw* nxt = lds + ((kt + 1) & 1) * STAGE_ELEMS;
            const bool more = kt + 1 < nk;
            if (more) {
#pragma unroll
              for (int t = 0; t < APT; ++t) pa[t] += step_a;
#pragma unroll
              for (int t = 0; t < BPTB; ++t) pb[t] += step_b;
              load_all(pa, pb, A.rs, A.ks, B.rs, B.ks, st);
            }
            consume(cur);
            if (more) lds_store(st, nxt);
            __syncthreads();
          }
        }
      }
      const int64_t kt0 = k_begin + nk * BK;
      if (kt0 < k_end) {
        fill_guarded(A, B, m0, n0, kt0, k_end, lds);
        __syncthreads();
        consume(lds);
        __syncthreads();
      }
    }
  }

  template <bool EDGE>
  static __device__ __forceinline__ void mainloop(const Operand<T>& A, const Operand<T>& B,
                                                  int64_t m0, int64_t n0, int64_t k_begin,
                                                  int64_t k_end, Raw* lds, AccV acc[FM][FN]) {
    mainloop_fn<EDGE>(A, B, m0, n0, k_begin, k_end, lds,
                      [&](const Raw* stage) { compute(stage, acc); });
  }

  static __device__ __forceinline__ void zero(AccV acc[FM][FN]) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int e = 0; e < Traits<T>::NACC; ++e) acc[i][j][e] = Acc(0);
  }

  // C[m, n] = alpha * acc + beta * C (beta == 0: C not read).  C element
  // (m, n) is at c[m * crs + n * ldc]; c_vec: 16-byte aligned with crs == 1.
  static __device__ __forceinline__ void epilogue(const AccV acc[FM][FN], Acc alpha, Acc beta,
                                                  Acc* c, int64_t crs, int64_t ldc, int64_t M,
                                                  int64_t N, int64_t m0, int64_t n0, bool c_vec) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = (wave % (BM / WM)) * WM;
    const int wn = (wave / (BM / WM)) * WN;
    const int col = lane % MR;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int64_t n = n0 + wn + j * MR + col;
        if (n >= N) continue;
#pragma unroll
        for (int e0 = 0; e0 < Traits<T>::NACC; e0 += 4) {
          const int64_t m = m0 + wm + i * MR + acc_row<T>(lane, e0);
          Acc* cp = c + m * crs + n * ldc;
          bool done = false;
          if constexpr (sizeof(Acc) == 4) {
            if (c_vec && m + 3 < M) {
              f32x4 o;
              if (beta != Acc(0)) {
                const f32x4 old = *reinterpret_cast<const f32x4*>(cp);
#pragma unroll
                for (int e = 0; e < 4; ++e) o[e] = alpha * acc[i][j][e0 + e] + beta * old[e];
              } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) o[e] = alpha * acc[i][j][e0 + e];
              }
              *reinterpret_cast<f32x4*>(cp) = o;
              done = true;
            }
          }
          if (!done) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int64_t me = m0 + wm + i * MR + acc_row<T>(lane, e0 + e);
              if (me < M) {
                Acc* q = c + me * crs + n * ldc;
                const Acc v = alpha * acc[i][j][e0 + e];
                *q = (beta != Acc(0)) ? v + beta * *q : v;
              }
            }
          }
        }
      }
  }

  // Split-K: write the raw accumulator tile (BM x BN, column-major) to `part`.
  static __device__ __forceinline__ void store_partial(const AccV acc[FM][FN], Acc* part) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = (wave % (BM / WM)) * WM;
    const int wn = (wave / (BM / WM)) * WN;
    const int col = lane % MR;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int e = 0; e < Traits<T>::NACC; ++e) {
          const int m = wm + i * MR + acc_row<T>(lane, e);
          const int n = wn + j * MR + col;
          part[m + n * BM] = acc[i][j][e];
        }
  }

  // Adds the partial tile at `part` into acc (same lane layout).
  static __device__ __forceinline__ void add_partial(AccV acc[FM][FN], const Acc* part) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = (wave % (BM / WM)) * WM;
    const int wn = (wave / (BM / WM)) * WN;
    const int col = lane % MR;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int e = 0; e < Traits<T>::NACC; ++e) {
          const int m = wm + i * MR + acc_row<T>(lane, e);
          const int n = wn + j * MR + col;
          acc[i][j][e] += part[m + n * BM];
        }
  }
};

// Device-scope "last arrival" for split-K folding; resets the counter.
__device__ __forceinline__ bool arrive_last(unsigned int* counter, unsigned int total) {
  __shared__ bool last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned int prev = atomicAdd(counter, 1u);
    last = (prev == total - 1);
    if (last) *counter = 0u;
  }
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// Maps a linear workgroup id onto (tile_m, tile_n) so that workgroups that
// run together on one XCD (ids are dealt round-robin over the 8 XCDs) cover a
// compact group of tiles: they share A row panels and B column panels in that
// XCD's L2.
// Tile counts stay below 2^31 (checked on the host), so 32-bit arithmetic.
__device__ __forceinline__ void tile_coords(int64_t lin64, int64_t tiles_m64, int64_t tiles_n64,
                                            int64_t& tm, int64_t& tn) {
  uint32_t lin = uint32_t(lin64);
  const uint32_t tiles_m = uint32_t(tiles_m64), tiles_n = uint32_t(tiles_n64);
  const uint32_t total = tiles_m * tiles_n;
  if (total % 8 == 0) lin = (lin % 8) * (total / 8) + lin / 8;
  constexpr uint32_t G = 8;  // tile rows per group
  const uint32_t group = lin / (G * tiles_n);
  const uint32_t first_m = group * G;
  const uint32_t gsize = (tiles_m - first_m) < G ? (tiles_m - first_m) : G;
  const uint32_t in_group = lin % (G * tiles_n);
  tm = first_m + in_group % gsize;
  tn = in_group / gsize;
}

}  // namespace pbgemm

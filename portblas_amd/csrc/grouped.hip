// Grouped (uneven) GEMM: many independent problems of different shapes in
// ONE persistent launch per operand layout.  portBLAS has no counterpart (its
// batches share one shape: src/interface/gemm_interface.hpp:202-240, with the
// in-kernel batch loop of src/operations/blas3/gemm_local.hpp:273-279,511-516).
//
// The plan uploads one descriptor per problem.
//  BF16: the plan also balances every 64 x 64 tile over a persistent grid
//    (as many workgroups per CU as are resident) by its K, longest first,
//    each problem on one XCD, and uploads each workgroup's tile list.  The
//    ring kernel (problems with dimensions multiple of 8, aligned operands)
//    walks its list with four K chunks in flight across tile boundaries; the
//    lean kernel (the rest) with the next chunk or tile in flight during the
//    current one's MFMAs.  No atomics, no search, nothing to reset: the
//    launch replays in a graph.
//  FP32: a persistent grid pulls tile indices from a device-scope work
//    queue, finds the owning problem by binary search over the prefix sums
//    of the tile counts, and runs the MFMA tile engine on that tile.  The
//    last workgroup to leave resets the queue, so the launch needs no memset
//    and can be captured in a graph and replayed.
#include "gemm_front.hpp"
#include "gemm_launch.hpp"

#include <algorithm>
#include <queue>
#include <utility>
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

// BF16 problems: a lean 64 x 64 tile.  10 KB of LDS and at most 96
// registers let five workgroups share a CU, so the latency of one tile's
// loads is hidden both by the other resident tiles and by issuing the next
// chunk's loads before the current chunk's MFMAs (most grouped problems are a
// handful of tiles and a few K chunks; more resident tiles beat longer
// chunks: K chunks of 64 at four workgroups per CU, or of 128 at two or
// three, were slower).  Per K chunk of 32 every thread issues one 16-byte
// load per operand, then writes it to LDS K-contiguous ([row][k], pitch
// 40): a K-contiguous operand with one 16-byte store, an M/N-contiguous one
// transposed with eight 2-byte stores.  Each wave computes a 32 x 32 quarter
// with v_mfma_f32_32x32x16_bf16.
constexpr int kLT = 64;          // tile edge (M and N)
constexpr int kLK = 32;          // K chunk staged per step
constexpr int kLP = kLK + 8;     // LDS row pitch ([row][k], elements)
constexpr int kLG = kLK / 8;     // 16-byte groups per K-contiguous row of a chunk
constexpr int kLIt = kLT * kLG / 256;  // 16-byte loads per thread and operand

template <bool KC>
__device__ __forceinline__ void lean_load(pbdev::gptr<const uint16_t> x, int64_t rs, int64_t ks,
                                          int64_t rows, int64_t K, int64_t r0, int64_t k0, bool vec,
                                          int e, u32x4& out) {
  // e in [0, kLT * kLG): KC -> row e / kLG, k group (e % kLG) * 8; else k e / 8,
  // row group (e % 8) * 8.
  const int64_t r = KC ? r0 + e / kLG : r0 + (e % 8) * 8;
  const int64_t k = KC ? k0 + (e % kLG) * 8 : k0 + e / 8;
  if (vec && (KC ? (r < rows && k + 8 <= K) : (k < K && r + 8 <= rows))) {
    out = *(pbdev::gptr<const u32x4>)(x + r * rs + k * ks);
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

// 16-byte chunk c (8 k values) of LDS row r sits at chunk c ^ lean_swz(r).
// The MFMA fragment reads (16 consecutive rows, one chunk) stay conflict-free
// (one swizzle value per 16-row block), and the transposing 2-byte stores of
// an M/N-contiguous operand (8 rows 8 apart, one chunk) spread over 8 bank
// groups instead of 2.
__device__ __forceinline__ int lean_swz(int r) { return (r >> 4) & (kLG - 1); }

template <bool KC>
__device__ __forceinline__ void lean_store(uint16_t (*dst)[kLP], int e, const u32x4& v) {
  if constexpr (KC) {
    const int r = e / kLG;
    *reinterpret_cast<u32x4*>(&dst[r][((e % kLG) ^ lean_swz(r)) * 8]) = v;
  } else {
    uint16_t h[8];
    __builtin_memcpy(h, &v, 16);
    const int k = e / 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = (e % 8) * 8 + i;
      dst[r][((k / 8) ^ lean_swz(r)) * 8 + k % 8] = h[i];
    }
  }
}

// One BF16 problem's operand sources, read once per tile so that the loads of
// the next tile can be issued without waiting on its descriptor.
struct LeanSrc {
  pbdev::gptr<const uint16_t> a;
  pbdev::gptr<const uint16_t> b;
  int64_t a_rs, a_ks, b_rs, b_ks, M, N, K;
  bool vec;
};

__device__ __forceinline__ LeanSrc lean_src(const DevProblem& P) {
  return {pbdev::to_global(static_cast<const uint16_t*>(P.a)),
          pbdev::to_global(static_cast<const uint16_t*>(P.b)),
          P.a_rs, P.a_ks, P.b_rs, P.b_ks, P.M, P.N, P.K, P.vec != 0};
}

template <bool KCA, bool KCB>
__device__ __forceinline__ void lean_chunk_load(const LeanSrc& S, int64_t m0, int64_t n0, int64_t k0,
                                                u32x4 (&va)[kLIt], u32x4 (&vb)[kLIt]) {
#pragma unroll
  for (int it = 0; it < kLIt; ++it) {
    const int e = threadIdx.x + 256 * it;
    lean_load<KCA>(S.a, S.a_rs, S.a_ks, S.M, S.K, m0, k0, S.vec, e, va[it]);
    lean_load<KCB>(S.b, S.b_rs, S.b_ks, S.N, S.K, n0, k0, S.vec, e, vb[it]);
  }
}

// Workgroup w runs tiles[starts[w] .. starts[w + 1]) (tile = {problem,
// tile row, tile column}), a list the plan balanced by each tile's K.  The
// loads of the next K chunk, or of the next tile's first chunk, are in flight
// during the current chunk's MFMAs and the tile's epilogue.
template <bool KCA, bool KCB>
__global__ __launch_bounds__(256) void grouped_lean_kernel(const DevProblem* __restrict__ probs,
                                                           const int4* __restrict__ tiles,
                                                           const int* __restrict__ starts) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kLT][kLP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kLT][kLP];
  const int begin = starts[blockIdx.x], end = starts[blockIdx.x + 1];
  if (begin >= end) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  int4 cur = tiles[begin];
  u32x4 va[kLIt], vb[kLIt];
  lean_chunk_load<KCA, KCB>(lean_src(probs[cur.x]), int64_t(cur.y) * kLT, int64_t(cur.z) * kLT, 0, va, vb);
  for (int i = begin; i < end; ++i) {
    const DevProblem& P = probs[cur.x];
    const LeanSrc S = lean_src(P);
    const bool has_next = i + 1 < end;
    const int4 nxt = has_next ? tiles[i + 1] : cur;
    const LeanSrc Sn = lean_src(probs[nxt.x]);
    const int64_t m0 = int64_t(cur.y) * kLT, n0 = int64_t(cur.z) * kLT;
    const int64_t kchunks = S.K > 0 ? (S.K + kLK - 1) / kLK : 1;
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.f;
    for (int64_t kc = 0; kc < kchunks; ++kc) {
      __syncthreads();  // the previous chunk (or tile) is consumed
#pragma unroll
      for (int it = 0; it < kLIt; ++it) {
        const int e = threadIdx.x + 256 * it;
        lean_store<KCA>(As, e, va[it]);
        lean_store<KCB>(Bs, e, vb[it]);
      }
      __syncthreads();
      if (kc + 1 < kchunks)
        lean_chunk_load<KCA, KCB>(S, m0, n0, (kc + 1) * kLK, va, vb);
      else if (has_next)
        lean_chunk_load<KCA, KCB>(Sn, int64_t(nxt.y) * kLT, int64_t(nxt.z) * kLT, 0, va, vb);
#pragma unroll
      for (int kk = 0; kk < kLK; kk += 16) {
        bf16x8 a, b;
        const int ra = wm + lane % 32, rb = wn + lane % 32, c = kk / 8 + lane / 32;
        __builtin_memcpy(&a, &As[ra][(c ^ lean_swz(ra)) * 8], 16);
        __builtin_memcpy(&b, &Bs[rb][(c ^ lean_swz(rb)) * 8], 16);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
    }
    // Epilogue: lane holds column wn + lane % 32, rows wm + 8 (e / 4) +
    // 4 (lane / 32) + e % 4.
    const int64_t n = n0 + wn + lane % 32;
    if (n < P.N) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t m = m0 + wm + 8 * q + 4 * (lane / 32);
        const pbdev::gptr<float> cp = pbdev::to_global(P.c) + m + n * P.ldc;
        if (P.c_vec && m + 3 < P.M) {
          f32x4 o;
          if (P.beta != 0.f) {
            const f32x4 old = *(pbdev::gptr<const f32x4>)cp;
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = P.alpha * acc[4 * q + e] + P.beta * old[e];
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = P.alpha * acc[4 * q + e];
          }
          *(pbdev::gptr<f32x4>)cp = o;
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
    cur = nxt;
  }
}

// ---------------------------------------------------------------------------
// Ring variant of the lean kernel for problems whose M, N and K are multiples
// of 8 with 16-byte aligned operands (every 16-byte group of 8 elements is
// then wholly inside or wholly outside the operand) and leading dimensions
// below 2^24.  Each workgroup walks the flattened sequence of its tiles' K
// chunks with kRD chunks in flight in a register ring, so a chunk waits for a
// fraction of one memory latency instead of a whole one, and chunks past the
// end of the list re-read the last one (never consumed).  Two LDS buffers
// take alternate chunks, so one barrier per chunk separates a buffer's reads
// from its next writes.
//
// The loads are raw buffer loads, one 16-byte load per thread and operand
// per chunk, so that a chunk costs few instructions (the kernel is otherwise
// bound by instruction issue, not by memory): each thread's byte offset
// inside the tile's operand panel is computed once per tile (a thread whose
// rows are outside the operand gets an offset past the buffer's range, which
// the hardware answers with zeros), a chunk moves the buffer's base by a
// scalar step, and in a tile's last chunk a thread whose K group is past K
// is sent past the range too.
// ---------------------------------------------------------------------------
// Ring geometry: K chunk kRK (64: every row piece of a K-contiguous operand,
// and every column piece of an M/N-contiguous one, is a whole 128-byte line;
// with 32 each 128-byte line was fetched twice, once per half), kRD chunks in
// flight per workgroup.
constexpr int kRK = 64;
constexpr int kRD = 4;
constexpr int kRP = kRK;                  // LDS row pitch ([row][k], elements; no pad, see ring_swz)
constexpr int kRG = kRK / 8;              // 16-byte groups per K-contiguous row of a chunk
constexpr int kRIt = kLT * kRG / 256;     // 16-byte loads per thread and operand
// Buffer range of every panel: in-panel offsets are below 63 * 2^24 * 2 +
// 128 bytes (leading dimensions below 2^24), so below the range; kRFar is
// above it, so a load at kRFar reads zeros.
constexpr int kRRange = 0x7ffffff0;
constexpr uint32_t kRFar = 0xfffffff0u;
constexpr int64_t kRMaxLd = int64_t(1) << 24;

// Group e of a chunk: KC -> row e / kRG, k group (e % kRG) * 8; else k
// e / 8, row group (e % 8) * 8.
template <bool KC>
__device__ __forceinline__ int ring_row(int e) { return KC ? e / kRG : (e % 8) * 8; }
template <bool KC>
__device__ __forceinline__ int ring_k(int e) { return KC ? (e % kRG) * 8 : e / 8; }

// The producer: the next chunk to load and where its groups are.
struct RingSrc {
  const char* a;            // operand panels at the current chunk
  const char* b;
  int64_t a_step, b_step;   // bytes from one chunk to the next
  uint32_t oa[kRIt], ob[kRIt];  // this thread's byte offsets in the panels
  int krem;                 // K left from the current chunk on
  int kc, nk;               // chunk, chunks of the tile
};

__device__ __forceinline__ int ring_chunks(int64_t K) { return K > 0 ? int((K + kRK - 1) / kRK) : 1; }

template <bool KC>
__device__ __forceinline__ uint32_t ring_offset(int64_t rs, int64_t ks, int64_t rows_left, int e) {
  const int r = ring_row<KC>(e), k = ring_k<KC>(e);
  if (r >= rows_left) return kRFar;
  return 2u * (__umul24(uint32_t(r), uint32_t(rs)) + __umul24(uint32_t(k), uint32_t(ks)));
}

template <bool KCA, bool KCB>
__device__ __forceinline__ void ring_src_set(RingSrc& s, const DevProblem* probs, const int4* tiles, int i) {
  const int4 t = tiles[i];
  const DevProblem& P = probs[t.x];
  const int64_t m0 = int64_t(t.y) * kLT, n0 = int64_t(t.z) * kLT;
  s.a = static_cast<const char*>(P.a) + 2 * m0 * P.a_rs;
  s.b = static_cast<const char*>(P.b) + 2 * n0 * P.b_rs;
  s.a_step = 2 * int64_t(kRK) * P.a_ks;
  s.b_step = 2 * int64_t(kRK) * P.b_ks;
#pragma unroll
  for (int it = 0; it < kRIt; ++it) {
    const int e = threadIdx.x + 256 * it;
    s.oa[it] = ring_offset<KCA>(P.a_rs, P.a_ks, P.M - m0, e);
    s.ob[it] = ring_offset<KCB>(P.b_rs, P.b_ks, P.N - n0, e);
  }
  s.krem = int(P.K);
  s.kc = 0;
  s.nk = ring_chunks(P.K);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t ring_rsrc(const char* base) {
  // Raw buffer (stride 0); third dword = default data format for gfx9 parts.
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), (short)0, kRRange, 0x00020000);
}

template <bool KCA, bool KCB>
__device__ __forceinline__ void ring_load(const RingSrc& s, u32x4 (&va)[kRIt], u32x4 (&vb)[kRIt]) {
  const __amdgpu_buffer_rsrc_t ra = ring_rsrc(s.a), rb = ring_rsrc(s.b);
#pragma unroll
  for (int it = 0; it < kRIt; ++it) {
    const int e = threadIdx.x + 256 * it;
    const uint32_t oa = ring_k<KCA>(e) < s.krem ? s.oa[it] : kRFar;
    const uint32_t ob = ring_k<KCB>(e) < s.krem ? s.ob[it] : kRFar;
    va[it] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(ra, oa, 0, 0));
    vb[it] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rb, ob, 0, 0));
  }
}

// 16-byte chunk c of LDS row r sits at chunk c ^ ring_swz(r).  LDS has 32
// 4-byte banks and a row is 128 bytes, so row r's chunk c always falls on
// the 16-byte bank slot c ^ ring_swz(r).  The 8 lanes of one cycle of a
// fragment read (8 consecutive rows, one chunk) see r % 8 take every value
// once, hence 8 different slots; the 32 lanes of one cycle of a transposing
// 2-byte store (rows 8 apart, one chunk) see r / 8 take every value once,
// hence 8 different slots again; a 16-byte store row covers all 8 chunks.
// No access pattern of the kernel has a bank conflict.
__device__ __forceinline__ int ring_swz(int r) { return (r ^ (r >> 3)) & (kRG - 1); }

template <bool KC>
__device__ __forceinline__ void ring_store(uint16_t (*dst)[kRP], int e, const u32x4& v) {
  if constexpr (KC) {
    const int r = e / kRG;
    *reinterpret_cast<u32x4*>(&dst[r][((e % kRG) ^ ring_swz(r)) * 8]) = v;
  } else {
    uint16_t h[8];
    __builtin_memcpy(h, &v, 16);
    const int k = e / 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = (e % 8) * 8 + i;
      dst[r][((k / 8) ^ ring_swz(r)) * 8 + k % 8] = h[i];
    }
  }
}

// Three workgroups per CU (three waves per SIMD: at most 168 registers per
// lane, AGPRs included).
// Epilogue staging: each wave's 32 x 32 FP32 quadrant, column-major with a
// 144-byte pitch (36 floats), in a region of its own.
constexpr int kRCP = 36;

// Writes a wave's finished 32 x 32 quadrant of C.  The MFMA result (lane:
// column lane % 32, rows 8 q + 4 (lane / 32) + 0..3) goes through the wave's
// LDS region so that each store instruction writes 8 whole 128-byte column
// pieces (8 lanes per column, 16 bytes each) instead of 32 pieces of 32
// bytes: the store requests to the caches drop four-fold.  Lanes of one
// 16-byte write or read fall on different banks in every cycle (pitch 36
// floats).  C rows are multiples of 8 in this kernel, so a 4-row piece is
// wholly inside or outside.
__device__ __forceinline__ void ring_epilogue(const f32x16& acc, float (*stage)[kRCP], const DevProblem& P,
                                              int64_t m_base, int64_t n_base, int lane) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    f32x4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = acc[4 * q + e];
    *reinterpret_cast<f32x4*>(&stage[lane % 32][8 * q + 4 * (lane / 32)]) = v;
  }
  // The quadrant is read back by other lanes of the same wave: LDS serves a
  // wave's requests in order, so only the compiler must keep the order.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int rp = (lane % 8) * 4;
  const int64_t m = m_base + rp;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cn = lane / 8 + 8 * j;
    const int64_t n = n_base + cn;
    const f32x4 v = *reinterpret_cast<const f32x4*>(&stage[cn][rp]);
    if (n < P.N && m < P.M) {
      const pbdev::gptr<float> cp = pbdev::to_global(P.c) + m + n * P.ldc;
      if (P.c_vec) {
        f32x4 o;
        if (P.beta != 0.f) {
          const f32x4 old = *(pbdev::gptr<const f32x4>)cp;
#pragma unroll
          for (int e = 0; e < 4; ++e) o[e] = P.alpha * v[e] + P.beta * old[e];
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) o[e] = P.alpha * v[e];
        }
        // Streaming store: C is not read again here, and keeping it out of
        // the caches' way of the operand panels cut the launch by 11 %.
        __builtin_nontemporal_store(o, (pbdev::gptr<f32x4>)cp);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float w = P.alpha * v[e];
          cp[e] = (P.beta != 0.f) ? w + P.beta * cp[e] : w;
        }
      }
    }
  }
  // The region is rewritten at the wave's next tile: its reads come first.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Workgroup w runs tiles[starts[w] .. starts[w + 1]) (tile = {problem, tile
// row, tile column, K chunks of this and every later tile of the list}), a
// list the plan balanced by each tile's K.
template <bool KCA, bool KCB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void grouped_ring_kernel(
    const DevProblem* __restrict__ probs, const int4* __restrict__ tiles, const int* __restrict__ starts) {
  __shared__ __attribute__((aligned(16))) uint16_t As[2][kLT][kRP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][kLT][kRP];
  __shared__ __attribute__((aligned(16))) float Cs[4][32][kRCP];
  const int begin = starts[blockIdx.x], end = starts[blockIdx.x + 1];
  if (begin >= end) return;
  const int total = tiles[begin].w;  // K chunks of the whole list
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  // Producer (kRD chunks ahead): pi is its tile; consumer: ci, ckc, cnk.
  RingSrc pc;
  int pi = begin;
  ring_src_set<KCA, KCB>(pc, probs, tiles, pi);
  int ci = begin, ckc = 0, cnk = pc.nk;
  auto advance_producer = [&]() {
    if (pc.kc + 1 < pc.nk) {
      ++pc.kc;
      pc.a += pc.a_step;
      pc.b += pc.b_step;
      pc.krem -= kRK;
    } else if (pi + 1 < end) {
      ring_src_set<KCA, KCB>(pc, probs, tiles, ++pi);
    }
    // else: stay on the last chunk (re-read, never consumed)
  };
  u32x4 va[kRD][kRIt], vb[kRD][kRIt];
#pragma unroll
  for (int s = 0; s < kRD; ++s) {
    ring_load<KCA, KCB>(pc, va[s], vb[s]);
    advance_producer();
  }
  f32x16 acc;
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = 0.f;
  // One chunk: stage slot S into LDS buffer S % 2, refill the slot kRD
  // chunks ahead, run the chunk's MFMAs, and finish the tile after its last
  // chunk.  kRD is even, so the buffer parity follows the slot.  The
  // producer moves to its next chunk after the MFMAs, so the descriptor
  // loads of a new tile do not hold up the fragment reads' waits.
  auto step = [&](auto slot) {
    constexpr int S = decltype(slot)::value;
    constexpr int BUF = S & 1;
#pragma unroll
    for (int it = 0; it < kRIt; ++it) {
      const int e = threadIdx.x + 256 * it;
      ring_store<KCA>(As[BUF], e, va[S][it]);
      ring_store<KCB>(Bs[BUF], e, vb[S][it]);
    }
    __syncthreads();
    ring_load<KCA, KCB>(pc, va[S], vb[S]);
#pragma unroll
    for (int kk = 0; kk < kRK; kk += 16) {
      bf16x8 a, b;
      const int ra = wm + lane % 32, rb = wn + lane % 32, c = kk / 8 + lane / 32;
      __builtin_memcpy(&a, &As[BUF][ra][(c ^ ring_swz(ra)) * 8], 16);
      __builtin_memcpy(&b, &Bs[BUF][rb][(c ^ ring_swz(rb)) * 8], 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    advance_producer();
    if (ckc + 1 < cnk) {
      ++ckc;
      return;
    }
    const int4 t = tiles[ci];
    ring_epilogue(acc, Cs[wave], probs[t.x], int64_t(t.y) * kLT + wm, int64_t(t.z) * kLT + wn, lane);
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.f;
    if (ci + 1 < end) {
      ++ci;
      ckc = 0;
      cnk = ring_chunks(probs[tiles[ci].x].K);
    }
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  static_assert(kRD == 4, "the step sequence below is written for four slots");
  // Whole rounds of four chunks in a loop with its only exit at the round's
  // end, then the last 0-3 chunks.  The compiler's wait counts at slot 0
  // merge every edge into the loop head: an early exit routed through the
  // head (as the structurizer does with breaks) would make them assume a
  // slot 0 refilled one step ago and wait for nearly every load in flight.
  int c = 0;
  for (; c + kRD <= total; c += kRD) {
    step(I0{});
    step(I1{});
    step(I2{});
    step(I3{});
  }
  const int rem = total - c;
  if (rem > 0) step(I0{});
  if (rem > 1) step(I1{});
  if (rem > 2) step(I2{});
  // The ring's last loads (re-reads, never consumed) land before the wave
  // ends and its registers go to another wave.
  __builtin_amdgcn_s_waitcnt(0);
}

using G64 = TileShape<64, 64, 64, 32, 32>;
using GF64 = TileShape<64, 64, 32, 32, 32>;

struct Launch {
  int layout;   // 2 * kca + kcb
  int count;
  int64_t tiles;
  DevProblem* dev;
  bool ring = false;  // BF16: the ring kernel (dimensions multiples of 8, aligned)
  // BF16: the balanced static schedule (grid workgroups, tile lists).
  int grid = 0;
  const int4* tile_list = nullptr;
  const int* starts = nullptr;
};

}  // namespace pbgemm

struct pb_grouped_plan_s {
  pb_gemm_precision_t precision;
  std::vector<pbgemm::Launch> launches;
  std::vector<void*> dev_mem;  // problem descriptors, tile lists, offsets
  int device = 0;
};

namespace pbgemm {

void launch_lean(pb_handle_t h, const Launch& L) {
  const dim3 grid(unsigned(L.grid));
  if (L.ring) {
    switch (L.layout) {
      case 3: hipLaunchKernelGGL((grouped_ring_kernel<true, true>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
      case 2: hipLaunchKernelGGL((grouped_ring_kernel<true, false>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
      case 1: hipLaunchKernelGGL((grouped_ring_kernel<false, true>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
      default: hipLaunchKernelGGL((grouped_ring_kernel<false, false>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
    }
    return;
  }
  switch (L.layout) {
    case 3: hipLaunchKernelGGL((grouped_lean_kernel<true, true>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
    case 2: hipLaunchKernelGGL((grouped_lean_kernel<true, false>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
    case 1: hipLaunchKernelGGL((grouped_lean_kernel<false, true>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
    default: hipLaunchKernelGGL((grouped_lean_kernel<false, false>), grid, dim3(256), 0, h->stream, L.dev, L.tile_list, L.starts); break;
  }
}

// The tiles of one BF16 launch, XCD by XCD.  Workgroup w runs on XCD w % 8,
// and each of the 8 XCDs has its own L2, so every problem goes whole to one
// XCD (longest first, to the least-loaded XCD) and its operand panels are
// fetched into one L2 only.  A problem worth more than 1/spread_div of the
// launch (built-in: a sixteenth) is spread over all XCDs, tile by tile,
// instead.  Cost of a tile: K
// chunks + 1 for the epilogue.  Each XCD's tiles come longest first, a
// problem's tiles together (the XCD's workgroups work through one problem's
// panels at a time).
constexpr int kXcds = 8;

struct TileItem {
  int64_t cost;
  int4 t;  // problem, tile row, tile column
};

std::vector<std::vector<TileItem>> xcd_tiles(const std::vector<DevProblem>& v, int xcds, int kchunk,
                                             int spread_div) {
  struct Item { int64_t cost; int prob; };
  std::vector<Item> probs_by_cost;
  int64_t total = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    probs_by_cost.push_back({pb::ceil_div(std::max<int64_t>(v[i].K, 1), kchunk) + 1, int(i)});
    total += probs_by_cost.back().cost * int64_t(v[i].tiles_m) * v[i].tiles_n;
  }
  std::stable_sort(probs_by_cost.begin(), probs_by_cost.end(), [&](const Item& x, const Item& y) {
    return x.cost * int64_t(v[x.prob].tiles_m) * v[x.prob].tiles_n >
           y.cost * int64_t(v[y.prob].tiles_m) * v[y.prob].tiles_n;
  });
  std::vector<int64_t> xload(xcds, 0);
  std::vector<std::vector<TileItem>> out(xcds);
  for (const Item& it : probs_by_cost) {
    const DevProblem& d = v[it.prob];
    const int64_t c = it.cost * int64_t(d.tiles_m) * d.tiles_n;
    if (xcds > 1 && spread_div > 1 && c * spread_div > total) {
      for (int32_t tm = 0; tm < d.tiles_m; ++tm)
        for (int32_t tn = 0; tn < d.tiles_n; ++tn) {
          const int x = int(std::min_element(xload.begin(), xload.end()) - xload.begin());
          xload[x] += it.cost;
          out[x].push_back({it.cost, int4{it.prob, tm, tn, 0}});
        }
      continue;
    }
    const int x = int(std::min_element(xload.begin(), xload.end()) - xload.begin());
    xload[x] += c;
    for (int32_t tm = 0; tm < d.tiles_m; ++tm)
      for (int32_t tn = 0; tn < d.tiles_n; ++tn) out[x].push_back({it.cost, int4{it.prob, tm, tn, 0}});
  }
  for (auto& q : out)
    std::stable_sort(q.begin(), q.end(), [](const TileItem& p1, const TileItem& p2) { return p1.cost > p2.cost; });
  return out;
}

// Relative speed of the three workgroups that share a CU when the ring
// kernel's grid holds exactly three per CU.  The dispatcher fills every CU
// once before it starts the next round, so workgroup w is the (w / CUs)-th
// on its CU, and the SIMDs' issue arbitration favours older waves: with
// equal work the three ended at 71, 84 and 98 us on average (512-problem
// set, MI355X; within each rank the ends lay within 4 us), i.e. they run at
// 1 : 0.84 : 0.72.  Other grids are balanced as if all ran alike.
constexpr double kRankSpeed[3] = {1.0, 0.78, 0.61};

// Static schedule: each XCD's tiles, in order, to the one of its workgroups
// (w % xcds == x) that would finish its list first (assigned cost over its
// speed).  Returns per-workgroup lists, concatenated, and their offsets;
// each tile's .w holds the K chunks (of kchunk) of that tile and of every
// later tile of its list (the ring kernel's chunk count), or false if a list
// has 2^31 chunks or more.  `per_cu`: workgroups per CU in the grid, `cus`:
// the device's CUs (rank speeds apply when grid == 3 * cus; per_cu 0: never).
bool balance_tiles(const std::vector<DevProblem>& v, int grid, int kchunk, int per_cu, int cus,
                   int spread_div, std::vector<int4>& list, std::vector<int>& starts) {
  const int xcds = grid >= kXcds ? kXcds : 1;
  const auto by_xcd = xcd_tiles(v, xcds, kchunk, spread_div);
  const bool ranked = per_cu == 3 && grid == 3 * cus;
  std::vector<double> speed(grid, 1.0);
  if (ranked)
    for (int w = 0; w < grid; ++w) speed[w] = kRankSpeed[w / cus];
  using Load = std::pair<double, int>;  // (finish time of the assigned cost, workgroup)
  std::vector<std::priority_queue<Load, std::vector<Load>, std::greater<Load>>> heaps(xcds);
  for (int w = 0; w < grid; ++w) heaps[w % xcds].push({0.0, w});
  std::vector<std::vector<int4>> per(grid);
  for (int x = 0; x < xcds; ++x)
    for (const TileItem& it : by_xcd[x]) {
      Load l = heaps[x].top();
      heaps[x].pop();
      per[l.second].push_back(it.t);
      l.first += double(it.cost) / speed[l.second];
      heaps[x].push(l);
    }
  for (auto& q : per) {
    int64_t suffix = 0;
    for (size_t i = q.size(); i-- > 0;) {
      suffix += std::max<int64_t>(1, pb::ceil_div(v[q[i].x].K, kchunk));
      if (suffix > INT32_MAX) return false;
      q[i].w = int(suffix);
    }
  }
  list.clear();
  starts.assign(1, 0);
  for (auto& q : per) {
    list.insert(list.end(), q.begin(), q.end());
    starts.push_back(int(list.size()));
  }
  return true;
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
  return pb_gemm_grouped_plan_create_ex(h, prec, problems, count, nullptr, out);
}

pb_status_t pb_gemm_grouped_plan_create_ex(pb_handle_t h, pb_gemm_precision_t prec,
                                           const pb_gemm_problem_t* problems, int64_t count,
                                           const pb_grouped_options_t* options, pb_grouped_plan_t* out) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(out != nullptr, "null plan pointer");
  const pb_grouped_options_t opt = options ? *options : pb_grouped_options_t{0, 0, 0, 0};
  PB_REQUIRE(opt.ring >= 0 && opt.ring <= 2 && opt.rank_weights >= 0 && opt.rank_weights <= 2 &&
                 opt.spread_div >= 0 && opt.per_cu >= 0,
             "bad grouped options (ring %d, rank_weights %d, spread_div %d, per_cu %d)", opt.ring,
             opt.rank_weights, opt.spread_div, opt.per_cu);
  const bool use_ring = opt.ring != 2;
  const bool rank_weights = opt.rank_weights != 2;
  const int spread_div = opt.spread_div > 0 ? opt.spread_div : 16;
  *out = nullptr;
  PB_REQUIRE(prec == PB_GEMM_BF16 || prec == PB_GEMM_FP32_MFMA,
             "grouped GEMM supports PB_GEMM_BF16 and PB_GEMM_FP32_MFMA inputs");
  PB_REQUIRE(count >= 0 && count <= (1 << 24), "bad problem count");
  PB_REQUIRE(count == 0 || problems != nullptr, "null problem array");
  const bool bf16 = prec == PB_GEMM_BF16;
  const int vw = bf16 ? 8 : 4;
  auto* plan = new pb_grouped_plan_s();
  plan->precision = prec;
  plan->device = h->device;
  auto destroy = [&]() {
    for (void* m : plan->dev_mem) (void)hipFree(m);
    delete plan;
  };
  constexpr int bm = 64, bn = 64;  // BF16: the lean kernel; FP32: the 64 x 64 tile engine
  // Layouts 0-3 (2 * kca + kcb); BF16 problems the ring kernel takes go to
  // the same layout + 4.
  std::vector<DevProblem> by_layout[8];
  for (int64_t i = 0; i < count; ++i) {
    const pb_gemm_problem_t& q = problems[i];
    const char ta = pb::norm_trans(q.transa), tb = pb::norm_trans(q.transb);
    auto fail = [&](const char* what) {
      destroy();
      return pb::set_error(PB_STATUS_INVALID_VALUE, "problem %lld: %s", (long long)i, what);
    };
    if (!ta || !tb) return fail("invalid transpose flag");
    if (q.m < 0 || q.n < 0 || q.k < 0) return fail("negative dimension");
    if (q.m == 0 || q.n == 0) continue;
    if (q.ldc < q.m || q.lda < (ta == 'n' ? q.m : q.k) || q.ldb < (tb == 'n' ? q.k : q.n))
      return fail("leading dimension too small");
    if (!q.c || (q.k > 0 && (!q.a || !q.b))) return fail("null pointer");
    if (pb::ceil_div(q.m, bm) > 0x7fffffff || pb::ceil_div(q.n, bn) > 0x7fffffff)
      return fail("too many tiles");
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
    const bool ring = use_ring && bf16 && d.vec && q.m % 8 == 0 && q.n % 8 == 0 && q.k % 8 == 0 &&
                      q.lda < kRMaxLd && q.ldb < kRMaxLd && q.k <= 0x7fffffffLL - kRK;
    by_layout[2 * (ta == 't') + (tb == 'n') + (ring ? 4 : 0)].push_back(d);
  }
  auto upload = [&](const void* src, size_t bytes, void** dst) -> pb_status_t {
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) {
      destroy();
      return pb::set_error(PB_STATUS_ALLOC_FAILED, "plan allocation failed: %s", hipGetErrorString(e));
    }
    plan->dev_mem.push_back(*dst);
    e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      destroy();
      return pb::set_error(PB_STATUS_DEVICE_ERROR, "plan upload failed: %s", hipGetErrorString(e));
    }
    return PB_STATUS_SUCCESS;
  };
  // The static schedule assumes every workgroup of the grid is resident at
  // once: size the grid by the lean kernel's occupancy (registers limit it).
  auto lean_per_cu = [&](int layout) -> int64_t {
    int n = 0;
    hipError_t e = hipSuccess;
    switch (layout) {
      case 7: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_ring_kernel<true, true>, 256, 0); break;
      case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_ring_kernel<true, false>, 256, 0); break;
      case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_ring_kernel<false, true>, 256, 0); break;
      case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_ring_kernel<false, false>, 256, 0); break;
      case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_lean_kernel<true, true>, 256, 0); break;
      case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_lean_kernel<true, false>, 256, 0); break;
      case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_lean_kernel<false, true>, 256, 0); break;
      default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_lean_kernel<false, false>, 256, 0); break;
    }
    if (e != hipSuccess || n < 1) {
      (void)hipGetLastError();
      n = 4;
    }
    return n;
  };
  for (int lay = 7; lay >= 0; --lay) {
    auto& v = by_layout[lay];
    if (v.empty()) continue;
    // Largest problems first: for the FP32 work queue their tiles are handed
    // out first, so the tail of the queue holds the small ones.
    std::stable_sort(v.begin(), v.end(), [](const DevProblem& x, const DevProblem& y) {
      return x.M * x.N * x.K > y.M * y.N * y.K;
    });
    int64_t t = 0;
    for (auto& d : v) {
      d.tile_begin = t;
      t += int64_t(d.tiles_m) * d.tiles_n;
    }
    if (bf16 && t > 0x7fffffffLL) {
      destroy();
      return pb::set_error(PB_STATUS_INVALID_VALUE, "grouped BF16 GEMM: more than 2^31 tiles");
    }
    Launch L{lay % 4, int(v.size()), t, nullptr};
    L.ring = lay >= 4;
    void* dev = nullptr;
    PB_TRY(upload(v.data(), v.size() * sizeof(DevProblem), &dev));
    L.dev = static_cast<DevProblem*>(dev);
    if (bf16) {
      const int resident = int(lean_per_cu(lay));
      const int per_cu = opt.per_cu > 0 ? std::min(opt.per_cu, resident) : resident;
      L.grid = int(std::max<int64_t>(1, std::min<int64_t>(t, int64_t(per_cu) * h->num_cus)));
      std::vector<int4> list;
      std::vector<int> starts;
      if (!balance_tiles(v, L.grid, L.ring ? kRK : kLK, L.ring && rank_weights ? per_cu : 0, h->num_cus,
                         spread_div, list, starts)) {
        destroy();
        return pb::set_error(PB_STATUS_INVALID_VALUE, "grouped BF16 GEMM: 2^31 K chunks or more on one workgroup");
      }
      void* dl = nullptr;
      void* ds = nullptr;
      PB_TRY(upload(list.data(), list.size() * sizeof(int4), &dl));
      PB_TRY(upload(starts.data(), starts.size() * sizeof(int), &ds));
      L.tile_list = static_cast<const int4*>(dl);
      L.starts = static_cast<const int*>(ds);
    }
    plan->launches.push_back(L);
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
    if (plan->precision == PB_GEMM_BF16) {
      launch_lean(h, L);
    } else {
      const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(L.tiles, 4 * h->num_cus));
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
  for (void* m : plan->dev_mem) (void)hipFree(m);
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

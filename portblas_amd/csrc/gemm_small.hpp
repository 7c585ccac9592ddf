// Many small GEMMs (M, N <= 64) of a strided batch: one problem per
// workgroup.  The batched path of portBLAS runs its tiled kernel over the
// batch (src/operations/blas3/gemm_local.hpp:273-279,511-516); a 64 x 64
// problem is one tile, and a batch of them is bound by HBM, so what matters
// is bytes in flight: every thread issues all of its global loads for a K
// chunk at once, the chunk goes through LDS, and four waves each compute a
// 32 x 32 quarter of C with v_mfma_f32_16x16x4_f32 (FP32) or
// v_mfma_f64_16x16x4_f64 (FP64).  Full 64 x 64 FP32 problems with K a
// multiple of 64 stage K in chunks of 64 without bounds checks (a 64^3
// problem is one memory latency; four workgroups per CU); the others use
// chunks of 32 (five vector-staged FP32 workgroups per CU).  The next
// chunk's loads are issued before the current chunk's MFMAs.  FP32 64^3
// problems with beta == 0 take a persistent kernel instead, which moves
// operands straight from HBM into LDS and keeps the next problems in flight
// while it computes the current one.  An operand
// keeps its contiguous dimension in LDS: row-contiguous as [k][r] (pitch 80
// elements, 16-byte stores), K-contiguous as [r][k] (pitch KC + 4, 16-byte
// stores); with 64 LDS banks both make the 16-lane x 4-k MFMA fragment reads
// conflict-free.
#pragma once

#include "gemm_launch.hpp"

namespace pbgemm {
namespace small {

constexpr int kTile = 64;    // M, N of one problem tile
constexpr int kPitchR = 80;  // [k][r] layout (row-contiguous operands)

// KC: K chunk staged through LDS (32, or 64 for a single-chunk FP32 problem).
template <int KC> struct ChunkShape {
  static constexpr int kPitchK = KC + 4;  // [r][k] layout (K-contiguous operands)
  static constexpr int kLdsElems = KC * kPitchR > kTile * kPitchK ? KC * kPitchR : kTile * kPitchK;
};

// LDS index of element (r, k) of a staged operand ([r][k] when KCONT).
template <bool KCONT, int KC>
__device__ __forceinline__ int lds_at(int r, int k) {
  return KCONT ? r * ChunkShape<KC>::kPitchK + k : k * kPitchR + r;
}

// Vector loads need 16-byte alignment of every element group they touch.
template <typename T>
__device__ __forceinline__ void load_row_vec(const T* p, T out[16 / sizeof(T)]) {
  if constexpr (sizeof(T) == 4) {
    const pbdev::f32x4_t v = *reinterpret_cast<const pbdev::f32x4_t*>(p);
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
  } else {
    const pbdev::f64x2_t v = *reinterpret_cast<const pbdev::f64x2_t*>(p);
    out[0] = v.x; out[1] = v.y;
  }
}

// One K chunk of op(X) (element (r, k) at x[r*rs + k*ks], zero outside
// rows x K) for LDS as dst[k][r]: load() issues every global load of the
// thread, store() writes them to LDS.  VEC: 16-byte vectors along the
// contiguous dimension (rs == 1 or ks == 1; aligned, with a pitch that is a
// multiple of the vector width); otherwise scalar loads.  FULL (with VEC):
// rows == 64 and the chunk lies inside K, so no load needs a bounds check.
template <typename T, bool VEC, int KC, bool FULL = false>
struct Chunk {
  static constexpr int kKc = KC;
  static constexpr int VN = 16 / sizeof(T);
  static constexpr int IT = VEC ? kTile * kKc / VN / 256 : kTile * kKc / 256;
  static constexpr int W = VEC ? VN : 1;
  T v[IT][W];

  __device__ __forceinline__ void load(const T* x, int64_t rs, int64_t ks, int rows, int64_t K,
                                       int64_t k0) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = threadIdx.x + 256 * it;
      if (VEC && rs == 1) {  // vectors along r: (r .. r+VN-1, k)
        const int r = (e % (kTile / VN)) * VN, k = e / (kTile / VN);
        if (FULL || (k0 + k < K && r + VN <= rows)) {
          load_row_vec<T>(x + r + (k0 + k) * ks, v[it]);
        } else {
#pragma unroll
          for (int i = 0; i < W; ++i)
            v[it][i] = (k0 + k < K && r + i < rows) ? x[(r + i) + (k0 + k) * ks] : T(0);
        }
      } else if (VEC) {  // ks == 1, vectors along k: (r, k .. k+VN-1)
        const int k = (e % (kKc / VN)) * VN, r = e / (kKc / VN);
        if (FULL || (r < rows && k0 + k + VN <= K)) {
          load_row_vec<T>(x + r * rs + (k0 + k), v[it]);
        } else {
#pragma unroll
          for (int i = 0; i < W; ++i)
            v[it][i] = (r < rows && k0 + k + i < K) ? x[r * rs + (k0 + k + i)] : T(0);
        }
      } else {
        const int r = e % kTile, k = e / kTile;
        v[it][0] = (r < rows && k0 + k < K) ? x[r * rs + (k0 + k) * ks] : T(0);
      }
    }
  }

  // KC: the operand is K-contiguous (stored [r][k]); else [k][r].
  __device__ __forceinline__ void store(int64_t rs, bool kc, T* dst) const {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = threadIdx.x + 256 * it;
      if (VEC && rs == 1) {
        const int r = (e % (kTile / VN)) * VN, k = e / (kTile / VN);
        T* q = dst + lds_at<false, KC>(r, k);
        if constexpr (sizeof(T) == 4)
          *reinterpret_cast<pbdev::f32x4_t*>(q) = pbdev::f32x4_t{v[it][0], v[it][1], v[it][2], v[it][3]};
        else
          *reinterpret_cast<pbdev::f64x2_t*>(q) = pbdev::f64x2_t{v[it][0], v[it][1]};
      } else if (VEC) {
        const int k = (e % (kKc / VN)) * VN, r = e / (kKc / VN);
        T* q = dst + lds_at<true, KC>(r, k);  // 16-byte aligned (pitch KC + 4, k a multiple of VN)
#pragma unroll
        for (int i = 0; i < W; i += 8 / sizeof(T)) {
          if constexpr (sizeof(T) == 4) {
            typedef float f32x2 __attribute__((ext_vector_type(2)));
            *reinterpret_cast<f32x2*>(q + i) = f32x2{v[it][i], v[it][i + 1]};
          } else {
            q[i] = v[it][i];
          }
        }
      } else {
        const int r = e % kTile, k = e / kTile;
        dst[kc ? lds_at<true, KC>(r, k) : lds_at<false, KC>(r, k)] = v[it][0];
      }
    }
  }
};

// C = alpha * acc + beta * C of one problem (beta == 0: C not read).  The
// wave holds the 32 x 32 quarter at (wm, wn); accumulator e of a lane holds
// column lane % 16 and row acc_row(lane, e).
template <typename T>
__device__ __forceinline__ void store_c(const GemmProblem<T>& p, T* C, int M, int N, int wm, int wn,
                                        int lane, const typename Traits<T>::AccV (&acc)[2][2]) {
  const T alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = wn + 16 * j + lane % 16;
      if (n >= N) continue;
      if constexpr (sizeof(T) == 4) {
        // FP32: accumulators e = 0..3 are 4 consecutive rows.
        const int m = wm + 16 * i + acc_row<T>(lane, 0);
        T* cp = C + m + int64_t(n) * p.ldc;
        if (p.c_vec && m + 3 < M) {
          pbdev::f32x4_t o;
          if (beta != T(0)) {
            const pbdev::f32x4_t old = *reinterpret_cast<const pbdev::f32x4_t*>(cp);
            o = pbdev::f32x4_t{alpha * acc[i][j][0] + beta * old.x, alpha * acc[i][j][1] + beta * old.y,
                               alpha * acc[i][j][2] + beta * old.z, alpha * acc[i][j][3] + beta * old.w};
          } else {
            o = pbdev::f32x4_t{alpha * acc[i][j][0], alpha * acc[i][j][1], alpha * acc[i][j][2],
                               alpha * acc[i][j][3]};
          }
          *reinterpret_cast<pbdev::f32x4_t*>(cp) = o;
          continue;
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = wm + 16 * i + acc_row<T>(lane, e);
        if (m >= M) continue;
        T* cp = C + m * p.c_rs + int64_t(n) * p.ldc;
        const T v = alpha * acc[i][j][e];
        *cp = (beta != T(0)) ? v + beta * *cp : v;
      }
    }
}

// FULL: M == N == 64 and K a multiple of KC (VEC only).  Workgroup b runs
// problem b.
template <typename T, bool VEC, int KC, bool FULL = false>
__global__ __launch_bounds__(256, (KC == 64 && FULL) ? 4 : (VEC && sizeof(T) == 4) ? 5 : 1)
void small_gemm_kernel(GemmProblem<T> p) {
  using Acc = typename Traits<T>::AccV;  // 4 accumulators per lane (f32x4 / f64x4)
  constexpr int kKc = KC;
  __shared__ __attribute__((aligned(16))) T As[ChunkShape<KC>::kLdsElems];
  __shared__ __attribute__((aligned(16))) T Bs[ChunkShape<KC>::kLdsElems];
  // Layout per operand: [r][k] when K-contiguous and vector-staged.
  const bool akc = VEC && p.a_rs != 1 && p.a_ks == 1;
  const bool bkc = VEC && p.b_rs != 1 && p.b_ks == 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int M = int(p.M), N = int(p.N);
  const int64_t b = blockIdx.x;
  const T* A = p.a[0] + b * p.a_stride;
  const T* B = p.b[0] + b * p.b_stride;
  Chunk<T, VEC, KC, FULL> ca, cb;
  if (p.K > 0) {
    ca.load(A, p.a_rs, p.a_ks, M, p.K, 0);
    cb.load(B, p.b_rs, p.b_ks, N, p.K, 0);
  }
  Acc acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[i][j][e] = T(0);
  for (int64_t k0 = 0; k0 < p.K; k0 += kKc) {
    __syncthreads();  // the previous chunk is fully consumed
    ca.store(p.a_rs, akc, As);
    cb.store(p.b_rs, bkc, Bs);
    __syncthreads();
    if (k0 + kKc < p.K) {  // the next chunk's loads fly during this chunk's MFMAs
      ca.load(A, p.a_rs, p.a_ks, M, p.K, k0 + kKc);
      cb.load(B, p.b_rs, p.b_ks, N, p.K, k0 + kKc);
    }
#pragma unroll 4
    for (int k = 0; k < kKc; k += 4) {
      const int kk = k + lane / 16;
      T a[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = wm + 16 * i + lane % 16;
        a[i] = As[akc ? lds_at<true, KC>(r, kk) : lds_at<false, KC>(r, kk)];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r = wn + 16 * j + lane % 16;
        bv[j] = Bs[bkc ? lds_at<true, KC>(r, kk) : lds_at<false, KC>(r, kk)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if constexpr (sizeof(T) == 4)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bv[j], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i][j], 0, 0, 0);
        }
    }
  }
  store_c<T>(p, p.c + b * p.c_stride, M, N, wm, wn, lane, acc);
}

// FP32 64^3 problems with vector-aligned operands move their operands from
// HBM straight into LDS (global_load_lds, 16 bytes a lane, no register
// staging and no LDS store instructions): 16 KB per operand as 64 rows of
// 256 bytes.  An operand keeps its contiguous dimension
// along the LDS row: M/N-contiguous ones as [k][r], K-contiguous ones (KC) as
// [r][k].  The 16-byte chunk c of row rho sits at slot c ^ swz(rho); the DMA
// applies the swizzle by the choice of global chunk each lane fetches.
//
// MFMA step s (0..15) of a problem takes k = 16 (s / 4) + 4 g + s % 4 from lane
// group g = lane / 16, so a [r][k] operand reads the 4 k of 4 consecutive
// steps with one 16-byte read (swz = row & 15: the 16 lanes of a group read 16
// different slots) and a [k][r] operand reads one element per step (swz =
// 4 g: the four groups read four different 4-slot blocks of their rows).
template <bool KC>
__device__ __forceinline__ int dma_swz(int rho) {
  return KC ? (rho & 15) : (((rho >> 2) & 3) << 2);
}

// One wave's share of a problem's DMA: waves 0, 1 load rows 32 h .. 32 h + 31
// of A, waves 2, 3 of B; DMA j covers 4 rows (1 KB), lane l row 4 j + l / 16,
// slot l % 16.
template <bool KCA, bool KCB>
struct DmaShare {
  int op, h;
  int64_t src[8];  // element offset of the lane's 16-byte chunk in the operand

  __device__ __forceinline__ DmaShare(const GemmProblem<float>& p, int wave, int lane) {
    op = wave >> 1;
    h = wave & 1;
    const bool kc = op == 0 ? KCA : KCB;
    const int64_t rs = op == 0 ? p.a_rs : p.b_rs;
    const int64_t ks = op == 0 ? p.a_ks : p.b_ks;
    const int64_t pitch = kc ? rs : ks;  // elements between LDS rows in memory
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int rho = 32 * h + 4 * j + lane / 16;
      const int c = (lane & 15) ^ (kc ? dma_swz<true>(rho) : dma_swz<false>(rho));
      src[j] = int64_t(rho) * pitch + 4 * c;
    }
  }

  // Problem q's operand base for this wave.
  __device__ __forceinline__ const float* base(const GemmProblem<float>& p, int64_t q) const {
    return op == 0 ? p.a[0] + q * p.a_stride : p.b[0] + q * p.b_stride;
  }

  // Issues the wave's 8 DMAs of operand X into `lds` (A at lds[0 .. 4095],
  // B at lds[4096 .. 8191]).
  __device__ __forceinline__ void issue(const float* X, float* lds) const {
    float* dst = lds + op * kTile * kTile + 32 * h * kTile;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      __builtin_amdgcn_global_load_lds(pbdev::to_global(X + src[j]),
                                       (__attribute__((address_space(3))) void*)(dst + 4 * j * kTile),
                                       16, 0, 0);
  }
};

// The 64 MFMAs of one problem for the wave's 32 x 32 quarter (wm, wn).
template <bool KCA, bool KCB>
struct DmaMfma {
  int aoff[2], boff[2], g, wm, wn, r16;

  __device__ __forceinline__ DmaMfma(int wave, int lane) {
    wm = (wave & 1) * 32;
    wn = (wave >> 1) * 32;
    g = lane >> 4;
    r16 = lane & 15;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int ra = wm + 16 * i + r16, rb = wn + 16 * i + r16;
      // [r][k]: row r (the slot is added per step).  [k][r]: slot (r / 4) ^ 4 g,
      // element r % 4 (the row k is added per step).
      aoff[i] = KCA ? ra * kTile : (((ra >> 2) ^ (g << 2)) << 2) + (ra & 3);
      boff[i] = KCB ? rb * kTile : (((rb >> 2) ^ (g << 2)) << 2) + (rb & 3);
    }
  }

  __device__ __forceinline__ void run(const float* As, const float* Bs,
                                      Traits<float>::AccV (&acc)[2][2]) const {
#pragma unroll
    for (int s1 = 0; s1 < 4; ++s1) {
      float a[2][4], bv[2][4];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if constexpr (KCA) {
          const int ra = wm + 16 * i + r16;
          const pbdev::f32x4_t v = *reinterpret_cast<const pbdev::f32x4_t*>(
              As + aoff[i] + (((4 * s1 + g) ^ dma_swz<true>(ra)) << 2));
          a[i][0] = v.x; a[i][1] = v.y; a[i][2] = v.z; a[i][3] = v.w;
        } else {
#pragma unroll
          for (int t = 0; t < 4; ++t) a[i][t] = As[(16 * s1 + 4 * g + t) * kTile + aoff[i]];
        }
        if constexpr (KCB) {
          const int rb = wn + 16 * i + r16;
          const pbdev::f32x4_t v = *reinterpret_cast<const pbdev::f32x4_t*>(
              Bs + boff[i] + (((4 * s1 + g) ^ dma_swz<true>(rb)) << 2));
          bv[i][0] = v.x; bv[i][1] = v.y; bv[i][2] = v.z; bv[i][3] = v.w;
        } else {
#pragma unroll
          for (int t = 0; t < 4; ++t) bv[i][t] = Bs[(16 * s1 + 4 * g + t) * kTile + boff[i]];
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][t], bv[j][t], acc[i][j], 0, 0, 0);
    }
  }
};

__device__ __forceinline__ void zero_acc(Traits<float>::AccV (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[i][j][e] = 0.f;
}

// 64^3 problems (K == 64), beta == 0, vector-aligned C: a persistent grid of
// one workgroup per CU whose NB LDS buffers of 32 KB hold the next NB - 1
// problems in flight while it computes the current one.  Workgroups that each
// load their problem, compute and store move in step (all of a CU's
// workgroups wait for HBM, then all compute while HBM idles); here every
// workgroup keeps HBM busy through its MFMAs.  Workgroup w runs problems
// w, w + G, w + 2 G, ... (G = the grid), so the grid streams through the
// batch in order.
//
// Per problem and wave: 8 DMAs, then 4 C stores.  Problem it's DMAs are
// followed by the stores of problem it - D (D = NB - 1) and, for each of the
// D problems after it, 8 DMAs (and, but for the last, 4 stores): the wait for
// them is vmcnt(12 D) (vector memory operations complete in order), vmcnt(8 D)
// in the first D steps (fewer stores behind them), vmcnt(0) in the last D.
// The counts assume no other vector memory operations in the loop: no
// register spills (the kernel needs about 150 of the 512 registers a lone
// wave per SIMD may use).
template <bool KCA, bool KCB, int NB>
__global__ __launch_bounds__(256) void small_stream_kernel(GemmProblem<float> p) {
  constexpr int kElems = 2 * kTile * kTile;
  // One array per buffer, indexed by constants only (the loop is unrolled by
  // NB), so the compiler's own LDS-DMA waits stay per buffer.
  __shared__ __attribute__((aligned(16))) float L0[kElems];
  __shared__ __attribute__((aligned(16))) float L1[kElems];
  __shared__ __attribute__((aligned(16))) float L2[NB > 2 ? kElems : 4];
  __shared__ __attribute__((aligned(16))) float L3[NB > 3 ? kElems : 4];
  __shared__ __attribute__((aligned(16))) float L4[NB > 4 ? kElems : 4];
  auto buf = [&](auto r) -> float* {
    if constexpr (decltype(r)::value == 0) return L0;
    else if constexpr (decltype(r)::value == 1) return L1;
    else if constexpr (decltype(r)::value == 2) return L2;
    else if constexpr (decltype(r)::value == 3) return L3;
    else return L4;
  };
  constexpr int D = NB - 1;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const DmaShare<KCA, KCB> d(p, wave, lane);
  const DmaMfma<KCA, KCB> mm(wave, lane);
  const int64_t G = gridDim.x;
  const int64_t w = blockIdx.x;
  const int64_t n = (p.batch - w + G - 1) / G;  // problems of this workgroup (>= 1)
  const float alpha = p.alpha;

  auto step = [&](auto r, int64_t it) {
    constexpr int R = decltype(r)::value;
    constexpr int RN = (R + D) % NB;  // buffer of problem it + D
    // __syncthreads() would also wait for every DMA in flight (its fence
    // drains the vector memory counter), so the barriers are written out.
    if (it + D < n) {
      d.issue(d.base(p, w + (it + D) * G), buf(std::integral_constant<int, RN>{}));
      if (it < D) {  // no stores of problems before it + 1 - D yet: 8 D or more after it
        if constexpr (D == 1) asm volatile("s_waitcnt vmcnt(8)\n\ts_barrier" ::: "memory");
        else if constexpr (D == 2) asm volatile("s_waitcnt vmcnt(16)\n\ts_barrier" ::: "memory");
        else if constexpr (D == 3) asm volatile("s_waitcnt vmcnt(24)\n\ts_barrier" ::: "memory");
        else asm volatile("s_waitcnt vmcnt(32)\n\ts_barrier" ::: "memory");
      } else {
        if constexpr (D == 1) asm volatile("s_waitcnt vmcnt(12)\n\ts_barrier" ::: "memory");
        else if constexpr (D == 2) asm volatile("s_waitcnt vmcnt(24)\n\ts_barrier" ::: "memory");
        else if constexpr (D == 3) asm volatile("s_waitcnt vmcnt(36)\n\ts_barrier" ::: "memory");
        else asm volatile("s_waitcnt vmcnt(48)\n\ts_barrier" ::: "memory");
      }
    } else {
      asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
    }
    Traits<float>::AccV acc[2][2];
    zero_acc(acc);
    const float* X = buf(r);
    mm.run(X, X + kTile * kTile, acc);
    float* C = p.c + (w + it * G) * p.c_stride;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int m = mm.wm + 16 * i + acc_row<float>(lane, 0);
        const int nn = mm.wn + 16 * j + lane % 16;
        *reinterpret_cast<pbdev::f32x4_t*>(C + m + int64_t(nn) * p.ldc) =
            pbdev::f32x4_t{alpha * acc[i][j][0], alpha * acc[i][j][1], alpha * acc[i][j][2],
                           alpha * acc[i][j][3]};
      }
    // Every wave's reads of buffer R are done: it is free for problem it + NB.
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };

  auto prologue = [&](auto r) {
    constexpr int R = decltype(r)::value;
    if constexpr (R < D)
      if (R < n) d.issue(d.base(p, w + R * G), buf(r));
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  prologue(I0{});
  prologue(I1{});
  prologue(I2{});
  prologue(I3{});
  for (int64_t it = 0; it < n; it += NB) {
    step(I0{}, it);
    if (it + 1 < n) step(I1{}, it + 1);
    if constexpr (NB > 2)
      if (it + 2 < n) step(I2{}, it + 2);
    if constexpr (NB > 3)
      if (it + 3 < n) step(I3{}, it + 3);
    if constexpr (NB > 4)
      if (it + 4 < n) step(I4{}, it + 4);
  }
}

// One problem per workgroup pays off for a batch of at least one problem
// per CU with M, N <= 64 (interleaved batches keep the tile engine).
inline bool eligible(pb_handle_t h, int64_t M, int64_t N, int64_t batch, bool interleaved) {
  return !interleaved && M <= kTile && N <= kTile && batch >= h->num_cus && batch <= 0x7fffffff;
}

// LDS buffers of the persistent 64^3 kernel (NB - 1 problems in flight).
// On 4096 x 64^3 (back-to-back calls / one call after a cache flush, us):
// NB = 2 (two workgroups per CU) 35.6 / 40.9, NB = 3 33.7 / 37.7, NB = 4
// 35.0 / 39.3, NB = 5 36.7 / 41.0; the one-problem-per-workgroup kernel
// 36.7 / 40.6.  Streaming (non-temporal) C stores: 35.0 / 44.7 at NB = 3.
// Four buffers with two problems in flight (no barrier after a step's reads,
// since the buffer refilled next was read two steps back): 33.6 / 37.7, the
// same as three buffers with one barrier more per problem.
constexpr int kStreamBuffers = 3;

template <typename T>
pb_status_t launch(pb_handle_t h, const GemmProblem<T>& p, bool vec) {
  const dim3 grid{unsigned(p.batch)};
  if constexpr (sizeof(T) == 4) {
    if (vec && p.M == kTile && p.N == kTile && p.K == 64 && p.beta == 0.f && p.c_vec) {
      const bool kca = p.a_ks == 1, kcb = p.b_ks == 1;
      constexpr int NB = kStreamBuffers;
      // Workgroups of NB x 32 KB that fit a CU's 160 KB of LDS (one for NB = 3).
      const dim3 sgrid{unsigned(std::min<int64_t>(p.batch, int64_t(h->num_cus) * (5 / NB)))};
#define PB_SMALL_STREAM(A, B) \
  hipLaunchKernelGGL((small_stream_kernel<A, B, NB>), sgrid, dim3(256), 0, h->stream, p)
      if (kca && kcb) PB_SMALL_STREAM(true, true);
      else if (kca) PB_SMALL_STREAM(true, false);
      else if (kcb) PB_SMALL_STREAM(false, true);
      else PB_SMALL_STREAM(false, false);
#undef PB_SMALL_STREAM
      PB_CHECK_LAUNCH();
      return PB_STATUS_SUCCESS;
    }
    if (vec && p.M == kTile && p.N == kTile && p.K % 64 == 0) {
      hipLaunchKernelGGL((small_gemm_kernel<T, true, 64, true>), grid, dim3(256), 0, h->stream, p);
      PB_CHECK_LAUNCH();
      return PB_STATUS_SUCCESS;
    }
  }
  if (vec) {
    hipLaunchKernelGGL((small_gemm_kernel<T, true, 32>), grid, dim3(256), 0, h->stream, p);
  } else {
    hipLaunchKernelGGL((small_gemm_kernel<T, false, 32>), grid, dim3(256), 0, h->stream, p);
  }
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // namespace small
}  // namespace pbgemm

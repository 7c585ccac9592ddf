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
// chunk's loads are issued before the current chunk's MFMAs.  An operand
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
  const T alpha = p.alpha, beta = p.beta;
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
  // C = alpha * acc + beta * C (beta == 0: C not read).  Accumulator e of a
  // lane holds column lane % 16 and row acc_row(lane, e).
  T* C = p.c + b * p.c_stride;
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

// One problem per workgroup pays off for a batch of at least one problem
// per CU with M, N <= 64 (interleaved batches keep the tile engine).
inline bool eligible(pb_handle_t h, int64_t M, int64_t N, int64_t batch, bool interleaved) {
  return !interleaved && M <= kTile && N <= kTile && batch >= h->num_cus && batch <= 0x7fffffff;
}

template <typename T>
pb_status_t launch(pb_handle_t h, const GemmProblem<T>& p, bool vec) {
  const dim3 grid{unsigned(p.batch)};
  if constexpr (sizeof(T) == 4) {
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

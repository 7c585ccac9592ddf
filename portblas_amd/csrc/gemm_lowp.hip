// Reduced-precision GEMMs with FP32 accumulation: BF16x3 (split precision),
// BF16 and per-tensor-scaled FP8-e4m3, plus the BF16-input GEMM and the tile
// configuration table the autotuner searches.
//
// FP32 inputs go through one pre-pass per operand that reads op(X) once and
// writes it K-contiguous (transposing through LDS when X is M/N-contiguous) in
// the low-precision format:
//   BF16x3:  hi = bf16(x), lo = bf16(x - hi); C = A_hi B_lo + A_lo B_hi + A_hi B_hi
//   BF16x4:  BF16x3 plus the A_lo B_lo term
//   BF16:    hi only
//   FP8:     q = e4m3(x * 448 / amax(X)); the epilogue multiplies by
//            amax(A) amax(B) / 448^2 (read on the device, no host round trip)
//   FP16x3:  x_s = x * 2^s(X) with amax(X) 2^s(X) in [2^14, 2^15) (amax over
//            the finite elements), hi = f16(x_s), lo = f16(x_s - hi);
//            C = 2^-(s(A) + s(B)) (A_hi B_lo + A_lo B_hi + A_hi B_hi), each
//            operand kept to 22 significand bits (BF16x3: 16)
// The GEMM then runs on the packed operands (every operand K-contiguous and
// 16-byte aligned): the 256 x 256 fused-term kernel (gemm_mt.hpp) when the
// problem fills the GPU with such tiles, else the MFMA tile engine.
#include "gemm_front.hpp"
#include "gemm_launch.hpp"
#include "gemm_mt.hpp"
#include "lowp.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace pbgemm {

pb_status_t sgemm_plain(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                        float alpha, const float* a, int64_t lda, int64_t sa, const float* b,
                        int64_t ldb, int64_t sb, float beta, float* c, int64_t ldc, int64_t sc,
                        int64_t batch, bool simt);
pb_status_t sgemm_alpha_zero(pb_handle_t h, int64_t M, int64_t N, float beta, float* C,
                             int64_t ldc, int64_t stridec, int64_t batch);

// ---------------------------------------------------------------------------
// Pre-pass kernels
// ---------------------------------------------------------------------------
// amax pre-pass.  Every workgroup writes the largest |x| it saw to its own
// slot of a partial array (no atomics, nothing to clear first); the scales
// kernel folds the partials.  NaN never wins: fmaxf returns the other
// operand.
constexpr int kAmaxBlock = 256;
constexpr int kAmaxMaxParts = 2048;  // partial slots per operand (grid cap)

// |x| as the amax passes count it: FIN (the FP16 split's scale) skips
// infinities, which the scaled parts carry through unchanged.
template <bool FIN>
__device__ __forceinline__ float amag(float x) {
  const float a = fabsf(x);
  if constexpr (FIN) return a < __builtin_huge_valf() ? a : 0.f;
  return a;
}

// Largest value over the workgroup, written by thread 0 to part[blockIdx.x].
__device__ __forceinline__ void amax_block_store(float best, float* part) {
  __shared__ float red[kAmaxBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_xor(best, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
#pragma unroll
    for (int w = 1; w < kAmaxBlock / 64; ++w) m = fmaxf(m, red[w]);
    part[blockIdx.x] = m;
  }
}

// amax over the logical rows x K elements of op(X) (all batches), any strides.
template <bool FIN>
__global__ __launch_bounds__(kAmaxBlock) void amax_kernel(const float* __restrict__ x, int64_t rs, int64_t ks,
                                                          int64_t rows, int64_t K, int64_t stride, int64_t batch,
                                                          float* part) {
  const int64_t per = rows * K;
  const int64_t total = per * batch;
  float best = 0.f;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = i / per, e = i % per;
    const int64_t r = rs == 1 ? e % rows : e / K;
    const int64_t k = rs == 1 ? e / rows : e % K;
    best = fmaxf(best, amag<FIN>(x[b * stride + r * rs + k * ks]));
  }
  amax_block_store(best, part);
}

// amax of a dense block of n floats (an operand stored without padding):
// 16-byte non-temporal loads, 8 in flight per lane.
template <bool FIN>
__global__ __launch_bounds__(kAmaxBlock) void amax_flat_kernel(const float* __restrict__ x, int64_t n,
                                                               float* part) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  constexpr int U = 8;
  const int64_t nvec = n / 4;
  const f4* xv = reinterpret_cast<const f4*>(x);
  float best = 0.f;
  const int64_t step = int64_t(gridDim.x) * kAmaxBlock * U;
  for (int64_t base = int64_t(blockIdx.x) * kAmaxBlock * U + threadIdx.x; base < nvec; base += step) {
    f4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = base + u * kAmaxBlock < nvec ? __builtin_nontemporal_load(xv + base + u * kAmaxBlock) : f4{0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < U; ++u)
      best = fmaxf(best, fmaxf(fmaxf(amag<FIN>(v[u].x), amag<FIN>(v[u].y)),
                               fmaxf(amag<FIN>(v[u].z), amag<FIN>(v[u].w))));
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nvec * 4) best = fmaxf(best, amag<FIN>(x[nvec * 4 + threadIdx.x]));
  amax_block_store(best, part);
}

// amax over `nlines` lines (per batch) of `len` contiguous floats, line j of
// batch b at x + b * stride + j * pitch: one workgroup per line at a time,
// coalesced along the line, no per-element index arithmetic.
template <bool FIN>
__global__ __launch_bounds__(kAmaxBlock) void amax_lines_kernel(const float* __restrict__ x, int64_t len,
                                                                int64_t pitch, int64_t nlines, int64_t stride,
                                                                int64_t batch, float* part) {
  float best = 0.f;
  for (int64_t l = blockIdx.x; l < nlines * batch; l += gridDim.x) {
    const float* line = x + (l / nlines) * stride + (l % nlines) * pitch;
    for (int64_t i = threadIdx.x; i < len; i += kAmaxBlock) best = fmaxf(best, amag<FIN>(line[i]));
  }
  amax_block_store(best, part);
}

// Folds each operand's partial maxima (na and nb slots) into amax[0..1]
// (thread 0 of the workgroup).
__device__ __forceinline__ void fold_amax(const float* pa, int na, const float* pb, int nb, float* amax) {
  __shared__ float red[2][kAmaxBlock / 64];
  float a = 0.f, b = 0.f;
  for (int i = threadIdx.x; i < na; i += kAmaxBlock) a = fmaxf(a, pa[i]);
  for (int i = threadIdx.x; i < nb; i += kAmaxBlock) b = fmaxf(b, pb[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = fmaxf(a, __shfl_xor(a, off, 64));
    b = fmaxf(b, __shfl_xor(b, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int w = 1; w < kAmaxBlock / 64; ++w) {
    red[0][0] = fmaxf(red[0][0], red[0][w]);
    red[1][0] = fmaxf(red[1][0], red[1][w]);
  }
  amax[0] = red[0][0];
  amax[1] = red[1][0];
}

// Power-of-two scale exponent of one FP8 e4m3 operand: the largest s with
// amax * 2^s <= 448 (e4m3's largest finite value), so no element saturates
// and the scaling itself is exact.  0 when the amax is 0 or not finite;
// clamped to [-63, 63] so that 2^-(s(A) + s(B)) stays a normal float.
__device__ __forceinline__ int fp8_scale_exp(float amax) {
  if (!(amax > 0.f && amax < __builtin_huge_valf())) return 0;
  int e;
  const float f = frexpf(amax, &e);  // amax = f 2^e, f in [0.5, 1); 448 = 0.875 2^9
  const int s = (f <= 0.875f ? 9 : 8) - e;
  return s < -63 ? -63 : (s > 63 ? 63 : s);
}

// FP8 scales: scales[0] = 2^s(A), scales[1] = 2^s(B), scales[2] =
// 2^-(s(A) + s(B)).
__global__ __launch_bounds__(kAmaxBlock) void fp8_scales_kernel(const float* pa, int na, const float* pb, int nb,
                                                                float* scales) {
  float am[2];
  fold_amax(pa, na, pb, nb, am);
  if (threadIdx.x != 0) return;
  const int ea = fp8_scale_exp(am[0]), eb = fp8_scale_exp(am[1]);
  scales[0] = ldexpf(1.0f, ea);
  scales[1] = ldexpf(1.0f, eb);
  scales[2] = ldexpf(1.0f, -(ea + eb));
}

// Power-of-two scale exponent of one FP16x3 operand: amax * 2^s in
// [2^14, 2^15), so |hi| <= 2^15 stays finite and elements down to 2^-28
// amax keep normal hi and lo parts.  Clamped to [-63, 74]: the unscaling
// factor 2^-(s(A) + s(B)) then stays a finite, exactly representable float.
__device__ __forceinline__ int f16_scale_exp(float amax) {
  if (!(amax > 0.f)) return 0;  // all zero (or no finite element)
  int e;
  frexpf(amax, &e);  // amax = f 2^e, f in [0.5, 1)
  int s = 15 - e;
  return s < -63 ? -63 : (s > 74 ? 74 : s);
}

// FP16x3 scales: scales[0] = 2^s(A), scales[1] = 2^s(B),
// scales[2] = 2^-(s(A) + s(B)), all exact.
// With `factor` (one operand, nb == 0), also factor[0] = 2^-s(A), the
// operand's share of the unscaling.
__global__ __launch_bounds__(kAmaxBlock) void f16_scales_kernel(const float* pa, int na, const float* pb, int nb,
                                                                float* scales, float* factor) {
  float am[2];
  fold_amax(pa, na, pb, nb, am);
  if (threadIdx.x != 0) return;
  const int ea = f16_scale_exp(am[0]), eb = f16_scale_exp(am[1]);
  scales[0] = ldexpf(1.0f, ea);
  scales[1] = ldexpf(1.0f, eb);
  scales[2] = ldexpf(1.0f, -(ea + eb));
  if (factor) factor[0] = ldexpf(1.0f, -ea);
}

// FP16x3 (FP8: FP8) scales from packs that used guessed exponents
// guess[0..1]: writes scales as f16_scales_kernel (fp8_scales_kernel) does,
// flags[i] = 1 where the true exponent differs from the guess (that operand
// is packed again), and the true exponents into guess for the next call.
// One workgroup of kCheckBlock threads, 8 independent loads in flight per
// thread and operand (the packs leave one partial per workgroup: 8192 per
// 8192^2 operand).
constexpr int kCheckBlock = 1024;
template <bool FP8>
__global__ __launch_bounds__(kCheckBlock) void scale_check_kernel(const float* pa, int na, const float* pb, int nb,
                                                                  int* guess, float* scales, int* flags) {
  __shared__ float red[2][kCheckBlock / 64];
  constexpr int U = 8;
  float m[2][U];
#pragma unroll
  for (int u = 0; u < U; ++u) m[0][u] = m[1][u] = 0.f;
  for (int64_t i0 = threadIdx.x; i0 < (na > nb ? na : nb); i0 += kCheckBlock * U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + u * kCheckBlock;
      if (i < na) m[0][u] = fmaxf(m[0][u], pa[i]);
      if (i < nb) m[1][u] = fmaxf(m[1][u], pb[i]);
    }
  }
  float a = m[0][0], b = m[1][0];
#pragma unroll
  for (int u = 1; u < U; ++u) {
    a = fmaxf(a, m[0][u]);
    b = fmaxf(b, m[1][u]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = fmaxf(a, __shfl_xor(a, off, 64));
    b = fmaxf(b, __shfl_xor(b, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  float am[2] = {red[0][0], red[1][0]};
#pragma unroll
  for (int w = 1; w < kCheckBlock / 64; ++w) {
    am[0] = fmaxf(am[0], red[0][w]);
    am[1] = fmaxf(am[1], red[1][w]);
  }
  const int ea = FP8 ? fp8_scale_exp(am[0]) : f16_scale_exp(am[0]);
  const int eb = FP8 ? fp8_scale_exp(am[1]) : f16_scale_exp(am[1]);
  scales[0] = ldexpf(1.0f, ea);
  scales[1] = ldexpf(1.0f, eb);
  scales[2] = ldexpf(1.0f, -(ea + eb));
  flags[0] = ea != guess[0];
  flags[1] = eb != guess[1];
  guess[0] = ea;
  guess[1] = eb;
}

enum PackMode { kPackBf16 = 0, kPackBf16Split = 1, kPackFp8 = 2, kPackF16Split = 3 };

// Modes that multiply by a per-tensor scale (scales[scale_idx]) first.
constexpr bool scaled_mode(int mode) { return mode == kPackFp8 || mode == kPackF16Split; }

// IEEE binary16 bits of x, rounded to nearest even (subnormals kept).
__device__ __forceinline__ uint16_t f16_bits(_Float16 x) { return __builtin_bit_cast(uint16_t, x); }

// Index of packed element (r, k) of batch b.  slab == 0: row-major
// [batch][rows][kp]; slab > 0: K-slab-major [batch][kp / slab][rows][slab],
// the layout the 256 x 256 kernel streams (one K step of 32 consecutive rows
// is one contiguous kilobyte).
__device__ __forceinline__ int64_t packed_index(int64_t b, int64_t r, int64_t k, int64_t rows,
                                                int64_t kp, int slab) {
  if (slab == 0) return (b * rows + r) * kp + k;
  return ((b * (kp / slab) + k / slab) * rows + r) * slab + (k % slab);
}

// Writes op(X) (rows x K, element (r, k) at x[r*rs + k*ks], batch b at
// + b*stride) in the packed layout (packed_index; zero for k in [K, kp)).
// 64 x 64 tiles through LDS: reads walk X's contiguous dimension, writes walk K.
template <int MODE>
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ x, int64_t rs,
                                                   int64_t ks, int64_t rows, int64_t K,
                                                   int64_t stride, int64_t kp, int slab,
                                                   void* dst_hi, void* dst_lo, const float* scales,
                                                   int scale_idx) {
  __shared__ float tile[64][65];
  const int64_t r0 = int64_t(blockIdx.x) * 64;
  const int64_t k0 = int64_t(blockIdx.y) * 64;
  const int64_t b = blockIdx.z;
  const float* xb = x + b * stride;
  const bool rows_contig = (rs == 1);
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int rr = rows_contig ? e % 64 : e / 64;
    const int kk = rows_contig ? e / 64 : e % 64;
    const int64_t r = r0 + rr, k = k0 + kk;
    tile[rr][kk] = (r < rows && k < K) ? xb[r * rs + k * ks] : 0.f;
  }
  __syncthreads();
  const float s = scaled_mode(MODE) ? scales[scale_idx] : 1.0f;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int rr = e / 64, kk = e % 64;
    const int64_t r = r0 + rr, k = k0 + kk;
    if (r >= rows || k >= kp) continue;
    const int64_t o = packed_index(b, r, k, rows, kp, slab);
    const float v = tile[rr][kk];
    if constexpr (MODE == kPackFp8) {
      static_cast<uint8_t*>(dst_hi)[o] = pblowp::f32_to_e4m3(v * s);
    } else if constexpr (MODE == kPackF16Split) {
      const float vs = v * s;
      const _Float16 hi = _Float16(vs);
      static_cast<uint16_t*>(dst_hi)[o] = f16_bits(hi);
      static_cast<uint16_t*>(dst_lo)[o] = f16_bits(_Float16(vs - float(hi)));
    } else {
      const uint16_t hi = pblowp::f32_to_bf16(v);
      static_cast<uint16_t*>(dst_hi)[o] = hi;
      if constexpr (MODE == kPackBf16Split)
        static_cast<uint16_t*>(dst_lo)[o] = pblowp::f32_to_bf16(v - pblowp::bf16_to_f32(hi));
    }
  }
}

// Converts v (already scaled for FP8) to the packed format(s).
template <int MODE>
__device__ __forceinline__ void pack_convert(float v, uint16_t& hi, uint16_t& lo, uint8_t& q) {
  if constexpr (MODE == kPackFp8) {
    q = pblowp::f32_to_e4m3(v);
  } else if constexpr (MODE == kPackF16Split) {
    const _Float16 h = _Float16(v);
    hi = f16_bits(h);
    lo = f16_bits(_Float16(v - float(h)));  // v - h is exact in FP32
  } else {
    hi = pblowp::f32_to_bf16(v);
    if constexpr (MODE == kPackBf16Split) lo = pblowp::f32_to_bf16(v - pblowp::bf16_to_f32(hi));
  }
}

// Where pack_vec_kernel takes its scale from.
enum PackScale { kScaleGiven = 0, kScaleGuess = 1, kScaleRepack = 2 };

// 64-wide K tiles that one workgroup of pack_vec_kernel converts in turn.
constexpr int kPackKTiles = 2;

// Vectorized pre-pass for 16-byte aligned X with 4-element aligned pitches.
// Same contract as pack_kernel, with grid.y covering kPackKTiles K tiles per
// workgroup.  A 64 x 64 tile of op(X) is read with 16-byte loads along X's
// contiguous dimension (KC: along k; else along the rows) into LDS as
// [row][k]; then thread t converts the 16 consecutive k 16 (t / 64) .. +15 of
// row t % 64, so 64 neighbouring threads write 64 consecutive rows of one K
// slab (contiguous in the slab-major layout).  The next tile's loads are in
// flight while the current one is converted and written.
//
// SC selects where the FP16x3 scale comes from.  kScaleGiven: scales[scale_idx].
// kScaleGuess: 2^(*ctl), a guess of the operand's scale exponent; the
// workgroup also writes the finite amax of its elements to
// amax_part[linear block index].  kScaleRepack: scales[scale_idx], and only
// when *ctl is nonzero (the guess was wrong); otherwise every workgroup exits.
template <int MODE, bool KC, int SC = kScaleGiven>
__global__ __launch_bounds__(256) void pack_vec_kernel(const float* __restrict__ x, int64_t rs,
                                                       int64_t ks, int64_t rows, int64_t K,
                                                       int64_t stride, int64_t kp, int slab,
                                                       void* dst_hi, void* dst_lo,
                                                       const float* scales, int scale_idx,
                                                       const int* ctl, float* amax_part) {
  if constexpr (SC == kScaleRepack)
    if (*ctl == 0) return;
  __shared__ float tile[64][65];  // [row][k]
  const int64_t r0 = int64_t(KC ? blockIdx.y : blockIdx.x) * 64;
  const int64_t kt0 = int64_t(KC ? blockIdx.x : blockIdx.y) * kPackKTiles;
  const int64_t b = blockIdx.z;
  const float* xb = x + b * stride;
  int64_t nkt = (kp + 63) / 64 - kt0;
  if (nkt > kPackKTiles) nkt = kPackKTiles;
  float v[4][4];
  auto load = [&](int64_t k0) {
    if (r0 + 64 <= rows && k0 + 64 <= K) {
      // Interior tile (the same branch for the whole workgroup): four
      // unconditional 16-byte loads in flight together.  With per-lane
      // conditions the compiler waits for each load before the next one that
      // writes the same registers on the other path.
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int idx = threadIdx.x + 256 * i;
        const int rr = KC ? idx / 16 : 4 * (idx % 16);
        const int kk = KC ? 4 * (idx % 16) : idx / 16;
        const float* src = KC ? xb + (r0 + rr) * rs + k0 + kk : xb + r0 + rr + (k0 + kk) * ks;
        const pbdev::f32x4_t t = pbdev::ld_stream(reinterpret_cast<const pbdev::f32x4_t*>(src));
        v[i][0] = t.x; v[i][1] = t.y; v[i][2] = t.z; v[i][3] = t.w;
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int idx = threadIdx.x + 256 * i;
      // KC: row idx / 16, k 4 (idx % 16); else k idx / 16, row 4 (idx % 16).
      const int rr = KC ? idx / 16 : 4 * (idx % 16);
      const int kk = KC ? 4 * (idx % 16) : idx / 16;
      const int64_t r = r0 + rr, k = k0 + kk;
#pragma unroll
      for (int e = 0; e < 4; ++e) v[i][e] = 0.f;
      if constexpr (KC) {
        if (r < rows) {
          if (k + 3 < K) {
            const pbdev::f32x4_t t =
                pbdev::ld_stream(reinterpret_cast<const pbdev::f32x4_t*>(xb + r * rs + k));
            v[i][0] = t.x; v[i][1] = t.y; v[i][2] = t.z; v[i][3] = t.w;
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[i][e] = (k + e < K) ? xb[r * rs + k + e] : 0.f;
          }
        }
      } else {
        if (k < K) {
          if (r + 3 < rows) {
            const pbdev::f32x4_t t =
                pbdev::ld_stream(reinterpret_cast<const pbdev::f32x4_t*>(xb + r + k * ks));
            v[i][0] = t.x; v[i][1] = t.y; v[i][2] = t.z; v[i][3] = t.w;
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[i][e] = (r + e < rows) ? xb[r + e + k * ks] : 0.f;
          }
        }
      }
    }
  };
  auto to_lds = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int idx = threadIdx.x + 256 * i;
      const int rr = KC ? idx / 16 : 4 * (idx % 16);
      const int kk = KC ? 4 * (idx % 16) : idx / 16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if constexpr (KC) tile[rr][kk + e] = v[i][e];
        else tile[rr + e][kk] = v[i][e];
      }
    }
  };
  const float s = SC == kScaleGuess ? ldexpf(1.0f, *ctl) : scaled_mode(MODE) ? scales[scale_idx] : 1.0f;
  float best = 0.f;  // finite amax of the converted elements (kScaleGuess)
  const int rr = threadIdx.x % 64, kq = 16 * (threadIdx.x / 64);
  auto convert = [&](int64_t k0) {
    const int64_t r = r0 + rr, k = k0 + kq;
    if (r >= rows || k >= kp) return;
    // 16 consecutive k of one row are contiguous in every packed layout (kp
    // and the slab width are multiples of 16).
    const int64_t o = packed_index(b, r, k, rows, kp, slab);
    uint16_t h[16], l[16];
    uint8_t q[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const float t = tile[rr][kq + e];
      // Every element of op(X) is converted by exactly one thread.
      if constexpr (SC == kScaleGuess) best = fmaxf(best, amag<MODE == kPackF16Split>(t));
      pack_convert<MODE>(t * s, h[e], l[e], q[e]);
    }
    if constexpr (MODE == kPackFp8) {
      uint4 w;
      __builtin_memcpy(&w, q, 16);
      *reinterpret_cast<uint4*>(static_cast<uint8_t*>(dst_hi) + o) = w;
    } else {
      uint4 w[2];
      __builtin_memcpy(w, h, 32);
      uint4* dh = reinterpret_cast<uint4*>(static_cast<uint16_t*>(dst_hi) + o);
      dh[0] = w[0];
      dh[1] = w[1];
      if constexpr (MODE == kPackBf16Split || MODE == kPackF16Split) {
        __builtin_memcpy(w, l, 32);
        uint4* dl = reinterpret_cast<uint4*>(static_cast<uint16_t*>(dst_lo) + o);
        dl[0] = w[0];
        dl[1] = w[1];
      }
    }
  };
  // The barriers are written out: __syncthreads() would also wait for the
  // previous tile's stores and the next tile's loads.
  load(kt0 * 64);
  for (int64_t t = 0; t < nkt; ++t) {
    if (t > 0) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");  // tile consumed
    to_lds();
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (t + 1 < nkt) load((kt0 + t + 1) * 64);
    convert((kt0 + t) * 64);
  }
  if constexpr (SC == kScaleGuess) {
    __shared__ float red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_xor(best, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0)
      amax_part[blockIdx.x + int64_t(gridDim.x) * (blockIdx.y + int64_t(gridDim.y) * blockIdx.z)] =
          fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  }
}

// Grid of pack_vec_kernel<·, KC> for op(X) with `rows` rows and packed width
// kp.  Consecutive workgroups take neighbouring tiles along X's contiguous
// dimension: row tiles on x for M/N-contiguous operands, K tiles on x for
// K-contiguous ones (KC).
inline dim3 pack_vec_grid(int64_t rows, int64_t kp, int64_t batch, bool kc) {
  const unsigned rt = unsigned(pb::ceil_div(rows, 64));
  const unsigned kt = unsigned(pb::ceil_div(pb::ceil_div(kp, 64), kPackKTiles));
  return kc ? dim3{kt, rt, unsigned(batch)} : dim3{rt, kt, unsigned(batch)};
}

// Whether that grid's y extent fits the launch limit.
inline bool pack_vec_fits(int64_t rows, int64_t kp, bool kc) {
  return (kc ? pb::ceil_div(rows, 64) : pb::ceil_div(pb::ceil_div(kp, 64), kPackKTiles)) <= 65535;
}

// ---------------------------------------------------------------------------
// Tile configurations (the autotuner's search space for the BF16 engine):
// 8 tile shapes x 5 split-K factors.
// ---------------------------------------------------------------------------
using B128 = TileShape<128, 128, 64, 64, 64>;
using B128x64 = TileShape<128, 64, 64, 64, 32>;
using B64x128 = TileShape<64, 128, 64, 32, 64>;
using B64 = TileShape<64, 64, 64, 32, 32>;
using B256x128 = TileShape<256, 128, 64, 128, 64>;
using B128x256 = TileShape<128, 256, 64, 64, 128>;
using B128k32 = TileShape<128, 128, 32, 64, 64>;
using B256 = TileShape<256, 256, 64, 128, 128>;
using Q128 = TileShape<128, 128, 128, 64, 64>;  // FP8 (K group of 128)

constexpr int kNumShapes = 8;
constexpr int kSplits[] = {1, 2, 4, 8, 16};
constexpr int kNumSplits = 5;
constexpr int kNumConfigs = kNumShapes * kNumSplits;
const char* kShapeNames[kNumShapes] = {"128x128x64", "128x64x64", "64x128x64", "64x64x64",
                                      "256x128x64", "128x256x64", "128x128x32", "256x256x64"};
const int kShapeBM[kNumShapes] = {128, 128, 64, 64, 256, 128, 128, 256};
const int kShapeBN[kNumShapes] = {128, 64, 128, 64, 128, 256, 128, 256};

std::string& config_name_storage(int id) {
  static std::vector<std::string> names = [] {
    std::vector<std::string> v;
    for (int s = 0; s < kNumShapes; ++s)
      for (int k = 0; k < kNumSplits; ++k)
        v.push_back(std::string(kShapeNames[s]) + "/splitk" + std::to_string(kSplits[k]));
    return v;
  }();
  return names[id];
}

template <typename T>
pb_status_t launch_config(pb_handle_t h, int id, const GemmProblem<T>& p, bool kca, bool kcb,
                          bool vec) {
  const int shape = id / kNumSplits, splits = kSplits[id % kNumSplits];
  switch (shape) {
    case 0: return launch_gemm<T, B128>(h, p, kca, kcb, vec, splits);
    case 1: return launch_gemm<T, B128x64>(h, p, kca, kcb, vec, splits);
    case 2: return launch_gemm<T, B64x128>(h, p, kca, kcb, vec, splits);
    case 3: return launch_gemm<T, B64>(h, p, kca, kcb, vec, splits);
    case 4: return launch_gemm<T, B256x128>(h, p, kca, kcb, vec, splits);
    case 5: return launch_gemm<T, B128x256>(h, p, kca, kcb, vec, splits);
    case 6: return launch_gemm<T, B128k32>(h, p, kca, kcb, vec, splits);
    case 7: return launch_gemm<T, B256>(h, p, kca, kcb, vec, splits);
  }
  return pb::set_error(PB_STATUS_INVALID_VALUE, "config id %d out of range", id);
}

// Split-K partial bytes config `id` needs for an M x N x K problem (bf16).
size_t config_splitk_bytes(int id, int64_t M, int64_t N, int64_t K, int64_t batch) {
  const int shape = id / kNumSplits, splits = kSplits[id % kNumSplits];
  switch (shape) {
    case 0: return splitk_bytes<bf16_t, B128>(M, N, K, batch, splits);
    case 1: return splitk_bytes<bf16_t, B128x64>(M, N, K, batch, splits);
    case 2: return splitk_bytes<bf16_t, B64x128>(M, N, K, batch, splits);
    case 3: return splitk_bytes<bf16_t, B64>(M, N, K, batch, splits);
    case 4: return splitk_bytes<bf16_t, B256x128>(M, N, K, batch, splits);
    case 5: return splitk_bytes<bf16_t, B128x256>(M, N, K, batch, splits);
    case 6: return splitk_bytes<bf16_t, B128k32>(M, N, K, batch, splits);
    case 7: return splitk_bytes<bf16_t, B256>(M, N, K, batch, splits);
  }
  return 0;
}

// Built-in rule (used when no learned selector is installed): the largest
// tile that still gives every CU two workgroups, else split K.
int default_config(pb_handle_t h, int64_t M, int64_t N, int64_t K, int64_t batch) {
  const int64_t cus = h->num_cus;
  auto tiles = [&](int bm, int bn) { return pb::ceil_div(M, bm) * pb::ceil_div(N, bn) * batch; };
  if (tiles(256, 256) >= cus && M >= 2048 && N >= 2048) return 7 * kNumSplits;
  if (tiles(128, 128) >= 2 * cus) return 0;
  if (tiles(64, 64) >= 2 * cus) return 3 * kNumSplits;
  const int64_t t = tiles(64, 64);
  int s = 0;
  while (s + 1 < kNumSplits && t * kSplits[s] < 2 * cus && K / kSplits[s + 1] >= 256) ++s;
  return 3 * kNumSplits + s;
}

int selector_choose(pb_handle_t h, char ta, char tb, int64_t M, int64_t N, int64_t K,
                    int64_t batch);

// amax of op(X) (rows x K, element (r, k) at x[r*rs + k*ks], batch b at +
// b*stride) into one partial per workgroup; returns the partial count.
// Dense operands stream as one block; others go line by line along their
// contiguous dimension.
template <bool FIN>
int launch_amax(pb_handle_t h, const float* x, int64_t rs, int64_t ks, int64_t rows, int64_t K, int64_t stride,
                int64_t batch, float* part) {
  const int64_t per = rows * K;
  const bool dense = ((rs == 1 && ks == rows) || (ks == 1 && rs == K)) && (batch == 1 || stride == per);
  int64_t grid;
  if (dense && aligned16(x)) {
    grid = std::min<int64_t>(pb::ceil_div(per * batch, 4 * 8 * kAmaxBlock), 8 * h->num_cus);
    grid = std::max<int64_t>(std::min<int64_t>(grid, kAmaxMaxParts), 1);
    hipLaunchKernelGGL(amax_flat_kernel<FIN>, dim3(unsigned(grid)), dim3(kAmaxBlock), 0, h->stream, x,
                       per * batch, part);
  } else if (rs == 1 || ks == 1) {
    const int64_t len = rs == 1 ? rows : K, nl = rs == 1 ? K : rows, pitch = rs == 1 ? ks : rs;
    grid = std::min<int64_t>(nl * batch, 8 * h->num_cus);
    grid = std::max<int64_t>(std::min<int64_t>(grid, kAmaxMaxParts), 1);
    hipLaunchKernelGGL(amax_lines_kernel<FIN>, dim3(unsigned(grid)), dim3(kAmaxBlock), 0, h->stream, x, len,
                       pitch, nl, stride, batch, part);
  } else {
    grid = std::min<int64_t>(pb::ceil_div(per * batch, kAmaxBlock), 4 * h->num_cus);
    grid = std::max<int64_t>(std::min<int64_t>(grid, kAmaxMaxParts), 1);
    hipLaunchKernelGGL(amax_kernel<FIN>, dim3(unsigned(grid)), dim3(kAmaxBlock), 0, h->stream, x, rs, ks, rows,
                       K, stride, batch, part);
  }
  return int(grid);
}

// Tile plan of the FP16 tile-engine GEMMs: 128 x 128 tiles when they give
// every CU two workgroups, else 64 x 64 tiles with K split until they do.
struct F16Plan { bool big = false; int splits = 1; };

inline F16Plan plan_f16(pb_handle_t h, int64_t M, int64_t N, int64_t K, int64_t batch) {
  const int64_t cus = h->num_cus;
  const int64_t t128 = pb::ceil_div(M, 128) * pb::ceil_div(N, 128) * batch;
  const int64_t t64 = pb::ceil_div(M, 64) * pb::ceil_div(N, 64) * batch;
  F16Plan pl;
  if (t128 >= 2 * cus) { pl.big = true; return pl; }
  while (t64 * pl.splits < 2 * cus && K / (pl.splits * 2) >= 256 && pl.splits < 16) pl.splits *= 2;
  return pl;
}

// ---------------------------------------------------------------------------
// Runners
// ---------------------------------------------------------------------------
struct Scratch {
  uint8_t* base = nullptr;
  size_t off = 0;
  template <typename U>
  U* take(size_t count) {
    U* p = reinterpret_cast<U*>(base + off);
    off = pb::align_up(off + count * sizeof(U), 256);
    return p;
  }
};

// Runs C = alpha * sum_t op(A_t) op(B_t) + beta * C on BF16 (or FP8)
// operands with config `id` (-1: pick one).
template <typename T>
pb_status_t run_lowp(pb_handle_t h, int id, GemmProblem<T> p, bool kca, bool kcb, bool vec) {
  if (id < 0) {
    id = selector_choose(h, 'n', 'n', p.M, p.N, p.K, p.batch);
    if (id < 0) id = default_config(h, p.M, p.N, p.K, p.batch);
  }
  PB_REQUIRE(id >= 0 && id < kNumConfigs, "config id %d out of range", id);
  return launch_config<T>(h, id, p, kca, kcb, vec);
}

// The 256 x 256 kernel pays off when its tiles fill every CU at least once
// and little of them is padding.
bool mt_eligible(pb_handle_t h, int64_t M, int64_t N, int64_t K, int64_t batch) {
  if (batch > 65535 || K <= 0) return false;
  const int64_t tiles = pb::ceil_div(M, mt::kBM) * pb::ceil_div(N, mt::kBN) * batch;
  const double waste = double(tiles) * mt::kBM * mt::kBN / (double(M) * double(N) * double(batch));
  return tiles >= h->num_cus && waste < 1.15 && mt::fits(M, N, pb::align_up(size_t(K), 128));
}

pb_status_t gemm_ex_lowp(pb_handle_t h, pb_gemm_precision_t prec, const GemmCall<float>& g) {
  const int64_t M = g.M, N = g.N, K = g.K, batch = g.batch;
  const bool fp8 = prec == PB_GEMM_FP8_E4M3;
  const bool f16s = prec == PB_GEMM_FP16X3;
  const bool split = prec == PB_GEMM_BF16X3 || prec == PB_GEMM_BF16X4 || f16s;
  const bool use_mt = mt_eligible(h, M, N, K, batch);
  // The 256 x 256 kernel's K step: 16 (BF16 / FP16 hi/lo), 32 (BF16) or 64 (FP8)
  // elements, 32 or 64 bytes per row.
  const int slab = use_mt ? (split ? 16 : fp8 ? 64 : 32) : 0;
  // Packed pitch: the 256 x 256 kernel's ring alignment (4 K steps), else 16
  // (the vectorized pre-pass converts 16 consecutive k per thread).
  const int64_t kp = pb::align_up(size_t(K > 0 ? K : 1), use_mt ? size_t(4 * slab) : 16);
  const size_t esz = fp8 ? 1 : 2;
  const int nparts = split ? 2 : 1;
  const size_t abytes = pb::align_up(size_t(M) * kp * esz * batch, 256);
  const size_t bbytes = pb::align_up(size_t(N) * kp * esz * batch, 256);
  // The GEMM config (and so its split-K partials) is fixed before the
  // workspace is taken: packed operands and partials share one allocation.
  int id = -1, fp8_splits = 1;
  size_t part_bytes = 0;
  F16Plan f16p{};
  if (f16s && !use_mt) {
    f16p = plan_f16(h, M, N, K, batch);
    part_bytes = f16p.big ? 0 : splitk_bytes<fp16_t, B64>(M, N, K, batch, f16p.splits);
  } else if (fp8 && !use_mt) {
    const int64_t t = pb::ceil_div(M, 128) * pb::ceil_div(N, 128) * batch;
    while (t * fp8_splits < 2 * h->num_cus && K / (fp8_splits * 2) >= 512 && fp8_splits < 16)
      fp8_splits *= 2;
    part_bytes = splitk_bytes<fp8_t, Q128>(M, N, K, batch, fp8_splits);
  } else if (!use_mt) {
    id = selector_choose(h, 'n', 'n', M, N, K, batch);
    if (id < 0) id = default_config(h, M, N, K, batch);
    PB_REQUIRE(id >= 0 && id < kNumConfigs, "config id %d out of range", id);
    part_bytes = config_splitk_bytes(id, M, N, K, batch);
  }
  // Operand strides of op(A) (rows M) and op(B) (rows N), as for the engine.
  const int64_t a_rs = g.ta == 'n' ? 1 : g.lda, a_ks = g.ta == 'n' ? g.lda : 1;
  const int64_t b_rs = g.tb == 'n' ? g.ldb : 1, b_ks = g.tb == 'n' ? 1 : g.ldb;
  PB_REQUIRE(batch <= 65535, "batch too large for the pre-pass");
  // 16-byte loads need an aligned base and 4-element pitches; the packed
  // pitch kp is a multiple of 8.
  auto vec_ok = [&](const float* x, int64_t rs, int64_t ks, int64_t rows, int64_t stride) {
    const int64_t pitch = ks == 1 ? rs : ks;
    return aligned16(x) && pitch % 4 == 0 && (batch == 1 || stride % 4 == 0) && (rs == 1 || ks == 1) &&
           pack_vec_fits(rows, kp, ks == 1);
  };
  // FP16x3 and FP8 with both operands on the vectorized pre-pass skip the
  // amax passes: each operand is packed with the scale exponent of the
  // handle's previous call of the same precision while the pack takes its
  // amax; a check kernel then sets the true scales and flags an operand whose
  // exponent changed, and that operand alone is packed again (the other
  // repack launch exits at once).  The packed planes and scales are those of
  // the two-pass order.
  const bool guess = (f16s || fp8) && vec_ok(g.a, a_rs, a_ks, M, g.sa) && vec_ok(g.b, b_rs, b_ks, N, g.sb);
  const dim3 ga = pack_vec_grid(M, kp, batch, a_ks == 1), gb = pack_vec_grid(N, kp, batch, b_ks == 1);
  const int64_t gna = int64_t(ga.x) * ga.y * ga.z, gnb = int64_t(gb.x) * gb.y * gb.z;
  PB_REQUIRE(!guess || (gna <= INT32_MAX && gnb <= INT32_MAX), "operand too large for the pre-pass");
  const size_t guess_bytes = guess ? pb::align_up(4 * size_t(gna), 256) + pb::align_up(4 * size_t(gnb), 256) + 256 : 0;
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, nparts * (abytes + bbytes) + part_bytes + guess_bytes + 1024, &ws));
  Scratch sc{static_cast<uint8_t*>(ws)};
  void* a_hi = sc.take<uint8_t>(abytes);
  void* a_lo = nparts == 2 ? sc.take<uint8_t>(abytes) : nullptr;
  void* b_hi = sc.take<uint8_t>(bbytes);
  void* b_lo = nparts == 2 ? sc.take<uint8_t>(bbytes) : nullptr;
  float* scales = sc.take<float>(4);
  float* amax_part = sc.take<float>(2 * kAmaxMaxParts);
  float* partial = part_bytes ? sc.take<float>(part_bytes / sizeof(float)) : nullptr;

  if (guess) {
    float* pa = sc.take<float>(size_t(gna));
    float* pbp = sc.take<float>(size_t(gnb));
    int* flags = sc.take<int>(2);
    const bool kca = a_ks == 1, kcb = b_ks == 1;
    int* state = h->scale_guess + (fp8 ? 2 : 0);  // this precision's exponents (A, B)
    auto packg = [&](auto sc_mode, bool kc, dim3 grid, const float* x, int64_t rs, int64_t ks, int64_t rows,
                     int64_t stride, void* hi, void* lo, int sidx, const int* ctl, float* part) {
      constexpr int SCM = decltype(sc_mode)::value;
      const float* sp = scales;
#define PB_PACKG(MODE, KC)                                                                                      \
  hipLaunchKernelGGL((pack_vec_kernel<MODE, KC, SCM>), grid, dim3(256), 0, h->stream, x, rs, ks, rows, K, stride, \
                     kp, slab, hi, lo, sp, sidx, ctl, part)
      if (fp8) {
        if (kc) PB_PACKG(kPackFp8, true);
        else PB_PACKG(kPackFp8, false);
      } else {
        if (kc) PB_PACKG(kPackF16Split, true);
        else PB_PACKG(kPackF16Split, false);
      }
#undef PB_PACKG
    };
    using Guess = std::integral_constant<int, kScaleGuess>;
    using Repack = std::integral_constant<int, kScaleRepack>;
    packg(Guess{}, kca, ga, g.a, a_rs, a_ks, M, g.sa, a_hi, a_lo, 0, state, pa);
    packg(Guess{}, kcb, gb, g.b, b_rs, b_ks, N, g.sb, b_hi, b_lo, 1, state + 1, pbp);
    if (fp8)
      hipLaunchKernelGGL(scale_check_kernel<true>, dim3(1), dim3(kCheckBlock), 0, h->stream, pa, int(gna), pbp,
                         int(gnb), state, scales, flags);
    else
      hipLaunchKernelGGL(scale_check_kernel<false>, dim3(1), dim3(kCheckBlock), 0, h->stream, pa, int(gna), pbp,
                         int(gnb), state, scales, flags);
    packg(Repack{}, kca, ga, g.a, a_rs, a_ks, M, g.sa, a_hi, a_lo, 0, flags, static_cast<float*>(nullptr));
    packg(Repack{}, kcb, gb, g.b, b_rs, b_ks, N, g.sb, b_hi, b_lo, 1, flags + 1, static_cast<float*>(nullptr));
    PB_CHECK_LAUNCH();
  } else {
    if (fp8 || f16s) {
      // amax of each operand over all batches (one partial per workgroup),
      // then the scales (device side).  Dense operands stream as one block;
      // others go line by line along their contiguous dimension.  The FP16
      // split's amax counts finite elements only.
      auto amax = [&](const float* x, int64_t rs, int64_t ks, int64_t rows, int64_t stride, float* part) -> int {
        return f16s ? launch_amax<true>(h, x, rs, ks, rows, K, stride, batch, part)
                    : launch_amax<false>(h, x, rs, ks, rows, K, stride, batch, part);
      };
      const int na = amax(g.a, a_rs, a_ks, M, g.sa, amax_part);
      const int nb = amax(g.b, b_rs, b_ks, N, g.sb, amax_part + kAmaxMaxParts);
      if (f16s)
        hipLaunchKernelGGL(f16_scales_kernel, dim3(1), dim3(kAmaxBlock), 0, h->stream, amax_part, na,
                           amax_part + kAmaxMaxParts, nb, scales, static_cast<float*>(nullptr));
      else
        hipLaunchKernelGGL(fp8_scales_kernel, dim3(1), dim3(kAmaxBlock), 0, h->stream, amax_part, na,
                           amax_part + kAmaxMaxParts, nb, scales);
      PB_CHECK_LAUNCH();
    }
    auto pack = [&](const float* x, int64_t rs, int64_t ks, int64_t rows, int64_t stride, void* hi,
                    void* lo, int sidx) {
      dim3 grid{unsigned(pb::ceil_div(rows, 64)), unsigned(pb::ceil_div(kp, 64)), unsigned(batch)};
      const bool kc = ks == 1;
      const int* none_i = nullptr;
      float* none_f = nullptr;
      if (vec_ok(x, rs, ks, rows, stride)) {
        const dim3 vgrid = pack_vec_grid(rows, kp, batch, kc);
#define PB_PACKV(MODE)                                                                       \
    do {                                                                                       \
      if (kc)                                                                                  \
        hipLaunchKernelGGL((pack_vec_kernel<MODE, true>), vgrid, dim3(256), 0, h->stream, x, rs, \
                           ks, rows, K, stride, kp, slab, hi, lo, scales, sidx, none_i, none_f); \
      else                                                                                     \
        hipLaunchKernelGGL((pack_vec_kernel<MODE, false>), vgrid, dim3(256), 0, h->stream, x,  \
                           rs, ks, rows, K, stride, kp, slab, hi, lo, scales, sidx, none_i, none_f); \
    } while (0)
        if (f16s) PB_PACKV(kPackF16Split);
        else if (split) PB_PACKV(kPackBf16Split);
        else if (prec == PB_GEMM_BF16) PB_PACKV(kPackBf16);
        else PB_PACKV(kPackFp8);
#undef PB_PACKV
        return;
      }
      if (f16s)
        hipLaunchKernelGGL(pack_kernel<kPackF16Split>, grid, dim3(256), 0, h->stream, x, rs, ks,
                           rows, K, stride, kp, slab, hi, lo, scales, sidx);
      else if (split)
        hipLaunchKernelGGL(pack_kernel<kPackBf16Split>, grid, dim3(256), 0, h->stream, x, rs, ks,
                           rows, K, stride, kp, slab, hi, lo, scales, sidx);
      else if (prec == PB_GEMM_BF16)
        hipLaunchKernelGGL(pack_kernel<kPackBf16>, grid, dim3(256), 0, h->stream, x, rs, ks, rows, K,
                           stride, kp, slab, hi, lo, scales, sidx);
      else
        hipLaunchKernelGGL(pack_kernel<kPackFp8>, grid, dim3(256), 0, h->stream, x, rs, ks, rows, K,
                           stride, kp, slab, hi, lo, scales, sidx);
    };
    pack(g.a, a_rs, a_ks, M, g.sa, a_hi, a_lo, 0);
    pack(g.b, b_rs, b_ks, N, g.sb, b_hi, b_lo, 1);
    PB_CHECK_LAUNCH();
  }

  const bool c_vec = aligned16(g.c) && g.ldc % 4 == 0 && (batch == 1 || g.sc % 4 == 0);
  if (fp8 && use_mt) {
    mt::Problem q{};
    q.M = int(M); q.N = int(N); q.kp = int(kp);
    q.alpha = g.alpha; q.beta = g.beta; q.alpha_dev = scales + 2;  // the dequantisation factor
    q.a[0] = a_hi; q.b[0] = b_hi;
    q.a_stride = M * kp; q.b_stride = N * kp;
    q.c = g.c; q.ldc = g.ldc; q.c_stride = g.sc; q.batch = int(batch); q.c_vec = c_vec;
    return mt::launch<64, 1, 1>(h, q);
  }
  if (fp8) {
    GemmProblem<fp8_t> p{};
    p.M = M; p.N = N; p.K = K; p.alpha = g.alpha; p.beta = g.beta;
    p.alpha_dev = scales + 2;
    p.nterms = 1;
    p.a[0] = static_cast<const uint8_t*>(a_hi);
    p.b[0] = static_cast<const uint8_t*>(b_hi);
    p.a_rs = kp; p.a_ks = 1; p.a_stride = M * kp;
    p.b_rs = kp; p.b_ks = 1; p.b_stride = N * kp;
    p.c = g.c; p.c_rs = 1; p.ldc = g.ldc; p.c_stride = g.sc; p.batch = batch; p.c_vec = c_vec;
    p.partial = partial;
    return launch_gemm<fp8_t, Q128>(h, p, true, true, true, fp8_splits);
  }
  const auto* ah = static_cast<const uint16_t*>(a_hi);
  const auto* al = static_cast<const uint16_t*>(a_lo);
  const auto* bh = static_cast<const uint16_t*>(b_hi);
  const auto* bl = static_cast<const uint16_t*>(b_lo);
  if (f16s && use_mt) {
    mt::Problem q{};
    q.M = int(M); q.N = int(N); q.kp = int(kp);
    q.alpha = g.alpha; q.beta = g.beta; q.alpha_dev = scales + 2;  // 2^-(s(A) + s(B))
    q.a[0] = ah; q.a[1] = al; q.b[0] = bh; q.b[1] = bl;
    q.a_stride = M * kp; q.b_stride = N * kp;
    q.c = g.c; q.ldc = g.ldc; q.c_stride = g.sc; q.batch = int(batch); q.c_vec = c_vec;
    return mt::launch<16, 3, 2, true>(h, q);
  }
  if (f16s) {
    GemmProblem<fp16_t> p{};
    p.M = M; p.N = N; p.K = K; p.alpha = g.alpha; p.beta = g.beta;
    p.alpha_dev = scales + 2;
    p.nterms = 3;  // small terms first, then the leading one
    p.a[0] = ah; p.b[0] = bl;
    p.a[1] = al; p.b[1] = bh;
    p.a[2] = ah; p.b[2] = bh;
    p.a_rs = kp; p.a_ks = 1; p.a_stride = M * kp;
    p.b_rs = kp; p.b_ks = 1; p.b_stride = N * kp;
    p.c = g.c; p.c_rs = 1; p.ldc = g.ldc; p.c_stride = g.sc; p.batch = batch; p.c_vec = c_vec;
    p.partial = partial;
    if (f16p.big) return launch_gemm<fp16_t, B128>(h, p, true, true, true, 1);
    return launch_gemm<fp16_t, B64>(h, p, true, true, true, f16p.splits);
  }
  if (use_mt) {
    mt::Problem q{};
    q.M = int(M); q.N = int(N); q.kp = int(kp);
    q.alpha = g.alpha; q.beta = g.beta; q.alpha_dev = nullptr;
    q.a[0] = ah; q.a[1] = al; q.b[0] = bh; q.b[1] = bl;
    q.a_stride = M * kp; q.b_stride = N * kp;
    q.c = g.c; q.ldc = g.ldc; q.c_stride = g.sc; q.batch = int(batch); q.c_vec = c_vec;
    if (prec == PB_GEMM_BF16X3) return mt::launch<16, 3>(h, q);
    if (prec == PB_GEMM_BF16X4) return mt::launch<16, 4>(h, q);
    return mt::launch<32, 1>(h, q);
  }
  GemmProblem<bf16_t> p{};
  p.M = M; p.N = N; p.K = K; p.alpha = g.alpha; p.beta = g.beta;
  if (split) {
    // Small terms first, then the leading term.
    p.nterms = prec == PB_GEMM_BF16X4 ? 4 : 3;
    p.a[0] = ah; p.b[0] = bl;
    p.a[1] = al; p.b[1] = bh;
    if (prec == PB_GEMM_BF16X4) {
      p.a[2] = al; p.b[2] = bl;
    }
    p.a[p.nterms - 1] = ah; p.b[p.nterms - 1] = bh;
  } else {
    p.nterms = 1;
    p.a[0] = ah; p.b[0] = bh;
  }
  p.a_rs = kp; p.a_ks = 1; p.a_stride = M * kp;
  p.b_rs = kp; p.b_ks = 1; p.b_stride = N * kp;
  p.c = g.c; p.c_rs = 1; p.ldc = g.ldc; p.c_stride = g.sc; p.batch = batch; p.c_vec = c_vec;
  p.partial = partial;
  return launch_config<bf16_t>(h, id, p, true, true, true);
}

// BF16-input GEMM on the caller's layout (no pre-pass).
pb_status_t gemm_bf16_direct(pb_handle_t h, int id, const GemmCall<uint16_t>& g, float alpha,
                             float beta, float* c) {
  GemmProblem<bf16_t> p{};
  p.M = g.M; p.N = g.N; p.K = g.K; p.alpha = alpha; p.beta = beta;
  p.nterms = 1;
  p.a[0] = g.a;
  p.b[0] = g.b;
  p.batch = g.batch;
  const bool vec = set_operands(p, g.ta, g.tb, g.lda, g.sa, g.ldb, g.sb, false);
  p.c = c; p.c_rs = 1; p.ldc = g.ldc; p.c_stride = g.sc;
  p.c_vec = aligned16(c) && g.ldc % 4 == 0 && (g.batch == 1 || g.sc % 4 == 0);
  if (id < 0) {
    id = selector_choose(h, g.ta, g.tb, g.M, g.N, g.K, g.batch);
    if (id < 0) id = default_config(h, g.M, g.N, g.K, g.batch);
  }
  PB_REQUIRE(id >= 0 && id < kNumConfigs, "config id %d out of range", id);
  return launch_config<bf16_t>(h, id, p, g.ta == 't', g.tb == 'n', vec);
}

pb_status_t gemm_bf16_entry(pb_handle_t h, int id, char ta, char tb, int64_t m, int64_t n,
                            int64_t k, float alpha, const uint16_t* a, int64_t lda, int64_t sa,
                            const uint16_t* b, int64_t ldb, int64_t sb, float beta, float* c,
                            int64_t ldc, int64_t sc, int64_t batch) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  GemmCall<uint16_t> g;
  bool done = false;
  PB_TRY((gemm_front<float, uint16_t>(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c,
                                      ldc, sc, batch, PB_BATCH_STRIDED, g, done)));
  if (done) return PB_STATUS_SUCCESS;
  if (alpha == 0.f) return sgemm_alpha_zero(h, m, n, beta, c, ldc, sc, batch);
  return gemm_bf16_direct(h, id, g, alpha, beta, c);
}

// ---------------------------------------------------------------------------
// FP16 GEMMs (the reference's _gemm on sycl::half, BLAS_ENABLE_HALF:
// src/interface/blas3/backend/default.hpp:150-246, amd_gpu.hpp:292-359).  FP16
// inputs through v_mfma_f32_32x32x16_f16 with FP32 accumulation; two of the
// BF16 engine's tile shapes (128 x 128 for problems that fill the GPU with
// them, else 64 x 64 with split-K), chosen like the FP32 plan.
// ---------------------------------------------------------------------------
pb_status_t gemm_f16_direct(pb_handle_t h, const GemmCall<uint16_t>& g, float alpha, float beta, float* c,
                            int64_t ldc, int64_t sc) {
  GemmProblem<fp16_t> p{};
  p.M = g.M; p.N = g.N; p.K = g.K; p.alpha = alpha; p.beta = beta;
  p.nterms = 1;
  p.a[0] = g.a;
  p.b[0] = g.b;
  p.batch = g.batch;
  const bool vec = set_operands(p, g.ta, g.tb, g.lda, g.sa, g.ldb, g.sb, false);
  p.c = c; p.c_rs = 1; p.ldc = ldc; p.c_stride = sc;
  p.c_vec = aligned16(c) && ldc % 4 == 0 && (g.batch == 1 || sc % 4 == 0);
  const F16Plan pl = plan_f16(h, g.M, g.N, g.K, g.batch);
  if (pl.big) return launch_gemm<fp16_t, B128>(h, p, g.ta == 't', g.tb == 'n', vec, 1);
  return launch_gemm<fp16_t, B64>(h, p, g.ta == 't', g.tb == 'n', vec, pl.splits);
}

// C (half) <- half(T + beta * C) with T already alpha * op(A) op(B) in FP32
// (t == nullptr: T = 0); beta == 0 never reads C.
__global__ void f16_store_kernel(int64_t M, int64_t N, const float* __restrict__ t, int64_t ldt,
                                 int64_t st, float beta, _Float16* c, int64_t ldc, int64_t sc, int64_t batch) {
  const int64_t per = M * N, total = per * batch;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = e / per, r = e % per, i = r % M, j = r / M;
    _Float16* cp = c + b * sc + i + j * ldc;
    float v = t ? t[b * st + i + j * ldt] : 0.f;
    if (beta != 0.f) v += beta * float(*cp);
    *cp = _Float16(v);
  }
}

}  // namespace pbgemm

using namespace pbgemm;

extern "C" {

pb_status_t pb_gemm_ex(pb_handle_t h, pb_gemm_precision_t prec, char ta, char tb, int64_t m,
                       int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                       int64_t sa, const float* b, int64_t ldb, int64_t sb, float beta, float* c,
                       int64_t ldc, int64_t sc, int64_t batch) {
  if (prec == PB_GEMM_FP32_MFMA || prec == PB_GEMM_FP32_SIMT)
    return sgemm_plain(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch,
                       prec == PB_GEMM_FP32_SIMT);
  pb::clear_error();
  PB_TRY(pb::enter(h));
  if (prec != PB_GEMM_BF16X3 && prec != PB_GEMM_BF16X4 && prec != PB_GEMM_BF16 &&
      prec != PB_GEMM_FP8_E4M3 && prec != PB_GEMM_FP16X3)
    return pb::set_error(PB_STATUS_INVALID_VALUE, "invalid precision %d", int(prec));
  GemmCall<float> g;
  bool done = false;
  PB_TRY(gemm_front<float>(h, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                           batch, PB_BATCH_STRIDED, g, done));
  if (done) return PB_STATUS_SUCCESS;
  if (alpha == 0.f) return sgemm_alpha_zero(h, m, n, beta, c, ldc, sc, batch);
  return gemm_ex_lowp(h, prec, g);
}

pb_status_t pb_pack_bf16x3(pb_handle_t h, int operand, char trans, int64_t rows, int64_t k,
                           const float* x, int64_t ldx, uint16_t* hi, uint16_t* lo,
                           int64_t ldp) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  const char t = pb::norm_trans(trans);
  PB_REQUIRE(t != 0, "invalid transpose flag");
  PB_REQUIRE(operand == 0 || operand == 1, "operand must be 0 (A) or 1 (B)");
  PB_REQUIRE(rows >= 0 && k >= 0, "negative dimension");
  if (rows == 0 || k == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(x && hi && lo, "null pointer");
  PB_REQUIRE(ldp >= k && ldp % mt::kKAlign<16> == 0, "ldp must be >= k and a multiple of %d",
             mt::kKAlign<16>);
  // A: op(A) is rows x k; B: op(B) is k x rows, addressed as (row, k).
  int64_t rs, ks;
  if (operand == 0) { rs = t == 'n' ? 1 : ldx; ks = t == 'n' ? ldx : 1; }
  else { rs = t == 'n' ? ldx : 1; ks = t == 'n' ? 1 : ldx; }
  PB_REQUIRE(ldx >= ((rs == 1) ? rows : k), "leading dimension too small");
  dim3 grid{unsigned(pb::ceil_div(rows, 64)), unsigned(pb::ceil_div(ldp, 64)), 1u};
  const bool vec = aligned16(x) && (rs == 1 ? ks : rs) % 4 == 0 && (rs == 1 || ks == 1) &&
                   pack_vec_fits(rows, ldp, ks == 1);
  void* dh = static_cast<void*>(hi);
  void* dl = static_cast<void*>(lo);
  const float* none = nullptr;
  constexpr int slab = 16;
  if (vec && ks == 1)
    hipLaunchKernelGGL((pack_vec_kernel<kPackBf16Split, true>), pack_vec_grid(rows, ldp, 1, true), dim3(256), 0,
                       h->stream, x, rs, ks, rows, k, int64_t(0), ldp, slab, dh, dl, none, 0,
                       static_cast<const int*>(nullptr), static_cast<float*>(nullptr));
  else if (vec)
    hipLaunchKernelGGL((pack_vec_kernel<kPackBf16Split, false>), pack_vec_grid(rows, ldp, 1, false), dim3(256), 0,
                       h->stream, x, rs, ks, rows, k, int64_t(0), ldp, slab, dh, dl, none, 0,
                       static_cast<const int*>(nullptr), static_cast<float*>(nullptr));
  else
    hipLaunchKernelGGL(pack_kernel<kPackBf16Split>, grid, dim3(256), 0, h->stream, x, rs, ks,
                       rows, k, int64_t(0), ldp, slab, dh, dl, none, 0);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_gemm_bf16x3_packed(pb_handle_t h, int64_t m, int64_t n, int64_t k, float alpha,
                                  const uint16_t* a_hi, const uint16_t* a_lo, int64_t ldpa,
                                  const uint16_t* b_hi, const uint16_t* b_lo, int64_t ldpb,
                                  float beta, float* c, int64_t ldc) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(m >= 0 && n >= 0 && k >= 0, "negative dimension");
  if (m == 0 || n == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(c != nullptr && ldc >= m, "bad C");
  PB_REQUIRE(k == 0 || (a_hi && a_lo && b_hi && b_lo), "null operand");
  PB_REQUIRE(ldpa == ldpb && ldpa >= k && ldpa % mt::kKAlign<16> == 0,
             "packed pitches must match, be >= k and a multiple of %d", mt::kKAlign<16>);
  // alpha == 0 or an empty contraction: C <- beta C.
  if (alpha == 0.f || k == 0) return sgemm_alpha_zero(h, m, n, beta, c, ldc, 0, 1);
  PB_REQUIRE(aligned16(a_hi) && aligned16(a_lo) && aligned16(b_hi) && aligned16(b_lo),
             "packed operands must be 16-byte aligned");
  PB_REQUIRE(mt::fits(m, n, ldpa), "packed operands too large for 32-bit offsets");
  // The packed format is the fused kernel's (K-slab-major, zero K padding,
  // written by pb_pack_bf16x3); it runs for every shape.
  mt::Problem q{};
  q.M = int(m); q.N = int(n); q.kp = int(ldpa);
  q.alpha = alpha; q.beta = beta; q.alpha_dev = nullptr;
  q.a[0] = a_hi; q.a[1] = a_lo; q.b[0] = b_hi; q.b[1] = b_lo;
  q.c = c; q.ldc = ldc; q.batch = 1; q.c_vec = aligned16(c) && ldc % 4 == 0;
  return mt::launch<16, 3>(h, q);
}

pb_status_t pb_pack_fp16x3(pb_handle_t h, int operand, char trans, int64_t rows, int64_t k,
                           const float* x, int64_t ldx, uint16_t* hi, uint16_t* lo, int64_t ldp,
                           float* factor) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  const char t = pb::norm_trans(trans);
  PB_REQUIRE(t != 0, "invalid transpose flag");
  PB_REQUIRE(operand == 0 || operand == 1, "operand must be 0 (A) or 1 (B)");
  PB_REQUIRE(rows >= 0 && k >= 0, "negative dimension");
  if (rows == 0 || k == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(x && hi && lo && factor, "null pointer");
  PB_REQUIRE(ldp >= k && ldp % mt::kKAlign<16> == 0, "ldp must be >= k and a multiple of %d",
             mt::kKAlign<16>);
  int64_t rs, ks;
  if (operand == 0) { rs = t == 'n' ? 1 : ldx; ks = t == 'n' ? ldx : 1; }
  else { rs = t == 'n' ? ldx : 1; ks = t == 'n' ? 1 : ldx; }
  PB_REQUIRE(ldx >= ((rs == 1) ? rows : k), "leading dimension too small");
  // The operand's amax over its finite elements, its scale, then the split.
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, sizeof(float) * (kAmaxMaxParts + 64), &ws));
  float* scales = static_cast<float*>(ws);
  float* part = scales + 64;
  const int na = launch_amax<true>(h, x, rs, ks, rows, k, 0, 1, part);
  hipLaunchKernelGGL(f16_scales_kernel, dim3(1), dim3(kAmaxBlock), 0, h->stream, part, na,
                     static_cast<const float*>(nullptr), 0, scales, factor);
  dim3 grid{unsigned(pb::ceil_div(rows, 64)), unsigned(pb::ceil_div(ldp, 64)), 1u};
  const bool vec = aligned16(x) && (rs == 1 ? ks : rs) % 4 == 0 && (rs == 1 || ks == 1) &&
                   pack_vec_fits(rows, ldp, ks == 1);
  void* dh = static_cast<void*>(hi);
  void* dl = static_cast<void*>(lo);
  constexpr int slab = 16;
  if (vec && ks == 1)
    hipLaunchKernelGGL((pack_vec_kernel<kPackF16Split, true>), pack_vec_grid(rows, ldp, 1, true), dim3(256), 0,
                       h->stream, x, rs, ks, rows, k, int64_t(0), ldp, slab, dh, dl,
                       static_cast<const float*>(scales), 0, static_cast<const int*>(nullptr),
                       static_cast<float*>(nullptr));
  else if (vec)
    hipLaunchKernelGGL((pack_vec_kernel<kPackF16Split, false>), pack_vec_grid(rows, ldp, 1, false), dim3(256), 0,
                       h->stream, x, rs, ks, rows, k, int64_t(0), ldp, slab, dh, dl,
                       static_cast<const float*>(scales), 0, static_cast<const int*>(nullptr),
                       static_cast<float*>(nullptr));
  else
    hipLaunchKernelGGL(pack_kernel<kPackF16Split>, grid, dim3(256), 0, h->stream, x, rs, ks,
                       rows, k, int64_t(0), ldp, slab, dh, dl, static_cast<const float*>(scales), 0);
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

pb_status_t pb_gemm_fp16x3_packed(pb_handle_t h, int64_t m, int64_t n, int64_t k, float alpha,
                                  const uint16_t* a_hi, const uint16_t* a_lo, const float* a_factor,
                                  int64_t ldpa, const uint16_t* b_hi, const uint16_t* b_lo,
                                  const float* b_factor, int64_t ldpb, float beta, float* c,
                                  int64_t ldc) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(m >= 0 && n >= 0 && k >= 0, "negative dimension");
  if (m == 0 || n == 0) return PB_STATUS_SUCCESS;
  PB_REQUIRE(c != nullptr && ldc >= m, "bad C");
  PB_REQUIRE(k == 0 || (a_hi && a_lo && b_hi && b_lo && a_factor && b_factor), "null operand");
  PB_REQUIRE(ldpa == ldpb && ldpa >= k && ldpa % mt::kKAlign<16> == 0,
             "packed pitches must match, be >= k and a multiple of %d", mt::kKAlign<16>);
  if (alpha == 0.f || k == 0) return sgemm_alpha_zero(h, m, n, beta, c, ldc, 0, 1);
  PB_REQUIRE(aligned16(a_hi) && aligned16(a_lo) && aligned16(b_hi) && aligned16(b_lo),
             "packed operands must be 16-byte aligned");
  PB_REQUIRE(mt::fits(m, n, ldpa), "packed operands too large for 32-bit offsets");
  mt::Problem q{};
  q.M = int(m); q.N = int(n); q.kp = int(ldpa);
  q.alpha = alpha; q.beta = beta; q.alpha_dev = a_factor; q.alpha_dev2 = b_factor;
  q.a[0] = a_hi; q.a[1] = a_lo; q.b[0] = b_hi; q.b[1] = b_lo;
  q.c = c; q.ldc = ldc; q.batch = 1; q.c_vec = aligned16(c) && ldc % 4 == 0;
  return mt::launch<16, 3, 2, true>(h, q);
}

pb_status_t pb_gemm_bf16(pb_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                         float alpha, const uint16_t* a, int64_t lda, int64_t sa,
                         const uint16_t* b, int64_t ldb, int64_t sb, float beta, float* c,
                         int64_t ldc, int64_t sc, int64_t batch) {
  return gemm_bf16_entry(h, -1, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                         batch);
}

pb_status_t pb_gemm_bf16_config(pb_handle_t h, int id, char ta, char tb, int64_t m, int64_t n,
                                int64_t k, float alpha, const uint16_t* a, int64_t lda,
                                int64_t sa, const uint16_t* b, int64_t ldb, int64_t sb,
                                float beta, float* c, int64_t ldc, int64_t sc, int64_t batch) {
  if (id < 0 || id >= kNumConfigs)
    return pb::set_error(PB_STATUS_INVALID_VALUE, "config id %d out of range", id);
  return gemm_bf16_entry(h, id, ta, tb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc,
                         batch);
}

int pb_tune_num_configs(void) { return kNumConfigs; }

const char* pb_tune_config_name(int id) {
  if (id < 0 || id >= kNumConfigs) return "";
  return config_name_storage(id).c_str();
}

pb_status_t pb_tune_time_config(pb_handle_t h, int id, char ta, char tb, int64_t m, int64_t n,
                                int64_t k, const uint16_t* a, int64_t lda, const uint16_t* b,
                                int64_t ldb, float* c, int64_t ldc, int64_t batch, int warmup,
                                int reps, float* usec_out) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  PB_REQUIRE(usec_out != nullptr && reps > 0, "bad timing arguments");
  PB_REQUIRE(id >= 0 && id < kNumConfigs, "config id %d out of range", id);
  const char nta = pb::norm_trans(ta), ntb = pb::norm_trans(tb);
  PB_REQUIRE(nta && ntb, "invalid transpose flag");
  const int64_t sa = (nta == 'n' ? k : m) * lda, sb = (ntb == 'n' ? n : k) * ldb, sc = n * ldc;
  for (int i = 0; i < warmup; ++i)
    PB_TRY(gemm_bf16_entry(h, id, ta, tb, m, n, k, 1.f, a, lda, sa, b, ldb, sb, 0.f, c, ldc, sc,
                           batch));
  hipEvent_t e0, e1;
  PB_CHECK_HIP(hipEventCreate(&e0));
  PB_CHECK_HIP(hipEventCreate(&e1));
  PB_CHECK_HIP(hipEventRecord(e0, h->stream));
  pb_status_t st = PB_STATUS_SUCCESS;
  for (int i = 0; i < reps && st == PB_STATUS_SUCCESS; ++i)
    st = gemm_bf16_entry(h, id, ta, tb, m, n, k, 1.f, a, lda, sa, b, ldb, sb, 0.f, c, ldc, sc,
                         batch);
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

// FP16 inputs, FP32 accumulate and output (the analogue of pb_gemm_bf16).
pb_status_t pb_gemm_f16(pb_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                        float alpha, const uint16_t* a, int64_t lda, int64_t stridea, const uint16_t* b,
                        int64_t ldb, int64_t strideb, float beta, float* c, int64_t ldc, int64_t stridec,
                        int64_t batch) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  GemmCall<uint16_t> g;
  bool done = false;
  PB_TRY((gemm_front<float, uint16_t>(h, transa, transb, m, n, k, alpha, a, lda, stridea, b, ldb, strideb,
                                      beta, c, ldc, stridec, batch, PB_BATCH_STRIDED, g, done)));
  if (done) return PB_STATUS_SUCCESS;
  if (alpha == 0.f) return sgemm_alpha_zero(h, m, n, beta, c, ldc, stridec, batch);
  return gemm_f16_direct(h, g, alpha, beta, c, ldc, stridec);
}

// All-FP16 GEMM (A, B, C half; alpha and beta given as float): the FP32
// product alpha op(A) op(B) goes to the workspace, then C <- half(T + beta C).
pb_status_t pb_hgemm(pb_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
                     float alpha, const uint16_t* a, int64_t lda, const uint16_t* b, int64_t ldb, float beta,
                     uint16_t* c, int64_t ldc) {
  pb::clear_error();
  PB_TRY(pb::enter(h));
  GemmCall<uint16_t> g;
  bool done = false;
  // Validation with C typed as the FP32 product (ld M) that is written first.
  PB_TRY((gemm_front<float, uint16_t>(h, transa, transb, m, n, k, alpha, a, lda, 0, b, ldb, 0, beta,
                                      reinterpret_cast<float*>(c), ldc, 0, 1, PB_BATCH_STRIDED, g, done)));
  if (done) return PB_STATUS_SUCCESS;
  PB_REQUIRE(c != nullptr, "null C");
  // Workspace: split-K partials (64 x 64 shape at most 16 splits), then T.
  const size_t part = pb::align_up(splitk_bytes<fp16_t, B64>(m, n, k, 1, 16), 256);
  void* ws = nullptr;
  PB_TRY(pb::workspace(h, part + size_t(m) * size_t(n) * sizeof(float), &ws));
  float* t = reinterpret_cast<float*>(static_cast<char*>(ws) + part);
  const bool no_product = alpha == 0.f || k == 0;
  if (!no_product) PB_TRY(gemm_f16_direct(h, g, alpha, 0.f, t, m, m * n));
  int64_t grid = std::min<int64_t>(pb::ceil_div(m * n, 256), 8 * h->num_cus);
  hipLaunchKernelGGL(f16_store_kernel, dim3(unsigned(std::max<int64_t>(grid, 1))), dim3(256), 0, h->stream, m,
                     n, no_product ? static_cast<const float*>(nullptr) : static_cast<const float*>(t), m,
                     m * n, beta, reinterpret_cast<_Float16*>(c), ldc, ldc * n, int64_t(1));
  PB_CHECK_LAUNCH();
  return PB_STATUS_SUCCESS;
}

}  // extern "C"

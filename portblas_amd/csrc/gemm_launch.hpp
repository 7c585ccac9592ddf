// Generic batched / split-K GEMM kernel over the tile engine, plus the host
// launcher every typed entry point uses.
#pragma once

#include "gemm_core.hpp"

namespace pbgemm {

// Which tiles a launch covers.
enum TileSet : int {
  kAllTiles = 0,       // every tile of the tiles_m x tiles_n grid
  kInteriorTiles = 1,  // tiles entirely inside M x N (full_m x full_n grid)
  kEdgeTiles = 2       // the right column strip, then the bottom row strip
};

template <typename T>
struct GemmProblem {
  using Raw = typename Traits<T>::Raw;
  using Acc = typename Traits<T>::Acc;
  int64_t M, N, K;
  Acc alpha, beta;
  const Acc* alpha_dev;  // optional device factor multiplied into alpha
  // Up to 4 (A, B) pairs accumulated into one result (split-precision terms).
  int nterms;
  const Raw* a[4];
  const Raw* b[4];
  int64_t a_rs, a_ks, a_stride;  // element (m, k) of op(A) at a[m*a_rs + k*a_ks]
  int64_t b_rs, b_ks, b_stride;  // element (n, k) of op(B) at b[n*b_rs + k*b_ks]
  Acc* c;
  int64_t c_rs, ldc, c_stride;   // element (m, n) of C at c[m*c_rs + n*ldc]
  int64_t batch;
  bool c_vec;
  // Filled in by the launcher.
  int64_t tiles_m, tiles_n, full_m, full_n;
  int tile_set;
  int splits;
  int64_t kchunk;
  Acc* partial;
  unsigned int* counters;
};

// Tile (tm, tn) of launch-local tile index `lin` for the launch's tile set.
template <typename T>
__device__ __forceinline__ void tile_of(const GemmProblem<T>& p, int64_t lin, int64_t& tm,
                                        int64_t& tn) {
  if (p.tile_set == kAllTiles) {
    tile_coords(lin, p.tiles_m, p.tiles_n, tm, tn);
  } else if (p.tile_set == kInteriorTiles) {
    tile_coords(lin, p.full_m, p.full_n, tm, tn);
  } else {
    const uint32_t l = uint32_t(lin), tiles_m = uint32_t(p.tiles_m);
    const uint32_t right = tiles_m * uint32_t(p.tiles_n - p.full_n);
    if (l < right) {
      tm = l % tiles_m;
      tn = p.full_n + l / tiles_m;
    } else {
      const uint32_t e = l - right, rows = tiles_m - uint32_t(p.full_m);
      tm = p.full_m + e % rows;
      tn = e / rows;
    }
  }
}

template <typename T, typename Shape, bool KCA, bool KCB, bool EDGE>
__global__ __launch_bounds__(256, (sizeof(T) == 8 && Shape::kBM == 128) ? 2 : 1) void gemm_kernel(GemmProblem<T> p) {
  using TL = Tile<T, Shape, KCA, KCB>;
  using Acc = typename Traits<T>::Acc;
  __shared__ __attribute__((aligned(16))) typename TL::Raw lds[2 * TL::STAGE_ELEMS];
  const int64_t tiles = p.tiles_m * p.tiles_n;
  const int s = int(blockIdx.x % unsigned(p.splits));
  const int64_t lin = blockIdx.x / unsigned(p.splits);
  int64_t tm, tn;
  tile_of(p, lin, tm, tn);
  const int64_t tid = tm * p.tiles_n + tn;  // id in the full tile grid
  const int64_t m0 = tm * TL::BM, n0 = tn * TL::BN;
  const int64_t kb = int64_t(s) * p.kchunk;
  const int64_t ke = (kb + p.kchunk < p.K) ? kb + p.kchunk : p.K;
  for (int64_t b = blockIdx.y; b < p.batch; b += gridDim.y) {
    typename TL::AccV acc[TL::FM][TL::FN];
    TL::zero(acc);
    for (int t = 0; t < p.nterms; ++t) {
      Operand<T> A{p.a[t] + b * p.a_stride, p.a_rs, p.a_ks, p.M};
      Operand<T> B{p.b[t] + b * p.b_stride, p.b_rs, p.b_ks, p.N};
      TL::template mainloop<EDGE>(A, B, m0, n0, kb, ke, lds, acc);
    }
    Acc* c = p.c + b * p.c_stride;
    const Acc alpha = p.alpha_dev ? p.alpha * *p.alpha_dev : p.alpha;
    if (p.splits == 1) {
      TL::epilogue(acc, alpha, p.beta, c, p.c_rs, p.ldc, p.M, p.N, m0, n0, p.c_vec);
      continue;
    }
    constexpr int64_t TE = int64_t(TL::BM) * TL::BN;
    Acc* part = p.partial + (b * tiles + tid) * p.splits * TE;
    TL::store_partial(acc, part + s * TE);
  }
}

// Split-K: sums the K-split partial tiles in a fixed order and applies the
// epilogue.  A separate launch (stream order publishes the partials; no
// device-scope fences, which would write back the L2 once per workgroup).
// Grid: (tiles, batch); element e of a tile is (e % BM, e / BM), so
// consecutive threads walk the contiguous rows of C.
template <typename T, int BM, int BN>
__global__ __launch_bounds__(256) void splitk_fold_kernel(GemmProblem<T> p) {
  using Acc = typename Traits<T>::Acc;
  constexpr int64_t TE = int64_t(BM) * BN;
  const int64_t tiles = p.tiles_m * p.tiles_n;
  const int64_t tid = blockIdx.x;
  const int64_t tm = tid / p.tiles_n, tn = tid % p.tiles_n;
  for (int64_t b = blockIdx.y; b < p.batch; b += gridDim.y) {
    const Acc* part = p.partial + (b * tiles + tid) * p.splits * TE;
    Acc* c = p.c + b * p.c_stride;
    const Acc alpha = p.alpha_dev ? p.alpha * *p.alpha_dev : p.alpha;
    for (int e = threadIdx.x; e < TE; e += 256) {
      const int64_t m = tm * BM + e % BM, n = tn * BN + e / BM;
      if (m >= p.M || n >= p.N) continue;
      Acc sum = part[e];
      for (int s = 1; s < p.splits; ++s) sum += part[s * TE + e];
      Acc* cp = c + m * p.c_rs + n * p.ldc;
      const Acc v = alpha * sum;
      *cp = (p.beta != Acc(0)) ? v + p.beta * *cp : v;
    }
  }
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <typename T, typename Shape, bool EDGE>
void dispatch_layout(pb_handle_t h, const GemmProblem<T>& p, dim3 grid, bool kca, bool kcb) {
#define PB_LAUNCH(KA, KB)                                                                 \
  hipLaunchKernelGGL((gemm_kernel<T, Shape, KA, KB, EDGE>), grid, dim3(256), 0, h->stream, p)
  if (kca && kcb) PB_LAUNCH(true, true);
  else if (kca) PB_LAUNCH(true, false);
  else if (kcb) PB_LAUNCH(false, true);
  else PB_LAUNCH(false, false);
#undef PB_LAUNCH
}

// K splits actually used for `splits` requested: every split owns at least
// one K tile.  Returns the split count and sets the K chunk per split.
template <int BK>
inline int clamp_splits(int64_t K, int splits, int64_t& kchunk) {
  if (splits < 1) splits = 1;
  const int64_t max_splits = pb::ceil_div(K, BK);
  if (splits > max_splits) splits = int(max_splits < 1 ? 1 : max_splits);
  kchunk = pb::ceil_div(pb::ceil_div(K, splits), BK) * BK;
  if (kchunk < BK) kchunk = BK;
  splits = int(pb::ceil_div(K, kchunk));
  return splits < 1 ? 1 : splits;
}

// Bytes of split-K partial tiles launch_gemm<T, Shape> needs (0 without a split).
template <typename T, typename Shape>
size_t splitk_bytes(int64_t M, int64_t N, int64_t K, int64_t batch, int splits) {
  int64_t kchunk;
  splits = clamp_splits<Shape::kBK>(K, splits, kchunk);
  if (splits <= 1) return 0;
  const int64_t tiles = pb::ceil_div(M, Shape::kBM) * pb::ceil_div(N, Shape::kBN);
  return sizeof(typename Traits<T>::Acc) * size_t(tiles * batch * splits) * Shape::kBM *
         Shape::kBN;
}

// Launches one GEMM problem with tile shape `Shape` and `splits` K splits
// (clamped so every split owns at least one K tile).  With vector-legal
// operands the interior tiles run the unguarded kernel and the edge strips a
// guarded one; otherwise every tile runs guarded.  Split-K partials go to
// p.partial when the caller provides it (sized by splitk_bytes; callers that
// keep their own data in the handle workspace must), else to the workspace.
template <typename T, typename Shape>
pb_status_t launch_gemm(pb_handle_t h, GemmProblem<T> p, bool kca, bool kcb, bool vec,
                        int splits) {
  constexpr int BM = Shape::kBM, BN = Shape::kBN, BK = Shape::kBK;
  p.tiles_m = pb::ceil_div(p.M, BM);
  p.tiles_n = pb::ceil_div(p.N, BN);
  p.full_m = p.M / BM;
  p.full_n = p.N / BN;
  const int64_t tiles = p.tiles_m * p.tiles_n;
  splits = clamp_splits<BK>(p.K, splits, p.kchunk);
  p.splits = splits;
  p.counters = pb::counters(h, 0);
  if (splits > 1 && p.partial == nullptr) {
    void* ws = nullptr;
    PB_TRY(pb::workspace(h, splitk_bytes<T, Shape>(p.M, p.N, p.K, p.batch, splits), &ws));
    p.partial = static_cast<typename Traits<T>::Acc*>(ws);
  }
  const unsigned gy = unsigned(p.batch < 65535 ? p.batch : 65535);
  const int64_t interior = vec ? p.full_m * p.full_n : 0;
  const int64_t edge = tiles - interior;
  PB_REQUIRE(tiles * splits <= 0x7fffffffLL, "problem too large for one launch");
  if (interior > 0) {
    p.tile_set = kInteriorTiles;
    dispatch_layout<T, Shape, false>(h, p, dim3{unsigned(interior * splits), gy}, kca, kcb);
    PB_CHECK_LAUNCH();
  }
  if (edge > 0) {
    p.tile_set = vec ? kEdgeTiles : kAllTiles;
    dispatch_layout<T, Shape, true>(h, p, dim3{unsigned(edge * splits), gy}, kca, kcb);
    PB_CHECK_LAUNCH();
  }
  if (splits > 1) {
    const dim3 fold_grid{unsigned(tiles), gy};
    hipLaunchKernelGGL((splitk_fold_kernel<T, BM, BN>), fold_grid, dim3(256), 0, h->stream, p);
    PB_CHECK_LAUNCH();
  }
  return PB_STATUS_SUCCESS;
}

// Fills the operand strides of `p` from BLAS arguments.  Interleaved batches
// (element (r, c) of batch b at ptr[b + batch * (r + ld * c)]) become unit
// batch strides with every in-matrix stride multiplied by the batch count.
// Returns whether 16-byte vector loads are legal for every operand.
template <typename T>
bool set_operands(GemmProblem<T>& p, char ta, char tb, int64_t lda, int64_t sa, int64_t ldb,
                  int64_t sb, bool interleaved) {
  const int64_t mul = interleaved ? p.batch : 1;
  // op(A) is M x K: 'n' -> (m, k) at a[m + k*lda]; 't' -> a[k + m*lda].
  if (ta == 'n') { p.a_rs = 1 * mul; p.a_ks = lda * mul; }
  else { p.a_rs = lda * mul; p.a_ks = 1 * mul; }
  // op(B) is K x N, addressed as (n, k): 'n' -> b[k + n*ldb]; 't' -> b[n + k*ldb].
  if (tb == 'n') { p.b_rs = ldb * mul; p.b_ks = 1 * mul; }
  else { p.b_rs = 1 * mul; p.b_ks = ldb * mul; }
  p.a_stride = interleaved ? 1 : sa;
  p.b_stride = interleaved ? 1 : sb;
  constexpr int VW = Traits<T>::VW;
  auto ok = [&](const void* base, int64_t ld, int64_t stride) {
    return !interleaved && aligned16(base) && ld % VW == 0 && (p.batch == 1 || stride % VW == 0);
  };
  bool vec = true;
  for (int t = 0; t < p.nterms; ++t) vec = vec && ok(p.a[t], lda, sa) && ok(p.b[t], ldb, sb);
  return vec;
}

}  // namespace pbgemm

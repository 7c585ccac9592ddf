// Host-side GEMM front end shared by every GEMM entry point: argument
// validation, default batch strides and the alpha == 0 rule, restated from
// src/interface/gemm_interface.hpp:105-240.
#pragma once

#include "pb_internal.hpp"

namespace pbgemm {

template <typename T>
struct GemmCall {
  char ta = 'n', tb = 'n';
  int64_t M = 0, N = 0, K = 0;
  T alpha{}, beta{};
  const T* a = nullptr;
  int64_t lda = 0, sa = 0;
  const T* b = nullptr;
  int64_t ldb = 0, sb = 0;
  T* c = nullptr;
  int64_t ldc = 0, sc = 0;
  int64_t batch = 1;
  bool interleaved = false;
};

// _gemm_batched's default strides (gemm_interface.hpp:209-219): the footprint
// of one matrix for strided batches, unused (0) for interleaved ones.
inline void default_strides(char ta, char tb, int64_t M, int64_t N, int64_t K, int64_t lda,
                            int64_t ldb, int64_t ldc, pb_batch_type_t bt, int64_t& sa,
                            int64_t& sb, int64_t& sc) {
  sa = sb = sc = 0;
  if (bt != PB_BATCH_STRIDED) return;
  auto lower = [](char t) { return (t >= 'A' && t <= 'Z') ? char(t - 'A' + 'a') : t; };
  sa = (lower(ta) != 'n') ? M * lda : K * lda;
  sb = (lower(tb) != 'n') ? ldb * K : N * ldb;
  sc = ldc * N;
}

// Validates and normalises one GEMM call.  `done` is set when nothing is left
// to launch (an empty problem).  The caller applies the alpha == 0 rule when
// g.alpha == 0 (checked before the transpose flags, as the reference does).
template <typename T, typename U = T>
pb_status_t gemm_front(pb_handle_t, char ta, char tb, int64_t M, int64_t N, int64_t K, T alpha,
                       const U* a, int64_t lda, int64_t sa, const U* b, int64_t ldb, int64_t sb,
                       T beta, T* c, int64_t ldc, int64_t sc, int64_t batch, pb_batch_type_t bt,
                       GemmCall<U>& g, bool& done) {
  done = false;
  PB_REQUIRE(M >= 0 && N >= 0 && K >= 0, "negative dimension (M=%lld N=%lld K=%lld)",
             (long long)M, (long long)N, (long long)K);
  PB_REQUIRE(batch >= 0, "negative batch size");
  PB_REQUIRE(bt == PB_BATCH_STRIDED || bt == PB_BATCH_INTERLEAVED, "invalid batch type");
  g.M = M; g.N = N; g.K = K;
  g.alpha = U(alpha); g.beta = U(beta);
  g.a = a; g.lda = lda; g.sa = sa;
  g.b = b; g.ldb = ldb; g.sb = sb;
  g.ldc = ldc; g.sc = sc;
  g.c = reinterpret_cast<U*>(c);
  g.batch = batch;
  g.interleaved = (bt == PB_BATCH_INTERLEAVED) && batch > 1;
  if (M == 0 || N == 0 || batch == 0) { done = true; return PB_STATUS_SUCCESS; }
  PB_REQUIRE(c != nullptr, "null C");
  PB_REQUIRE(ldc >= (M > 1 ? M : 1), "ldc (%lld) < max(1, M)", (long long)ldc);
  if (alpha == T(0)) return PB_STATUS_SUCCESS;  // gemm_interface.hpp:112 (before trans checks)
  const char nta = pb::norm_trans(ta), ntb = pb::norm_trans(tb);
  if (!nta) return pb::set_error(PB_STATUS_INVALID_VALUE, "invalid _TransA");
  if (!ntb) return pb::set_error(PB_STATUS_INVALID_VALUE, "invalid _TransB");
  g.ta = nta; g.tb = ntb;
  if (batch > 1 && bt == PB_BATCH_STRIDED) {  // gemm_interface.hpp:158-166
    if (sc < ldc * N || sc < 0) return pb::set_error(PB_STATUS_INVALID_VALUE, "invalid _stridec");
    if (sa < 0) return pb::set_error(PB_STATUS_INVALID_VALUE, "invalid _stridea");
    if (sb < 0) return pb::set_error(PB_STATUS_INVALID_VALUE, "invalid _strideb");
  }
  const int64_t arows = nta == 'n' ? M : K;
  const int64_t brows = ntb == 'n' ? K : N;
  PB_REQUIRE(lda >= (arows > 1 ? arows : 1), "lda (%lld) < max(1, %lld)", (long long)lda,
             (long long)arows);
  PB_REQUIRE(ldb >= (brows > 1 ? brows : 1), "ldb (%lld) < max(1, %lld)", (long long)ldb,
             (long long)brows);
  if (K > 0) PB_REQUIRE(a != nullptr && b != nullptr, "null A or B");
  return PB_STATUS_SUCCESS;
}

}  // namespace pbgemm

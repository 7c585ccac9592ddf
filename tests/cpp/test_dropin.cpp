// Drop-in check of the C++ API (include/portblas_amd.hpp): the calls a
// portBLAS user makes (blas::SB_Handle, _gemm, _gemm_strided_batched, _symm,
// _trsm, _gemv, _dot / _asum / _nrm2 / _iamax / _iamin async and sync, _axpy,
// _scal, _copy, _swap, _sdsdot, extension::_reduction, the reference's
// exception types) compiled against this library, checked
// against FP64 host loops.  Built by portblas_amd/build.py; run by
// tests/test_gpu_dropin.py.  Exit status 0 = every check passed.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "portblas_amd.hpp"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  std::fflush(stdout);
  if (!ok) ++failures;
}

template <typename T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(T) + 16) != hipSuccess) throw std::runtime_error("hipMalloc");
  if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("hipMemcpy");
  return d;
}

template <typename T>
std::vector<T> download(const T* d, size_t n) {
  std::vector<T> h(n);
  if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("hipMemcpy");
  return h;
}

std::vector<float> rand_vec(size_t n, unsigned seed) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  std::vector<float> v(n);
  for (auto& x : v) x = u(g);
  return v;
}

}  // namespace

int main() {
  blas::SB_Handle sb(nullptr);
  expect(sb.get_num_compute_units() > 0, "get_num_compute_units");

  // _gemm 'n','t' with alpha, beta, padded leading dimensions.
  {
    const int M = 67, N = 45, K = 33, lda = 70, ldb = 50, ldc = 71;
    auto A = rand_vec(size_t(lda) * K, 1), B = rand_vec(size_t(ldb) * K, 2),
         C = rand_vec(size_t(ldc) * N, 3);
    float *dA = upload(A), *dB = upload(B), *dC = upload(C);
    auto ev = blas::_gemm(sb, 'n', 't', M, N, K, 1.5f, dA, lda, dB, ldb, 0.5f, dC, ldc);
    sb.wait(ev);
    auto got = download(dC, C.size());
    double worst = 0;
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < M; ++i) {
        double s = 0, sa = 0;
        for (int k = 0; k < K; ++k) {
          s += double(A[i + size_t(k) * lda]) * B[j + size_t(k) * ldb];
          sa += std::fabs(double(A[i + size_t(k) * lda]) * B[j + size_t(k) * ldb]);
        }
        const double ref = 1.5 * s + 0.5 * C[i + size_t(j) * ldc];
        const double tol = 1.01 * (K + 2) * 0x1p-24 * (1.5 * sa + 0.5 * std::fabs(C[i + size_t(j) * ldc]));
        worst = std::fmax(worst, std::fabs(got[i + size_t(j) * ldc] - ref) / (tol + 1e-30));
      }
    bool pad_ok = true;
    for (int j = 0; j < N; ++j)
      for (int i = M; i < ldc; ++i) pad_ok &= got[i + size_t(j) * ldc] == C[i + size_t(j) * ldc];
    expect(worst <= 1.0, "_gemm n,t within the FP32 bound");
    expect(pad_ok, "_gemm leaves ldc padding untouched");
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
  }

  // _gemm_strided_batched, double.
  {
    const int M = 16, N = 12, K = 9, batch = 5;
    std::vector<double> A(size_t(M) * K * batch), B(size_t(K) * N * batch), C(size_t(M) * N * batch, 0.0);
    for (size_t i = 0; i < A.size(); ++i) A[i] = std::sin(0.1 * i);
    for (size_t i = 0; i < B.size(); ++i) B[i] = std::cos(0.07 * i);
    double *dA = upload(A), *dB = upload(B), *dC = upload(C);
    blas::_gemm_strided_batched(sb, 'n', 'n', M, N, K, 1.0, dA, M, int64_t(M) * K, dB, K,
                                int64_t(K) * N, 0.0, dC, M, int64_t(M) * N, batch);
    sb.wait();
    auto got = download(dC, C.size());
    double worst = 0;
    for (int b = 0; b < batch; ++b)
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
          double s = 0;
          for (int k = 0; k < K; ++k)
            s += A[b * M * K + i + k * M] * B[b * K * N + k + j * K];
          worst = std::fmax(worst, std::fabs(got[b * M * N + i + j * M] - s));
        }
    expect(worst < 1e-12, "_gemm_strided_batched double");
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
  }

  // _gemv 't' with a negative x increment.
  {
    const int M = 300, N = 200, lda = 301;
    auto A = rand_vec(size_t(lda) * N, 4), x = rand_vec(size_t(M) * 2, 5), y = rand_vec(N, 6);
    float *dA = upload(A), *dx = upload(x), *dy = upload(y);
    blas::_gemv(sb, 't', M, N, 2.0f, dA, lda, dx, -2, -1.0f, dy, 1);
    auto got = download(dy, N);
    double worst = 0;
    for (int j = 0; j < N; ++j) {
      double s = 0, sa = 0;
      for (int i = 0; i < M; ++i) {
        const double xv = x[size_t(M - 1 - i) * 2];
        s += A[i + size_t(j) * lda] * xv;
        sa += std::fabs(A[i + size_t(j) * lda] * xv);
      }
      const double ref = 2.0 * s - y[j];
      const double tol = 1.01 * (M + 2) * 0x1p-24 * (2.0 * sa + std::fabs(y[j]));
      worst = std::fmax(worst, std::fabs(got[j] - ref) / tol);
    }
    expect(worst <= 1.0, "_gemv t, incx = -2");
    (void)hipFree(dA); (void)hipFree(dx); (void)hipFree(dy);
  }

  // BLAS-1: async accumulate into *rs, sync forms, ties and NaN for iamax.
  {
    const int n = 100003;
    auto x = rand_vec(n, 7), y = rand_vec(n, 8);
    x[777] = std::nanf("");
    x[5000] = 1.0f;
    x[90000] = -1.0f;  // same |x| as x[5000]: the lower index wins
    float *dx = upload(x), *dy = upload(y);
    double dot = 0, asum = 0, ss = 0;
    for (int i = 0; i < n; ++i) {
      if (i == 777) continue;
      dot += double(x[i]) * y[i];
      asum += std::fabs(x[i]);
      ss += double(x[i]) * x[i];
    }
    x[777] = 0.f;  // exclude the NaN from the sums above only
    std::vector<float> rs0 = {0.25f};
    float* drs = upload(rs0);
    blas::_dot(sb, n, dy, 1, dy, 1, drs);
    double yy = 0;
    for (int i = 0; i < n; ++i) yy += double(y[i]) * y[i];
    const float acc = download(drs, 1)[0];
    expect(std::fabs(acc - (0.25 + yy)) <= 1e-4 * (0.25 + yy), "_dot async accumulates into rs");
    const float s_dot = blas::_dot(sb, n, dy, 1, dy, 1);
    expect(std::fabs(s_dot - yy) <= 1e-4 * yy, "_dot sync");
    const int32_t imax = blas::_iamax(sb, n, dx, 1);
    expect(imax == 5000, "_iamax sync: NaN never wins, lowest index on ties");
    int32_t* didx = nullptr;
    (void)hipMalloc(&didx, sizeof(int32_t));
    blas::_iamin(sb, n, dy, 1, didx);
    int32_t want = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(y[i]) < std::fabs(y[want])) want = i;
    expect(download(didx, 1)[0] == want, "_iamin async");
    std::vector<float> x2 = {3.f, -4.f};
    float* dx2 = upload(x2);
    expect(std::fabs(blas::_nrm2(sb, 2, dx2, 1) - 5.f) < 1e-6f, "_nrm2 sync closed form 3-4-5");
    expect(std::fabs(blas::_asum(sb, 2, dx2, 1) - 7.f) < 1e-6f, "_asum sync");
    (void)dot; (void)asum; (void)ss;
    (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(drs); (void)hipFree(didx); (void)hipFree(dx2);
  }

  // extension::_reduction on a 3 x 3 column-major matrix: inner = column
  // sums, outer = row sums.
  {
    std::vector<float> in = {1, 2, 3, 4, 5, 6, 7, 8, 9};
    float* din = upload(in);
    float* dout = nullptr;
    (void)hipMalloc(&dout, 3 * sizeof(float));
    blas::extension::_reduction<blas::AddOperator, float>(sb, din, 3, dout, 3, 3,
                                                          blas::reduction_dim_t::inner);
    auto inner = download(dout, 3);
    blas::extension::_reduction<blas::AddOperator, float>(sb, din, 3, dout, 3, 3,
                                                          blas::reduction_dim_t::outer);
    auto outer = download(dout, 3);
    expect(inner[0] == 6 && inner[1] == 15 && inner[2] == 24, "_reduction inner");
    expect(outer[0] == 12 && outer[1] == 15 && outer[2] == 18, "_reduction outer");
    (void)hipFree(din); (void)hipFree(dout);
  }

  // Streaming BLAS-1 and sdsdot (blas1_interface.h:45-70,116-125,193-211).
  {
    std::vector<float> x = {1.f, 2.f, 3.f, 4.f, 5.f}, y = {10.f, 20.f, 30.f, 40.f, 50.f};
    float* dx = upload(x);
    float* dy = upload(y);
    blas::_axpy(sb, 5, 2.f, dx, 1, dy, 1);
    auto r = download(dy, 5);
    expect(r[0] == 12.f && r[4] == 60.f, "_axpy y <- y + alpha x");
    blas::_scal(sb, 3, -1.f, dx, 2);  // elements 0, 2, 4
    r = download(dx, 5);
    expect(r[0] == -1.f && r[1] == 2.f && r[2] == -3.f && r[4] == -5.f, "_scal with incx 2");
    blas::_copy(sb, 5, dy, 1, dx, -1);  // reversed view of x
    r = download(dx, 5);
    expect(r[4] == 12.f && r[0] == 60.f, "_copy into a negative-increment view");
    blas::_swap(sb, 5, dx, 1, dy, 1);
    auto ry = download(dy, 5);
    expect(ry[0] == 60.f && ry[4] == 12.f, "_swap");
    const float sd = blas::_sdsdot(sb, 2, 0.5f, dx, 1, dx, 1);  // x is now {12, 24, ...}
    expect(std::fabs(sd - (0.5f + 144.f + 576.f)) < 1e-3f, "_sdsdot sync");
    (void)hipFree(dx); (void)hipFree(dy);
  }

  // _symm (side 'l', upper triangle stored) and _trsm (left, lower, no
  // transpose, non-unit) against host loops; the unreferenced triangle holds NaN.
  {
    const int m = 40, n = 24;
    std::vector<float> a(m * m, NAN), b = rand_vec(size_t(m) * n, 7), c(size_t(m) * n, 0.f);
    std::vector<float> lower(m * m, NAN);
    auto fa = rand_vec(size_t(m) * m, 8);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        if (i <= j) a[i + j * m] = fa[i + j * m];
        if (i >= j) lower[i + j * m] = (i == j) ? 4.f : 0.1f * fa[i + j * m];
      }
    float *da = upload(a), *db = upload(b), *dc = upload(c), *dl = upload(lower);
    blas::_symm(sb, 'l', 'u', m, n, 1.f, da, m, db, m, 0.f, dc, m);
    auto got = download(dc, c.size());
    double worst = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) {
        double s = 0;
        for (int k = 0; k < m; ++k) s += double(i <= k ? a[i + k * m] : a[k + i * m]) * b[k + j * m];
        worst = std::fmax(worst, std::fabs(got[i + j * m] - s));
      }
    expect(worst < 1e-4, "_symm side l, upper");
    blas::_trsm(sb, 'l', 'l', 'n', 'n', m, n, 2.f, dl, m, db, m);  // B <- X, L X = 2 B
    auto x = download(db, b.size());
    worst = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) {
        double s = 0;
        for (int k = 0; k <= i; ++k) s += double(lower[i + k * m]) * x[k + j * m];
        worst = std::fmax(worst, std::fabs(s - 2.0 * b[i + j * m]));
      }
    expect(worst < 1e-4, "_trsm left, lower: residual");
    bool threw = false;
    try {
      blas::_trsm(sb, 'x', 'l', 'n', 'n', m, n, 1.f, dl, m, db, m);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    expect(threw, "_trsm bad side throws std::invalid_argument");
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dc); (void)hipFree(dl);
  }

  // Exceptions: the reference throws std::invalid_argument on a bad trans.
  {
    bool threw = false;
    try {
      blas::_gemm<float>(sb, 'x', 'n', 4, 4, 4, 1.f, nullptr, 4, nullptr, 4, 0.f, nullptr, 4);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    expect(threw, "_gemm bad trans throws std::invalid_argument");
  }

  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}

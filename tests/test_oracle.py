"""Pins the CPU oracle (oracle/pb_oracle.c) before it is trusted as the
checker of the GPU kernels: against OpenBLAS through scipy (the CBLAS the
reference's tests use), the committed golden fixtures, the reference's
known-answer patterns, closed forms and brute force.  CPU only."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.linalg import blas as sblas

import oracle
from conftest import U32, U64, cfg1_inputs, check_cfg1, colmajor

RNG = np.random.default_rng(1234)


def test_generator_pinned(golden):
    head = oracle.uniform(0, 16)
    np.testing.assert_array_equal(head, golden["uniform_seed0_head"])
    x = oracle.uniform(7, 100_000, lo=-1.0, hi=1.0)
    assert x.min() >= -1.0 and x.max() < 1.0
    assert abs(float(x.mean())) < 0.01  # uniform on [-1, 1)
    # Distinct streams per seed, reproducible per seed.
    assert not np.array_equal(oracle.uniform(1, 64), oracle.uniform(2, 64))
    np.testing.assert_array_equal(oracle.uniform(3, 64), oracle.uniform(3, 64))


def gemm_bound(K, alpha, absprod, beta, cabs, u):
    """Deterministic forward-error bound of any FP summation order:
    |c - c_exact| <= 1.01 K u |alpha| (|A||B|) + 2u |beta C| + tiny."""
    return 1.01 * (K + 2) * u * abs(alpha) * absprod + 2 * u * abs(beta) * cabs + 1e-30


def _gemm_inputs(ci, golden):
    p = f"gemm{ci}_"
    ta, tb, m, n, k, lda, ldb, ldc = (int(v) for v in golden[p + "args"])
    ta, tb = chr(ta), chr(tb)
    alpha, beta = (float(v) for v in golden[p + "scal"])
    sa, sb, sc = (int(v) for v in golden[p + "seeds"])
    ac = k if ta == "n" else m
    bc = n if tb == "n" else k
    A = oracle.uniform(sa, lda * ac)
    B = oracle.uniform(sb, ldb * bc)
    C = oracle.uniform(sc, ldc * n)
    return ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, A, B, C


@pytest.mark.parametrize("ci", range(5))
def test_gemm_oracle_vs_cblas_golden(golden, ci):
    ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, A, B, C = _gemm_inputs(ci, golden)
    ref = oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    cblas = golden[f"gemm{ci}_out"].astype(np.float64)
    absprod = oracle.gemm_absprod_f64(ta, tb, m, n, k, A, lda, B, ldb)
    tol = gemm_bound(k, alpha, colmajor(absprod, m, n, m), beta, np.abs(colmajor(C, m, n, ldc)), U32)
    err = np.abs(colmajor(ref, m, n, ldc) - colmajor(cblas, m, n, ldc))
    assert np.all(err <= tol)
    # Rows beyond m (ldc padding) are untouched.
    if ldc > m:
        np.testing.assert_array_equal(ref.reshape(n, ldc)[:, m:], C.reshape(n, ldc)[:, m:])


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
def test_gemm_oracle_vs_dgemm(ta, tb):
    m, n, k = 23, 19, 37
    lda = (m if ta == "n" else k) + 3
    ldb = (k if tb == "n" else n) + 1
    ldc = m + 2
    A = RNG.standard_normal(lda * (k if ta == "n" else m))
    B = RNG.standard_normal(ldb * (n if tb == "n" else k))
    C = RNG.standard_normal(ldc * n)
    alpha, beta = 1.25, -0.5
    ref = oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    Am = colmajor(A, *((m, k) if ta == "n" else (k, m)), lda)
    Bm = colmajor(B, *((k, n) if tb == "n" else (n, k)), ldb)
    Cm = np.asfortranarray(colmajor(C, m, n, ldc))
    want = sblas.dgemm(alpha, Am, Bm, beta=beta, c=Cm, trans_a=ta == "t", trans_b=tb == "t")
    np.testing.assert_allclose(colmajor(ref, m, n, ldc), want, rtol=1e-12, atol=1e-12)


def test_gemm_check_catches_planted_mistakes(golden):
    """The tolerance used everywhere rejects the plausible bugs: a transposed
    operand, a dropped beta term, a sign error in alpha, a k-off-by-one."""
    ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, A, B, C = _gemm_inputs(0, golden)
    ref = colmajor(oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc), m, n, ldc)
    absprod = colmajor(oracle.gemm_absprod_f64(ta, tb, m, n, k, A, lda, B, ldb), m, n, m)
    tol = gemm_bound(k, alpha, absprod, beta, np.abs(colmajor(C, m, n, ldc)), U32)

    def rejected(bad):
        return np.any(np.abs(colmajor(bad, m, n, ldc) - ref) > tol)

    # beta dropped
    assert rejected(oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, 0.0, C, ldc))
    # alpha sign
    assert rejected(oracle.gemm_f64(ta, tb, m, n, k, -alpha, A, lda, B, ldb, beta, C, ldc))
    # one k term short
    assert rejected(oracle.gemm_f64(ta, tb, m, n, k - 1, alpha, A, lda, B, ldb, beta, C, ldc))
    # B read transposed (square sub-block so shapes stay legal)
    s = min(m, n, k)
    Z = np.zeros(s * s)
    refs = colmajor(oracle.gemm_f64("n", "n", s, s, s, 1.0, A, s, B, s, 0.0, Z, s), s, s, s)
    bad = colmajor(oracle.gemm_f64("n", "t", s, s, s, 1.0, A, s, B, s, 0.0, Z, s), s, s, s)
    tol_s = gemm_bound(s, 1.0, colmajor(oracle.gemm_absprod_f64("n", "n", s, s, s, A, s, B, s), s, s, s),
                       0.0, 0.0, U32)
    assert np.any(np.abs(bad - refs) > tol_s)


def test_portblas_cpu_gemm_restatement(golden):
    """The FP32 restatement of portBLAS's CPU schedule stays within the FP32
    error bound of the FP64 reference (it is the CPU baseline)."""
    for ci in range(5):
        ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, A, B, C = _gemm_inputs(ci, golden)
        got = oracle.pb_sgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        ref = oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        absprod = colmajor(oracle.gemm_absprod_f64(ta, tb, m, n, k, A, lda, B, ldb), m, n, m)
        # The (beta/alpha) C pre-scale adds one rounding of |beta C| per element.
        tol = gemm_bound(k, alpha, absprod, beta, np.abs(colmajor(C, m, n, ldc)), U32) * 2
        assert np.all(np.abs(colmajor(got, m, n, ldc) - colmajor(ref, m, n, ldc)) <= tol)


def test_cfg1_oracle_and_cpu_path_vs_cblas(golden):
    """cfg1 (1024^3, alpha 1.5, beta 0.5, seed 0) checked against CBLAS
    sgemm: the FP64 oracle and the restated portBLAS CPU schedule both agree
    with the committed CBLAS entries and checksums; a transposed operand in
    the oracle call is caught."""
    ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, A, B, C = cfg1_inputs(golden)
    ref = oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    check_cfg1(golden, ref, 1.0)
    cpu = oracle.pb_sgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    check_cfg1(golden, cpu, 3.0)
    wrong = oracle.gemm_f64(ta, "t", m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    with pytest.raises(AssertionError):
        check_cfg1(golden, wrong, 1.0)


def test_gemm_closed_forms():
    # alpha = 0 -> beta C; beta = 0 never reads C (NaN C must not leak).
    m = n = k = 8
    A = RNG.standard_normal(m * k)
    B = RNG.standard_normal(k * n)
    C = np.full(m * n, np.nan)
    out = oracle.gemm_f64("n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    assert np.all(np.isfinite(out))
    I = np.eye(8).reshape(-1, order="F")
    out = oracle.gemm_f64("n", "n", 8, 8, 8, 2.0, A, 8, I, 8, 0.0, np.zeros(64), 8)
    np.testing.assert_allclose(out, 2.0 * A, rtol=0, atol=0)
    Apos, Bpos = np.abs(A), np.abs(B)
    np.testing.assert_allclose(
        oracle.gemm_absprod_f64("n", "n", m, n, k, A, m, B, k),
        oracle.gemm_f64("n", "n", m, n, k, 1.0, Apos, m, Bpos, k, 0.0, np.zeros(m * n), m), rtol=1e-14)


@pytest.mark.parametrize("ci", range(4))
def test_gemv_oracle_vs_cblas_golden(golden, ci):
    p = f"gemv{ci}_"
    tr, m, n, lda, incx, incy = (int(v) for v in golden[p + "args"])
    tr = chr(tr)
    alpha, beta = (float(v) for v in golden[p + "scal"])
    sa, sx, sy = (int(v) for v in golden[p + "seeds"])
    xl, yl = (n, m) if tr == "n" else (m, n)
    A = oracle.uniform(sa, lda * n)
    x = oracle.uniform(sx, 1 + (xl - 1) * incx)
    y = oracle.uniform(sy, 1 + (yl - 1) * incy)
    ref = oracle.gemv_f64(tr, m, n, alpha, A, lda, x, incx, beta, y, incy)
    cb = golden[p + "out"].astype(np.float64)
    Am = np.abs(colmajor(A, m, n, lda).astype(np.float64))
    xs = np.abs(x[::incx].astype(np.float64))
    scale = (Am @ xs if tr == "n" else Am.T @ xs)
    tol = 1.01 * (xl + 2) * U32 * abs(alpha) * scale + 2 * U32 * abs(beta) * np.abs(y[::incy])
    assert np.all(np.abs(ref[::incy] - cb[::incy]) <= tol + 1e-30)
    # pb_sgemv restatement: same bound (single rounding chain per output).
    got = oracle.pb_sgemv(tr, m, n, alpha, A, lda, x, incx, beta, y, incy)
    assert np.all(np.abs(got[::incy].astype(np.float64) - ref[::incy]) <= 2 * tol + 1e-30)


def test_gemv_negative_increments():
    m, n = 5, 4
    A = RNG.standard_normal(m * n)
    x = RNG.standard_normal(n)
    y = RNG.standard_normal(m)
    ref = oracle.gemv_f64("n", m, n, 1.0, A, m, x, -1, 0.0, y, 1)
    np.testing.assert_allclose(ref, colmajor(A, m, n, m) @ x[::-1], rtol=1e-13)
    ref = oracle.gemv_f64("n", m, n, 1.0, A, m, x, 1, 0.0, y, -1)
    np.testing.assert_allclose(ref[::-1], colmajor(A, m, n, m) @ x, rtol=1e-13)


@pytest.mark.parametrize("ci", range(4))
def test_blas1_oracle_vs_cblas_golden(golden, ci):
    p = f"blas1_{ci}_"
    n, incx, incy = (int(v) for v in golden[p + "args"])
    sx, sy = (int(v) for v in golden[p + "seeds"])
    x = oracle.uniform(sx, 1 + (n - 1) * incx)
    y = oracle.uniform(sy, 1 + (n - 1) * incy)
    xs, ys = x[::incx].astype(np.float64), y[::incy].astype(np.float64)
    dot = oracle.dot_f64(n, x, incx, y, incy)
    assert abs(dot - float(golden[p + "dot"])) <= 1.01 * (n + 2) * U32 * np.sum(np.abs(xs * ys))
    asum = oracle.asum_f64(n, x, incx)
    assert abs(asum - float(golden[p + "asum"])) <= 1.01 * (n + 2) * U32 * asum
    nrm = oracle.nrm2_f64(n, x, incx)
    assert abs(nrm - float(golden[p + "nrm2"])) <= 1.01 * (n + 2) * U32 * nrm
    assert oracle.iamax(n, x, incx, True) == int(golden[p + "iamax"])
    assert oracle.iamax(n, x, incx, False) == int(golden[p + "iamin"])
    # The FP32 restatement of portBLAS's CPU schedule.
    for op, want in (("dot", dot), ("asum", asum), ("nrm2", nrm)):
        got = oracle.pb_sreduce(op, n, x, incx, y, incy)
        assert abs(got - want) <= 1.01 * (n + 2) * U32 * max(abs(want), np.sum(np.abs(xs * ys)), 1e-30)


def test_blas1_accumulate_semantics():
    x = np.array([3.0, 4.0])
    assert oracle.nrm2_f64(2, x, 1, rs0=0.0) == 5.0
    assert oracle.nrm2_f64(2, x, 1, rs0=11.0) == 6.0  # sqrt(rs + sum x^2)
    assert oracle.nrm2_f64(0, x, 1, rs0=16.0) == 4.0  # n == 0 still takes the root
    assert oracle.dot_f64(0, x, 1, x, 1, rs0=7.0) == 7.0  # n == 0 leaves rs
    assert oracle.dot_f64(2, x, 1, x, 1, rs0=1.0) == 26.0
    assert oracle.asum_f64(2, -x, 1, rs0=1.0) == 8.0
    assert oracle.asum_f64(2, x, -1) == 7.0
    # dot with one reversed vector pairs x_i with y_{n-1-i}
    y = np.array([1.0, 10.0])
    assert oracle.dot_f64(2, x, 1, y, -1) == 3.0 * 10.0 + 4.0 * 1.0


@pytest.mark.parametrize("name", ["limit", "incrementing", "decrementing"])
@pytest.mark.parametrize("n", [11, 65, 1000])
def test_iaminmax_reference_patterns(golden, name, n):
    x = golden[f"pat_{name}_{n}_x"]
    assert oracle.iamax(n, x, 1, True) == int(golden[f"pat_{name}_{n}_iamax"])
    xm = golden[f"pat_{name}_{n}_xmin"]
    assert oracle.iamax(n, xm, 1, False) == int(golden[f"pat_{name}_{n}_iamin"])


def _brute(x, is_max):
    """The same rule written directly: among non-NaN |x_i|, the best value
    (min for iamin, with values > FLT_MAX unable to beat the FLT_MAX start),
    lowest index on ties; sentinel when nothing qualifies."""
    ax = np.abs(np.asarray(x, dtype=np.float64))
    best, bi = (0.0 if is_max else float(np.finfo(np.float32).max)), oracle.INT32_MAX
    for i, v in enumerate(ax):
        if np.isnan(v):
            continue
        if (is_max and (v > best or (v == best and i < bi))) or \
           (not is_max and (v < best or (v == best and i < bi))):
            best, bi = v, i
    return bi


def test_iaminmax_brute_force_tiny():
    vals = np.array([0.0, -0.0, 1.0, -1.0, 2.0, np.nan, np.inf, -np.inf, 0.5], dtype=np.float32)
    rng = np.random.default_rng(5)
    for _ in range(400):
        n = int(rng.integers(0, 7))
        x = rng.choice(vals, size=n).astype(np.float32)
        for is_max in (True, False):
            assert oracle.iamax(n, x, 1, is_max) == _brute(x, is_max), (x, is_max)


def test_iaminmax_special_cases():
    S = oracle.INT32_MAX
    x = np.array([1.0, -3.0, 3.0, 2.0], dtype=np.float32)
    assert oracle.iamax(4, x, 1) == 1  # tie |-3| == |3|: lowest index
    assert oracle.iamax(4, x, -1) == 0  # negative increment -> 0
    assert oracle.iamax(-1, x, 1) == 0  # negative n -> 0
    assert oracle.iamax(0, x, 1) == S  # empty -> the initial tuple's index
    nan = np.array([np.nan, np.nan], dtype=np.float32)
    assert oracle.iamax(2, nan, 1) == S and oracle.iamax(2, nan, 1, False) == S
    withnan = np.array([0.5, np.nan, 2.0], dtype=np.float32)
    assert oracle.iamax(3, withnan, 1) == 2  # NaN never wins
    assert oracle.iamax(3, withnan, 1, False) == 0
    assert oracle.iamax(2, np.array([np.inf, np.inf], np.float32), 1, False) == S  # inf never beats FLT_MAX
    assert oracle.iamax(3, np.array([0.0, 0.0, 0.0], np.float32), 1) == 0
    assert oracle.iamax(3, x, 2) == 1  # elements 1.0, 3.0 -> logical index 1
    # 16 exact ties at random positions: the first one wins.
    rng = np.random.default_rng(11)
    y = oracle.uniform(42, 4096) * 0.5
    pos = np.sort(rng.choice(4096, size=16, replace=False))
    y[pos] = np.where(rng.random(16) < 0.5, -1.0, 1.0)
    assert oracle.iamax(4096, y, 1) == pos[0]
    # CBLAS agrees when there are no NaNs.
    assert int(sblas.isamax(y)) == pos[0]


def test_reduction_doc_example(golden):
    A = golden["red_doc_A"]
    np.testing.assert_array_equal(oracle.reduction_f64("add", "inner", A, 3, 3, 3), golden["red_doc_inner"])
    np.testing.assert_array_equal(oracle.reduction_f64("add", "outer", A, 3, 3, 3), golden["red_doc_outer"])


@pytest.mark.parametrize("ci", range(3))
@pytest.mark.parametrize("op", ["add", "product", "max", "min", "abs_add", "mean"])
@pytest.mark.parametrize("dim", ["inner", "outer"])
def test_reduction_oracle_vs_golden(golden, ci, op, dim):
    rows, cols, ld = (int(v) for v in golden[f"red{ci}_args"])
    A = oracle.uniform(int(golden[f"red{ci}_seed"]), ld * cols, lo=-2.0, hi=2.0)
    got = oracle.reduction_f64(op, dim, A, ld, rows, cols)
    np.testing.assert_allclose(got, golden[f"red{ci}_{op}_{dim}"], rtol=1e-12, atol=1e-300)

"""FP32 / FP64 GEMMs large enough for the 128 x 128 MFMA tiles, and the FP32
SIMT kernel with interior tiles, on the GPU.

The plain FP32/FP64 path takes its 128 x 128 tiles (F128, D128) when there
are at least two per compute unit and the 128-rounded area is within 15% of
the problem's; ``_takes_128_tiles`` restates that rule and every test below
asserts it first, so a shape that stops reaching those kernels fails loudly
instead of silently testing the 64 x 64 ones.

Each case is checked three ways, none of which needs a full FP64 product:
* sampled entries (64 rows x 64 columns, tile edges included) against the
  FP64 oracle on the sub-problem, under the same accumulation bound as the
  small tests;
* row and column checksums of the whole of C against the FP64 products
  op(A) (op(B) 1) and (1^T op(A)) op(B), under the summed bound;
* a full-matrix bit-exact check on small-integer inputs, whose every partial
  sum is exact in FP32 (|sum| < 2^24), so any correct summation order gives
  the integer result exactly.

Semantics follow test/unittest/blas3/blas3_gemm_test.cpp:29-141 (alpha,
beta, leading-dimension multipliers, transposes) at sizes its own list does
not reach (its largest is 2200, 324 tiles of 128^2).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import U32, U64, colmajor, to_dev

pytestmark = pytest.mark.gpu


def _takes_128_tiles(m, n, batch, cus):
    t128 = -(-m // 128) * -(-n // 128) * batch
    pad = (-(-m // 128) * 128) * (-(-n // 128) * 128) / (m * n)
    return t128 >= 2 * cus and pad < 1.15


def _takes_interior_tiles(m, n):
    return m >= 128 and n >= 128


def _sample(dim, rng, count=64):
    edges = [0, 1, 63, 64, 127, 128, 255, 256, dim - 129, dim - 128, dim - 2, dim - 1]
    pick = {e for e in edges if 0 <= e < dim}
    pick.update(rng.choice(dim, size=min(dim, count - len(pick)), replace=False).tolist())
    return np.array(sorted(pick))[:count]


def _ops(ta, tb, m, n, k, A, lda, B, ldb):
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    opA = colmajor(A, ar, ac, lda)
    opB = colmajor(B, br, bc, ldb)
    return (opA if ta == "n" else opA.T), (opB if tb == "n" else opB.T)


def _check(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, got, u, seed):
    """Sampled entries against the oracle plus row/column checksums."""
    rng = np.random.default_rng(seed)
    opA, opB = _ops(ta, tb, m, n, k, A, lda, B, ldb)
    Cin = colmajor(C, m, n, ldc).astype(np.float64)
    G = colmajor(got, m, n, ldc).astype(np.float64)
    assert np.all(np.isfinite(G)), "non-finite output"
    # sampled entries vs the oracle on the sub-problem
    rows, cols = _sample(m, rng), _sample(n, rng)
    sa = np.asfortranarray(opA[rows, :], dtype=np.float64)
    sb = np.asfortranarray(opB[:, cols], dtype=np.float64)
    sc = np.asfortranarray(Cin[np.ix_(rows, cols)])
    r, c = rows.size, cols.size
    ref = colmajor(oracle.gemm_f64("n", "n", r, c, k, alpha, sa.ravel("F"), r, sb.ravel("F"), k, beta,
                                   sc.ravel("F"), r), r, c, r)
    ap = colmajor(oracle.gemm_absprod_f64("n", "n", r, c, k, sa.ravel("F"), r, sb.ravel("F"), k), r, c, r)
    tol = 1.01 * (k + 2) * u * abs(alpha) * ap + 2 * u * abs(beta) * np.abs(sc) + 1e-300
    err = np.abs(G[np.ix_(rows, cols)] - ref)
    assert np.all(err <= tol), f"sampled entries: max err/tol {np.max(err / tol):.3g}"
    # checksums over every entry: column sums and row sums of C
    A64, B64 = opA.astype(np.float64), opB.astype(np.float64)
    aA, aB = np.abs(A64), np.abs(B64)
    # bound: the sum of the entries' bounds, plus the FP64 rounding of the
    # checksums themselves (at most (m + n + k) U64 times the summed magnitudes)
    col_ref = alpha * (A64.sum(axis=0) @ B64) + beta * Cin.sum(axis=0)
    col_ap = aA.sum(axis=0) @ aB
    col_c = np.abs(Cin).sum(axis=0)
    col_tol = (1.01 * (k + 2) * u * abs(alpha) * col_ap + 2 * u * abs(beta) * col_c
               + 4 * (m + n + k) * U64 * (abs(alpha) * col_ap + abs(beta) * col_c + np.abs(G).sum(axis=0)))
    row_ref = alpha * (A64 @ B64.sum(axis=1)) + beta * Cin.sum(axis=1)
    row_ap = aA @ aB.sum(axis=1)
    row_c = np.abs(Cin).sum(axis=1)
    row_tol = (1.01 * (k + 2) * u * abs(alpha) * row_ap + 2 * u * abs(beta) * row_c
               + 4 * (m + n + k) * U64 * (abs(alpha) * row_ap + abs(beta) * row_c + np.abs(G).sum(axis=1)))
    ce = np.abs(G.sum(axis=0) - col_ref)
    re = np.abs(G.sum(axis=1) - row_ref)
    assert np.all(ce <= col_tol), f"column checksums: max err/tol {np.max(ce / col_tol):.3g}"
    assert np.all(re <= row_tol), f"row checksums: max err/tol {np.max(re / row_tol):.3g}"


def _run(handle, torch, ta, tb, m, n, k, alpha, beta, dtype, seed, lda_pad=0, ldb_pad=0, ldc_pad=0,
         simt=False):
    import portblas_amd as pb

    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    lda, ldb, ldc = ar + lda_pad, br + ldb_pad, m + ldc_pad
    A = oracle.uniform(seed * 3 + 1, lda * ac).astype(dtype)
    B = oracle.uniform(seed * 3 + 2, ldb * bc).astype(dtype)
    C = oracle.uniform(seed * 3 + 3, ldc * n).astype(dtype)
    dC = to_dev(torch, C)
    if simt:
        pb._gemm_ex(handle, pb.gemm_precision_t.fp32_simt, ta, tb, m, n, k, alpha, to_dev(torch, A), lda,
                    to_dev(torch, B), ldb, beta, dC, ldc)
    else:
        pb._gemm(handle, ta, tb, m, n, k, alpha, to_dev(torch, A), lda, to_dev(torch, B), ldb, beta, dC, ldc)
    got = dC.cpu().numpy()
    u = U32 if dtype == np.float32 else U64
    _check(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, got, u, seed)
    if ldc_pad:  # padding rows untouched
        np.testing.assert_array_equal(got.reshape(n, ldc)[:, m:], C.reshape(n, ldc)[:, m:])


def _run_exact(handle, torch, ta, tb, m, n, k, dtype, seed, simt=False):
    """Small-integer inputs: C = op(A) op(B) + 2 C exactly, compared bit for bit."""
    import portblas_amd as pb

    rng = np.random.default_rng(seed)
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    A = rng.integers(-2, 3, size=ar * ac).astype(dtype)
    B = rng.integers(-2, 3, size=br * bc).astype(dtype)
    C = rng.integers(-64, 65, size=m * n).astype(dtype)
    assert 4 * k + 2 * 64 < 2 ** 24
    dC = to_dev(torch, C)
    if simt:
        pb._gemm_ex(handle, pb.gemm_precision_t.fp32_simt, ta, tb, m, n, k, 1.0, to_dev(torch, A), ar,
                    to_dev(torch, B), br, 2.0, dC, m)
    else:
        pb._gemm(handle, ta, tb, m, n, k, 1.0, to_dev(torch, A), ar, to_dev(torch, B), br, 2.0, dC, m)
    got = colmajor(dC.cpu().numpy(), m, n, m)
    opA, opB = _ops(ta, tb, m, n, k, A, ar, B, br)
    # every product and partial sum is an integer below 2^24: exact in FP32
    # and FP64, so the FP64 matrix product is the exact result
    want = opA.astype(np.float64) @ opB.astype(np.float64) + 2.0 * colmajor(C, m, n, m)
    mism = np.argwhere(got.astype(np.float64) != want)
    assert mism.size == 0, f"{len(mism)} entries differ, first at {mism[0].tolist()}"


@pytest.fixture(scope="module")
def cus(handle):
    return handle.get_num_compute_units()


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
def test_sgemm_128_tiles(handle, torch_cuda, cus, ta, tb):
    """SGEMM 3072^3 on the 128 x 128 FP32 MFMA tiles, alpha 1.5, beta 0.5."""
    m = n = k = 3072
    assert _takes_128_tiles(m, n, 1, cus)
    _run(handle, torch_cuda, ta, tb, m, n, k, 1.5, 0.5, np.float32, seed=41)


def test_sgemm_128_tiles_ragged_padded(handle, torch_cuda, cus):
    """Ragged edges (3000 x 2900, padding 4% of the tiled area) with padded
    leading dimensions, beta != 0, and a K that is not a tile multiple."""
    m, n, k = 3000, 2900, 1000
    assert _takes_128_tiles(m, n, 1, cus)
    for ta, tb in (("n", "n"), ("t", "t"), ("n", "t"), ("t", "n")):
        _run(handle, torch_cuda, ta, tb, m, n, k, -2.0, 0.25, np.float32, seed=43, lda_pad=3, ldb_pad=5,
             ldc_pad=7)


@pytest.mark.parametrize("ta,tb", [("n", "n"), ("t", "t")])
def test_sgemm_128_tiles_exact_integers(handle, torch_cuda, cus, ta, tb):
    m, n, k = 3072, 2944, 2048
    assert _takes_128_tiles(m, n, 1, cus)
    _run_exact(handle, torch_cuda, ta, tb, m, n, k, np.float32, seed=47)


def test_sgemm_8192_plain_fp32(handle, torch_cuda, cus):
    """The plain FP32 MFMA GEMM of the 8192^3 precision study (the
    denominator of the split-precision speedup), 'n','n', alpha 1, beta 0."""
    m = n = k = 8192
    assert _takes_128_tiles(m, n, 1, cus)
    _run(handle, torch_cuda, "n", "n", m, n, k, 1.0, 0.0, np.float32, seed=53)


def test_dgemm_4096_128_tiles(handle, torch_cuda, cus):
    """DGEMM 4096^3 'n','n' (the FP64 configuration of the precision study)."""
    m = n = k = 4096
    assert _takes_128_tiles(m, n, 1, cus)
    _run(handle, torch_cuda, "n", "n", m, n, k, 1.0, 0.0, np.float64, seed=59)


@pytest.mark.parametrize("ta,tb", [("t", "n"), ("n", "t"), ("t", "t")])
def test_dgemm_128_tiles_transposes(handle, torch_cuda, cus, ta, tb):
    m = n = 3072
    k = 1536
    assert _takes_128_tiles(m, n, 1, cus)
    _run(handle, torch_cuda, ta, tb, m, n, k, 1.25, -0.5, np.float64, seed=61, lda_pad=1, ldc_pad=2)


def test_dgemm_128_tiles_exact_integers(handle, torch_cuda, cus):
    m, n, k = 3072, 3000, 1024
    assert _takes_128_tiles(m, n, 1, cus)
    _run_exact(handle, torch_cuda, "n", "t", m, n, k, np.float64, seed=67)


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
def test_simt_interior_tiles(handle, torch_cuda, ta, tb):
    """FP32 SIMT kernel at 1024 x 1152 x 512: 72 unguarded interior tiles
    plus ragged edge tiles in a second case."""
    assert _takes_interior_tiles(1024, 1152)
    _run(handle, torch_cuda, ta, tb, 1024, 1152, 512, 1.5, 0.5, np.float32, seed=71, simt=True)
    _run(handle, torch_cuda, ta, tb, 1100, 1030, 300, 1.5, 0.5, np.float32, seed=73, simt=True, ldc_pad=3)


def test_simt_exact_integers(handle, torch_cuda):
    _run_exact(handle, torch_cuda, "n", "n", 1024, 1024, 1024, np.float32, seed=79, simt=True)

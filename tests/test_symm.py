"""SYMM (blas::_symm, src/interface/symm_interface.hpp:36-80): the CPU oracle
(oracle/symm.py) pinned against OpenBLAS's dsymm, then the GPU path (the
stored triangle mirrored in the workspace + the MFMA GEMM engine) against
the oracle with the FP32 / FP64 accumulation bound; the unreferenced
triangle holds NaN so that reading it would show."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import symm as osy

U32 = 2.0 ** -24
U64 = 2.0 ** -53


def _dense(flat, rows, cols, ld):
    return np.asarray(flat, np.float64)[: ld * cols].reshape(cols, ld).T[:rows, :]


@pytest.mark.parametrize("side", ["l", "r", "L", "R"])
@pytest.mark.parametrize("uplo", ["u", "l", "U"])
def test_oracle_matches_openblas(side, uplo):
    from scipy.linalg import blas

    rng = np.random.default_rng(3)
    M, N = 7, 5
    k = M if side in "lL" else N
    lda, ldb, ldc = k + 2, M + 1, M + 3
    a = rng.standard_normal(lda * k)
    b = rng.standard_normal(ldb * N)
    c = rng.standard_normal(ldc * N)
    got = _dense(osy.symm_f64(side, uplo, M, N, 1.5, a, lda, b, ldb, -0.5, c, ldc), M, N, ldc)
    want = blas.dsymm(1.5, _dense(a, k, k, lda), _dense(b, M, N, ldb), beta=-0.5, c=_dense(c, M, N, ldc),
                      side=int(side in "rR"), lower=int(uplo in "lL"))
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
    # planted mistake: mirroring the other triangle disagrees with OpenBLAS
    bad = _dense(osy.symm_f64(side, "l" if uplo in "uU" else "u", M, N, 1.5, a, lda, b, ldb, -0.5, c, ldc),
                 M, N, ldc)
    assert not np.allclose(bad, want)


def _run(handle, torch, dtype, side, uplo, M, N, alpha, beta, pad=0, seed=0):
    import portblas_amd as pb

    rng = np.random.default_rng(seed)
    k = M if side.lower() == "l" else N
    lda, ldb, ldc = max(1, k) + pad, max(1, M) + pad, max(1, M) + pad
    a = (rng.random(lda * k) * 2 - 1).astype(dtype)
    A = a[: lda * k].reshape(k, lda).T
    i, j = np.indices((k, k))
    unref = (i > j) if uplo.lower() == "u" else (i < j)
    A[:k, :k][unref] = np.nan  # never read
    b = (rng.random(ldb * N) * 2 - 1).astype(dtype)
    c = (rng.random(ldc * N) * 2 - 1).astype(dtype)
    ta, tb, tc = (torch.from_numpy(v.copy()).cuda() for v in (a, b, c))
    pb._symm(handle, side, uplo, M, N, alpha, ta, lda, tb, ldb, beta, tc, ldc)
    handle.wait()
    got = _dense(tc.cpu().numpy(), M, N, ldc)
    want = _dense(osy.symm_f64(side, uplo, M, N, alpha, a, lda, b, ldb, beta, c, ldc), M, N, ldc)
    absA = np.abs(np.nan_to_num(osy.mirror(k, a, lda, uplo)))
    if side.lower() == "l":
        scale = oracle.gemm_f64("n", "n", M, N, k, 1.0, absA, k, np.abs(b.astype(np.float64)), ldb, 0.0,
                                np.zeros(ldc * N), ldc)
    else:
        scale = oracle.gemm_f64("n", "n", M, N, k, 1.0, np.abs(b.astype(np.float64)), ldb, absA, k, 0.0,
                                np.zeros(ldc * N), ldc)
    u = U32 if dtype == np.float32 else U64
    tol = 1.01 * (k + 2) * u * abs(alpha) * _dense(scale, M, N, ldc) + 2 * u * abs(beta) * np.abs(
        _dense(c, M, N, ldc)) + 1e-300
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - want) <= tol), float(np.max(np.abs(got - want) / tol))
    if pad:  # padding rows of C untouched
        cc = tc.cpu().numpy().reshape(N, ldc)
        assert np.array_equal(cc[:, M:], c.reshape(N, ldc)[:, M:])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("side,uplo", [("l", "u"), ("l", "l"), ("r", "u"), ("R", "L")])
@pytest.mark.parametrize("M,N", [(1, 1), (77, 130), (300, 65), (64, 1000)])
def test_gpu_symm(handle, torch_cuda, dtype, side, uplo, M, N):
    _run(handle, torch_cuda, dtype, side, uplo, M, N, 1.25, 0.5, pad=3, seed=M + N)


@pytest.mark.gpu
@pytest.mark.parametrize("side", ["l", "r"])
def test_gpu_symm_large_and_beta_zero(handle, torch_cuda, side):
    _run(handle, torch_cuda, np.float32, side, "l", 2048, 1536, -0.75, 0.0, seed=9)


@pytest.mark.gpu
def test_gpu_symm_alpha_zero_and_errors(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    c = torch.ones(16, device="cuda")
    a = torch.full((16,), float("nan"), device="cuda")
    pb._symm(handle, "l", "u", 4, 4, 0.0, a, 4, a, 4, 2.0, c, 4)  # A, B not read
    handle.wait()
    assert torch.equal(c, torch.full((16,), 2.0, device="cuda"))
    with pytest.raises(ValueError):
        pb._symm(handle, "x", "u", 4, 4, 1.0, c, 4, c, 4, 0.0, c, 4)
    with pytest.raises(ValueError):
        pb._symm(handle, "l", "q", 4, 4, 1.0, c, 4, c, 4, 0.0, c, 4)
    with pytest.raises(ValueError):
        pb._symm(handle, "l", "u", 4, 4, 1.0, c, 3, c, 4, 0.0, c, 4)  # lda < M

"""TRSM (blas::_trsm, src/interface/trsm_interface.hpp:36-420): the CPU
oracle (oracle/trsm.py) pinned against OpenBLAS's dtrsm, then the GPU path
(64 x 64 diagonal blocks inverted in LDS + MFMA GEMM steps) on the reference
test's matrices (test/unittest/blas3/blas3_trsm_test.cpp: m, n in {7, 513,
1027}, lda and ldb twice the minimum, alpha 2, unused triangle 0 or NaN)
with a componentwise residual bound."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle import trsm as ot

U32 = 2.0 ** -24
U64 = 2.0 ** -53
COMBOS = list(itertools.product("lr", "ul", "nt", "un"))  # side, uplo, trans, diag


def _dense(flat, rows, cols, ld):
    return np.asarray(flat, np.float64)[: ld * cols].reshape(cols, ld).T[:rows, :]


@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_oracle_matches_openblas(side, uplo, trans, diag):
    from scipy.linalg import blas

    rng = np.random.default_rng(5)
    M, N = 9, 6
    k = M if side == "l" else N
    lda, ldb = k + 2, M + 1
    a = ot.trsm_matrix(rng, k, lda, uplo, diag, 3.0, 0.0, np.float64)
    b = rng.standard_normal(ldb * N)
    got = _dense(ot.trsm_f64(side, uplo, trans, diag, M, N, 2.0, a, lda, b, ldb), M, N, ldb)
    want = blas.dtrsm(2.0, _dense(a, k, k, lda), _dense(b, M, N, ldb), side=int(side == "r"),
                      lower=int(uplo == "l"), trans_a=int(trans == "t"), diag=int(diag == "u"))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # planted mistakes: the wrong transpose, or a read diagonal under diag 'u'
    wrong_t = _dense(ot.trsm_f64(side, uplo, "n" if trans == "t" else "t", diag, M, N, 2.0, a, lda, b, ldb),
                     M, N, ldb)
    assert not np.allclose(wrong_t, want)
    if diag == "u":
        wrong_d = _dense(ot.trsm_f64(side, uplo, trans, "n", M, N, 2.0, a, lda, b, ldb), M, N, ldb)
        assert not np.allclose(wrong_d, want)


def test_reference_matrix_is_diagonally_dominant_and_keeps_unused():
    rng = np.random.default_rng(1)
    k, lda = 40, 80
    a = ot.trsm_matrix(rng, k, lda, "l", "n", 4.0, np.nan, np.float64)
    A = _dense(a, k, k, lda)
    assert np.all(np.isnan(A[np.triu_indices(k, 1)]))
    low = np.tril(A, -1)
    assert np.all(np.abs(low).sum(axis=1) < 4.0) and np.all(np.diag(A) == 4.0)


def _unit_triangle(rng, k, lda, uplo, diag, c, unused, dtype):
    """Ones on the diagonal (stored only for diag 'n'), off-diagonal entries
    uniform in [-c, c] in the stored triangle, `unused` elsewhere: not
    diagonally dominant, and for c of order one badly conditioned (cond 1e4 to
    1e8 at the sizes tested)."""
    A = np.full((k, lda), unused, np.float64)  # A[col, row], column-major storage
    for j in range(k):
        for i in (range(j + 1, k) if uplo == "l" else range(j)):
            A[j, i] = rng.uniform(-c, c)
        A[j, j] = 1.0 if diag == "n" else unused
    return A.reshape(-1).astype(dtype)


def _check(handle, torch, dtype, side, uplo, trans, diag, M, N, unused, seed, offdiag=None):
    import portblas_amd as pb

    rng = np.random.default_rng(seed)
    k = M if side == "l" else N
    lda, ldb = 2 * k, 2 * M
    if offdiag is None:
        dval = rng.uniform(1.0, 10.0)
        a = ot.trsm_matrix(rng, k, lda, uplo, diag, dval, unused, dtype)
    else:
        a = _unit_triangle(rng, k, lda, uplo, diag, offdiag, unused, dtype)
    b = (rng.random(ldb * N) * 2 - 1).astype(dtype)
    ta, tb = torch.from_numpy(a.copy()).cuda(), torch.from_numpy(b.copy()).cuda()
    pb._trsm(handle, side, uplo, trans, diag, M, N, 2.0, ta, lda, tb, ldb)
    handle.wait()
    out = tb.cpu().numpy()
    X = _dense(out, M, N, ldb)
    assert np.all(np.isfinite(X)), "non-finite result"
    # padding rows of B are untouched
    assert np.array_equal(out.reshape(N, ldb)[:, M:], b.reshape(N, ldb)[:, M:])
    T = ot.effective(k, a, lda, uplo, trans, diag)
    B = _dense(b, M, N, ldb)
    if side == "l":
        resid, scale = T @ X - 2.0 * B, np.abs(T) @ np.abs(X) + 2.0 * np.abs(B)
    else:
        resid, scale = X @ T - 2.0 * B, np.abs(X) @ np.abs(T) + 2.0 * np.abs(B)
    u = U32 if dtype == np.float32 else U64
    r = np.max(np.abs(resid) / (16 * (k + 2) * u * scale + 1e-300))
    assert r <= 1.0, f"residual / bound {r:.3g}"


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
@pytest.mark.parametrize("M,N", [(7, 7), (513, 200), (100, 1027)])
def test_gpu_trsm_reference_cases(handle, torch_cuda, side, uplo, trans, diag, M, N):
    _check(handle, torch_cuda, np.float32, side, uplo, trans, diag, M, N, np.nan, seed=M * 7 + N)


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans,diag", [("l", "l", "n", "n"), ("r", "u", "t", "u"), ("l", "u", "t", "n"),
                                                  ("r", "l", "n", "n")])
def test_gpu_dtrsm_and_unused_zero(handle, torch_cuda, side, uplo, trans, diag):
    _check(handle, torch_cuda, np.float64, side, uplo, trans, diag, 513, 129, 0.0, seed=3)
    _check(handle, torch_cuda, np.float32, side, uplo, trans, diag, 1027, 64, 0.0, seed=4)


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans,diag", COMBOS)
def test_gpu_trsm_panels(handle, torch_cuda, side, uplo, trans, diag):
    """Order 2100 (33 diagonal blocks, the last of 52): the solve walks panels
    of several blocks, each followed by one wider trailing update."""
    M, N = (2100, 40) if side == "l" else (40, 2100)
    _check(handle, torch_cuda, np.float32, side, uplo, trans, diag, M, N, np.nan, seed=11)
    if diag == "n":
        _check(handle, torch_cuda, np.float64, side, uplo, trans, diag, M, N, 0.0, seed=12)


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans,diag,M,N,dtype", [
    ("l", "l", "n", "n", 2048, 512, np.float32), ("l", "u", "n", "n", 2048, 256, np.float32),
    ("r", "u", "n", "u", 512, 2048, np.float32), ("r", "l", "t", "n", 256, 2048, np.float32),
    ("l", "l", "t", "n", 2048, 512, np.float64), ("r", "l", "n", "n", 256, 2048, np.float64)])
def test_gpu_trsm_split_trailing_updates(handle, torch_cuda, side, uplo, trans, diag, M, N, dtype):
    """Shapes whose narrower trailing updates split K while the full-width
    update does not (n or m = 512: 256 tiles of 64 x 64 at full width, 192 at
    1536 rows) or splits less (256: 2-way at full width, 3-way at 1536 rows)."""
    _check(handle, torch_cuda, dtype, side, uplo, trans, diag, M, N, np.nan, seed=M + 3 * N)


@pytest.mark.gpu
@pytest.mark.parametrize("side,uplo,trans,diag", [("l", "l", "n", "n"), ("l", "u", "t", "u"), ("r", "u", "n", "n"),
                                                  ("r", "l", "t", "u")])
@pytest.mark.parametrize("k,c", [(513, 0.3), (200, 0.5), (128, 1.0)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gpu_trsm_ill_conditioned(handle, torch_cuda, side, uplo, trans, diag, k, c, dtype):
    """Unit-diagonal triangles with off-diagonal entries of order c (cond(A)
    about 4e4, 6e4 and 5e8; the 64 x 64 diagonal blocks about 15, 130 and 1e5):
    the explicitly inverted diagonal blocks must still give the componentwise
    residual bound that substitution guarantees."""
    M, N = (k, 96) if side == "l" else (96, k)
    _check(handle, torch_cuda, dtype, side, uplo, trans, diag, M, N, np.nan, seed=k + int(10 * c), offdiag=c)


@pytest.mark.gpu
def test_gpu_trsm_argument_errors(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    a = torch.eye(4, device="cuda").reshape(-1)
    b = torch.ones(16, device="cuda")
    for args in [("x", "u", "n", "n", 4, 4), ("l", "x", "n", "n", 4, 4), ("l", "u", "c", "n", 4, 4),
                 ("l", "u", "n", "x", 4, 4), ("l", "u", "n", "n", 0, 4)]:
        with pytest.raises(ValueError):
            pb._trsm(handle, *args, 1.0, a, 4, b, 4)
    pb._trsm(handle, "L", "U", "N", "U", 4, 4, 3.0, a, 4, b, 4)  # identity, unit: B <- 3 B
    handle.wait()
    assert torch.equal(b, torch.full((16,), 3.0, device="cuda"))

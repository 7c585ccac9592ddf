"""Complex GEMM (the reference's _gemm on std::complex, BLAS_ENABLE_COMPLEX):
the CPU oracle pinned against OpenBLAS's zgemm, then the GPU path (one real
MFMA GEMM of twice the size + a complex epilogue) on the reference test's
parameters (test/unittest/blas3/blas3_gemm_test.cpp:150-235: alpha 1.5+1i,
beta 1.5+3i or 0, alpha 0, 'n'/'t', leading dimensions multiplied, an
offset into the buffers) with the accumulation bound."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import cgemm as oc

U32 = 2.0 ** -24
U64 = 2.0 ** -53


def _crand(rng, n, dtype):
    return (rng.random(n) * 2 - 1 + 1j * (rng.random(n) * 2 - 1)).astype(dtype)


@pytest.mark.parametrize("ta,tb", [("n", "n"), ("n", "t"), ("t", "n"), ("t", "t")])
def test_oracle_matches_openblas(ta, tb):
    from scipy.linalg import blas

    rng = np.random.default_rng(2)
    M, N, K = 5, 4, 6
    lda = (M if ta == "n" else K) + 1
    ldb = (K if tb == "n" else N) + 2
    ldc = M + 1
    a = _crand(rng, lda * (K if ta == "n" else M), np.complex128)
    b = _crand(rng, ldb * (N if tb == "n" else K), np.complex128)
    c = _crand(rng, ldc * N, np.complex128)
    al, be = 1.5 + 1.0j, 1.5 + 3.0j
    got = oc.dense(oc.cgemm_c128(ta, tb, M, N, K, al, a, lda, b, ldb, be, c, ldc), M, N, ldc)
    A = oc.dense(a, *((M, K) if ta == "n" else (K, M)), lda)
    B = oc.dense(b, *((K, N) if tb == "n" else (N, K)), ldb)
    want = blas.zgemm(al, A, B, beta=be, c=oc.dense(c, M, N, ldc), trans_a=int(ta == "t"), trans_b=int(tb == "t"))
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
    if ta == "t":  # planted mistake: a conjugate transpose
        bad = blas.zgemm(al, A, B, beta=be, c=oc.dense(c, M, N, ldc), trans_a=2, trans_b=int(tb == "t"))
        assert not np.allclose(bad, want)


def _run(handle, torch, cdtype, ta, tb, m, n, k, alpha, beta, mul=2, offset=3, seed=0):
    import portblas_amd as pb

    rng = np.random.default_rng(seed)
    lda = max(1, (m if ta == "n" else k) * mul)
    ldb = max(1, (k if tb == "n" else n) * mul)
    ldc = max(1, m * mul)
    sa, sb, sc = lda * (k if ta == "n" else m), ldb * (n if tb == "n" else k), ldc * n
    a = _crand(rng, sa + offset, cdtype)
    b = _crand(rng, sb + offset, cdtype)
    c = _crand(rng, sc + offset, cdtype)
    ta_, tb_, tc_ = (torch.from_numpy(v.copy()).cuda() for v in (a, b, c))
    pb._gemm(handle, ta, tb, m, n, k, alpha, ta_[offset:], lda, tb_[offset:], ldb, beta, tc_[offset:], ldc)
    handle.wait()
    got = tc_.cpu().numpy()
    assert np.array_equal(got[:offset], c[:offset])  # before the offset: untouched
    want = oc.cgemm_c128(ta, tb, m, n, k, alpha, a[offset:], lda, b[offset:], ldb, beta, c[offset:], ldc)
    G, W = oc.dense(got[offset:], m, n, ldc), oc.dense(want, m, n, ldc)
    A = np.abs(oc.dense(a[offset:], *((m, k) if ta == "n" else (k, m)), lda))
    B = np.abs(oc.dense(b[offset:], *((k, n) if tb == "n" else (n, k)), ldb))
    A = A if ta == "n" else A.T
    B = B if tb == "n" else B.T
    u = U32 if cdtype == np.complex64 else U64
    scale = abs(alpha) * (A @ B if k > 0 else np.zeros((m, n)))
    tol = 2.02 * (2 * k + 4) * u * scale + 4 * u * abs(beta) * np.abs(oc.dense(c[offset:], m, n, ldc)) + 1e-300
    assert np.all(np.abs(G - W) <= tol), float(np.max(np.abs(G - W) / tol))
    if mul > 1:  # padding rows of C untouched
        cc = got[offset:offset + sc].reshape(n, ldc)
        assert np.array_equal(cc[:, m:], c[offset:offset + sc].reshape(n, ldc)[:, m:])


@pytest.mark.gpu
@pytest.mark.parametrize("cdtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("ta,tb", [("n", "n"), ("n", "t"), ("t", "n"), ("t", "T")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (11, 11, 11), (16, 63, 33), (257, 130, 200)])
@pytest.mark.parametrize("alpha,beta", [(1.5 + 1.0j, 1.5 + 3.0j), (1.5 + 3.0j, 0.0)])
def test_gpu_cgemm(handle, torch_cuda, cdtype, ta, tb, m, n, k, alpha, beta):
    _run(handle, torch_cuda, cdtype, ta.lower(), tb.lower(), m, n, k, alpha, beta, seed=m + n + k)


@pytest.mark.gpu
def test_gpu_cgemm_alpha_zero_and_large(handle, torch_cuda):
    _run(handle, torch_cuda, np.complex64, "n", "n", 64, 32, 16, 0.0, 1.0 + 0.0j)
    _run(handle, torch_cuda, np.complex64, "n", "n", 64, 32, 16, 0.0, 0.0)
    _run(handle, torch_cuda, np.complex64, "t", "n", 1024, 1024, 512, 1.0 - 0.5j, 0.25j, mul=1, offset=0, seed=9)

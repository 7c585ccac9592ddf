"""FP16 GEMMs (the reference's _gemm on sycl::half, BLAS_ENABLE_HALF):
pb_hgemm (half A, B, C) and pb_gemm_f16 (half A, B, FP32 C) on the GPU
against the FP64 GEMM oracle of the same (exactly representable) inputs.
Bounds: FP32 accumulation, (K + 2) 2^-24 |A||B|, plus one rounding to half
(2^-11 relative) for pb_hgemm's result and its beta C term."""
from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
U32 = 2.0 ** -24
UH = 2.0 ** -11


def _dense(flat, rows, cols, ld):
    return np.asarray(flat, np.float64)[: ld * cols].reshape(cols, ld).T[:rows, :]


def _case(rng, ta, tb, m, n, k, pad):
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    lda, ldb, ldc = max(ar, 1) + pad, max(br, 1) + pad, max(m, 1) + pad
    a = (rng.random(lda * ac) * 2 - 1).astype(np.float16)
    b = (rng.random(ldb * bc) * 2 - 1).astype(np.float16)
    return a, b, lda, ldb, ldc


@pytest.mark.parametrize("ta,tb", [("n", "n"), ("n", "t"), ("t", "n"), ("t", "t")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (77, 130, 200), (300, 260, 64), (1024, 1024, 512), (96, 80, 4096)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_hgemm(handle, torch_cuda, ta, tb, m, n, k, beta):
    import portblas_amd as pb

    torch = torch_cuda
    rng = np.random.default_rng(m + n + k)
    a, b, lda, ldb, ldc = _case(rng, ta, tb, m, n, k, pad=2)
    c = (rng.random(ldc * n) * 2 - 1).astype(np.float16)
    tc = torch.from_numpy(c.copy()).cuda()
    pb._gemm(handle, ta, tb, m, n, k, 1.25, torch.from_numpy(a).cuda(), lda, torch.from_numpy(b).cuda(), ldb,
             beta, tc, ldc)
    handle.wait()
    got = tc.cpu().numpy()
    ref = oracle.gemm_f64(ta, tb, m, n, k, 1.25, a.astype(np.float64), lda, b.astype(np.float64), ldb, beta,
                          c.astype(np.float64), ldc)
    scale = oracle.gemm_f64(ta, tb, m, n, k, 1.25, np.abs(a.astype(np.float64)), lda,
                            np.abs(b.astype(np.float64)), ldb, 0.0, np.zeros(ldc * n), ldc)
    R, G = _dense(ref, m, n, ldc), _dense(got, m, n, ldc)
    tol = (1.01 * (k + 2) * U32 + UH) * _dense(scale, m, n, ldc) + UH * np.abs(R) + \
        2 * UH * beta * np.abs(_dense(c, m, n, ldc)) + 2.0 ** -24
    assert np.all(np.abs(G - R) <= tol), float(np.max(np.abs(G - R) / tol))
    assert np.array_equal(got.reshape(n, ldc)[:, m:], c.reshape(n, ldc)[:, m:])  # padding untouched


def test_hgemm_alpha_zero_and_k_zero(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    c = torch.full((64,), 2.0, dtype=torch.float16, device="cuda")
    z = torch.zeros(64, dtype=torch.float16, device="cuda")
    pb._gemm(handle, "n", "n", 8, 8, 8, 0.0, z, 8, z, 8, 0.5, c, 8)
    handle.wait()
    assert torch.equal(c, torch.full((64,), 1.0, dtype=torch.float16, device="cuda"))
    pb._gemm(handle, "n", "n", 8, 8, 0, 1.0, z, 8, z, 8, 0.0, c, 8)  # K == 0, beta 0: C <- 0
    handle.wait()
    assert torch.equal(c, z)


@pytest.mark.parametrize("ta,tb,m,n,k,batch", [("n", "n", 200, 130, 77, 1), ("t", "n", 2048, 2048, 256, 1),
                                               ("n", "t", 64, 48, 1000, 3)])
def test_gemm_f16_fp32_output(handle, torch_cuda, ta, tb, m, n, k, batch):
    import portblas_amd as pb

    torch = torch_cuda
    rng = np.random.default_rng(7)
    a, b, lda, ldb, ldc = _case(rng, ta, tb, m, n, k, pad=0)
    sa, sb, sc = a.size, b.size, ldc * n
    A = np.concatenate([a] * batch)
    B = np.concatenate([b] * batch)
    C = (rng.random(sc * batch) * 2 - 1).astype(np.float32)
    tC = torch.from_numpy(C.copy()).cuda()
    pb._gemm_f16(handle, ta, tb, m, n, k, -0.75, torch.from_numpy(A).cuda(), lda, torch.from_numpy(B).cuda(),
                 ldb, 0.25, tC, ldc, sa, sb, sc, batch)
    handle.wait()
    got = tC.cpu().numpy()
    for i in range(batch):
        ref = oracle.gemm_f64(ta, tb, m, n, k, -0.75, a.astype(np.float64), lda, b.astype(np.float64), ldb, 0.25,
                              C[i * sc:(i + 1) * sc].astype(np.float64), ldc)
        scale = oracle.gemm_f64(ta, tb, m, n, k, 0.75, np.abs(a.astype(np.float64)), lda,
                                np.abs(b.astype(np.float64)), ldb, 0.0, np.zeros(sc), ldc)
        tol = 1.01 * (k + 2) * U32 * _dense(scale, m, n, ldc) + 2 * U32 * 0.25 * np.abs(
            _dense(C[i * sc:(i + 1) * sc], m, n, ldc)) + 1e-30
        err = np.abs(_dense(got[i * sc:(i + 1) * sc], m, n, ldc) - _dense(ref, m, n, ldc))
        assert np.all(err <= tol), float(np.max(err / tol))

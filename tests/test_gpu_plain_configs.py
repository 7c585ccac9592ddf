"""Every configuration of the plain FP32 / FP64 GEMM engine (the autotuner's
search space: 128 x 128 and 64 x 64 MFMA tiles with split-K 1 to 128, and the
small-problem batch kernel) run explicitly through pb_gemm_plain_config and
checked against the FP64 oracle, so every kernel the learned selector may pick
has parity evidence whatever it picks (src/interface/blas3/backend/
amd_gpu.hpp:52-288 makes this choice in the reference)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import U32, U64, colmajor

pytestmark = pytest.mark.gpu

# (ta, tb, m, n, k, batch, lda_pad): ragged edges and split K for the tile
# configs; the last shape also admits the small-batch kernel (256 >= CUs).
SHAPES = [
    ("n", "t", 200, 130, 1000, 1, 3),
    ("t", "n", 96, 72, 257, 2, 0),
    ("n", "n", 64, 48, 130, 256, 0),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("si", range(len(SHAPES)))
def test_every_plain_config(handle, torch_cuda, dtype, si):
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    ta, tb, m, n, k, batch, pad = SHAPES[si]
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    lda, ldb, ldc = ar + pad, br, m
    rng = np.random.default_rng(si)
    A = (rng.random(lda * ac * batch) * 2 - 1).astype(dtype)
    B = (rng.random(ldb * bc * batch) * 2 - 1).astype(dtype)
    C0 = (rng.random(ldc * n * batch) * 2 - 1).astype(dtype)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dt = 1 if dtype == np.float32 else 2
    u = U32 if dtype == np.float32 else U64
    alpha, beta = 1.5, -0.5
    nconf = L.pb_tune_num_plain_configs()
    assert nconf == 17 and L.pb_tune_plain_config_name(16) == b"small-batch"
    refs, bounds = [], []
    for b in range(batch):
        Ab, Bb, Cb = A[b * lda * ac:], B[b * ldb * bc:], C0[b * ldc * n:(b + 1) * ldc * n]
        refs.append(oracle.gemm_f64(ta, tb, m, n, k, alpha, Ab, lda, Bb, ldb, beta, Cb, ldc))
        absprod = oracle.gemm_absprod_f64(ta, tb, m, n, k, Ab, lda, Bb, ldb)
        bounds.append(1.01 * (k + 2) * u * abs(alpha) * absprod + 2 * u * abs(beta) * np.abs(Cb) + 1e-300)
    small_ok = m <= 64 and n <= 64 and batch >= handle.get_num_compute_units()
    for cid in range(-1, nconf):
        dC = torch.from_numpy(C0.copy()).cuda()
        st = L.pb_gemm_plain_config(handle.handle, dt, cid, ta.encode(), tb.encode(), m, n, k, alpha,
                                    dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc, batch)
        if cid == nconf - 1 and not small_ok:
            assert st == 2, "the small-batch config must refuse a shape it cannot run"
            continue
        pb.check(st)
        got = dC.cpu().numpy().astype(np.float64)
        for b in range(batch):
            e = np.abs(colmajor(got[b * ldc * n:], m, n, ldc) - colmajor(refs[b], m, n, ldc))
            tol = colmajor(bounds[b], m, n, ldc)
            assert np.all(e <= tol), f"config {cid} batch {b}: max err/tol {np.max(e / tol):.3g}"

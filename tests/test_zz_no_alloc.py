"""The allocation-free SGEMM (pb_sgemm_noalloc) against pb_sgemm, through the
C ABI.  Both run the same plan and kernels, so their results must be
bit-identical (with the built-in plan on both handles); the split-K partials
of the allocation-free call live in the
library's per-GPU scratch array instead of the handle workspace, which must
stay unallocated.  Two split-K calls in a row, the second smaller (it reuses
the array the first one left behind), then a call without a split, each also
held to the FP32 error bound against the FP64 oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import U32, colmajor, to_dev

pytestmark = pytest.mark.gpu

# (ta, tb, m, n, k, alpha, beta): the first two split K (fewer 64 x 64 tiles
# than compute units), the second smaller than the first; the third does not.
CASES = [
    ("n", "t", 192, 160, 8192, 1.5, 0.5),
    ("t", "n", 70, 50, 3000, -0.75, 0.0),
    ("n", "n", 1024, 1024, 256, 1.0, 2.0),
]


def _inputs(ta, tb, m, n, k, seed):
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    A = oracle.uniform(seed * 3 + 1, ar * ac)
    B = oracle.uniform(seed * 3 + 2, br * bc)
    C = oracle.uniform(seed * 3 + 3, m * n)
    return A, ar, B, br, C


def test_sgemm_noalloc_matches_sgemm(torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    # The built-in plan on both handles: an installed FP32 selector may pick a
    # split whose partials exceed the array, and then only pb_sgemm uses it.
    h_ref = pb.SB_Handle(torch.cuda.Stream(), plain_selector=None)
    h_na = pb.SB_Handle(torch.cuda.Stream(), plain_selector=None)
    assert h_na.workspace_bytes() == 0
    for i, (ta, tb, m, n, k, alpha, beta) in enumerate(CASES):
        print(f"pb_sgemm_noalloc case {i}: {ta}{tb} {m}x{n}x{k}", flush=True)
        A, lda, B, ldb, C = _inputs(ta, tb, m, n, k, seed=20 + i)
        dA, dB = to_dev(torch, A), to_dev(torch, B)
        c_ref, c_na = to_dev(torch, C), to_dev(torch, C)
        pb.check(L.pb_sgemm(h_ref.handle, ta.encode(), tb.encode(), m, n, k, alpha, dA.data_ptr(), lda,
                            dB.data_ptr(), ldb, beta, c_ref.data_ptr(), m))
        h_ref.wait()
        pb.check(L.pb_sgemm_noalloc(h_na.handle, ta.encode(), tb.encode(), m, n, k, alpha, dA.data_ptr(),
                                    lda, dB.data_ptr(), ldb, beta, c_na.data_ptr(), m))
        h_na.wait()
        got_ref, got = c_ref.cpu().numpy(), c_na.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), got_ref.view(np.uint32)), \
            f"case {i}: differs from pb_sgemm in {np.count_nonzero(got != got_ref)} entries"
        ref = oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, m)
        absprod = oracle.gemm_absprod_f64(ta, tb, m, n, k, A, lda, B, ldb)
        tol = 1.01 * (k + 2) * U32 * abs(alpha) * absprod + 2 * U32 * abs(beta) * np.abs(C) + 1e-30
        err = np.abs(got.astype(np.float64) - ref)
        assert np.all(err <= tol), f"case {i}: max err/tol {np.max(err / tol):.3g}"
        if i < 2:  # the split-K cases did use partials: the reference handle grew its workspace
            assert h_ref.workspace_bytes() > 0
        assert h_na.workspace_bytes() == 0, "pb_sgemm_noalloc grew the handle workspace"
    h_ref.close()
    h_na.close()


def test_sgemm_noalloc_argument_rules(torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    h = pb.SB_Handle(torch.cuda.Stream(), plain_selector=None)
    c = torch.full((16,), 3.0, device="cuda")
    # invalid transpose flag, as pb_sgemm reports it
    st = L.pb_sgemm_noalloc(h.handle, b"x", b"n", 4, 4, 4, 1.0, c.data_ptr(), 4, c.data_ptr(), 4, 0.0,
                            c.data_ptr(), 4)
    assert st == 1 and b"_TransA" in L.pb_last_error()
    # alpha == 0 scales C by beta and reads neither A nor B
    pb.check(L.pb_sgemm_noalloc(h.handle, b"n", b"n", 4, 4, 4, 0.0, None, 4, None, 4, 0.5, c.data_ptr(), 4))
    h.wait()
    assert torch.all(c == 1.5)
    # empty problem: nothing to do
    pb.check(L.pb_sgemm_noalloc(h.handle, b"n", b"n", 0, 4, 4, 1.0, None, 1, None, 4, 0.0, c.data_ptr(), 1))
    assert h.workspace_bytes() == 0
    h.close()

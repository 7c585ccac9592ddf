"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and run through the C ABI
of libportblas_amd.so; everything else runs on the CPU."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden" / "golden_cases.npz"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


def _gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def torch_cuda():
    if not _gpu_available():
        pytest.skip("no GPU visible in this process")
    import torch

    return torch


@pytest.fixture(scope="session")
def handle(torch_cuda):
    import portblas_amd as pb

    h = pb.SB_Handle()
    yield h
    h.wait()


def colmajor(flat, rows, cols, ld):
    """View of (the first ld * cols elements of) a flat column-major buffer
    as a rows x cols numpy array."""
    return np.asarray(flat)[: cols * ld].reshape(cols, ld).T[:rows, :]


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


U32 = 2.0 ** -24  # FP32 unit roundoff
U64 = 2.0 ** -53  # FP64 unit roundoff


def cfg1_inputs(golden):
    import oracle

    """The 1024^3 CPU-runnable configuration: A, B, C consecutive on stream 0."""
    ta, tb, m, n, k, lda, ldb, ldc = (int(v) for v in golden["cfg1_args"])
    alpha, beta = (float(v) for v in golden["cfg1_scal"])
    u = oracle.uniform(int(golden["cfg1_seed"]), m * k + k * n + m * n)
    return chr(ta), chr(tb), m, n, k, lda, ldb, ldc, alpha, beta, u[: m * k], u[m * k: m * k + k * n], \
        u[m * k + k * n:]


def check_cfg1(golden, got, slack):
    """Entries and checksums of a 1024^3 result against the CBLAS fixture:
    both carry at most one FP32 forward error, so each compare allows
    `slack` x the bound."""
    import oracle

    ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, A, B, C = cfg1_inputs(golden)
    G = colmajor(got, m, n, ldc).astype(np.float64)
    absprod = colmajor(oracle.gemm_absprod_f64(ta, tb, m, n, k, A, lda, B, ldb), m, n, m)
    tol = slack * (1.01 * (k + 2) * U32 * abs(alpha) * absprod + 2 * U32 * abs(beta) * np.abs(colmajor(C, m, n, ldc)) + 1e-30)
    rows, cols = golden["cfg1_rows"], golden["cfg1_cols"]
    e = np.abs(G[np.ix_(rows, cols)] - golden["cfg1_sampled"].astype(np.float64))
    assert np.all(e <= tol[np.ix_(rows, cols)]), f"sampled: max err/tol {np.max(e / tol[np.ix_(rows, cols)]):.3g}"
    assert np.all(np.abs(G.sum(axis=0) - golden["cfg1_colsum"]) <= tol.sum(axis=0))
    assert np.all(np.abs(G.sum(axis=1) - golden["cfg1_rowsum"]) <= tol.sum(axis=1))

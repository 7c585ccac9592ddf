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

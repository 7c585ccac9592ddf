"""Runs the C++ drop-in check (tests/cpp/test_dropin.cpp, built against
include/portblas_amd.hpp and libportblas_amd.so by portblas_amd/build.py) as a
child process on the GPU."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "test_dropin"


def test_cpp_dropin_api(torch_cuda):
    if not EXE.exists():
        pytest.fail("tests/cpp/test_dropin is not built (run portblas_amd/build.py)")
    proc = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "OK: 0 failure(s)" in proc.stdout

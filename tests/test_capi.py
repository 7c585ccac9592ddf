"""The C ABI library loads, exports every symbol include/portblas_amd.h
declares, and rejects calls without a handle — CPU only (no compute calls)."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

import portblas_amd as pb
from portblas_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def test_library_built_and_loads():
    assert _lib.LIB_PATH.exists(), "run `python -m portblas_amd.build` first"
    L = pb.load()
    assert L.pb_abi_version() == 1


def test_no_unresolved_internal_symbols():
    """Every kernel launch stub and internal function the library references
    is defined in it (an unresolved kernel stub links as a shared library but
    fails at load or launch time)."""
    out = subprocess.run(["nm", "-D", "-C", "--undefined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    bad = [line for line in out.splitlines() if "pbgemm" in line or "__device_stub" in line or " pb::" in line]
    assert not bad, bad


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = _lib.declared_symbols()
    assert len(declared) >= 50
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    # ... and the Python binding declares a signature for each of them.
    assert set(declared) <= set(_lib.SIGNATURES), set(declared) - set(_lib.SIGNATURES)


def test_declarations_cite_the_reference():
    text = (ROOT / "include" / "portblas_amd.h").read_text()
    assert len(re.findall(r"\.h(pp)?:\d+", text)) >= 15


def test_null_handle_is_rejected():
    L = pb.load()
    st = L.pb_sgemm(None, b"n", b"n", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4)
    assert st == 5  # PB_STATUS_NOT_INITIALIZED
    assert b"handle" in L.pb_last_error()
    assert L.pb_sdot(None, 4, None, 1, None, 1, None) == 5
    assert L.pb_status_string(1) == b"PB_STATUS_INVALID_VALUE"
    assert L.pb_tune_num_configs() == 40
    assert L.pb_tune_config_name(0).startswith(b"128x128")
    assert L.pb_tune_config_name(99) == b""


def test_handle_create_without_gpu_fails_cleanly():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    L = pb.load()
    h = ctypes.c_void_p()
    st = L.pb_handle_create(None, ctypes.byref(h))
    assert st != 0 and not h.value
    with pytest.raises((RuntimeError, NotImplementedError)):
        pb.SB_Handle(0)


def test_status_exceptions_map_like_the_reference():
    # INVALID_VALUE -> ValueError (std::invalid_argument), NOT_SUPPORTED ->
    # NotImplementedError (blas::unsupported_exception), others -> RuntimeError.
    with pytest.raises(ValueError):
        pb.check(1)
    with pytest.raises(NotImplementedError):
        pb.check(2)
    with pytest.raises(RuntimeError):
        pb.check(3)
    pb.check(0)


def test_symm_trsm_reject_mixed_or_unsupported_element_types():
    """_symm and _trsm take one real element type for every operand: a float64
    A with float32 B, or bfloat16 operands, raise TypeError before any call
    into the library (no handle or GPU needed)."""
    import types

    import torch

    fake = types.SimpleNamespace(_L=None, handle=None)
    f32, f64 = torch.zeros(4), torch.zeros(4, dtype=torch.float64)
    bf = torch.zeros(4, dtype=torch.bfloat16)
    with pytest.raises(TypeError):
        pb._trsm(fake, "l", "l", "n", "n", 2, 2, 1.0, f64, 2, f32, 2)
    with pytest.raises(TypeError):
        pb._trsm(fake, "l", "l", "n", "n", 2, 2, 1.0, bf, 2, bf, 2)
    with pytest.raises(TypeError):
        pb._symm(fake, "l", "u", 2, 2, 1.0, f32, 2, f32, 2, 0.0, f64, 2)

"""CPU oracle for the portBLAS hot path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline), never as a code path of the library.  ``portblas_amd`` never
imports it and has no CPU fallback.

Contents (C sources in ``pb_oracle.c``, built by ``oracle/Makefile``):

* FP64-accumulated references of GEMM, GEMV, dot, asum, nrm2 and the
  extension reduction (``*_f64``), the exact iamax/iamin rule, and
  ``absprod`` (|op(A)||op(B)|, the scale of GEMM rounding error);
* restatements of portBLAS's own CPU schedules in FP32 (``pb_sgemm``,
  ``pb_sgemv``, ``pb_sreduce``) used as the timed CPU baseline;
* the seeded input generator (splitmix64) shared by tests and benchmarks.

Pinning (see tests/test_oracle.py): the references are checked against
OpenBLAS through scipy.linalg.blas (the CBLAS the reference's own tests use,
common/include/common/system_reference_blas.hpp), the committed golden
fixtures under tests/golden/, closed forms, and brute force.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB_PATH = _DIR / "liboracle.so"
_lib = None

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_f32 = ctypes.c_float
_ch = ctypes.c_char


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not _LIB_PATH.exists() or _LIB_PATH.stat().st_mtime < (_DIR / "pb_oracle.c").stat().st_mtime:
        subprocess.run(["make", "-B", "-C", str(_DIR)], check=True, capture_output=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        L.orc_gemm_f64.argtypes = [_ch, _ch, _i64, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p, _i64]
        L.orc_gemm_absprod_f64.argtypes = [_ch, _ch, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _i64]
        L.orc_pb_sgemm.argtypes = [_ch, _ch, _i64, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64]
        L.orc_gemv_f64.argtypes = [_ch, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p, _i64]
        L.orc_pb_sgemv.argtypes = [_ch, _i64, _i64, _f32, _p, _i64, _p, _i64, _f32, _p, _i64]
        L.orc_saxpy.argtypes = [_i64, _f32, _p, _i64, _p, _i64]
        L.orc_daxpy.argtypes = [_i64, _f64, _p, _i64, _p, _i64]
        for name in ("orc_dot_f64",):
            getattr(L, name).argtypes = [_i64, _p, _i64, _p, _i64, _f64]
            getattr(L, name).restype = _f64
        for name in ("orc_asum_f64", "orc_nrm2_f64"):
            getattr(L, name).argtypes = [_i64, _p, _i64, _f64]
            getattr(L, name).restype = _f64
        L.orc_pb_sreduce.argtypes = [ctypes.c_int, _i64, _p, _i64, _p, _i64, _f32]
        L.orc_pb_sreduce.restype = _f32
        L.orc_iamax_f64.argtypes = [ctypes.c_int, _i64, _p, _i64, _i64, ctypes.c_int]
        L.orc_iamax_f64.restype = _i64
        L.orc_reduction_f64.argtypes = [ctypes.c_int, ctypes.c_int, _p, _i64, _p, _i64, _i64]
        L.orc_uniform_f32.argtypes = [ctypes.c_uint64, _i64, _f32, _f32, _p]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _c(t: str) -> bytes:
    return t.encode()[:1]


# ---------------------------------------------------------------------------
# Inputs
# ---------------------------------------------------------------------------
def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """n float32 values, uniform in [lo, hi), splitmix64 stream `seed`."""
    out = np.empty(int(n), dtype=np.float32)
    lib().orc_uniform_f32(ctypes.c_uint64(seed & (2**64 - 1)), int(n), lo, hi, _ptr(out))
    return out


# ---------------------------------------------------------------------------
# References (FP64 accumulation).  Arrays are flat column-major buffers.
# ---------------------------------------------------------------------------
def gemm_f64(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc) -> np.ndarray:
    """Returns a float64 copy of C after C = alpha op(A) op(B) + beta C."""
    A64 = np.ascontiguousarray(A, dtype=np.float64)
    B64 = np.ascontiguousarray(B, dtype=np.float64)
    C64 = np.array(C, dtype=np.float64, copy=True)
    lib().orc_gemm_f64(_c(ta), _c(tb), M, N, K, float(alpha), _ptr(A64), lda, _ptr(B64), ldb,
                       float(beta), _ptr(C64), ldc)
    return C64


def gemm_absprod_f64(ta, tb, M, N, K, A, lda, B, ldb) -> np.ndarray:
    """|op(A)||op(B)| as an M x N column-major float64 buffer (ld = M)."""
    A64 = np.ascontiguousarray(A, dtype=np.float64)
    B64 = np.ascontiguousarray(B, dtype=np.float64)
    out = np.zeros(max(M * N, 1), dtype=np.float64)
    lib().orc_gemm_absprod_f64(_c(ta), _c(tb), M, N, K, _ptr(A64), lda, _ptr(B64), ldb, _ptr(out), max(M, 1))
    return out[: M * N]


def pb_sgemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc) -> np.ndarray:
    """portBLAS CPU schedule restated in FP32; returns the new C (float32)."""
    A32 = np.ascontiguousarray(A, dtype=np.float32)
    B32 = np.ascontiguousarray(B, dtype=np.float32)
    C32 = np.array(C, dtype=np.float32, copy=True)
    lib().orc_pb_sgemm(_c(ta), _c(tb), M, N, K, float(alpha), _ptr(A32), lda, _ptr(B32), ldb,
                       float(beta), _ptr(C32), ldc)
    return C32


def gemv_f64(trans, M, N, alpha, A, lda, x, incx, beta, y, incy) -> np.ndarray:
    A64 = np.ascontiguousarray(A, dtype=np.float64)
    x64 = np.ascontiguousarray(x, dtype=np.float64)
    y64 = np.array(y, dtype=np.float64, copy=True)
    lib().orc_gemv_f64(_c(trans), M, N, float(alpha), _ptr(A64), lda, _ptr(x64), incx, float(beta),
                       _ptr(y64), incy)
    return y64


def pb_sgemv(trans, M, N, alpha, A, lda, x, incx, beta, y, incy) -> np.ndarray:
    A32 = np.ascontiguousarray(A, dtype=np.float32)
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    y32 = np.array(y, dtype=np.float32, copy=True)
    lib().orc_pb_sgemv(_c(trans), M, N, float(alpha), _ptr(A32), lda, _ptr(x32), incx, float(beta),
                       _ptr(y32), incy)
    return y32


def dot_f64(n, x, incx, y, incy, rs0=0.0) -> float:
    x64 = np.ascontiguousarray(x, dtype=np.float64)
    y64 = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_dot_f64(n, _ptr(x64), incx, _ptr(y64), incy, float(rs0))


def asum_f64(n, x, incx, rs0=0.0) -> float:
    x64 = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_asum_f64(n, _ptr(x64), incx, float(rs0))


def nrm2_f64(n, x, incx, rs0=0.0) -> float:
    x64 = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_nrm2_f64(n, _ptr(x64), incx, float(rs0))


def pb_sreduce(op: str, n, x, incx, y=None, incy=1, rs0=0.0) -> float:
    """portBLAS CPU reduction schedule in FP32; op in {'dot','asum','nrm2'}."""
    code = {"dot": 0, "asum": 1, "nrm2": 2}[op]
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    y32 = np.ascontiguousarray(y if y is not None else x, dtype=np.float32)
    return lib().orc_pb_sreduce(code, n, _ptr(x32), incx, _ptr(y32), incy, float(rs0))


INT32_MAX = 2**31 - 1


def iamax(n, x, incx, is_max=True, sentinel=INT32_MAX) -> int:
    """Exact iamax/iamin index (portBLAS rule); x is float32 or float64."""
    elem_is_float = int(np.asarray(x).dtype == np.float32)
    x64 = np.ascontiguousarray(x, dtype=np.float64)
    return int(lib().orc_iamax_f64(int(is_max), n, _ptr(x64) if x64.size else 0, incx, sentinel,
                                   elem_is_float))


REDUCTION_OPS = {"add": 0, "product": 1, "max": 2, "min": 3, "abs_add": 4, "mean": 5}


def reduction_f64(op: str, dim: str, A, ld, rows, cols) -> np.ndarray:
    A64 = np.ascontiguousarray(A, dtype=np.float64)
    n = cols if dim == "inner" else rows
    out = np.zeros(max(n, 1), dtype=np.float64)
    lib().orc_reduction_f64(REDUCTION_OPS[op], 0 if dim == "inner" else 1, _ptr(A64), ld, _ptr(out), rows, cols)
    return out[:n]


def num_threads() -> int:
    return int(lib().orc_num_threads())

"""Python mirror of portBLAS's hot-path API over the C ABI.

Names, argument order and meaning follow the reference
(include/interface/blas3_interface.h:86-123, blas2_interface.h:629-654,
blas1_interface.h:582-718 and 905-1000, extension_interface.h:453-462).
Containers are device tensors (torch, the stand-in for USM pointers) or raw
device addresses (int).  Errors raise ValueError where portBLAS throws
std::invalid_argument, NotImplementedError for blas::unsupported_exception and
RuntimeError for device failures.  Every call runs the HIP kernels of
libportblas_amd.so; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes
import enum
from pathlib import Path
from typing import Any

from . import _lib

# The learned tile-configuration selector trained on MI355X timings.
DEFAULT_SELECTOR = Path(__file__).resolve().parent / "selector_mi355x.npz"
DEFAULT_PLAIN_SELECTOR = Path(__file__).resolve().parent / "selector_plain_mi355x.npz"

# ---------------------------------------------------------------------------
# Enums (include/operations/blas3_trees.h:59, extension/reduction.h:40)
# ---------------------------------------------------------------------------


class gemm_batch_type_t(enum.IntEnum):
    strided = 0
    interleaved = 1


class reduction_dim_t(enum.IntEnum):
    inner = 0
    outer = 1


class gemm_precision_t(enum.IntEnum):
    fp32_mfma = 0
    fp32_simt = 1
    bf16x3 = 2
    bf16 = 3
    fp8_e4m3 = 4
    bf16x4 = 5
    fp16x3 = 6


# Operators of extension::_reduction (src/operations/blas_operators.hpp).
AddOperator = 0
ProductOperator = 1
MaxOperator = 2
MinOperator = 3
AbsoluteAddOperator = 4
MeanOperator = 5

_STATUS_EXC = {1: ValueError, 2: NotImplementedError}


def check(status: int) -> None:
    if status != 0:
        L = _lib.load()
        msg = (L.pb_last_error() or b"").decode()
        name = (L.pb_status_string(status) or b"").decode()
        raise _STATUS_EXC.get(status, RuntimeError)(f"{name}: {msg}")


def _addr(x: Any) -> int:
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    return int(x.data_ptr())


def _kind(x: Any, default: str = "f32") -> str:
    """'f32' / 'f64' / 'bf16' from a tensor's dtype."""
    dt = getattr(x, "dtype", None)
    if dt is None:
        return default
    s = str(dt)
    if s.endswith("float32"):
        return "f32"
    if s.endswith("float64"):
        return "f64"
    if s.endswith("bfloat16") or s.endswith("int16") or s.endswith("uint16"):
        return "bf16"
    raise TypeError(f"unsupported element type {dt}")


def _same_real_kind(*xs: Any) -> str:
    """'f32' or 'f64' when every operand has that element type; TypeError
    otherwise (the reference's SYMM and TRSM take one element type)."""
    kinds = {_kind(x) for x in xs}
    if len(kinds) != 1 or not kinds <= {"f32", "f64"}:
        raise TypeError(f"operands must all be float32 or all float64, got "
                        f"{[str(getattr(x, 'dtype', type(x).__name__)) for x in xs]}")
    return kinds.pop()


def _t(c: str) -> bytes:
    return c.encode()[:1]


class SB_Handle:
    """Replaces blas::SB_Handle (include/sb_handle/portblas_handle.h:46-200).

    ``stream`` is a torch.cuda.Stream, a raw hipStream_t (int) or None for the
    current torch stream (the null stream without torch).

    ``selector``: "default" installs the learned tile-configuration selector
    trained on MI355X (``selector_mi355x.npz`` next to this module) for the
    BF16 GEMM engine; None keeps the built-in rule; a path installs that
    network (``portblas_amd.tuning.Selector.save`` format).

    ``plain_selector``: the same for the FP32 / FP64 engine behind _gemm
    (``selector_plain_mi355x.npz``)."""

    def __init__(self, stream: Any = None, selector: Any = "default", plain_selector: Any = "default"):
        self._L = _lib.load()
        self._h = ctypes.c_void_p()
        check(self._L.pb_handle_create(self._stream_ptr(stream), ctypes.byref(self._h)))
        if selector is not None or plain_selector is not None:
            from . import tuning

            if selector is not None:
                path = DEFAULT_SELECTOR if selector == "default" else Path(selector)
                if selector != "default" or path.exists():
                    tuning.Selector.load(path).install(self)
            if plain_selector is not None:
                path = DEFAULT_PLAIN_SELECTOR if plain_selector == "default" else Path(plain_selector)
                if plain_selector != "default" or path.exists():
                    tuning.install_plain(tuning.Selector.load(path), self)

    @staticmethod
    def _stream_ptr(stream: Any) -> int:
        if stream is None:
            try:
                import torch

                return int(torch.cuda.current_stream().cuda_stream)
            except Exception:
                return 0
        if isinstance(stream, int):
            return stream
        return int(stream.cuda_stream)

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def set_stream(self, stream: Any) -> None:
        check(self._L.pb_handle_set_stream(self._h, self._stream_ptr(stream)))

    def get_num_compute_units(self) -> int:
        return int(self._L.pb_handle_get_num_compute_units(self._h))

    def wait(self, events: Any = None) -> None:
        check(self._L.pb_handle_wait(self._h))

    def reserve_workspace(self, nbytes: int) -> None:
        check(self._L.pb_handle_reserve_workspace(self._h, int(nbytes)))

    def workspace_bytes(self) -> int:
        return int(self._L.pb_handle_workspace_bytes(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            check(self._L.pb_handle_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# BLAS-3
# ---------------------------------------------------------------------------
def _gemm(sb_handle, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc):
    L = sb_handle._L
    dt = str(getattr(c, "dtype", ""))
    if dt.endswith("float16") and "b" not in dt.split(".")[-1]:  # torch.float16 / numpy float16
        check(L.pb_hgemm(sb_handle.handle, _t(transa), _t(transb), M, N, K, float(alpha), _addr(a), lda,
                         _addr(b), ldb, float(beta), _addr(c), ldc))
        return
    if "complex" in dt:  # interleaved complex64 / complex128 tensors
        real = ctypes.c_double if "128" in dt else ctypes.c_float
        fn = L.pb_zgemm if "128" in dt else L.pb_cgemm
        al = (real * 2)(complex(alpha).real, complex(alpha).imag)
        be = (real * 2)(complex(beta).real, complex(beta).imag)
        check(fn(sb_handle.handle, _t(transa), _t(transb), M, N, K, al, _addr(a), lda, _addr(b), ldb, be,
                 _addr(c), ldc))
        return
    k = _kind(c)
    fn = L.pb_dgemm if k == "f64" else L.pb_sgemm
    check(fn(sb_handle.handle, _t(transa), _t(transb), M, N, K, alpha, _addr(a), lda, _addr(b), ldb,
             beta, _addr(c), ldc))


def _gemm_batched(sb_handle, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc,
                  batch_size, batch_type=gemm_batch_type_t.strided):
    L = sb_handle._L
    fn = L.pb_dgemm_batched if _kind(c) == "f64" else L.pb_sgemm_batched
    check(fn(sb_handle.handle, _t(transa), _t(transb), M, N, K, alpha, _addr(a), lda, _addr(b), ldb,
             beta, _addr(c), ldc, batch_size, int(batch_type)))


def _gemm_strided_batched(sb_handle, transa, transb, M, N, K, alpha, a, lda, stridea, b, ldb,
                          strideb, beta, c, ldc, stridec, batch_size):
    L = sb_handle._L
    fn = L.pb_dgemm_strided_batched if _kind(c) == "f64" else L.pb_sgemm_strided_batched
    check(fn(sb_handle.handle, _t(transa), _t(transb), M, N, K, alpha, _addr(a), lda, stridea,
             _addr(b), ldb, strideb, beta, _addr(c), ldc, stridec, batch_size))


def _gemm_ex(sb_handle, precision, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc,
             stridea=0, strideb=0, stridec=0, batch_size=1):
    """FP32-in/FP32-out GEMM with a selectable arithmetic (gemm_precision_t)."""
    check(sb_handle._L.pb_gemm_ex(sb_handle.handle, int(precision), _t(transa), _t(transb), M, N, K,
                                  alpha, _addr(a), lda, stridea, _addr(b), ldb, strideb, beta,
                                  _addr(c), ldc, stridec, batch_size))


def _gemm_f16(sb_handle, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc,
              stridea=0, strideb=0, stridec=0, batch_size=1):
    """FP16 inputs, FP32 accumulate and output."""
    check(sb_handle._L.pb_gemm_f16(sb_handle.handle, _t(transa), _t(transb), M, N, K, alpha, _addr(a), lda,
                                   stridea, _addr(b), ldb, strideb, beta, _addr(c), ldc, stridec, batch_size))


def _gemm_bf16(sb_handle, transa, transb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc,
               stridea=0, strideb=0, stridec=0, batch_size=1):
    """BF16 inputs, FP32 accumulate and output."""
    check(sb_handle._L.pb_gemm_bf16(sb_handle.handle, _t(transa), _t(transb), M, N, K, alpha,
                                    _addr(a), lda, stridea, _addr(b), ldb, strideb, beta, _addr(c),
                                    ldc, stridec, batch_size))


def _symm(sb_handle, side, uplo, M, N, alpha, a, lda, b, ldb, beta, c, ldc):
    """C <- alpha A B + beta C (side 'l') or alpha B A + beta C (side 'r'),
    A symmetric with its `uplo` triangle stored (symm_interface.hpp:36-80)."""
    L = sb_handle._L
    fn = L.pb_dsymm if _same_real_kind(a, b, c) == "f64" else L.pb_ssymm
    check(fn(sb_handle.handle, _t(side), _t(uplo), M, N, alpha, _addr(a), lda, _addr(b), ldb, beta,
             _addr(c), ldc))


def _trsm(sb_handle, side, uplo, trans, diag, M, N, alpha, a, lda, b, ldb):
    """Solves op(A) X = alpha B (side 'l') or X op(A) = alpha B (side 'r'),
    overwriting B with X (trsm_interface.hpp:36-420)."""
    L = sb_handle._L
    fn = L.pb_dtrsm if _same_real_kind(a, b) == "f64" else L.pb_strsm
    check(fn(sb_handle.handle, _t(side), _t(uplo), _t(trans), _t(diag), M, N, alpha, _addr(a), lda,
             _addr(b), ldb))


# ---------------------------------------------------------------------------
# BLAS-2
# ---------------------------------------------------------------------------
def _gemv(sb_handle, trans, M, N, alpha, a, lda, x, incx, beta, y, incy):
    L = sb_handle._L
    fn = L.pb_dgemv if _kind(y) == "f64" else L.pb_sgemv
    check(fn(sb_handle.handle, _t(trans), M, N, alpha, _addr(a), lda, _addr(x), incx, beta, _addr(y),
             incy))


# ---------------------------------------------------------------------------
# BLAS-1.  With `rs` (a device tensor) the call is asynchronous and the
# result lands in rs exactly as portBLAS's atomics leave it (dot/asum/nrm2
# accumulate into *rs); without `rs` the synchronous form returns the value.
# ---------------------------------------------------------------------------
def _dot(sb_handle, N, x, incx, y, incy, rs=None):
    L = sb_handle._L
    if rs is None:
        if _kind(x) == "f64":
            return _sync_via_device(sb_handle, "f64", lambda d: check(
                L.pb_ddot(sb_handle.handle, N, _addr(x), incx, _addr(y), incy, d)))
        out = ctypes.c_float()
        check(L.pb_sdot_sync(sb_handle.handle, N, _addr(x), incx, _addr(y), incy, ctypes.byref(out)))
        return out.value
    fn = L.pb_ddot if _kind(x) == "f64" else L.pb_sdot
    check(fn(sb_handle.handle, N, _addr(x), incx, _addr(y), incy, _addr(rs)))


def _asum(sb_handle, N, x, incx, rs=None):
    L = sb_handle._L
    if rs is None:
        if _kind(x) == "f64":
            return _sync_via_device(sb_handle, "f64", lambda d: check(
                L.pb_dasum(sb_handle.handle, N, _addr(x), incx, d)))
        out = ctypes.c_float()
        check(L.pb_sasum_sync(sb_handle.handle, N, _addr(x), incx, ctypes.byref(out)))
        return out.value
    fn = L.pb_dasum if _kind(x) == "f64" else L.pb_sasum
    check(fn(sb_handle.handle, N, _addr(x), incx, _addr(rs)))


def _nrm2(sb_handle, N, x, incx, rs=None):
    L = sb_handle._L
    if rs is None:
        if _kind(x) == "f64":
            return _sync_via_device(sb_handle, "f64", lambda d: check(
                L.pb_dnrm2(sb_handle.handle, N, _addr(x), incx, d)))
        out = ctypes.c_float()
        check(L.pb_snrm2_sync(sb_handle.handle, N, _addr(x), incx, ctypes.byref(out)))
        return out.value
    fn = L.pb_dnrm2 if _kind(x) == "f64" else L.pb_snrm2
    check(fn(sb_handle.handle, N, _addr(x), incx, _addr(rs)))


def _sdsdot(sb_handle, N, sb, x, incx, y, incy, rs=None):
    """rs <- sb + (rs + dot(x, y)) in FP32 (blas1_interface.hpp:148-177); the
    synchronous form (no rs) returns sb + dot(x, y)."""
    L = sb_handle._L
    if rs is None:
        out = ctypes.c_float()
        check(L.pb_sdsdot_sync(sb_handle.handle, N, sb, _addr(x), incx, _addr(y), incy, ctypes.byref(out)))
        return out.value
    check(L.pb_sdsdot(sb_handle.handle, N, sb, _addr(x), incx, _addr(y), incy, _addr(rs)))


def _axpy(sb_handle, N, alpha, x, incx, y, incy):
    """y <- y + alpha x (blas1_interface.hpp:46-60)."""
    L = sb_handle._L
    fn = L.pb_daxpy if _kind(y) == "f64" else L.pb_saxpy
    check(fn(sb_handle.handle, N, alpha, _addr(x), incx, _addr(y), incy))


def _scal(sb_handle, N, alpha, x, incx):
    """x <- alpha x; alpha == 0 assigns zero (blas1_interface.hpp:440-456)."""
    L = sb_handle._L
    fn = L.pb_dscal if _kind(x) == "f64" else L.pb_sscal
    check(fn(sb_handle.handle, N, alpha, _addr(x), incx))


def _copy(sb_handle, N, x, incx, y, incy):
    """y <- x (blas1_interface.hpp:62-74)."""
    L = sb_handle._L
    fn = L.pb_dcopy if _kind(y) == "f64" else L.pb_scopy
    check(fn(sb_handle.handle, N, _addr(x), incx, _addr(y), incy))


def _swap(sb_handle, N, x, incx, y, incy):
    """x <-> y (blas1_interface.hpp:419-430)."""
    L = sb_handle._L
    fn = L.pb_dswap if _kind(y) == "f64" else L.pb_sswap
    check(fn(sb_handle.handle, N, _addr(x), incx, _addr(y), incy))


def _iamax(sb_handle, N, x, incx, rs=None):
    return _iamaxmin(sb_handle, N, x, incx, rs, True)


def _iamin(sb_handle, N, x, incx, rs=None):
    return _iamaxmin(sb_handle, N, x, incx, rs, False)


def _iamaxmin(sb_handle, N, x, incx, rs, is_max):
    L = sb_handle._L
    k = _kind(x)
    if rs is None:
        if k == "f32":
            out = ctypes.c_int32()
            fn = L.pb_isamax_sync if is_max else L.pb_isamin_sync
            check(fn(sb_handle.handle, N, _addr(x), incx, ctypes.byref(out)))
            return out.value
        fn = L.pb_idamax if is_max else L.pb_idamin
        return _sync_via_device(sb_handle, "i32", lambda d: check(fn(sb_handle.handle, N, _addr(x), incx, d)))
    idx64 = str(getattr(rs, "dtype", "")).endswith("int64")
    if k == "f64":
        if idx64:
            raise NotImplementedError("64-bit index results are provided for float vectors")
        fn = L.pb_idamax if is_max else L.pb_idamin
    elif idx64:
        fn = L.pb_isamax_64 if is_max else L.pb_isamin_64
    else:
        fn = L.pb_isamax if is_max else L.pb_isamin
    check(fn(sb_handle.handle, N, _addr(x), incx, _addr(rs)))


def _sync_via_device(sb_handle, kind, launch):
    import torch

    dt = {"f64": torch.float64, "f32": torch.float32, "i32": torch.int32}[kind]
    d = torch.zeros(1, dtype=dt, device="cuda") if kind != "i32" else torch.full((1,), -1, dtype=dt, device="cuda")
    launch(d.data_ptr())
    sb_handle.wait()
    return d.item()


# ---------------------------------------------------------------------------
# Extension
# ---------------------------------------------------------------------------
class extension:
    @staticmethod
    def _reduction(operator, sb_handle, buffer_in, ld, buffer_out, rows, cols, reduction_dim):
        L = sb_handle._L
        fn = L.pb_dreduction if _kind(buffer_in) == "f64" else L.pb_sreduction
        check(fn(sb_handle.handle, int(operator), int(reduction_dim), _addr(buffer_in), ld,
                 _addr(buffer_out), rows, cols))

"""Streaming BLAS-1 and sdsdot restated on the CPU (numpy) — TEST
INFRASTRUCTURE ONLY, imported by tests as the checker.

Follows src/interface/blas1_interface.hpp:
  _axpy   :46-60    y <- alpha * x + y with one rounding (a fused multiply-add,
                    the C oracle's fmaf / fma): how OpenBLAS, the reference's
                    test oracle, and contracting device compilers evaluate it
  _copy   :62-74    y <- x
  _sdsdot :148-177  rs <- sb + (rs + sum x_i y_i); n == 0 writes sb
  _swap   :419-430  x <-> y
  _scal   :440-456  x <- alpha * x; alpha == 0 assigns zero
and the vector views of src/views/view_sycl.hpp:95-100 (element i at
x[i * inc]; a negative increment walks from x[(n - 1) * |inc|] down).

Pinned in tests/test_blas1_stream.py against OpenBLAS through
scipy.linalg.blas (saxpy, sscal, scopy, sswap with positive and negative
increments) and closed forms (scal by zero on NaN and inf, sdsdot's n == 0
and accumulation).
"""
from __future__ import annotations

import numpy as np


def view_index(n: int, inc: int) -> np.ndarray:
    """Storage indices of the n elements of a BLAS vector view."""
    i = np.arange(max(n, 0), dtype=np.int64)
    return i * inc if inc >= 0 else (n - 1 - i) * (-inc)


def axpy(n, alpha, x, incx, y, incy):
    import ctypes

    from . import lib

    x = np.ascontiguousarray(x)
    y = np.array(y, copy=True)
    if n <= 0:
        return y
    assert x.dtype == y.dtype and x.dtype in (np.float32, np.float64)
    fn = lib().orc_saxpy if x.dtype == np.float32 else lib().orc_daxpy
    fn(n, float(alpha), x.ctypes.data_as(ctypes.c_void_p), incx, y.ctypes.data_as(ctypes.c_void_p), incy)
    return y


def scal(n, alpha, x, incx):
    x = np.array(x, copy=True)
    if n <= 0:
        return x
    ix = view_index(n, incx)
    t = x.dtype.type
    x[ix] = t(0) if alpha == 0 else t(alpha) * x[ix]
    return x


def copy(n, x, incx, y, incy):
    y = np.array(y, copy=True)
    if n > 0:
        y[view_index(n, incy)] = x[view_index(n, incx)]
    return y


def swap(n, x, incx, y, incy):
    x, y = np.array(x, copy=True), np.array(y, copy=True)
    if n > 0:
        ix, iy = view_index(n, incx), view_index(n, incy)
        tx = x[ix].copy()
        x[ix] = y[iy]
        y[iy] = tx
    return x, y


def sdsdot_f64(n, sb, x, incx, y, incy, rs0=0.0) -> float:
    """The exact value sb + (rs0 + sum x_i y_i) (FP64 sum of exact FP32
    products), and sb alone for n == 0."""
    if n == 0:
        return float(np.float32(sb))
    if n < 0:
        return float(sb) + float(rs0)
    ix, iy = view_index(n, incx), view_index(n, incy)
    return float(sb) + (float(rs0) + float(np.dot(x[ix].astype(np.float64), y[iy].astype(np.float64))))

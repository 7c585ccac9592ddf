"""Streaming BLAS-1 (axpy, scal, copy, swap) and sdsdot: the CPU oracle
(oracle/blas1x.py) pinned against OpenBLAS through scipy.linalg.blas and
closed forms, then the GPU kernels against the oracle — bit-exact for the
elementwise operations, within the FP32 accumulation bound for sdsdot."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import blas1x as ox

CASES = [(0, 1, 1), (1, 1, 1), (7, 1, 1), (1000, 1, 1), (1003, 1, 1), (333, 2, 3), (257, -1, 1), (129, 3, -2),
         (64, -2, -3)]


def _vec(rng, n, inc, dtype=np.float32):
    return (rng.random(max(1, 1 + (max(n, 1) - 1) * abs(inc))) * 2 - 1).astype(dtype)


@pytest.mark.parametrize("n,incx,incy", [c for c in CASES if c[0] > 0])
def test_oracle_matches_openblas(n, incx, incy):
    from scipy.linalg import blas

    rng = np.random.default_rng(abs(n + 10 * incx + 100 * incy))
    x, y = _vec(rng, n, incx), _vec(rng, n, incy)
    np.testing.assert_array_equal(ox.axpy(n, 1.75, x, incx, y, incy),
                                  blas.saxpy(x, y.copy(), n=n, a=1.75, incx=incx, incy=incy))
    np.testing.assert_array_equal(ox.copy(n, x, incx, y, incy), blas.scopy(x, y.copy(), n=n, incx=incx, incy=incy))
    sx, sy = blas.sswap(x.copy(), y.copy(), n=n, incx=incx, incy=incy)
    ox_x, ox_y = ox.swap(n, x, incx, y, incy)
    np.testing.assert_array_equal(ox_x, sx)
    np.testing.assert_array_equal(ox_y, sy)
    if incx > 0:  # scipy's scal takes positive increments only
        np.testing.assert_array_equal(ox.scal(n, -0.5, x, incx), blas.sscal(-0.5, x.copy(), n=n, incx=incx))


def test_oracle_closed_forms():
    x = np.array([np.nan, np.inf, -2.0, 3.0], np.float32)
    assert np.array_equal(ox.scal(4, 0.0, x, 1), np.zeros(4, np.float32))  # zero, not NaN
    assert np.array_equal(ox.scal(2, 0.0, x, 2), np.array([0, np.inf, 0, 3], np.float32))
    assert ox.sdsdot_f64(0, 2.5, x, 1, x, 1, rs0=7.0) == 2.5  # n == 0 overwrites with sb
    a = np.array([1, 2, 3], np.float32)
    assert ox.sdsdot_f64(3, 0.5, a, 1, a, 1, rs0=1.0) == 0.5 + 1.0 + 14.0
    assert ox.sdsdot_f64(3, 0.5, a, -1, a, 1) == 0.5 + 3 * 1 + 2 * 2 + 1 * 3
    # a planted mistake (alpha applied to y) fails the OpenBLAS comparison
    from scipy.linalg import blas

    y = np.array([1, 1, 1], np.float32)
    assert not np.array_equal(a + 2 * y, blas.saxpy(a, y.copy(), a=2.0))


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("n,incx,incy", CASES + [(1 << 20, 1, 1), ((1 << 20) + 5, 1, 1)])
def test_gpu_stream_ops_bit_exact(handle, torch_cuda, dtype, n, incx, incy):
    import portblas_amd as pb

    torch = torch_cuda
    npdt = np.dtype(dtype)
    rng = np.random.default_rng(abs(n % 997 + incx + 7 * incy))
    x, y = _vec(rng, n, incx, npdt), _vec(rng, n, incy, npdt)
    tx, ty = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    pb._axpy(handle, n, -1.25, tx, incx, ty, incy)
    handle.wait()
    np.testing.assert_array_equal(ty.cpu().numpy(), ox.axpy(n, -1.25, x, incx, y, incy))
    ty = torch.from_numpy(y).cuda()
    pb._copy(handle, n, tx, incx, ty, incy)
    handle.wait()
    np.testing.assert_array_equal(ty.cpu().numpy(), ox.copy(n, x, incx, y, incy))
    tx, ty = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    pb._swap(handle, n, tx, incx, ty, incy)
    handle.wait()
    ex, ey = ox.swap(n, x, incx, y, incy)
    np.testing.assert_array_equal(tx.cpu().numpy(), ex)
    np.testing.assert_array_equal(ty.cpu().numpy(), ey)
    for alpha in (0.375, 0.0):
        tx = torch.from_numpy(x).cuda()
        pb._scal(handle, n, alpha, tx, incx)
        handle.wait()
        np.testing.assert_array_equal(tx.cpu().numpy(), ox.scal(n, alpha, x, incx))


@pytest.mark.gpu
def test_gpu_scal_zero_clears_nan_and_inf(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    x = torch.tensor([float("nan"), float("inf"), -1.0, 2.0, float("nan")], device="cuda")
    pb._scal(handle, 5, 0.0, x, 1)
    handle.wait()
    assert torch.equal(x, torch.zeros(5, device="cuda"))


@pytest.mark.gpu
@pytest.mark.parametrize("n,incx,incy", [(0, 1, 1), (-3, 1, 1), (1, 1, 1), (1000, 1, 1), (1 << 22, 1, 1),
                                         (4099, 2, -3)])
def test_gpu_sdsdot(handle, torch_cuda, n, incx, incy):
    import portblas_amd as pb

    torch = torch_cuda
    rng = np.random.default_rng(abs(n) % 1000)
    m = max(n, 1)
    x, y = _vec(rng, m, incx), _vec(rng, m, incy)
    tx, ty = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    rs = torch.tensor([0.75], device="cuda")
    pb._sdsdot(handle, n, 1.5, tx, incx, ty, incy, rs)
    handle.wait()
    want = ox.sdsdot_f64(n, 1.5, x, incx, y, incy, rs0=0.75)
    absd = 1.5 + 0.75 + (float(np.abs(x[ox.view_index(n, incx)]).astype(np.float64)
                               @ np.abs(y[ox.view_index(n, incy)]).astype(np.float64)) if n > 0 else 0.0)
    tol = (max(n, 0) + 3) * 2.0 ** -24 * absd
    assert abs(rs.item() - want) <= tol, (rs.item(), want, tol)
    # the synchronous form starts from zero
    if n > 0:
        got = pb._sdsdot(handle, n, 1.5, tx, incx, ty, incy)
        assert abs(got - ox.sdsdot_f64(n, 1.5, x, incx, y, incy)) <= tol

"""GEMV and extension::_reduction on the GPU vs the oracle and golden fixtures.
Sweeps follow test/unittest/blas2/blas2_gemv_test.cpp:139-151 and
test/unittest/extension/reduction_test.cpp:43-53."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import U32, U64, colmajor, to_dev

pytestmark = pytest.mark.gpu


def _gemv_check(handle, torch, tr, m, n, alpha, beta, lda, incx, incy, seed, dtype=np.float32,
                poison_y=False, offset=0):
    import portblas_amd as pb

    xl, yl = (n, m) if tr == "n" else (m, n)
    A = oracle.uniform(seed, offset + lda * n).astype(dtype)
    x = oracle.uniform(seed + 1, offset + 1 + (xl - 1) * abs(incx)).astype(dtype)
    y = oracle.uniform(seed + 2, offset + 1 + (yl - 1) * abs(incy)).astype(dtype)
    if poison_y:
        y[:] = np.nan
    dA, dx, dy = to_dev(torch, A), to_dev(torch, x), to_dev(torch, y)
    pb._gemv(handle, tr, m, n, alpha, dA[offset:], lda, dx[offset:], incx, beta, dy[offset:], incy)
    got = dy.cpu().numpy().astype(np.float64)[offset:]
    ref = oracle.gemv_f64(tr, m, n, alpha, A[offset:], lda, x[offset:], incx, beta,
                          np.nan_to_num(y[offset:]) if poison_y else y[offset:], incy)
    u = U32 if dtype == np.float32 else U64
    Am = np.abs(colmajor(A[offset:], m, n, lda).astype(np.float64))
    xs = np.abs(x[offset:][:: abs(incx)][:xl].astype(np.float64))
    if incx < 0:
        xs = xs[::-1]  # logical order (negative increments walk backwards)
    scale = Am @ xs if tr == "n" else Am.T @ xs
    if incy < 0:
        scale = scale[::-1]
    tol = 1.01 * (xl + 2) * u * abs(alpha) * scale
    yy = np.abs(np.nan_to_num(y[offset:][:: abs(incy)][:yl].astype(np.float64)))
    tol = tol + 2 * u * abs(beta) * yy + 1e-30
    gv = got[:: abs(incy)][:yl]
    rv = ref[:: abs(incy)][:yl]
    assert np.all(np.isfinite(gv))
    assert np.all(np.abs(gv - rv) <= tol), np.max(np.abs(gv - rv) / tol)
    # Elements between the strided y entries are untouched.
    if abs(incy) > 1:
        mask = np.ones(got.size, bool)
        mask[:: abs(incy)] = False
        np.testing.assert_array_equal(got[mask], np.nan_to_num(y[offset:])[mask] if not poison_y else got[mask])


@pytest.mark.parametrize("ci", range(4))
def test_gemv_golden(handle, torch_cuda, golden, ci):
    import portblas_amd as pb

    torch = torch_cuda
    p = f"gemv{ci}_"
    tr, m, n, lda, incx, incy = (int(v) for v in golden[p + "args"])
    tr = chr(tr)
    alpha, beta = (float(v) for v in golden[p + "scal"])
    sa, sx, sy = (int(v) for v in golden[p + "seeds"])
    xl, yl = (n, m) if tr == "n" else (m, n)
    A = oracle.uniform(sa, lda * n)
    x = oracle.uniform(sx, 1 + (xl - 1) * incx)
    y = oracle.uniform(sy, 1 + (yl - 1) * incy)
    dy = to_dev(torch, y)
    pb._gemv(handle, tr, m, n, alpha, to_dev(torch, A), lda, to_dev(torch, x), incx, beta, dy, incy)
    got = dy.cpu().numpy().astype(np.float64)
    cb = golden[p + "out"].astype(np.float64)
    Am = np.abs(colmajor(A, m, n, lda).astype(np.float64))
    xs = np.abs(x[::incx].astype(np.float64))
    scale = Am @ xs if tr == "n" else Am.T @ xs
    tol = 2.02 * (xl + 2) * U32 * abs(alpha) * scale + 4 * U32 * abs(beta) * np.abs(y[::incy]) + 1e-30
    assert np.all(np.abs(got[::incy] - cb[::incy]) <= tol)


@pytest.mark.parametrize("tr", ["n", "t"])
@pytest.mark.parametrize("m,n", [(11, 14), (1023, 1010), (5, 3000), (3000, 5), (257, 2049)])
@pytest.mark.parametrize("incx,incy,lda_mul", [(1, 1, 1), (2, 3, 2), (-1, 1, 1), (1, -2, 1)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (1.5, 1.5)])
def test_gemv_sweep(handle, torch_cuda, tr, m, n, incx, incy, lda_mul, alpha, beta):
    _gemv_check(handle, torch_cuda, tr, m, n, alpha, beta, m * lda_mul, incx, incy,
                seed=m * 31 + n, poison_y=False)


@pytest.mark.parametrize("tr", ["n", "t"])
def test_gemv_beta_zero_never_reads_y(handle, torch_cuda, tr):
    _gemv_check(handle, torch_cuda, tr, 300, 500, 2.0, 0.0, 304, 1, 1, seed=5, poison_y=True)


@pytest.mark.parametrize("tr", ["n", "t"])
def test_gemv_unaligned_and_double(handle, torch_cuda, tr):
    _gemv_check(handle, torch_cuda, tr, 301, 259, 1.0, 0.5, 301, 1, 1, seed=9, offset=1)
    _gemv_check(handle, torch_cuda, tr, 1000, 777, -1.0, 0.25, 1004, 1, 1, seed=11, dtype=np.float64)


def test_gemv_t_row_chunks(handle, torch_cuda):
    """'t' with M around and beyond a few 2048-row chunks (partial last
    chunk, several chunks folded), FP32 and FP64."""
    for m, dt in ((8192, np.float32), (8195, np.float32), (20000, np.float32), (4096, np.float64),
                  (5000, np.float64)):
        _gemv_check(handle, torch_cuda, "t", m, 130, 1.25, 0.5, m, 1, 1, seed=m, dtype=dt)


@pytest.mark.parametrize("m,n,lda,incx,incy,alpha,beta,offset,dt", [
    (5000, 4100, 5000, 1, 1, 1.25, 0.5, 0, np.float32),   # partial last row chunk and column group
    (8192, 4096, 8200, 2, -1, -1.0, 0.0, 0, np.float32),  # padded lda, strided x, reversed y
    (6145, 4099, 6147, 1, 3, 1.0, 2.0, 1, np.float32),    # unaligned A and x: the scalar path
    (4100, 4104, 4100, 1, 1, 0.5, 0.25, 0, np.float64),
])
def test_gemv_t_wide(handle, torch_cuda, m, n, lda, incx, incy, alpha, beta, offset, dt):
    """'t' with at least 4096 columns and several 2048-row chunks: each
    workgroup walks all row chunks of its columns (no partials, no fold
    launch); every output entry against the FP64 oracle."""
    _gemv_check(handle, torch_cuda, "t", m, n, alpha, beta, lda, incx, incy, seed=m + n, offset=offset, dtype=dt)


@pytest.mark.parametrize("m,n,lda,incx,incy,alpha,beta,offset,dt", [
    (8200, 513, 8200, 1, 1, 1.25, 0.5, 0, np.float32),    # partial last 32-row block
    (8192, 700, 8196, 3, -2, -1.0, 0.0, 0, np.float32),   # padded lda, strided x, reversed y
    (9001, 301, 9003, 1, 1, 1.0, 2.0, 1, np.float32),     # unaligned A: the scalar path
    (4100, 260, 4100, -1, 1, 0.5, 0.25, 0, np.float64),   # FP64: 16-row blocks
])
def test_gemv_n_row_blocks(handle, torch_cuda, m, n, lda, incx, incy, alpha, beta, offset, dt):
    """'n' on tall matrices with a partial last row tile, padded lda, strided
    and reversed vectors, unaligned data and FP64; every output entry against
    the FP64 oracle."""
    _gemv_check(handle, torch_cuda, "n", m, n, alpha, beta, lda, incx, incy, seed=m + n, offset=offset, dtype=dt)


@pytest.mark.parametrize("m,n,lda,incx,incy,alpha,beta,poison", [
    (8224, 600, 8224, 1, 1, 1.5, 0.5, False),    # a partial last group of tiles, beta reads y
    (8192, 1000, 8200, 2, -1, -1.0, 0.0, True),  # padded lda, strided x, reversed y, y never read
    (16384, 37, 16384, -3, 3, 1.0, 2.0, False),  # fewer columns than waves in a workgroup
])
def test_gemv_n_tall(handle, torch_cuda, m, n, lda, incx, incy, alpha, beta, poison):
    """Tall FP32 'n' problems (32 row tiles per CU and more, few or many
    columns): vs the FP64 oracle with increments, padding, alpha/beta and the
    beta == 0 no-read rule."""
    _gemv_check(handle, torch_cuda, "n", m, n, alpha, beta, lda, incx, incy, seed=m + n, poison_y=poison)


@pytest.mark.parametrize("tr", ["n", "t"])
def test_gemv_full_size_rows(handle, torch_cuda, tr):
    """The benchmark's 8192 x 8192 SGEMV, alpha 1 beta 0; 64 output entries are
    checked against the FP64 oracle (the remaining ones by linearity)."""
    import portblas_amd as pb

    torch = torch_cuda
    M = N = 8192
    g = torch.Generator(device="cuda").manual_seed(3)
    A = torch.rand(M * N, device="cuda", generator=g) * 2 - 1
    x = torch.rand(N, device="cuda", generator=g) * 2 - 1
    y = torch.empty(M, device="cuda")
    pb._gemv(handle, tr, M, N, 1.0, A, M, x, 1, 0.0, y, 1)
    # linearity: op(A)(2x) = 2 op(A) x exactly (power-of-two scaling)
    y2 = torch.empty(M, device="cuda")
    pb._gemv(handle, tr, M, N, 1.0, A, M, 2 * x, 1, 0.0, y2, 1)
    assert torch.equal(y2, 2 * y)
    rows = np.random.default_rng(0).choice(M, 64, replace=False)
    Am = A.view(N, M).t()  # column-major view: Am[i, j] = A[i + j*M]
    sub = (Am[rows, :] if tr == "n" else Am[:, rows].t()).double().cpu().numpy()
    xs = x.double().cpu().numpy()
    ref = sub @ xs
    tol = 1.01 * (N + 2) * U32 * (np.abs(sub) @ np.abs(xs))
    assert np.all(np.abs(y.cpu().numpy()[rows] - ref) <= tol)


OPS = {"add": 0, "product": 1, "max": 2, "min": 3, "abs_add": 4, "mean": 5}


def _red_tol(op, A, dim, u):
    absM = np.abs(A)
    n = A.shape[0] if dim == "inner" else A.shape[1]
    axis = 0 if dim == "inner" else 1
    if op in ("max", "min"):
        return 0.0
    if op == "product":
        return 1.01 * (n + 2) * u * np.abs(np.prod(A, axis=axis))
    s = absM.sum(axis=axis)
    if op == "mean":
        s = s / n
    return 1.01 * (n + 2) * u * s + 1e-30


@pytest.mark.parametrize("ci", range(3))
@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("dim", ["inner", "outer"])
def test_reduction_golden(handle, torch_cuda, golden, ci, op, dim):
    import portblas_amd as pb

    torch = torch_cuda
    rows, cols, ld = (int(v) for v in golden[f"red{ci}_args"])
    A = oracle.uniform(int(golden[f"red{ci}_seed"]), ld * cols, lo=-2.0, hi=2.0)
    out = torch.full((cols if dim == "inner" else rows,), float("nan"), device="cuda")
    pb.extension._reduction(OPS[op], handle, to_dev(torch, A), ld, out, rows, cols,
                            pb.reduction_dim_t[dim])
    got = out.cpu().numpy().astype(np.float64)
    want = golden[f"red{ci}_{op}_{dim}"]
    M = colmajor(A, rows, cols, ld).astype(np.float64)
    tol = _red_tol(op, M, dim, U32)
    assert np.all(np.abs(got - want) <= tol + 1e-37)


def test_reduction_doc_example(handle, torch_cuda, golden):
    import portblas_amd as pb

    torch = torch_cuda
    A = to_dev(torch, golden["red_doc_A"])
    for dim, key in (("inner", "red_doc_inner"), ("outer", "red_doc_outer")):
        out = torch.zeros(3, device="cuda")
        pb.extension._reduction(pb.AddOperator, handle, A, 3, out, 3, 3, pb.reduction_dim_t[dim])
        np.testing.assert_array_equal(out.cpu().numpy(), golden[key])


@pytest.mark.parametrize("rows", [1, 7, 513, 70000])
@pytest.mark.parametrize("cols", [1, 65, 8195])
@pytest.mark.parametrize("op", ["add", "max", "mean", "abs_add"])
@pytest.mark.parametrize("dim", ["inner", "outer"])
def test_reduction_sweep(handle, torch_cuda, rows, cols, op, dim):
    import portblas_amd as pb

    if rows * cols > 40_000_000:
        pytest.skip("too large for the sweep")
    torch = torch_cuda
    ld = rows * (2 if rows < 100 else 1)
    A = oracle.uniform(rows * 13 + cols, ld * cols, lo=-2.0, hi=2.0)
    n_out = cols if dim == "inner" else rows
    out = torch.full((n_out,), float("nan"), device="cuda")
    pb.extension._reduction(OPS[op], handle, to_dev(torch, A), ld, out, rows, cols, pb.reduction_dim_t[dim])
    got = out.cpu().numpy().astype(np.float64)
    want = oracle.reduction_f64(op, dim, A, ld, rows, cols)
    tol = _red_tol(op, colmajor(A, rows, cols, ld).astype(np.float64), dim, U32)
    assert np.all(np.abs(got - want) <= tol + 1e-37)


def test_reduction_double(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    rows, cols = 513, 300
    A = oracle.uniform(77, rows * cols).astype(np.float64)
    for dim in ("inner", "outer"):
        out = torch.zeros(cols if dim == "inner" else rows, dtype=torch.float64, device="cuda")
        pb.extension._reduction(pb.AddOperator, handle, to_dev(torch, A), rows, out, rows, cols,
                                pb.reduction_dim_t[dim])
        want = oracle.reduction_f64("add", dim, A, rows, rows, cols)
        tol = _red_tol("add", colmajor(A, rows, cols, rows), dim, U64)
        assert np.all(np.abs(out.cpu().numpy() - want) <= tol)

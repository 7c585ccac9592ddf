"""dot / asum / nrm2 / iamax / iamin on the GPU vs the oracle and the CBLAS
golden fixtures.  Sweeps follow test/unittest/blas1/blas1_{dot,asum,nrm2}_test.cpp
and blas1_iaminmax_common.hpp; full-size cases use the benchmark's n = 2^26."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import U32, U64, to_dev

pytestmark = pytest.mark.gpu

INT32_MAX = 2**31 - 1


def _bound(n, scale, u=U32):
    # Any summation order: |err| <= (n + 2) u sum|terms|; the GPU's tree is far
    # shallower, so use the tighter (per-thread chain + tree) depth when large.
    depth = min(n, 4096) + 64
    return 1.01 * (depth + 2) * u * scale + 1e-30


@pytest.mark.parametrize("ci", range(4))
def test_blas1_golden(handle, torch_cuda, golden, ci):
    import portblas_amd as pb

    torch = torch_cuda
    p = f"blas1_{ci}_"
    n, incx, incy = (int(v) for v in golden[p + "args"])
    sx, sy = (int(v) for v in golden[p + "seeds"])
    x = oracle.uniform(sx, 1 + (n - 1) * incx)
    y = oracle.uniform(sy, 1 + (n - 1) * incy)
    dx, dy = to_dev(torch, x), to_dev(torch, y)
    xs, ys = x[::incx].astype(np.float64), y[::incy].astype(np.float64)
    # asynchronous API, result accumulated into rs (starts at 0.25)
    rs = torch.full((1,), 0.25, dtype=torch.float32, device="cuda")
    pb._dot(handle, n, dx, incx, dy, incy, rs)
    want = oracle.dot_f64(n, x, incx, y, incy, 0.25)
    assert abs(rs.item() - want) <= _bound(n, np.sum(np.abs(xs * ys)) + 0.25)
    assert abs(rs.item() - 0.25 - float(golden[p + "dot"])) <= 2 * _bound(n, np.sum(np.abs(xs * ys)) + 0.25)
    # synchronous API
    got = pb._dot(handle, n, dx, incx, dy, incy)
    assert abs(got - (want - 0.25)) <= _bound(n, np.sum(np.abs(xs * ys)))
    got = pb._asum(handle, n, dx, incx)
    assert abs(got - float(golden[p + "asum"])) <= 2 * _bound(n, np.sum(np.abs(xs)))
    got = pb._nrm2(handle, n, dx, incx)
    assert abs(got - float(golden[p + "nrm2"])) <= 2 * _bound(n, float(golden[p + "nrm2"]))
    assert pb._iamax(handle, n, dx, incx) == int(golden[p + "iamax"])
    assert pb._iamin(handle, n, dx, incx) == int(golden[p + "iamin"])


@pytest.mark.parametrize("n", [1, 11, 1002, 65536 + 3, 1_002_400])
@pytest.mark.parametrize("incx,incy", [(1, 1), (4, 3), (-1, 1), (-3, 2)])
def test_blas1_sweep(handle, torch_cuda, n, incx, incy):
    import portblas_amd as pb

    torch = torch_cuda
    x = oracle.uniform(n * 7 + 1, 1 + (n - 1) * abs(incx)) / n  # asum test scales by 1/n
    y = oracle.uniform(n * 7 + 2, 1 + (n - 1) * abs(incy))
    dx, dy = to_dev(torch, x), to_dev(torch, y)
    xv = x[::abs(incx)][: n].astype(np.float64)
    yv = y[::abs(incy)][: n].astype(np.float64)
    if incx < 0:
        xv = xv[::-1]
    if incy < 0:
        yv = yv[::-1]
    rs = torch.zeros(3, dtype=torch.float32, device="cuda")
    pb._dot(handle, n, dx, incx, dy, incy, rs[0:1])
    pb._asum(handle, n, dx, incx, rs[1:2])
    pb._nrm2(handle, n, dx, incx, rs[2:3])
    r = rs.cpu().numpy().astype(np.float64)
    assert abs(r[0] - oracle.dot_f64(n, x, incx, y, incy)) <= _bound(n, np.sum(np.abs(xv * yv)))
    assert abs(r[1] - oracle.asum_f64(n, x, incx)) <= _bound(n, np.sum(np.abs(xv)))
    assert abs(r[2] - oracle.nrm2_f64(n, x, incx)) <= _bound(n, oracle.nrm2_f64(n, x, incx))
    if incx > 0:
        assert pb._iamax(handle, n, dx, incx) == oracle.iamax(n, x, incx, True)
        assert pb._iamin(handle, n, dx, incx) == oracle.iamax(n, x, incx, False)
    else:
        assert pb._iamax(handle, n, dx, incx) == 0
        assert pb._iamin(handle, n, dx, incx) == 0


def test_blas1_edge_rules(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    x = torch.tensor([3.0, 4.0], device="cuda")
    rs = torch.tensor([7.0], device="cuda")
    pb._dot(handle, 0, x, 1, x, 1, rs)  # n == 0 leaves rs
    assert rs.item() == 7.0
    rs = torch.tensor([16.0], device="cuda")
    pb._nrm2(handle, 0, x, 1, rs)  # n == 0 still takes the square root
    assert rs.item() == 4.0
    rs = torch.tensor([11.0], device="cuda")
    pb._nrm2(handle, 2, x, 1, rs)  # sqrt(rs + sum x^2)
    assert rs.item() == 6.0
    idx = torch.full((1,), -5, dtype=torch.int32, device="cuda")
    pb._iamax(handle, 0, x, 1, idx)
    assert idx.item() == INT32_MAX
    pb._iamax(handle, 2, x, -1, idx)
    assert idx.item() == 0
    pb._iamin(handle, -3, x, 1, idx)
    assert idx.item() == 0
    nan = torch.tensor([float("nan")] * 5, device="cuda")
    assert pb._iamax(handle, 5, nan, 1) == INT32_MAX
    assert pb._iamin(handle, 5, nan, 1) == INT32_MAX
    inf = torch.tensor([float("inf")] * 3, device="cuda")
    assert pb._iamin(handle, 3, inf, 1) == INT32_MAX  # inf never beats the FLT_MAX start
    z = torch.zeros(9, device="cuda")
    assert pb._iamax(handle, 9, z, 1) == 0 and pb._iamin(handle, 9, z, 1) == 0
    i64 = torch.zeros(1, dtype=torch.int64, device="cuda")
    pb._iamax(handle, 0, x, 1, i64)
    assert i64.item() == 2**63 - 1


@pytest.mark.parametrize("name", ["limit", "incrementing", "decrementing"])
@pytest.mark.parametrize("n", [11, 65, 1000])
def test_iaminmax_patterns(handle, torch_cuda, golden, name, n):
    import portblas_amd as pb

    torch = torch_cuda
    x = to_dev(torch, golden[f"pat_{name}_{n}_x"])
    xm = to_dev(torch, golden[f"pat_{name}_{n}_xmin"])
    assert pb._iamax(handle, n, x, 1) == int(golden[f"pat_{name}_{n}_iamax"])
    assert pb._iamin(handle, n, xm, 1) == int(golden[f"pat_{name}_{n}_iamin"])


def test_blas1_double(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    n = 100_003
    x = oracle.uniform(3, n).astype(np.float64) * np.pi
    y = oracle.uniform(4, n).astype(np.float64)
    dx, dy = to_dev(torch, x), to_dev(torch, y)
    rs = torch.zeros(3, dtype=torch.float64, device="cuda")
    pb._dot(handle, n, dx, 1, dy, 1, rs[0:1])
    pb._asum(handle, n, dx, 1, rs[1:2])
    pb._nrm2(handle, n, dx, 1, rs[2:3])
    r = rs.cpu().numpy()
    assert abs(r[0] - oracle.dot_f64(n, x, 1, y, 1)) <= _bound(n, np.sum(np.abs(x * y)), U64)
    assert abs(r[1] - oracle.asum_f64(n, x, 1)) <= _bound(n, np.sum(np.abs(x)), U64)
    assert abs(r[2] - oracle.nrm2_f64(n, x, 1)) <= _bound(n, r[2], U64)
    assert pb._iamax(handle, n, dx, 1) == oracle.iamax(n, x, 1, True)
    assert pb._iamin(handle, n, dx, 1) == oracle.iamax(n, x, 1, False)


def test_iamax_full_size_ties_and_nan(handle, torch_cuda):
    """n = 2^26 with 16 exact-tie maxima (+-1.0 above every other value) at
    random positions: the lowest one wins; a NaN before the first tie is
    ignored (portBLAS semantics: NaN never wins)."""
    import portblas_amd as pb

    torch = torch_cuda
    n = 1 << 26
    g = torch.Generator(device="cuda").manual_seed(0)
    x = (torch.rand(n, device="cuda", generator=g) * 2 - 1) * 0.999
    rng = np.random.default_rng(2026)
    pos = np.sort(rng.choice(n, size=16, replace=False))
    sign = np.where(rng.random(16) < 0.5, -1.0, 1.0).astype(np.float32)
    x[torch.from_numpy(pos).cuda()] = torch.from_numpy(sign).cuda()
    assert pb._iamax(handle, n, x, 1) == int(pos[0])
    nan_at = int(pos[0]) // 2
    x[nan_at] = float("nan")
    assert pb._iamax(handle, n, x, 1) == int(pos[0])
    # iamin: plant 16 exact zeros; the first wins.
    x[torch.from_numpy(pos).cuda()] = 0.0
    assert pb._iamin(handle, n, x, 1) == int(pos[0])
    # strided view: every other element
    y = torch.ones(2 * 1000, device="cuda")
    y[2 * 777] = 5.0
    y[2 * 778] = -5.0
    assert pb._iamax(handle, 1000, y, 2) == 777


def test_reductions_full_size_properties(handle, torch_cuda):
    """n = 2^26: asum(x) == asum(|x|) bit-for-bit (same kernel order),
    nrm2(x)^2 vs dot(x, x), and both vs an FP64 sum."""
    import portblas_amd as pb

    torch = torch_cuda
    n = 1 << 26
    g = torch.Generator(device="cuda").manual_seed(1)
    x = torch.rand(n, device="cuda", generator=g) * 2 - 1
    a1 = pb._asum(handle, n, x, 1)
    a2 = pb._asum(handle, n, x.abs(), 1)
    assert a1 == a2
    a3 = pb._asum(handle, n, x, 1)
    assert a1 == a3  # deterministic
    ref = x.double().abs().sum().item()
    assert abs(a1 - ref) <= 1e-5 * ref
    d = pb._dot(handle, n, x, 1, x, 1)
    nr = pb._nrm2(handle, n, x, 1)
    assert abs(nr * nr - d) <= 1e-5 * d
    refd = (x.double() * x.double()).sum().item()
    assert abs(d - refd) <= 1e-5 * refd

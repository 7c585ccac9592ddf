"""Grouped (uneven) GEMM in one persistent launch, and hipGraph capture of
library calls.  Grouped problems have no reference counterpart (portBLAS
batches share one shape); each problem is checked against an FP64 product
of its own inputs with the FP32-accumulation bound (BF16 inputs are exact in
FP32).  Graph replays are checked against the eager result after every
replay, with eager work and synchronisation in between."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


class Prob(ctypes.Structure):
    _fields_ = [("transa", ctypes.c_char), ("transb", ctypes.c_char), ("m", ctypes.c_int64),
                ("n", ctypes.c_int64), ("k", ctypes.c_int64), ("alpha", ctypes.c_float),
                ("beta", ctypes.c_float), ("a", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("b", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("c", ctypes.c_void_p),
                ("ldc", ctypes.c_int64)]


def cm(flat, rows, cols, ld):
    return flat[: cols * ld].view(cols, ld).t()[:rows, :]


def _make_problems(torch, dims, dtype, seed, pad=0, beta_zero=False):
    g = torch.Generator(device="cuda").manual_seed(seed)
    rng = np.random.default_rng(seed)
    bufs, probs = [], (Prob * len(dims))()
    for i, (m, n, k) in enumerate(dims):
        ta = "n" if rng.random() < 0.5 else "t"
        tb = "n" if rng.random() < 0.5 else "t"
        ar, ac = (m, k) if ta == "n" else (k, m)
        br, bc = (k, n) if tb == "n" else (n, k)
        lda, ldb, ldc = ar + pad, br + pad, m + pad
        a = (torch.rand(lda * ac, generator=g, device="cuda") * 2 - 1).to(dtype)
        b = (torch.rand(ldb * bc, generator=g, device="cuda") * 2 - 1).to(dtype)
        c = torch.rand(ldc * n, generator=g, device="cuda") * 2 - 1
        alpha, beta = float(rng.choice([1.0, -0.5])), float(rng.choice([0.0, 0.75]))
        beta = 0.0 if beta_zero else beta
        probs[i] = Prob(ta.encode(), tb.encode(), m, n, k, alpha, beta, a.data_ptr(), lda, b.data_ptr(), ldb,
                        c.data_ptr(), ldc)
        bufs.append((ta, tb, m, n, k, alpha, beta, a, lda, b, ldb, c, c.clone(), ldc))
    return probs, bufs


def _check(bufs):
    for ta, tb, m, n, k, alpha, beta, a, lda, b, ldb, c, c0, ldc in bufs:
        ar, ac = (m, k) if ta == "n" else (k, m)
        br, bc = (k, n) if tb == "n" else (n, k)
        oa = cm(a.float(), ar, ac, lda).double()
        ob = cm(b.float(), br, bc, ldb).double()
        oa = oa if ta == "n" else oa.t()
        ob = ob if tb == "n" else ob.t()
        c0d = cm(c0, m, n, ldc).double()
        ref = alpha * (oa @ ob) + beta * c0d
        tol = 1.01 * (k + 2) * U * abs(alpha) * (oa.abs() @ ob.abs()) + 2 * U * abs(beta) * c0d.abs() + 1e-30
        err = (cm(c, m, n, ldc).double() - ref).abs()
        assert (err <= tol).all(), (ta, tb, m, n, k, (err / tol).max().item())
        if ldc > m:  # padding rows untouched
            assert ldc * n == c.numel()
            assert (c.view(n, ldc)[:, m:] == c0.view(n, ldc)[:, m:]).all()


@pytest.mark.parametrize("prec,dtype_name", [("bf16", "bfloat16"), ("fp32_mfma", "float32")])
def test_grouped_uneven_problems(handle, torch_cuda, prec, dtype_name):
    import portblas_amd as pb

    torch = torch_cuda
    rng = np.random.default_rng(7)
    dims = [tuple(int(v) for v in np.maximum(1, np.round(np.exp(rng.uniform(np.log(1), np.log(700), 3)))))
            for _ in range(60)]
    dims += [(64, 64, 64), (1, 1, 1), (513, 257, 129), (8, 1000, 3)]
    probs, bufs = _make_problems(torch, dims, getattr(torch, dtype_name), 11, pad=3)
    L = pb.load()
    pb.check(L.pb_gemm_grouped(handle.handle, int(pb.gemm_precision_t[prec]), probs, len(dims)))
    torch.cuda.synchronize()
    _check(bufs)


def test_grouped_ring_kernel_shapes(handle, torch_cuda):
    """Problems whose M, N, K are multiples of 8 with aligned operands go to
    the ring-pipelined kernel (4 K chunks in flight per workgroup), the others
    in the same plan to the lean kernel: both checked per problem, every
    layout, alpha and beta rules, tile edges, K = 8 and padded leading
    dimensions (ldc padding untouched)."""
    import portblas_amd as pb

    torch = torch_cuda
    rng = np.random.default_rng(21)
    dims = [tuple(int(v) for v in np.maximum(8, np.round(np.exp(rng.uniform(np.log(8), np.log(700), 3)) / 8) * 8))
            for _ in range(48)]
    dims += [(8, 8, 8), (64, 64, 8), (72, 136, 264), (1024, 32, 1024), (8, 1000, 16), (520, 8, 40)]
    probs, bufs = _make_problems(torch, dims, torch.bfloat16, 23, pad=8)
    odd = [(77, 130, 41), (64, 64, 63), (9, 9, 9)]
    probs2, bufs2 = _make_problems(torch, odd, torch.bfloat16, 24, pad=1)
    allp = (Prob * (len(dims) + len(odd)))(*(list(probs) + list(probs2)))
    L = pb.load()
    pb.check(L.pb_gemm_grouped(handle.handle, int(pb.gemm_precision_t.bf16), allp, len(allp)))
    torch.cuda.synchronize()
    _check(bufs)
    _check(bufs2)


def test_grouped_benchmark_set_full_size(handle, torch_cuda):
    """The benchmark's 512-problem set (M, N, K log-uniform in [32, 1024],
    multiples of 8, BF16 in, FP32 out, alpha 1, beta 0), every problem checked."""
    import portblas_amd as pb

    torch = torch_cuda
    rng = np.random.default_rng(4)
    dims = np.exp(rng.uniform(np.log(32), np.log(1024), size=(512, 3)))
    dims = [tuple(int(v) for v in row) for row in np.maximum(8, np.round(dims / 8) * 8).astype(np.int64)]
    g = torch.Generator(device="cuda").manual_seed(9)
    bufs, probs = [], (Prob * len(dims))()
    for i, (m, n, k) in enumerate(dims):
        a = (torch.rand(m * k, generator=g, device="cuda") * 2 - 1).to(torch.bfloat16)
        b = (torch.rand(k * n, generator=g, device="cuda") * 2 - 1).to(torch.bfloat16)
        c = torch.full((m * n,), float("nan"), device="cuda")
        probs[i] = Prob(b"n", b"n", m, n, k, 1.0, 0.0, a.data_ptr(), m, b.data_ptr(), k, c.data_ptr(), m)
        bufs.append(("n", "n", m, n, k, 1.0, 0.0, a, m, b, k, c, torch.zeros_like(c), m))
    L = pb.load()
    pb.check(L.pb_gemm_grouped(handle.handle, int(pb.gemm_precision_t.bf16), probs, len(dims)))
    torch.cuda.synchronize()
    _check(bufs)


def test_grouped_plan_reuse_and_graph_replay(handle, torch_cuda):
    """A plan runs many times; captured in a graph it replays to the eager
    result each time (the queue resets itself, no memset node)."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    rng = np.random.default_rng(3)
    dims = [tuple(int(v) for v in np.maximum(8, np.round(np.exp(rng.uniform(np.log(32), np.log(600), 3)) / 8) * 8))
            for _ in range(40)]
    probs, bufs = _make_problems(torch, dims, torch.bfloat16, 5, beta_zero=True)
    plan = ctypes.c_void_p()
    pb.check(L.pb_gemm_grouped_plan_create(handle.handle, int(pb.gemm_precision_t.bf16), probs, len(dims),
                                           ctypes.byref(plan)))
    try:
        pb.check(L.pb_gemm_grouped_plan_run(handle.handle, plan))
        torch.cuda.synchronize()
        eager = [b[11].clone() for b in bufs]
        pb.check(L.pb_gemm_grouped_plan_run(handle.handle, plan))  # second run, same result
        torch.cuda.synchronize()
        for e, b in zip(eager, bufs):
            assert torch.equal(e, b[11])
        s = torch.cuda.Stream()
        hs = pb.SB_Handle(s)
        graph = ctypes.c_void_p()
        with torch.cuda.stream(s):
            pb.check(L.pb_graph_capture_begin(hs.handle))
            pb.check(L.pb_gemm_grouped_plan_run(hs.handle, plan))
            pb.check(L.pb_graph_capture_end(hs.handle, ctypes.byref(graph)))
        try:
            for rep in range(4):
                for b in bufs:
                    b[11].fill_(float("nan"))
                torch.cuda.synchronize()
                pb.check(L.pb_graph_launch(hs.handle, graph))
                hs.wait()
                x = torch.ones(1 << 20, device="cuda").sum()  # eager work between replays
                torch.cuda.synchronize()
                assert x.item() == float(1 << 20)
                for e, b in zip(eager, bufs):
                    assert torch.equal(e, b[11]), f"replay {rep} differs"
        finally:
            L.pb_graph_destroy(graph)
    finally:
        L.pb_gemm_grouped_plan_destroy(plan)
    _check(bufs)


def test_graph_of_blas_calls(handle, torch_cuda):
    """GEMM + GEMV + dot captured together replay to the eager results."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    s = torch.cuda.Stream()
    hs = pb.SB_Handle(s)
    n = 512
    A = torch.rand(n * n, device="cuda") * 2 - 1
    B = torch.rand(n * n, device="cuda") * 2 - 1
    x = torch.rand(n, device="cuda") * 2 - 1
    C = torch.zeros(n * n, device="cuda")
    y = torch.zeros(n, device="cuda")
    rs = torch.zeros(1, device="cuda")

    def work(h):
        pb._gemm(h, "n", "t", n, n, n, 1.0, A, n, B, n, 0.0, C, n)
        pb._gemv(h, "n", n, n, 1.0, C, n, x, 1, 0.0, y, 1)
        pb._dot(h, n, y, 1, y, 1, rs)

    with torch.cuda.stream(s):
        rs.zero_()
        work(hs)
    hs.wait()
    C_e, y_e, rs_e = C.clone(), y.clone(), rs.clone()
    hs.reserve_workspace(1 << 24)
    graph = ctypes.c_void_p()
    with torch.cuda.stream(s):
        pb.check(L.pb_graph_capture_begin(hs.handle))
        work(hs)
        pb.check(L.pb_graph_capture_end(hs.handle, ctypes.byref(graph)))
    try:
        for rep in range(3):
            C.zero_(); y.zero_(); rs.zero_()
            torch.cuda.synchronize()
            pb.check(L.pb_graph_launch(hs.handle, graph))
            hs.wait()
            assert torch.equal(C, C_e) and torch.equal(y, y_e) and torch.equal(rs, rs_e), rep
    finally:
        L.pb_graph_destroy(graph)


def test_grouped_schedule_options_and_autotuner(handle, torch_cuda):
    """Every schedule the grouped autotuner times (ring or lean kernel, CU-
    position weights, XCD spread threshold, workgroups per CU) gives C within
    the FP32 bound of the FP64 product, identical across schedules (each tile
    accumulates its K chunks in the same order whatever workgroup runs it);
    the tuner returns one of its candidates, refuses beta != 0 problems, and
    the library refuses bad option values."""
    import portblas_amd as pb
    from portblas_amd import tuning

    torch = torch_cuda
    rng = np.random.default_rng(11)
    dims = [tuple(int(v) for v in np.maximum(8, np.round(np.exp(rng.uniform(np.log(24), np.log(700), 3)) / 8) * 8))
            for _ in range(40)]
    dims += [(37, 45, 53), (100, 3, 77)]  # not multiples of 8: the lean kernel
    probs, bufs = _make_problems(torch, dims, torch.bfloat16, seed=21, beta_zero=True)
    tuner = tuning.GroupedAutotuner(handle, pb.gemm_precision_t.bf16)
    L = pb.load()
    outs = []
    for opts in tuning.grouped_candidates():
        plan = tuner.plan(probs, len(dims), opts)
        pb.check(L.pb_gemm_grouped_plan_run(handle.handle, plan))
        torch.cuda.synchronize()
        pb.check(L.pb_gemm_grouped_plan_destroy(plan))
        _check(bufs)
        outs.append([b[11].clone() for b in bufs])
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert torch.equal(x, y)
    best, table = tuner.tune(probs, len(dims), reps=2, warm=1)
    assert best in table and set(table) == set(tuning.grouped_candidates())
    assert all(t > 0 for t in table.values())
    probs[0].beta = 0.5
    with pytest.raises(ValueError):
        tuner.tune(probs, len(dims))
    probs[0].beta = 0.0
    bad = tuning.GroupedOptions(3, 0, 0, 0)
    plan = ctypes.c_void_p()
    st = L.pb_gemm_grouped_plan_create_ex(handle.handle, int(pb.gemm_precision_t.bf16), probs, len(dims),
                                          ctypes.byref(bad), ctypes.byref(plan))
    assert st == 1 and not plan.value  # PB_STATUS_INVALID_VALUE


def test_grouped_ring_edges(handle, torch_cuda):
    """K = 0 on both kernels (C = beta C), a plan too small to fill the XCD
    groups (one and three problems), and a ring problem whose leading
    dimension is just under the ring kernel's 2^24 limit, so that a thread's
    byte offset in its panel comes within 2 % of 2^31."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    for dims in ([(64, 64, 0), (8, 16, 0), (37, 21, 0), (128, 64, 64)], [(72, 40, 24)],
                 [(8, 8, 8), (16, 24, 40), (64, 8, 0)]):
        probs, bufs = _make_problems(torch, dims, torch.bfloat16, seed=31 + len(dims))
        pb.check(L.pb_gemm_grouped(handle.handle, int(pb.gemm_precision_t.bf16), probs, len(dims)))
        torch.cuda.synchronize()
        _check(bufs)
    # A = op(A)^T stored K-contiguous ('t'): row r of the 64-row tile starts
    # r * lda elements in; lda = 2^24 - 8 puts row 63 at 2.11e9 bytes.
    g = torch.Generator(device="cuda").manual_seed(5)
    m, n, k, lda = 64, 64, 16, (1 << 24) - 8
    a = torch.zeros((m - 1) * lda + k, dtype=torch.bfloat16, device="cuda")
    rows = torch.arange(m, device="cuda") * lda
    vals = (torch.rand(m, k, generator=g, device="cuda") * 2 - 1).to(torch.bfloat16)
    a.view(-1)[(rows[:, None] + torch.arange(k, device="cuda")[None, :]).reshape(-1)] = vals.reshape(-1)
    b = (torch.rand(k * n, generator=g, device="cuda") * 2 - 1).to(torch.bfloat16)
    c = torch.zeros(m * n, device="cuda")
    p = (Prob * 1)(Prob(b"t", b"n", m, n, k, 1.0, 0.0, a.data_ptr(), lda, b.data_ptr(), k, c.data_ptr(), m))
    pb.check(L.pb_gemm_grouped(handle.handle, int(pb.gemm_precision_t.bf16), p, 1))
    torch.cuda.synchronize()
    ref = vals.double() @ b.double().view(n, k).t()
    tol = 1.01 * (k + 2) * U * (vals.double().abs() @ b.double().view(n, k).t().abs()) + 1e-30
    got = c.view(n, m).t().double()
    assert ((got - ref).abs() <= tol).all(), ((got - ref).abs() / tol).max().item()
    del a

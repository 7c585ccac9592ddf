"""Handle workspace, graphs and stream switching on the GPU.

The reference's Temp_Mem_Pool hands out scratch under a mutex and frees it
only when no kernel can still use it (include/sb_handle/temp_memory_pool.h:
33-114).  Here a graph captured on a handle keeps the workspace its kernels
were recorded with alive, so an eager call that grows the handle's workspace
afterwards does not leave the graph addressing freed memory."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from conftest import U32

pytestmark = pytest.mark.gpu


def test_graph_replay_after_workspace_growth(torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    s = torch.cuda.Stream()
    hs = pb.SB_Handle(s)
    # split-K GEMM (partials in the workspace) and a GEMV 'n' (partials too)
    m, n, k = 64, 64, 4096
    g = torch.Generator(device="cuda").manual_seed(11)
    A = torch.rand(m * k, device="cuda", generator=g) * 2 - 1
    B = torch.rand(k * n, device="cuda", generator=g) * 2 - 1
    C = torch.zeros(m * n, device="cuda")
    M2 = 4096
    G = torch.rand(M2 * 512, device="cuda", generator=g) * 2 - 1
    x = torch.rand(512, device="cuda", generator=g) * 2 - 1
    y = torch.zeros(M2, device="cuda")

    def work(h):
        pb._gemm(h, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        pb._gemv(h, "t", 512, M2, 1.0, G, 512, x, 1, 0.0, y, 1)

    with torch.cuda.stream(s):
        work(hs)
    hs.wait()
    C_e, y_e = C.clone(), y.clone()
    before = hs.workspace_bytes()
    assert before > 0, "the captured calls use no workspace; the test would prove nothing"
    graph = ctypes.c_void_p()
    with torch.cuda.stream(s):
        pb.check(L.pb_graph_capture_begin(hs.handle))
        work(hs)
        pb.check(L.pb_graph_capture_end(hs.handle, ctypes.byref(graph)))
    try:
        # grow the handle's workspace (the old buffer is released by the handle)
        hs.reserve_workspace(4 * before + (1 << 24))
        assert hs.workspace_bytes() > before
        # an eager call on the grown workspace, then replays of the old graph
        C2 = torch.zeros(m * n, device="cuda")
        with torch.cuda.stream(s):
            pb._gemm(hs, "n", "n", m, n, k, 2.0, A, m, B, k, 0.0, C2, m)
        hs.wait()
        assert torch.equal(C2, 2 * C_e)
        for rep in range(3):
            C.fill_(float("nan"))
            y.fill_(float("nan"))
            torch.cuda.synchronize()
            pb.check(L.pb_graph_launch(hs.handle, graph))
            hs.wait()
            z = torch.ones(1 << 20, device="cuda").sum()  # eager work between replays
            torch.cuda.synchronize()
            assert z.item() == float(1 << 20)
            assert torch.equal(C, C_e) and torch.equal(y, y_e), f"replay {rep} differs"
    finally:
        pb.check(L.pb_graph_destroy(graph))
    # the handle still works after the graph (and its buffer) is gone
    with torch.cuda.stream(s):
        work(hs)
    hs.wait()
    assert torch.equal(C, C_e) and torch.equal(y, y_e)
    hs.close()


def test_capture_needing_more_workspace_is_refused(torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    s = torch.cuda.Stream()
    hs = pb.SB_Handle(s)
    m, n, k = 64, 64, 8192
    A = torch.ones(m * k, device="cuda")
    B = torch.ones(k * n, device="cuda")
    C = torch.zeros(m * n, device="cuda")
    graph = ctypes.c_void_p()
    with torch.cuda.stream(s):
        pb.check(L.pb_graph_capture_begin(hs.handle))
        with pytest.raises(RuntimeError, match="PB_STATUS_ALLOC_FAILED.*pb_handle_reserve_workspace"):
            pb._gemm(hs, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        L.pb_graph_capture_end(hs.handle, ctypes.byref(graph))
    if graph.value:
        L.pb_graph_destroy(graph)
        graph = ctypes.c_void_p()
    # after a reservation the same capture works and replays exactly
    hs.reserve_workspace(1 << 24)
    with torch.cuda.stream(s):
        pb.check(L.pb_graph_capture_begin(hs.handle))
        pb._gemm(hs, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        pb.check(L.pb_graph_capture_end(hs.handle, ctypes.byref(graph)))
    try:
        pb.check(L.pb_graph_launch(hs.handle, graph))
        hs.wait()
        assert torch.all(C == float(k))
    finally:
        L.pb_graph_destroy(graph)
    hs.close()


def test_set_stream_orders_after_old_stream(torch_cuda):
    """Work queued on the old stream finishes before the new stream's first
    call touches the shared workspace: a split-K GEMM on stream 1, a switch,
    then a split-K GEMM of other data on stream 2, both checked."""
    import portblas_amd as pb

    torch = torch_cuda
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    hs = pb.SB_Handle(s1)
    m, n, k = 64, 64, 8192
    A1 = oracle.uniform(1, m * k)
    B1 = oracle.uniform(2, k * n)
    A2 = oracle.uniform(3, m * k)
    B2 = oracle.uniform(4, k * n)
    d = [torch.from_numpy(v).cuda() for v in (A1, B1, A2, B2)]
    C1 = torch.zeros(m * n, device="cuda")
    C2 = torch.zeros(m * n, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(s1):
        pb._gemm(hs, "n", "n", m, n, k, 1.0, d[0], m, d[1], k, 0.0, C1, m)
    hs.set_stream(s2)
    with torch.cuda.stream(s2):
        pb._gemm(hs, "n", "n", m, n, k, 1.0, d[2], m, d[3], k, 0.0, C2, m)
    torch.cuda.synchronize()
    for (A, B, C) in ((A1, B1, C1), (A2, B2, C2)):
        ref = oracle.gemm_f64("n", "n", m, n, k, 1.0, A, m, B, k, 0.0, np.zeros(m * n), m)
        ap = oracle.gemm_absprod_f64("n", "n", m, n, k, A, m, B, k)
        got = C.cpu().numpy()
        assert np.all(np.abs(got - ref) <= 1.01 * (k + 2) * U32 * ap + 1e-30)
    hs.close()


def test_fp16x3_graph_replay_follows_scale_changes(torch_cuda):
    """An FP16x3 call packs with the scale exponents of the handle's previous
    FP16x3 call and packs again when they change.  Replays of a captured call
    must follow changes of the input's magnitude between replays: each replay
    equals the eager product of the current inputs, and scaling A by 2^12
    scales C by exactly 2^12."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    s = torch.cuda.Stream()
    hs = pb.SB_Handle(s)
    m = n = 4096
    k = 256
    g = torch.Generator(device="cuda").manual_seed(13)
    A = torch.rand(m * k, device="cuda", generator=g) * 2 - 1
    B = torch.rand(k * n, device="cuda", generator=g) * 2 - 1
    Ab = A.clone()
    C = torch.zeros(m * n, device="cuda")
    P = pb.gemm_precision_t.fp16x3

    def call():
        pb._gemm_ex(hs, P, "n", "n", m, n, k, 1.0, Ab, m, B, k, 0.0, C, m)

    with torch.cuda.stream(s):
        call()
    hs.wait()
    C1 = C.clone()
    Ab.mul_(4096.0)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        call()
    hs.wait()
    C2 = C.clone()
    assert torch.equal(C2, C1 * 4096.0)
    graph = ctypes.c_void_p()
    with torch.cuda.stream(s):
        pb.check(L.pb_graph_capture_begin(hs.handle))
        call()
        pb.check(L.pb_graph_capture_end(hs.handle, ctypes.byref(graph)))
    try:
        for rep, (scale, want) in enumerate([(1.0, C1), (4096.0, C2), (4096.0, C2), (1.0, C1), (1.0, C1)]):
            Ab.copy_(A * scale)
            C.fill_(float("nan"))
            torch.cuda.synchronize()
            pb.check(L.pb_graph_launch(hs.handle, graph))
            hs.wait()
            z = torch.ones(1 << 20, device="cuda").sum()  # eager work between replays
            torch.cuda.synchronize()
            assert z.item() == float(1 << 20)
            assert torch.equal(C, want), f"replay {rep} differs"
    finally:
        pb.check(L.pb_graph_destroy(graph))

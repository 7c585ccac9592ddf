"""The selector MLP: the FP64 oracle (oracle/mlp.py) pinned by finite
differences on the CPU; the HIP training / inference kernels checked against
it on the GPU; the handle's host-side dispatch checked against the GPU BF16
inference."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as omlp

SIZES = (8, 16, 12, 10, 5)


def _net(seed=0, sizes=SIZES):
    rng = np.random.default_rng(seed)
    ws = [rng.standard_normal((o, i)) * np.sqrt(2.0 / i) for i, o in zip(sizes[:-1], sizes[1:])]
    bs = [rng.standard_normal(o) * 0.1 for o in sizes[1:]]
    return ws, bs


def test_oracle_gradients_match_finite_differences():
    ws, bs = _net(1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((7, SIZES[0]))
    t = rng.standard_normal((7, SIZES[-1]))
    m = (rng.random((7, SIZES[-1])) < 0.8).astype(float)
    dws, dbs, _ = omlp.backward(ws, bs, x, t, m)
    eps = 1e-6
    for l in range(len(ws)):
        for idx in [(0, 0), (ws[l].shape[0] - 1, ws[l].shape[1] - 1), (1, 2)]:
            w0 = ws[l][idx]
            ws[l][idx] = w0 + eps
            lp = omlp.loss(ws, bs, x, t, m)
            ws[l][idx] = w0 - eps
            lm = omlp.loss(ws, bs, x, t, m)
            ws[l][idx] = w0
            assert abs((lp - lm) / (2 * eps) - dws[l][idx]) <= 1e-6 * (1 + abs(dws[l][idx]))
        for j in (0, bs[l].size - 1):
            b0 = bs[l][j]
            bs[l][j] = b0 + eps
            lp = omlp.loss(ws, bs, x, t, m)
            bs[l][j] = b0 - eps
            lm = omlp.loss(ws, bs, x, t, m)
            bs[l][j] = b0
            assert abs((lp - lm) / (2 * eps) - dbs[l][j]) <= 1e-6 * (1 + abs(dbs[l][j]))


def test_oracle_forward_closed_form():
    # one linear layer = affine map; ReLU hidden layer clamps negatives
    w = [np.array([[1.0, -2.0]]), np.array([[3.0]])]
    b = [np.array([0.5]), np.array([-1.0])]
    acts = omlp.forward(w, b, np.array([[1.0, 1.0], [0.0, -1.0]]))
    np.testing.assert_allclose(acts[1], [[0.0], [2.5]])  # relu(-0.5)=0, relu(2.5)
    np.testing.assert_allclose(acts[2], [[-1.0], [6.5]])
    assert omlp.bf16_round(np.float32(1.0 + 2 ** -9)) == np.float32(1.0)  # ties to even
    assert omlp.bf16_round(np.float32(1.0 + 3 * 2 ** -9)) == np.float32(1.0 + 2 ** -7)


@pytest.mark.gpu
def test_gpu_mlp_forward_backward_adam(handle, torch_cuda):
    import ctypes

    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    ws, bs = _net(3)
    rng = np.random.default_rng(4)
    B = 37
    x = rng.standard_normal((B, SIZES[0]))
    t = rng.standard_normal((B, SIZES[-1]))
    m = (rng.random((B, SIZES[-1])) < 0.7).astype(float)
    f32 = lambda a: torch.tensor(np.asarray(a, np.float32), device="cuda")
    W = [f32(w) for w in ws]
    Bv = [f32(b) for b in bs]
    acts = [f32(x)] + [torch.empty(B, d, device="cuda") for d in SIZES[1:]]
    dW = [torch.zeros_like(w) for w in W]
    dB = [torch.zeros_like(b) for b in Bv]
    loss = torch.zeros(1, device="cuda")
    scratch = torch.empty(2 * B * max(SIZES), device="cuda")
    sizes = (ctypes.c_int * len(SIZES))(*SIZES)
    P = lambda ts: (ctypes.c_void_p * len(ts))(*[q.data_ptr() for q in ts])
    n = len(SIZES) - 1
    pb.check(L.pb_mlp_forward(handle.handle, n, sizes, P(W), P(Bv), P(acts), B))
    td, md = f32(t), f32(m)  # held alive: a temporary's block would be reused by the next allocation
    pb.check(L.pb_mlp_backward(handle.handle, n, sizes, P(W), P(acts), td.data_ptr(), md.data_ptr(),
                               P(dW), P(dB), loss.data_ptr(), scratch.data_ptr(), B))
    handle.wait()
    # oracle on the FP32-rounded parameters
    ws32 = [np.asarray(w, np.float32).astype(np.float64) for w in ws]
    bs32 = [np.asarray(b, np.float32).astype(np.float64) for b in bs]
    x32 = np.asarray(x, np.float32).astype(np.float64)
    ref_acts = omlp.forward(ws32, bs32, x32)
    for l in range(1, len(SIZES)):
        np.testing.assert_allclose(acts[l].cpu().numpy(), ref_acts[l], rtol=1e-4, atol=1e-5)
    rdw, rdb, rloss = omlp.backward(ws32, bs32, x32, np.asarray(t, np.float32), m)
    assert abs(loss.item() - rloss) <= 1e-4 * (1 + rloss)
    for l in range(n):
        np.testing.assert_allclose(dW[l].cpu().numpy(), rdw[l], rtol=2e-4, atol=2e-5)
        np.testing.assert_allclose(dB[l].cpu().numpy(), rdb[l], rtol=2e-4, atol=2e-5)
    # one Adam step matches the textbook update
    p0 = W[0].clone()
    g = dW[0]
    mm = torch.zeros_like(p0)
    vv = torch.zeros_like(p0)
    pb.check(L.pb_mlp_adam(handle.handle, W[0].data_ptr(), g.data_ptr(), mm.data_ptr(), vv.data_ptr(),
                           W[0].numel(), 1e-2, 0.9, 0.999, 1e-8, 1))
    handle.wait()
    gn = g.cpu().numpy().astype(np.float64)
    m1, v1 = 0.1 * gn, 0.001 * gn * gn
    want = p0.cpu().numpy() - 1e-2 * (m1 / 0.1) / (np.sqrt(v1 / 0.001) + 1e-8)
    np.testing.assert_allclose(W[0].cpu().numpy(), want, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_gpu_bf16_inference_and_host_dispatch_agree(handle, torch_cuda):
    from portblas_amd import tuning

    sel = tuning.Selector.init(sizes=(8, 64, 64, 64, tuning.num_configs()), seed=5)
    shapes = tuning.sample_shapes(300, seed=9)
    X = np.stack([tuning.features(*s) for s in shapes])
    sel.mean = X.mean(axis=0).astype(np.float32)
    sel.std = np.maximum(X.std(axis=0), 1e-3).astype(np.float32)
    pred = sel.predict_bf16(handle, X)
    ref = omlp.forward_bf16(sel.ws, sel.bs, sel.normalize(X))
    np.testing.assert_allclose(pred, ref, rtol=2e-3, atol=2e-3)
    sel.install(handle)
    try:
        host = np.array([tuning.select(handle, *s) for s in shapes])
        gpu = np.argmin(pred, axis=1)
        # identical arithmetic; a near-tie may still flip, so demand 99%
        assert np.mean(host == gpu) >= 0.99
    finally:
        tuning.uninstall(handle)


@pytest.mark.gpu
def test_gpu_autotune_pipeline_small(handle, torch_cuda):
    """End to end on a few shapes: exhaustive timing, GPU training, BF16
    evaluation; the trained network must pick configurations no worse than
    a random choice on its own training set."""
    from portblas_amd import tuning

    shapes = tuning.sample_shapes(24, seed=3, max_flop=2.0 ** 30, max_elems=2 ** 22)
    tuner = tuning.Autotuner(handle, max_elems=2 ** 22)
    X, T = tuning.collect(tuner, shapes)
    assert np.all(np.isfinite(T)) and np.all(T > 0)
    sel = tuning.Selector.init(sizes=(8, 64, 64, 64, T.shape[1]), seed=0)
    sel.train(handle, X, tuning.targets(T), steps=300, batch=24, lr=3e-3)
    chosen = np.argmin(sel.predict_bf16(handle, X), axis=1)
    got = tuning.evaluate(chosen, T)
    rnd = tuning.evaluate(np.random.default_rng(0).integers(0, T.shape[1], size=len(T)), T)
    assert got["geomean_slowdown"] <= rnd["geomean_slowdown"]
    # every config computes the same product (parity of the whole search space)
    import torch

    import portblas_amd as pb

    L = pb.load()
    m, n, k = 77, 130, 200
    A = torch.randn(m * k, device="cuda").to(torch.bfloat16)
    B = torch.randn(k * n, device="cuda").to(torch.bfloat16)
    ref = (A.float().view(k, m).t() @ B.float().view(n, k).t()).t().contiguous().view(-1)
    for cid in range(T.shape[1]):
        C = torch.zeros(m * n, device="cuda")
        pb.check(L.pb_gemm_bf16_config(handle.handle, cid, b"n", b"n", m, n, k, 1.0, A.data_ptr(), m, 0,
                                       B.data_ptr(), k, 0, 0.0, C.data_ptr(), m, 0, 1))
        err = (C - ref).abs().max().item()
        assert err <= 1e-3 * k ** 0.5, (cid, err)

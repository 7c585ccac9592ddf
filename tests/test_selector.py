"""The trained tile-configuration selector (portblas_amd/selector_mi355x.npz)
against exhaustive autotuning on held-out shapes.

The fixture tests/golden/selector_heldout_mi355x.npz was written by
``python -m portblas_amd.autotune ... --fixture`` on an MI355X: 2000 shapes
the network never saw (M, N, K log-uniform in [16, 16384], batch 1..256,
random transposes), the measured time of every one of the 40 configurations
for each (exhaustive search, microseconds), and the configuration the
library's built-in rule picks.  The CPU tests evaluate the network with the
oracle's BF16 forward pass; the GPU tests use the library's BF16 inference
kernel and the handle's own dispatch, and re-time a few shapes live.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import mlp as omlp

ROOT = Path(__file__).resolve().parents[1]
NET = ROOT / "portblas_amd" / "selector_mi355x.npz"
HELDOUT = ROOT / "tests" / "golden" / "selector_heldout_mi355x.npz"


def _evaluate(chosen, T):
    rows = np.arange(T.shape[0])
    ratio = T[rows, chosen].astype(np.float64) / T.min(axis=1)
    return {"geomean": float(np.exp(np.mean(np.log(ratio)))),
            "total": float(T[rows, chosen].sum() / T.min(axis=1).sum()),
            "within5": float(np.mean(ratio <= 1.05))}


def _load():
    z = np.load(NET, allow_pickle=False)
    n = len(z["sizes"]) - 1
    net = ([z[f"w{l}"] for l in range(n)], [z[f"b{l}"] for l in range(n)], z["mean"], z["std"],
           [int(v) for v in z["sizes"]])
    return net, np.load(HELDOUT, allow_pickle=False)


def _oracle_choice(net, X):
    ws, bs, mean, std, _ = net
    return np.argmin(omlp.forward_bf16(ws, bs, ((X - mean) / std).astype(np.float32)), axis=1)


def test_selector_within_5pct_of_exhaustive_on_heldout_shapes():
    """The learned choice costs at most 5% over exhaustive autotuning
    (geometric mean over the held-out shapes) and beats both the built-in
    rule and the best single configuration."""
    net, d = _load()
    assert net[4] == [8, 64, 64, 64, d["T"].shape[1]]
    T = d["T"]
    mlp = _evaluate(_oracle_choice(net, d["X"]), T)
    rule = _evaluate(d["rule_choice"], T)
    single = _evaluate(np.full(len(T), int(d["best_single"])), T)
    assert mlp["geomean"] <= 1.05, mlp
    # what the network gains over the method without it
    assert mlp["geomean"] < rule["geomean"] - 0.05, (mlp, rule)
    assert mlp["total"] < rule["total"] - 0.05, (mlp, rule)
    assert mlp["within5"] > rule["within5"] + 0.2, (mlp, rule)
    assert mlp["geomean"] < single["geomean"] and mlp["total"] < single["total"]


def test_exhaustive_fixture_is_consistent():
    """Exhaustive search is the yardstick: its timings are positive and finite
    for every configuration, and the features encode the recorded shapes."""
    from portblas_amd import tuning

    _, d = _load()
    T, X, shapes = d["T"], d["X"], d["shapes"]
    assert T.shape == (2000, 40) and np.all(np.isfinite(T)) and np.all(T > 0)
    for row, x in zip(shapes[:50], X[:50]):
        ta, tb, m, n, k, batch = chr(int(row[0])), chr(int(row[1])), *(int(v) for v in row[2:])
        np.testing.assert_allclose(tuning.features(ta, tb, m, n, k, batch), x, rtol=1e-6)


@pytest.mark.gpu
def test_gpu_selector_inference_matches_oracle_and_dispatch(handle, torch_cuda):
    """The committed network through the library's BF16 inference kernel and
    through the handle's host-side dispatch picks what the oracle picks."""
    from portblas_amd import tuning

    net, d = _load()
    sel = tuning.Selector.load(NET)
    X = d["X"]
    gpu = np.argmin(sel.predict_bf16(handle, X), axis=1)
    ref = _oracle_choice(net, X)
    assert np.mean(gpu == ref) >= 0.99
    shapes = [(chr(int(r[0])), chr(int(r[1])), *(int(v) for v in r[2:])) for r in d["shapes"][:300]]
    sel.install(handle)
    try:
        host = np.array([tuning.select(handle, *s) for s in shapes])
    finally:
        tuning.uninstall(handle)
    assert np.mean(host == gpu[:300]) >= 0.99
    # with the network uninstalled the handle falls back to the built-in rule
    rule = np.array([tuning.select(handle, *s) for s in shapes])
    assert np.array_equal(rule, d["rule_choice"][:300])


@pytest.mark.gpu
def test_gpu_selector_choice_live_timing(handle, torch_cuda):
    """Re-time a few median-sized held-out shapes now: the network's choice is
    within 10% of the exhaustive optimum of this run (geometric mean; live
    timings carry a few percent of noise)."""
    from portblas_amd import tuning

    net, d = _load()
    sizes = d["shapes"][:, 2:5].astype(np.float64).prod(axis=1) * d["shapes"][:, 5]
    pick = np.argsort(sizes)[1000:1016]  # median-sized shapes: well above timing noise, about a second in all
    shapes = [(chr(int(r[0])), chr(int(r[1])), *(int(v) for v in r[2:])) for r in d["shapes"][pick]]
    tuner = tuning.Autotuner(handle)
    T = np.stack([tuner.time_all(s, reps=5) for s in shapes])
    chosen = _oracle_choice(net, d["X"][pick])
    assert _evaluate(chosen, T)["geomean"] <= 1.10


# ---------------------------------------------------------------------------
# The FP32 / FP64 plain engine's network (portblas_amd/selector_plain_mi355x.npz)
# and its held-out fixture (tests/golden/selector_plain_heldout_mi355x.npz,
# written by ``python -m portblas_amd.autotune --family plain ... --fixture``:
# 2000 FP32/FP64 shapes never trained on, M, N, K log-uniform in [16, 16384],
# batch 1..256, random transposes and lda padding, work capped at 2^35 flop;
# each timed in all 17 plain configurations and with the built-in rule, the
# last column).
# ---------------------------------------------------------------------------
PLAIN_NET = ROOT / "portblas_amd" / "selector_plain_mi355x.npz"
PLAIN_HELDOUT = ROOT / "tests" / "golden" / "selector_plain_heldout_mi355x.npz"


def _load_plain():
    z = np.load(PLAIN_NET, allow_pickle=False)
    n = len(z["sizes"]) - 1
    net = ([z[f"w{l}"] for l in range(n)], [z[f"b{l}"] for l in range(n)], z["mean"], z["std"],
           [int(v) for v in z["sizes"]])
    return net, np.load(PLAIN_HELDOUT, allow_pickle=False)


def _plain_times(chosen, T):
    """Time of each shape with the chosen configs; an untimed config (which
    the library would not run, falling back to its rule) costs the rule's."""
    t = T[np.arange(len(T)), chosen].astype(np.float64)
    return np.where(np.isfinite(t), t, T[:, -1])


def test_plain_selector_within_5pct_of_exhaustive_and_beats_rule():
    """FP32/FP64 _gemm: the learned choice is within 5% of exhaustive search
    over the 17 plain configurations (geometric mean and total time over the
    held-out shapes) and faster than the built-in rule the library uses
    without it."""
    net, d = _load_plain()
    T = d["T"].astype(np.float64)
    nconf = T.shape[1] - 1
    assert net[4] == [8, 64, 64, 64, nconf] and nconf == 17
    best = T[:, :nconf].min(axis=1)
    mlp = _plain_times(_oracle_choice(net, d["X"]), T)
    rule = T[:, -1]
    geo = lambda t: float(np.exp(np.mean(np.log(t / best))))
    assert geo(mlp) <= 1.05 and mlp.sum() / best.sum() <= 1.05, (geo(mlp), mlp.sum() / best.sum())
    assert geo(mlp) < geo(rule) - 0.03, (geo(mlp), geo(rule))
    assert mlp.sum() < 0.95 * rule.sum(), (mlp.sum(), rule.sum())
    # both element types and the padded lda shapes take part
    assert set(np.unique(d["X"][:, 5])) == {1.0, 2.0} and np.mean(d["X"][:, 7] > 0) > 0.2


PLAIN_ANCHORS = ROOT / "tests" / "golden" / "selector_plain_anchors_mi355x.npz"


def test_plain_selector_on_anchor_shapes():
    """The structured shapes the network was also trained on (squares and
    K-stretched squares of 256 .. 2048, small-problem batches; timed
    exhaustively on MI355X): its picks stay within 5% of the fastest config
    at every one of them, and at the reference's CPU-runnable configuration
    (1024^3 FP32 'n','n', one 64 x 64 tile per CU) it picks the fastest."""
    from portblas_amd import tuning

    net, _ = _load_plain()
    a = np.load(PLAIN_ANCHORS, allow_pickle=False)
    X, T, S = a["X"], a["T"].astype(np.float64), a["shapes"]
    assert len(S) == len(tuning.anchor_plain_shapes())
    for row, x in zip(S, X):
        dt, ta, tb = int(row[0]), chr(int(row[1])), chr(int(row[2]))
        np.testing.assert_allclose(tuning.plain_features(dt, ta, tb, *(int(v) for v in row[3:])), x, rtol=1e-6)
    nconf = T.shape[1] - 1
    chosen = _oracle_choice(net, X)
    best = T[:, :nconf].min(axis=1)
    ratio = _plain_times(chosen, T) / best
    assert ratio.max() <= 1.05, (ratio.max(), S[ratio.argmax()])
    cfg1 = [i for i, r in enumerate(S) if list(r) == [1, ord("n"), ord("n"), 1024, 1024, 1024, 1, 1024]]
    assert len(cfg1) == 1 and chosen[cfg1[0]] == int(np.argmin(T[cfg1[0], :nconf]))


def test_plain_fixture_is_consistent():
    from portblas_amd import tuning

    _, d = _load_plain()
    T, X, shapes = d["T"], d["X"], d["shapes"]
    assert T.shape == (2000, 18) and np.all(np.isfinite(T[:, -1])) and np.all(T[np.isfinite(T)] > 0)
    for row, x in zip(shapes[:50], X[:50]):
        dt, ta, tb = int(row[0]), chr(int(row[1])), chr(int(row[2]))
        m, n, k, batch, lda = (int(v) for v in row[3:])
        np.testing.assert_allclose(tuning.plain_features(dt, ta, tb, m, n, k, batch, lda), x, rtol=1e-6)


@pytest.mark.gpu
def test_gpu_plain_selector_inference_and_dispatch(handle, torch_cuda):
    """The plain network through the BF16 inference kernel and through the
    handle's host-side dispatch picks what the oracle picks."""
    import portblas_amd as pb
    from portblas_amd import tuning

    net, d = _load_plain()
    sel = tuning.Selector.load(PLAIN_NET)
    gpu = np.argmin(sel.predict_bf16(handle, d["X"]), axis=1)
    assert np.mean(gpu == _oracle_choice(net, d["X"])) >= 0.99
    shapes = [(int(r[0]), chr(int(r[1])), chr(int(r[2])), *(int(v) for v in r[3:])) for r in d["shapes"][:300]]
    h = pb.SB_Handle(plain_selector=None)
    assert tuning.select_plain(h, *shapes[0]) == -1  # none installed: the built-in rule
    tuning.install_plain(sel, h)
    host = np.array([tuning.select_plain(h, *s) for s in shapes])
    assert np.mean(host == gpu[:300]) >= 0.99
    # SB_Handle() installs the committed network by default
    assert np.array_equal(np.array([tuning.select_plain(handle, *s) for s in shapes[:50]]), host[:50])
    # outside the training domain the built-in rule decides
    for s in [(1, "n", "n", 8192, 8192, 8192, 1, 8192), (1, "n", "n", 64, 64, 64, 4096, 64),
              (2, "n", "n", 4096, 4096, 4096, 1, 4096)]:
        assert tuning.select_plain(h, *s) == -1
    h.close()


@pytest.mark.gpu
def test_gpu_plain_selector_choice_live_timing(handle, torch_cuda):
    """Re-time 16 median-sized held-out shapes: the network's choice is within
    10% of this run's exhaustive optimum (geometric mean)."""
    from portblas_amd import tuning

    net, d = _load_plain()
    sizes = d["shapes"][:, 3:6].astype(np.float64).prod(axis=1) * d["shapes"][:, 6]
    pick = np.argsort(sizes)[1000:1016]
    shapes = [(int(r[0]), chr(int(r[1])), chr(int(r[2])), *(int(v) for v in r[3:])) for r in d["shapes"][pick]]
    tuner = tuning.PlainAutotuner(handle)
    T = np.stack([tuner.time_all(s, reps=5) for s in shapes])
    chosen = _oracle_choice(net, d["X"][pick])
    t = _plain_times(chosen, T)
    best = T[:, :-1].min(axis=1)
    assert float(np.exp(np.mean(np.log(t / best)))) <= 1.10

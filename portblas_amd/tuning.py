"""Autotuner and learned tile-configuration selector.

The reference picks GEMM tile configurations with hand-written arith_ratio
buckets (src/interface/blas3/backend/amd_gpu.hpp:52-288) and ships an
exhaustive offline tuner (tools/auto_tuner/include/gemm_tuner.hpp:99-180:
time every config, keep the fastest).  Here:

* ``Autotuner`` times every one of the BF16 engine's configurations
  (pb_tune_time_config: 8 tile shapes x 5 split-K factors) for a shape — the
  exhaustive search, also the yardstick of the learned choice;
* ``collect`` builds the timing dataset the selector learns from;
* ``Selector`` is an MLP (8 features -> 3 x 64 hidden -> one predicted
  relative log-runtime per configuration) trained on the GPU by the
  library's own HIP kernels (pb_mlp_forward / pb_mlp_backward / pb_mlp_adam)
  and evaluated in BF16 (pb_mlp_forward_bf16); ``install`` hands it to a
  handle so BF16 GEMMs dispatch through it (pb_tune_set_selector).

Features (same order as the library's host-side evaluation): log2 M, log2 N,
log2 K, transa == 't', transb == 't', element type (0 = BF16), log2 batch,
log2(lda / rows of A).  Targets: log(t_config / t_best) for each config.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .blas import check

NUM_FEATURES = 8


def num_configs() -> int:
    return int(_lib.load().pb_tune_num_configs())


def config_name(i: int) -> str:
    return _lib.load().pb_tune_config_name(i).decode()


def features(ta: str, tb: str, M: int, N: int, K: int, batch: int, lda_pad: float = 1.0) -> np.ndarray:
    lg = lambda v: math.log2(max(int(v), 1))
    return np.array([lg(M), lg(N), lg(K), float(ta in "tTcC"), float(tb in "tTcC"), 0.0, lg(batch),
                     math.log2(max(lda_pad, 1.0))], dtype=np.float32)


def sample_shapes(n: int, seed: int, lo: int = 16, hi: int = 16384, max_batch: int = 256,
                  max_flop: float = 2.0 ** 36, max_elems: int = 2 ** 26):
    """M, N, K log-uniform in [lo, hi], batch uniform in {1..max_batch},
    transposes uniform; shapes whose work or operand footprint exceeds the
    caps are redrawn (keeps one timing sweep bounded)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        M, N, K = (int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(3))
        batch = int(rng.integers(1, max_batch + 1))
        if 2.0 * M * N * K * batch > max_flop:
            continue
        if max(M * K, K * N, M * N) * batch > max_elems:
            continue
        ta = "t" if rng.random() < 0.5 else "n"
        tb = "t" if rng.random() < 0.5 else "n"
        out.append((ta, tb, M, N, K, batch))
    return out


class Autotuner:
    """Exhaustive timing of every configuration (BF16 inputs, FP32 output)."""

    def __init__(self, handle, max_elems: int = 2 ** 26):
        import torch

        self.h = handle
        self.L = _lib.load()
        self.A = torch.zeros(max_elems, dtype=torch.bfloat16, device="cuda").uniform_(-1, 1)
        self.B = torch.zeros(max_elems, dtype=torch.bfloat16, device="cuda").uniform_(-1, 1)
        self.C = torch.zeros(max_elems, dtype=torch.float32, device="cuda")
        self.max_elems = max_elems

    def time_all(self, shape, warmup: int = 1, reps: int = 3) -> np.ndarray:
        ta, tb, M, N, K, batch = shape
        assert max(M * K, K * N, M * N) * batch <= self.max_elems
        lda = M if ta == "n" else K
        ldb = K if tb == "n" else N
        out = np.empty(num_configs(), dtype=np.float64)
        us = ctypes.c_float()
        for cid in range(out.size):
            check(self.L.pb_tune_time_config(self.h.handle, cid, ta.encode(), tb.encode(), M, N, K,
                                             self.A.data_ptr(), lda, self.B.data_ptr(), ldb,
                                             self.C.data_ptr(), M, batch, warmup, reps, ctypes.byref(us)))
            out[cid] = us.value
        return out

    def best(self, shape):
        t = self.time_all(shape)
        i = int(np.argmin(t))
        return i, float(t[i])


def collect(tuner: Autotuner, shapes, progress=None, reps: int = 3):
    """Feature matrix X (n x 8) and timing matrix T (n x configs, us)."""
    X = np.stack([features(*s) for s in shapes])
    T = np.empty((len(shapes), num_configs()))
    for i, s in enumerate(shapes):
        T[i] = tuner.time_all(s, reps=reps)
        if progress and (i + 1) % progress == 0:
            print(f"[tune] {i + 1}/{len(shapes)} shapes timed", flush=True)
    return X, T


def targets(T: np.ndarray, clip: float = math.log(4.0)) -> np.ndarray:
    """Relative log runtime log(t / min over configs), clipped at `clip`:
    configurations more than 4x slower than the best are all "bad", so the
    network spends its capacity ranking the near-best ones."""
    return np.minimum(np.log(T / T.min(axis=1, keepdims=True)), clip).astype(np.float32)


def weights(Y: np.ndarray, sharpness: float = 4.0) -> np.ndarray:
    """Per-output loss weights exp(-sharpness * target): an error on a
    configuration near the best of its shape decides the choice, one on a
    configuration twice as slow rarely does (weight 1 at the best, 0.06 at
    2x).  Passed to the training kernels as the loss mask."""
    return np.exp(-sharpness * Y).astype(np.float32)


@dataclass
class Selector:
    sizes: list
    ws: list  # numpy float32 [d_out x d_in]
    bs: list
    mean: np.ndarray
    std: np.ndarray

    @staticmethod
    def init(sizes=(NUM_FEATURES, 64, 64, 64, 40), seed=0) -> "Selector":
        rng = np.random.default_rng(seed)
        ws, bs = [], []
        for din, dout in zip(sizes[:-1], sizes[1:]):
            ws.append((rng.standard_normal((dout, din)) * math.sqrt(2.0 / din)).astype(np.float32))
            bs.append(np.zeros(dout, dtype=np.float32))
        return Selector(list(sizes), ws, bs, np.zeros(NUM_FEATURES, np.float32), np.ones(NUM_FEATURES, np.float32))

    # -- persistence ------------------------------------------------------
    def save(self, path):
        d = {"sizes": np.array(self.sizes), "mean": self.mean, "std": self.std}
        for l, (w, b) in enumerate(zip(self.ws, self.bs)):
            d[f"w{l}"], d[f"b{l}"] = w, b
        np.savez(path, **d)

    @staticmethod
    def load(path) -> "Selector":
        z = np.load(path, allow_pickle=False)
        sizes = [int(v) for v in z["sizes"]]
        n = len(sizes) - 1
        return Selector(sizes, [z[f"w{l}"] for l in range(n)], [z[f"b{l}"] for l in range(n)], z["mean"], z["std"])

    def normalize(self, X):
        return ((X - self.mean) / self.std).astype(np.float32)

    # -- training on the GPU with the library's kernels --------------------
    def train(self, handle, X, Y, mask=None, steps=3000, batch=1024, lr=2e-3, seed=0, log_every=0):
        import torch

        L = _lib.load()
        h = handle.handle
        self.mean = X.mean(axis=0).astype(np.float32)
        self.std = X.std(axis=0).astype(np.float32)
        self.std[self.std < 1e-6] = 1.0
        Xn = torch.from_numpy(self.normalize(X)).cuda()
        Yd = torch.from_numpy(np.ascontiguousarray(Y, dtype=np.float32)).cuda()
        Md = None if mask is None else torch.from_numpy(np.ascontiguousarray(mask, dtype=np.float32)).cuda()
        n_layers = len(self.sizes) - 1
        W = [torch.from_numpy(w.copy()).cuda() for w in self.ws]
        B = [torch.from_numpy(b.copy()).cuda() for b in self.bs]
        dW = [torch.zeros_like(w) for w in W]
        dB = [torch.zeros_like(b) for b in B]
        mW = [torch.zeros_like(w) for w in W]
        vW = [torch.zeros_like(w) for w in W]
        mB = [torch.zeros_like(b) for b in B]
        vB = [torch.zeros_like(b) for b in B]
        bsz = min(batch, X.shape[0])
        acts = [torch.empty(bsz, d, device="cuda") for d in self.sizes]
        tgt = torch.empty(bsz, self.sizes[-1], device="cuda")
        msk = torch.empty(bsz, self.sizes[-1], device="cuda") if Md is not None else None
        scratch = torch.empty(2 * bsz * max(self.sizes), device="cuda")
        loss = torch.zeros(1, device="cuda")
        sizes_c = (ctypes.c_int * len(self.sizes))(*self.sizes)
        ptrs = lambda ts: (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])
        w_p, b_p, a_p, dw_p, db_p = ptrs(W), ptrs(B), ptrs(acts), ptrs(dW), ptrs(dB)
        rng = np.random.default_rng(seed)
        history = []
        for step in range(1, steps + 1):
            idx = torch.from_numpy(rng.integers(0, X.shape[0], size=bsz)).cuda()
            acts[0].copy_(Xn[idx])
            tgt.copy_(Yd[idx])
            if msk is not None:
                msk.copy_(Md[idx])
            check(L.pb_mlp_forward(h, n_layers, sizes_c, w_p, b_p, a_p, bsz))
            check(L.pb_mlp_backward(h, n_layers, sizes_c, w_p, a_p, tgt.data_ptr(),
                                    msk.data_ptr() if msk is not None else None, dw_p, db_p,
                                    loss.data_ptr(), scratch.data_ptr(), bsz))
            cur_lr = lr * 0.5 * (1 + math.cos(math.pi * step / steps))
            for l in range(n_layers):
                check(L.pb_mlp_adam(h, W[l].data_ptr(), dW[l].data_ptr(), mW[l].data_ptr(), vW[l].data_ptr(),
                                    W[l].numel(), cur_lr, 0.9, 0.999, 1e-8, step))
                check(L.pb_mlp_adam(h, B[l].data_ptr(), dB[l].data_ptr(), mB[l].data_ptr(), vB[l].data_ptr(),
                                    B[l].numel(), cur_lr, 0.9, 0.999, 1e-8, step))
            if log_every and step % log_every == 0:
                history.append((step, loss.item()))
                print(f"[mlp] step {step} loss {history[-1][1]:.5f}", flush=True)
        handle.wait()
        self.ws = [w.cpu().numpy() for w in W]
        self.bs = [b.cpu().numpy() for b in B]
        return history

    def predict_bf16(self, handle, X) -> np.ndarray:
        """BF16 inference on the GPU (pb_mlp_forward_bf16)."""
        import torch

        L = _lib.load()
        Xn = torch.from_numpy(self.normalize(X)).cuda()
        Wb = [torch.from_numpy(w).cuda().to(torch.bfloat16) for w in self.ws]
        Bd = [torch.from_numpy(b).cuda() for b in self.bs]
        Y = torch.empty(X.shape[0], self.sizes[-1], device="cuda")
        sizes_c = (ctypes.c_int * len(self.sizes))(*self.sizes)
        wp = (ctypes.c_void_p * len(Wb))(*[w.data_ptr() for w in Wb])
        bp = (ctypes.c_void_p * len(Bd))(*[b.data_ptr() for b in Bd])
        check(L.pb_mlp_forward_bf16(handle.handle, len(self.ws), sizes_c, wp, bp, Xn.data_ptr(), Y.data_ptr(),
                                    X.shape[0]))
        handle.wait()
        return Y.cpu().numpy()

    def install(self, handle) -> None:
        """Make the handle pick BF16 GEMM configs with this network."""
        L = _lib.load()
        sizes_c = (ctypes.c_int * len(self.sizes))(*self.sizes)
        ws = [np.ascontiguousarray(w, np.float32) for w in self.ws]
        bs = [np.ascontiguousarray(b, np.float32) for b in self.bs]
        wp = (ctypes.c_void_p * len(ws))(*[w.ctypes.data for w in ws])
        bp = (ctypes.c_void_p * len(bs))(*[b.ctypes.data for b in bs])
        mean = np.ascontiguousarray(self.mean, np.float32)
        std = np.ascontiguousarray(self.std, np.float32)
        check(L.pb_tune_set_selector(handle.handle, len(ws), sizes_c, wp, bp, mean.ctypes.data, std.ctypes.data))


def uninstall(handle) -> None:
    check(_lib.load().pb_tune_set_selector(handle.handle, 0, None, None, None, None, None))


def select(handle, ta, tb, M, N, K, batch) -> int:
    """The configuration the handle would use for this shape."""
    return int(_lib.load().pb_tune_select(handle.handle, ta.encode(), tb.encode(), M, N, K, batch))


# ---------------------------------------------------------------------------
# The plain FP32 / FP64 engine (pb_sgemm / pb_dgemm and their batched forms)
# ---------------------------------------------------------------------------
PLAIN_DTYPES = {1: "float32", 2: "float64"}
PLAIN_SPLITS = (1, 2, 4, 8, 16, 32, 64, 128)


def num_plain_configs() -> int:
    return int(_lib.load().pb_tune_num_plain_configs())


def plain_config_name(i: int) -> str:
    return _lib.load().pb_tune_plain_config_name(i).decode()


def plain_features(dtype: int, ta: str, tb: str, M: int, N: int, K: int, batch: int, lda: int) -> np.ndarray:
    """The library's feature order (gemm.hip: plain_features)."""
    lg = lambda v: math.log2(max(float(v), 1.0))
    arows = M if ta == "n" else K
    return np.array([lg(M), lg(N), lg(K), float(ta == "t"), float(tb == "t"), float(dtype), lg(batch),
                     lg(lda / max(arows, 1))], dtype=np.float32)


def plain_split_tiles(M: int, N: int, batch: int, cid: int) -> int:
    """Tiles x splits of plain config `cid`, as the library counts them."""
    bm = 128 if cid < len(PLAIN_SPLITS) else 64
    tiles = -(-M // bm) * -(-N // bm) * batch
    splits = PLAIN_SPLITS[cid % len(PLAIN_SPLITS)]
    return tiles * splits if splits > 1 else 0


def sample_plain_shapes(n: int, seed: int, lo: int = 16, hi: int = 16384, max_batch: int = 256,
                        max_flop: float = 2.0 ** 35, max_elems: int = 2 ** 26):
    """(dtype, ta, tb, M, N, K, batch, lda): dtype FP32 or FP64 at random, M,
    N, K log-uniform in [lo, hi], batch uniform in {1..max_batch}, random
    transposes, lda padded beyond the rows of A in a third of the shapes;
    redrawn while work or footprint exceeds the caps."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        M, N, K = (int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(3))
        batch = int(rng.integers(1, max_batch + 1))
        if 2.0 * M * N * K * batch > max_flop:
            continue
        ta = "t" if rng.random() < 0.5 else "n"
        tb = "t" if rng.random() < 0.5 else "n"
        arows, acols = (M, K) if ta == "n" else (K, M)
        lda = arows + (int(rng.integers(1, 65)) if rng.random() < 1 / 3 else 0)
        if max(lda * acols, K * N, M * N) * batch > max_elems:
            continue
        out.append((int(rng.integers(1, 3)), ta, tb, M, N, K, batch, lda))
    return out


def anchor_plain_shapes(max_flop: float = 2.0 ** 35, max_elems: int = 2 ** 26):
    """Structured shapes added to the random training set: square and
    K-stretched GEMMs of 256 .. 2048 (powers of two and 1.5 x powers of two),
    both element types, all four transposes, batch 1, packed leading
    dimensions; and 64^3 / 32^3 batches of 64 .. 256.  Random log-uniform
    shapes rarely land on exact multiples of the tile edge and the CU count,
    where the best plan changes abruptly (1024^2 is exactly one 64 x 64 tile
    per CU on 256 CUs), and which the benchmark configurations sit on."""
    out = []
    for dt in PLAIN_DTYPES:
        for ta in "nt":
            for tb in "nt":
                for s in (256, 384, 512, 768, 1024, 1536, 2048):
                    for K in (s // 2, s, 2 * s):
                        if 2.0 * s * s * K > max_flop or max(s * K, K * s, s * s) > max_elems:
                            continue
                        out.append((dt, ta, tb, s, s, K, 1, s if ta == "n" else K))
                for e, b in ((64, 64), (64, 128), (64, 256), (32, 256)):
                    out.append((dt, ta, tb, e, e, e, b, e))
    return out


class PlainAutotuner:
    """Exhaustive timing of every plain configuration for FP32 and FP64
    shapes, plus the built-in rule (the last column).  Configurations whose
    split tiles exceed 16 per CU, and the small-batch kernel where it cannot
    run, are not timed (inf)."""

    def __init__(self, handle, max_elems: int = 2 ** 26):
        import torch

        self.h = handle
        self.L = _lib.load()
        self.cus = handle.get_num_compute_units()
        self.buf = {}
        for dt, name in PLAIN_DTYPES.items():
            tdt = getattr(torch, name)
            self.buf[dt] = (torch.empty(max_elems, dtype=tdt, device="cuda").uniform_(-1, 1),
                            torch.empty(max_elems, dtype=tdt, device="cuda").uniform_(-1, 1),
                            torch.zeros(max_elems, dtype=tdt, device="cuda"))
        self.max_elems = max_elems

    def time_all(self, shape, warmup: int = 1, reps: int = 3) -> np.ndarray:
        dt, ta, tb, M, N, K, batch, lda = shape
        A, B, C = self.buf[dt]
        ldb = K if tb == "n" else N
        nconf = num_plain_configs()
        out = np.full(nconf + 1, np.inf)
        us = ctypes.c_float()
        small_ok = M <= 64 and N <= 64 and batch >= self.cus
        for cid in list(range(nconf)) + [-1]:
            if 0 <= cid < nconf - 1 and plain_split_tiles(M, N, batch, cid) > 16 * self.cus:
                continue
            if cid == nconf - 1 and not small_ok:
                continue
            check(self.L.pb_tune_time_plain(self.h.handle, dt, cid, ta.encode(), tb.encode(), M, N, K,
                                            A.data_ptr(), lda, B.data_ptr(), ldb, C.data_ptr(), M, batch,
                                            warmup, reps, ctypes.byref(us)))
            out[cid if cid >= 0 else nconf] = us.value
        return out


def collect_plain(tuner: PlainAutotuner, shapes, progress=None, reps: int = 3):
    """Features (n x 8) and timings (n x configs + 1: the rule last, us)."""
    X = np.stack([plain_features(*s) for s in shapes])
    T = np.stack([tuner.time_all(s, reps=reps) for s in _progress(shapes, progress)])
    return X, T


def _progress(items, every):
    for i, it in enumerate(items):
        yield it
        if every and (i + 1) % every == 0:
            print(f"[tune] {i + 1}/{len(items)} shapes timed", flush=True)


def install_plain(selector: "Selector", handle) -> None:
    """Make the handle pick FP32 / FP64 plain configs with this network."""
    L = _lib.load()
    sizes_c = (ctypes.c_int * len(selector.sizes))(*selector.sizes)
    ws = [np.ascontiguousarray(w, np.float32) for w in selector.ws]
    bs = [np.ascontiguousarray(b, np.float32) for b in selector.bs]
    wp = (ctypes.c_void_p * len(ws))(*[w.ctypes.data for w in ws])
    bp = (ctypes.c_void_p * len(bs))(*[b.ctypes.data for b in bs])
    mean = np.ascontiguousarray(selector.mean, np.float32)
    std = np.ascontiguousarray(selector.std, np.float32)
    check(L.pb_tune_set_plain_selector(handle.handle, len(ws), sizes_c, wp, bp, mean.ctypes.data,
                                       std.ctypes.data))


def uninstall_plain(handle) -> None:
    check(_lib.load().pb_tune_set_plain_selector(handle.handle, 0, None, None, None, None, None))


def select_plain(handle, dt, ta, tb, M, N, K, batch, lda) -> int:
    return int(_lib.load().pb_tune_select_plain(handle.handle, dt, ta.encode(), tb.encode(), M, N, K, batch, lda))


def plain_time_of_choice(chosen: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Time each shape takes with the given configs: where a config was not
    timed (the library would not run it and falls back to the rule), the
    rule's time."""
    rows = np.arange(T.shape[0])
    t = T[rows, chosen]
    return np.where(np.isfinite(t), t, T[:, -1])


def evaluate(chosen: np.ndarray, T: np.ndarray) -> dict:
    """Quality of a choice of config per shape against exhaustive search."""
    rows = np.arange(T.shape[0])
    ratio = T[rows, chosen] / T.min(axis=1)
    return {
        "geomean_slowdown": float(np.exp(np.mean(np.log(ratio)))),
        "median_slowdown": float(np.median(ratio)),
        "frac_within_5pct": float(np.mean(ratio <= 1.05)),
        "p90_slowdown": float(np.quantile(ratio, 0.9)),
        "total_time_ratio": float(T[rows, chosen].sum() / T.min(axis=1).sum()),
    }


# ---------------------------------------------------------------------------
# Grouped BF16 GEMM: the schedule choices of a plan (pb_grouped_options_t),
# timed on the caller's own problems; the fastest is kept.  The reference has
# no grouped GEMM; its tuner's method (time every candidate, keep the
# fastest: tools/auto_tuner/include/gemm_tuner.hpp:99-180) is applied to the
# plan's schedule.
# ---------------------------------------------------------------------------
class GroupedOptions(ctypes.Structure):
    """pb_grouped_options_t: 0 in a field = the built-in choice."""
    _fields_ = [("ring", ctypes.c_int), ("rank_weights", ctypes.c_int), ("spread_div", ctypes.c_int),
                ("per_cu", ctypes.c_int)]

    def as_tuple(self):
        return (self.ring, self.rank_weights, self.spread_div, self.per_cu)


class GroupedProblem(ctypes.Structure):
    """pb_gemm_problem_t."""
    _fields_ = [("transa", ctypes.c_char), ("transb", ctypes.c_char), ("m", ctypes.c_int64),
                ("n", ctypes.c_int64), ("k", ctypes.c_int64), ("alpha", ctypes.c_float),
                ("beta", ctypes.c_float), ("a", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("b", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("c", ctypes.c_void_p),
                ("ldc", ctypes.c_int64)]


def grouped_candidates():
    """(ring, rank_weights, spread_div, per_cu): the built-in schedule first,
    then the ring or lean kernel, CU-position weights on or off, the XCD
    spread threshold (1/16, 1/4, never) and 3 or 2 workgroups per CU."""
    out = [(0, 0, 0, 0)]
    for ring in (1, 2):
        for rank in (1, 2):
            for spread in (16, 4, 1):
                for per_cu in (3, 2):
                    if ring == 2 and rank == 2:
                        continue  # the lean kernel never weighs ranks
                    out.append((ring, rank, spread, per_cu))
    return out


class GroupedAutotuner:
    """Times grouped plans built with each candidate schedule on one problem
    list (HIP events around pb_gemm_grouped_plan_run on the handle's stream)
    and keeps the fastest.  The plans run repeatedly, so every problem must
    have beta == 0 (C is then overwritten with the same result each time)."""

    def __init__(self, handle, precision: int):
        self.h = handle
        self.L = _lib.load()
        self.precision = int(precision)

    def plan(self, problems, count: int, opts) -> ctypes.c_void_p:
        o = GroupedOptions(*opts)
        plan = ctypes.c_void_p()
        check(self.L.pb_gemm_grouped_plan_create_ex(self.h.handle, self.precision, problems, count,
                                                    ctypes.byref(o), ctypes.byref(plan)))
        return plan

    def time(self, problems, count: int, opts, warm: int = 2, reps: int = 5) -> float:
        import torch

        plan = self.plan(problems, count, opts)
        try:
            for _ in range(warm):
                check(self.L.pb_gemm_grouped_plan_run(self.h.handle, plan))
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                check(self.L.pb_gemm_grouped_plan_run(self.h.handle, plan))
            e1.record()
            e1.synchronize()
            return e0.elapsed_time(e1) * 1e3 / reps  # us
        finally:
            check(self.L.pb_gemm_grouped_plan_destroy(plan))

    def tune(self, problems, count: int, candidates=None, warm: int = 2, reps: int = 5):
        """Returns (best options tuple, {options: us})."""
        if any(problems[i].beta != 0.0 for i in range(count)):
            raise ValueError("grouped autotuning runs each plan repeatedly: every problem needs beta == 0")
        table = {}
        for opts in candidates or grouped_candidates():
            table[tuple(opts)] = self.time(problems, count, opts, warm, reps)
        best = min(table, key=table.get)
        return best, table

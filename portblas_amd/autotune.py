"""Autotuning + learned-selector pipeline (GPU).

    python -m portblas_amd.autotune --train 20000 --heldout 2000 \\
        --out portblas_amd/selector_mi355x.npz --report tune_report.json

1. draws training and held-out shapes (M, N, K log-uniform in [16, 16384],
   batch in {1..256}, random transposes; work and footprint capped so one
   sweep stays bounded);
2. times every configuration of every shape (exhaustive autotuning,
   pb_tune_time_config);
3. trains the 8 -> 64 -> 64 -> 64 -> configs MLP on the GPU with the
   library's own kernels (FP32; squared error on log(t / t_best), weighted
   towards the near-best configurations of each shape), then evaluates it in
   BF16 on the held-out shapes against the exhaustive optimum and against
   the built-in rule;
4. saves the network, a JSON report and (``--fixture``) the held-out set.

Timing datasets can be collected in pieces (``--collect-only`` with
``--train-files`` / ``--heldout-file`` naming where to save them) and reused:
a file that exists is loaded instead of re-timed.
The precedent is tools/auto_tuner (gemm_tuner.hpp:99-180: time every config,
keep the best); the network replaces the hand-written buckets of
src/interface/blas3/backend/amd_gpu.hpp:52-288.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def _enc(s):
    return [ord(s[0]), ord(s[1]), *s[2:]]


def _dec(row):
    return (chr(int(row[0])), chr(int(row[1])), *(int(v) for v in row[2:]))


def _timed_set(tuning, tuner_box, n, seed, max_flop, reps, path, progress_every):
    """(X, T, shapes) for n shapes drawn with `seed`, loaded from `path` when
    it exists, else timed (and saved to `path` when given)."""
    if path and Path(path).exists():
        z = np.load(path, allow_pickle=False)
        print(f"[tune] loaded {path}: {len(z['X'])} shapes", flush=True)
        return z["X"], z["T"], [_dec(r) for r in z["shapes"]]
    shapes = tuning.sample_shapes(n, seed=seed, max_flop=max_flop)
    t0 = time.time()
    X, T = tuning.collect(tuner_box(), shapes, progress=progress_every, reps=reps)
    print(f"[tune] timed {n} shapes x {T.shape[1]} configs in {time.time() - t0:.0f} s", flush=True)
    if path:
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, X=X, T=T, shapes=np.array([_enc(s) for s in shapes], dtype=np.int64))
    return X, T, shapes


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=20000, help="training shapes per train file")
    ap.add_argument("--heldout", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--lr", type=float, default=5e-3)
    ap.add_argument("--sharpness", type=float, default=4.0, help="loss weight exp(-sharpness * target)")
    ap.add_argument("--candidates", type=int, default=4,
                    help="networks trained from different seeds; the best on a validation split is kept")
    ap.add_argument("--val-frac", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--max-flop", type=float, default=2.0 ** 35)
    ap.add_argument("--reps", type=int, default=3, help="timed repetitions per config")
    ap.add_argument("--train-files", nargs="*", default=[],
                    help="timing datasets of training shapes (file i uses seed + i)")
    ap.add_argument("--heldout-file", default=None)
    ap.add_argument("--anchors", default=None,
                    help="plain family: timing file of the anchor shapes (timed and written if missing)")
    ap.add_argument("--anchor-repeat", type=int, default=8, help="copies of the anchor shapes in the training set")
    ap.add_argument("--collect-only", action="store_true")
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    ap.add_argument("--fixture", default=None,
                    help="also save the held-out set, its timings and the built-in rule's choices (float32)")
    ap.add_argument("--family", choices=("bf16", "plain"), default="bf16",
                    help="bf16: the BF16-input engine (40 configs); plain: the FP32/FP64 engine behind "
                         "pb_sgemm/pb_dgemm (17 configs, timed with the built-in rule beside them)")
    args = ap.parse_args(argv)
    if args.family == "plain":
        return main_plain(args)

    import torch

    import portblas_amd as pb
    from portblas_amd import tuning

    h = pb.SB_Handle()
    nconf = tuning.num_configs()
    tuner = []

    def tuner_box():
        if not tuner:
            tuner.append(tuning.Autotuner(h))
        return tuner[0]

    t0 = time.time()
    train_files = args.train_files or [None]
    parts = [_timed_set(tuning, tuner_box, args.train, args.seed + i, args.max_flop, args.reps, f,
                        max(1, args.train // 20)) for i, f in enumerate(train_files)]
    Xtr = np.concatenate([p[0] for p in parts])
    Ttr = np.concatenate([p[1] for p in parts])
    if args.collect_only and not args.heldout_file:
        return {"train_shapes": int(len(Xtr))}
    Xte, Tte, te_shapes = _timed_set(tuning, tuner_box, args.heldout, args.seed + 1000, args.max_flop,
                                     args.reps, args.heldout_file, max(1, args.heldout // 10))
    if args.collect_only:
        return {"train_shapes": int(len(Xtr)), "heldout_shapes": int(len(Xte))}
    tuner.clear()
    torch.cuda.empty_cache()
    t_collect = time.time() - t0
    print(f"[tune] datasets: {len(Xtr)} + {len(Xte)} shapes x {nconf} configs ({t_collect:.0f} s)", flush=True)

    # Model selection on a validation split of the training shapes (never on
    # the held-out set): several seeds, keep the network whose BF16 choices
    # are fastest on the validation shapes.
    Ytr = tuning.targets(Ttr)
    perm = np.random.default_rng(args.seed).permutation(len(Xtr))
    n_val = int(round(args.val_frac * len(Xtr))) if args.candidates > 1 else 0
    val, fit = perm[:n_val], perm[n_val:]
    t1 = time.time()
    sel, hist, val_scores = None, [], []
    for c in range(args.candidates):
        cand = tuning.Selector.init(sizes=(tuning.NUM_FEATURES, 64, 64, 64, nconf), seed=args.seed + c)
        ch = cand.train(h, Xtr[fit], Ytr[fit], mask=tuning.weights(Ytr[fit], args.sharpness), steps=args.steps,
                        batch=1024, lr=args.lr, seed=args.seed + c, log_every=max(1, args.steps // 4))
        if n_val:
            ev = tuning.evaluate(np.argmin(cand.predict_bf16(h, Xtr[val]), axis=1), Ttr[val])
            score = ev["geomean_slowdown"] * ev["total_time_ratio"]
        else:
            ev, score = None, 0.0
        val_scores.append({"seed": args.seed + c, "validation": ev})
        print(f"[tune] candidate {c}: validation {ev}", flush=True)
        if sel is None or score < best_score:
            sel, hist, best_score = cand, ch, score
    t_train = time.time() - t1

    pred = sel.predict_bf16(h, Xte)
    chosen_mlp = np.argmin(pred, axis=1)
    # built-in rule (no learned selector) for the same shapes
    tuning.uninstall(h)
    chosen_rule = np.array([tuning.select(h, *s) for s in te_shapes])
    # the handle's own dispatch with the network installed agrees with the GPU inference
    sel.install(h)
    chosen_host = np.array([tuning.select(h, *s) for s in te_shapes])
    tuning.uninstall(h)
    agree = float(np.mean(chosen_host == chosen_mlp))
    best_single = int(np.argmin(np.log(Ttr).mean(axis=0)))
    rep = {
        "train_shapes": int(len(Xtr)), "heldout_shapes": int(len(Xte)), "configs": nconf,
        "timing_seconds": t_collect, "train_seconds": t_train, "final_loss": hist[-1][1] if hist else None,
        "training": {"steps": args.steps, "lr": args.lr, "loss_weight_sharpness": args.sharpness,
                     "fit_shapes": int(len(fit)), "validation_shapes": int(n_val), "candidates": val_scores},
        "mlp_bf16": tuning.evaluate(chosen_mlp, Tte),
        "builtin_rule": tuning.evaluate(chosen_rule, Tte),
        "best_single_config": tuning.evaluate(np.full(len(Tte), best_single), Tte),
        "best_single_config_name": tuning.config_name(best_single),
        "host_dispatch_agrees_with_gpu_inference": agree,
    }
    print(json.dumps(rep, indent=1), flush=True)
    if args.out:
        sel.save(args.out)
    if args.fixture:
        Path(args.fixture).parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(args.fixture, X=Xte.astype(np.float32), T=Tte.astype(np.float32),
                            shapes=np.array([_enc(s) for s in te_shapes], dtype=np.int64),
                            rule_choice=chosen_rule.astype(np.int32), best_single=np.int32(best_single))
    if args.report:
        Path(args.report).parent.mkdir(parents=True, exist_ok=True)
        Path(args.report).write_text(json.dumps(rep, indent=1))
    return rep


def _ratio_stats(t: np.ndarray, best: np.ndarray) -> dict:
    r = t / best
    return {"geomean_slowdown": float(np.exp(np.mean(np.log(r)))), "median_slowdown": float(np.median(r)),
            "frac_within_5pct": float(np.mean(r <= 1.05)), "p90_slowdown": float(np.quantile(r, 0.9)),
            "total_time_ratio": float(t.sum() / best.sum())}


def _plain_set(tuning, tuner_box, n, seed, max_flop, reps, path, progress_every):
    if path and Path(path).exists():
        z = np.load(path, allow_pickle=False)
        print(f"[tune] loaded {path}: {len(z['X'])} shapes", flush=True)
        return z["X"], z["T"], [_dec_plain(r) for r in z["shapes"]]
    shapes = tuning.sample_plain_shapes(n, seed=seed, max_flop=max_flop)
    t0 = time.time()
    X, T = tuning.collect_plain(tuner_box(), shapes, progress=progress_every, reps=reps)
    print(f"[tune] timed {n} shapes x {T.shape[1]} configs in {time.time() - t0:.0f} s", flush=True)
    if path:
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, X=X, T=T, shapes=np.array([_enc_plain(s) for s in shapes], dtype=np.int64))
    return X, T, shapes


def _plain_anchor_set(tuning, tuner_box, reps, path):
    """The anchor shapes (tuning.anchor_plain_shapes), timed once and kept in `path`."""
    if Path(path).exists():
        z = np.load(path, allow_pickle=False)
        print(f"[tune] loaded {path}: {len(z['X'])} anchor shapes", flush=True)
        return z["X"], z["T"], [_dec_plain(r) for r in z["shapes"]]
    shapes = tuning.anchor_plain_shapes()
    t0 = time.time()
    X, T = tuning.collect_plain(tuner_box(), shapes, progress=max(1, len(shapes) // 4), reps=reps)
    print(f"[tune] timed {len(shapes)} anchor shapes x {T.shape[1]} configs in {time.time() - t0:.0f} s", flush=True)
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, X=X, T=T, shapes=np.array([_enc_plain(s) for s in shapes], dtype=np.int64))
    return X, T, shapes


def _enc_plain(s):
    return [s[0], ord(s[1]), ord(s[2]), *s[3:]]


def _dec_plain(row):
    return (int(row[0]), chr(int(row[1])), chr(int(row[2])), *(int(v) for v in row[3:]))


def main_plain(args):
    """The FP32/FP64 plain engine: exhaustive timing of its 17 configurations
    and of the built-in rule, the MLP trained on the GPU, and its held-out
    quality against both."""
    import torch

    import portblas_amd as pb
    from portblas_amd import tuning

    h = pb.SB_Handle(selector=None, plain_selector=None)
    nconf = tuning.num_plain_configs()
    tuner = []

    def tuner_box():
        if not tuner:
            tuner.append(tuning.PlainAutotuner(h))
        return tuner[0]

    t0 = time.time()
    train_files = args.train_files or [None]
    parts = [_plain_set(tuning, tuner_box, args.train, args.seed + i, args.max_flop, args.reps, f,
                        max(1, args.train // 10)) for i, f in enumerate(train_files)]
    Xtr = np.concatenate([p[0] for p in parts])
    Ttr = np.concatenate([p[1] for p in parts])
    if args.anchors:
        Xa, Ta, _ = _plain_anchor_set(tuning, tuner_box, args.reps, args.anchors)
        Xtr = np.concatenate([Xtr] + [Xa] * args.anchor_repeat)
        Ttr = np.concatenate([Ttr] + [Ta] * args.anchor_repeat)
    Xte, Tte, te_shapes = _plain_set(tuning, tuner_box, args.heldout, args.seed + 1000, args.max_flop, args.reps,
                                     args.heldout_file, max(1, args.heldout // 5))
    if args.collect_only:
        return {"train_shapes": int(len(Xtr)), "heldout_shapes": int(len(Xte))}
    tuner.clear()
    torch.cuda.empty_cache()
    t_collect = time.time() - t0
    Ytr = tuning.targets(Ttr[:, :nconf])
    perm = np.random.default_rng(args.seed).permutation(len(Xtr))
    n_val = int(round(args.val_frac * len(Xtr))) if args.candidates > 1 else 0
    val, fit = perm[:n_val], perm[n_val:]
    best_tr = Ttr[:, :nconf].min(axis=1)
    t1 = time.time()
    sel, best_score, val_scores = None, None, []
    for c in range(args.candidates):
        cand = tuning.Selector.init(sizes=(tuning.NUM_FEATURES, 64, 64, 64, nconf), seed=args.seed + c)
        cand.train(h, Xtr[fit], Ytr[fit], mask=tuning.weights(Ytr[fit], args.sharpness), steps=args.steps,
                   batch=1024, lr=args.lr, seed=args.seed + c, log_every=max(1, args.steps // 4))
        if n_val:
            ch = np.argmin(cand.predict_bf16(h, Xtr[val]), axis=1)
            ev = _ratio_stats(tuning.plain_time_of_choice(ch, Ttr[val]), best_tr[val])
            score = ev["geomean_slowdown"] * ev["total_time_ratio"]
        else:
            ev, score = None, 0.0
        val_scores.append({"seed": args.seed + c, "validation": ev})
        print(f"[tune] candidate {c}: validation {ev}", flush=True)
        if sel is None or score < best_score:
            sel, best_score = cand, score
    t_train = time.time() - t1
    chosen = np.argmin(sel.predict_bf16(h, Xte), axis=1)
    best_te = Tte[:, :nconf].min(axis=1)
    t_mlp = tuning.plain_time_of_choice(chosen, Tte)
    t_rule = Tte[:, nconf]
    tuning.install_plain(sel, h)
    host = np.array([tuning.select_plain(h, *s) for s in te_shapes])
    tuning.uninstall_plain(h)
    rep = {
        "family": "plain (FP32/FP64 _gemm)", "configs": nconf,
        "train_shapes": int(len(Xtr)), "heldout_shapes": int(len(Xte)),
        "timing_seconds": t_collect, "train_seconds": t_train,
        "training": {"steps": args.steps, "lr": args.lr, "loss_weight_sharpness": args.sharpness,
                     "fit_shapes": int(len(fit)), "validation_shapes": int(n_val), "candidates": val_scores},
        "mlp_bf16_vs_exhaustive": _ratio_stats(t_mlp, best_te),
        "builtin_rule_vs_exhaustive": _ratio_stats(t_rule, best_te),
        "mlp_gain_over_rule_geomean": float(np.exp(np.mean(np.log(t_rule / t_mlp)))),
        "mlp_gain_over_rule_total": float(t_rule.sum() / t_mlp.sum()),
        "host_dispatch_agrees_with_gpu_inference": float(np.mean(host == chosen)),
        "max_flop": args.max_flop,
    }
    print(json.dumps(rep, indent=1), flush=True)
    if args.out:
        sel.save(args.out)
    if args.fixture:
        Path(args.fixture).parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(args.fixture, X=Xte.astype(np.float32), T=Tte.astype(np.float32),
                            shapes=np.array([_enc_plain(s) for s in te_shapes], dtype=np.int64))
    if args.report:
        Path(args.report).parent.mkdir(parents=True, exist_ok=True)
        Path(args.report).write_text(json.dumps(rep, indent=1))
    return rep


if __name__ == "__main__":
    sys.exit(0 if main() else 1)

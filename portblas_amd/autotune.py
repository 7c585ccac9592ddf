"""Autotuning + learned-selector pipeline (GPU).

    python -m portblas_amd.autotune --train 20000 --heldout 2000 \
        --out portblas_amd/selector_mi355x.npz --report gpurun_out/tune_report.json

1. draws training and held-out shapes (M, N, K log-uniform in [16, 16384],
   batch in {1..256}, random transposes; work and footprint capped so one
   sweep stays bounded);
2. times every configuration of every shape (exhaustive autotuning,
   pb_tune_time_config);
3. trains the 8 -> 64 -> 64 -> 64 -> configs MLP on the GPU with the
   library's own kernels (FP32), then evaluates it in BF16 on the held-out
   shapes against the exhaustive optimum and against the built-in rule;
4. saves the network and a JSON report.
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--heldout", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--max-flop", type=float, default=2.0 ** 35)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    ap.add_argument("--dataset", default=None, help="save/load the timing dataset (npz)")
    args = ap.parse_args(argv)

    import torch

    import portblas_amd as pb
    from portblas_amd import tuning

    h = pb.SB_Handle()
    nconf = tuning.num_configs()
    t0 = time.time()
    if args.dataset and Path(args.dataset).exists():
        z = np.load(args.dataset, allow_pickle=False)
        Xtr, Ttr, Xte, Tte = z["Xtr"], z["Ttr"], z["Xte"], z["Tte"]
        te_shapes = [tuple(s) for s in z["te_shapes"].tolist()]
        print(f"[tune] loaded dataset {args.dataset}", flush=True)
    else:
        tr_shapes = tuning.sample_shapes(args.train, seed=args.seed, max_flop=args.max_flop)
        te_shapes = tuning.sample_shapes(args.heldout, seed=args.seed + 1000, max_flop=args.max_flop)
        tuner = tuning.Autotuner(h)
        prog = max(1, args.train // 20)
        Xtr, Ttr = tuning.collect(tuner, tr_shapes, progress=prog)
        Xte, Tte = tuning.collect(tuner, te_shapes, progress=max(1, args.heldout // 10))
        del tuner
        torch.cuda.empty_cache()
        if args.dataset:
            enc = lambda s: [ord(s[0]), ord(s[1]), *s[2:]]
            np.savez(args.dataset, Xtr=Xtr, Ttr=Ttr, Xte=Xte, Tte=Tte,
                     te_shapes=np.array([enc(s) for s in te_shapes], dtype=np.int64))
    t_collect = time.time() - t0
    print(f"[tune] timing: {len(Xtr)} + {len(Xte)} shapes x {nconf} configs in {t_collect:.0f} s", flush=True)

    sel = tuning.Selector.init(sizes=(tuning.NUM_FEATURES, 64, 64, 64, nconf), seed=args.seed)
    t1 = time.time()
    hist = sel.train(h, Xtr, tuning.targets(Ttr), steps=args.steps, batch=1024, lr=2e-3, seed=args.seed,
                     log_every=max(1, args.steps // 10))
    t_train = time.time() - t1

    pred = sel.predict_bf16(h, Xte)
    chosen_mlp = np.argmin(pred, axis=1)
    # built-in rule (no learned selector) for the same shapes
    tuning.uninstall(h)
    def enc_shape(s):
        ta, tb = (chr(s[0]), chr(s[1])) if isinstance(s[0], (int, np.integer)) else (s[0], s[1])
        return ta, tb, int(s[2]), int(s[3]), int(s[4]), int(s[5])
    shapes = [enc_shape(s) for s in te_shapes]
    chosen_rule = np.array([tuning.select(h, *s) for s in shapes])
    # the handle's own dispatch with the network installed agrees with the GPU inference
    sel.install(h)
    chosen_host = np.array([tuning.select(h, *s) for s in shapes])
    agree = float(np.mean(chosen_host == chosen_mlp))
    rep = {
        "train_shapes": int(len(Xtr)), "heldout_shapes": int(len(Xte)), "configs": nconf,
        "timing_seconds": t_collect, "train_seconds": t_train, "final_loss": hist[-1][1] if hist else None,
        "mlp_bf16": tuning.evaluate(chosen_mlp, Tte),
        "builtin_rule": tuning.evaluate(chosen_rule, Tte),
        "host_dispatch_agrees_with_gpu_inference": agree,
        "best_single_config": tuning.evaluate(np.full(len(Tte), int(np.argmin(np.log(Tte).mean(axis=0)))), Tte),
    }
    print(json.dumps(rep, indent=1), flush=True)
    if args.out:
        sel.save(args.out)
    if args.report:
        Path(args.report).parent.mkdir(parents=True, exist_ok=True)
        Path(args.report).write_text(json.dumps(rep, indent=1))
    return rep


if __name__ == "__main__":
    sys.exit(0 if main() else 1)

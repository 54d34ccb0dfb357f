"""Train and evaluate the learned heterogeneity correction on the GPU.

    python -m pyvoxeldosimetry_amd.learned.train --n-train 2000 --n-eval 200 --shape 128

Training triples (activity, density, Monte Carlo dose) are generated on the
GPU from seeds 10000.. and kept in HBM; held-out phantoms use seeds 20000..
Prints one JSON report (losses, errors vs Monte Carlo for plain and corrected
dose, timing) and saves the weights with ``torch.save``.
"""
from __future__ import annotations

import argparse
import json
import math
import time

TRAIN_SEED0 = 10000
EVAL_SEED0 = 20000


def build_training_set(corr, n, shape, histories, log=None):
    import torch

    from .pipeline import synth_phantom
    data = []
    t0 = time.perf_counter()
    for i in range(n):
        act, rho = synth_phantom(TRAIN_SEED0 + i, shape, corr.device)
        tgt = corr.mc_dose(act, rho, histories, seed=TRAIN_SEED0 + i)
        data.append((act, rho, tgt))
        if log and (i + 1) % max(1, n // 10) == 0:
            torch.cuda.synchronize()
            log(f"  training set {i + 1}/{n} ({time.perf_counter() - t0:.1f} s)")
    return data


def train(corr, data, epochs, lr, log=None):
    import torch
    history = []
    step = 0
    total = epochs * len(data)
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = torch.randperm(len(data), generator=torch.Generator().manual_seed(ep)).tolist()
        acc = torch.zeros(2, dtype=torch.float64, device=corr.device)
        for i in order:
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, total)))     # cosine schedule
            sums = corr.train_step(*data[i], lr=lr_t)
            acc += sums[:2].double()
            step += 1
        history.append(float((acc[0] / acc[1]).item()))
        if log:
            torch.cuda.synchronize()
            log(f"  epoch {ep + 1}/{epochs} done ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return history, time.perf_counter() - t0


def main(argv=None):
    import sys

    import torch

    from .pipeline import LearnedCorrector, evaluate, time_per_volume

    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=2000)
    ap.add_argument("--n-eval", type=int, default=200)
    ap.add_argument("--shape", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--train-histories", type=int, default=16)
    ap.add_argument("--eval-histories", type=int, default=128)
    ap.add_argument("--save", default="gpurun_out/cnn_y90_2mm.pt")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    shape = (args.shape,) * 3
    corr = LearnedCorrector("Y90", voxel_mm=2.0)
    log(f"building {args.n_train} training triples {shape}")
    data = build_training_set(corr, args.n_train, shape, args.train_histories, log)
    log("training")
    hist, t_train = train(corr, data, args.epochs, args.lr, log)
    del data
    log("evaluating")
    ev = evaluate(corr, range(EVAL_SEED0, EVAL_SEED0 + args.n_eval), shape, args.eval_histories)
    tm = time_per_volume(corr, shape)
    torch.save(corr.net.state_dict(), args.save)
    rep = {"n_train": args.n_train, "n_eval": args.n_eval, "shape": list(shape), "epochs": args.epochs,
           "train_s": round(t_train, 2), "loss_history": hist, "eval": ev, "timing": tm,
           "params": corr.net.trainable_parameters()}
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()

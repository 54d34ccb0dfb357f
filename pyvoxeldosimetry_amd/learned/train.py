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


def build_training_set(corr, n, shape, histories, log=None, seed0=TRAIN_SEED0):
    import torch

    from .pipeline import synth_phantom
    data = []
    t0 = time.perf_counter()
    for i in range(n):
        act, rho = synth_phantom(seed0 + i, shape, corr.device)
        tgt = corr.mc_dose(act, rho, histories, seed=seed0 + i)
        data.append((act, rho, tgt))
        if log and (i + 1) % max(1, n // 10) == 0:
            torch.cuda.synchronize()
            log(f"  training set {i + 1}/{n} ({time.perf_counter() - t0:.1f} s)")
    return data


def augment(triple, g):
    """Random axis flips and (for cubic grids) axis permutations: the transport is
    symmetric under both when voxels are isotropic and the kernel is symmetrised."""
    import torch
    dims = [d for d in range(3) if torch.rand(1, generator=g).item() < 0.5]
    out = [torch.flip(x, dims) if dims else x for x in triple]
    if len(set(out[0].shape)) == 1:
        perm = torch.randperm(3, generator=g).tolist()
        if perm != [0, 1, 2]:
            out = [x.permute(*perm) for x in out]
    return [x.contiguous() for x in out]


def train(corr, data, epochs, lr, log=None, aug=True, on_epoch=None):
    import torch
    history = []
    step = 0
    total = epochs * len(data)
    t0 = time.perf_counter()
    for ep in range(epochs):
        g = torch.Generator().manual_seed(ep)
        order = torch.randperm(len(data), generator=g).tolist()
        acc = torch.zeros(2, dtype=torch.float64, device=corr.device)
        for i in order:
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, total)))     # cosine schedule
            sums = corr.train_step(*(augment(data[i], g) if aug else data[i]), lr=lr_t)
            acc += sums[:2].double()
            step += 1
        history.append(float((acc[0] / acc[1]).item()))
        if log:
            torch.cuda.synchronize()
            log(f"  epoch {ep + 1}/{epochs} done ({time.perf_counter() - t0:.1f} s) loss {history[-1]:.3e}")
        if on_epoch is not None:
            on_epoch(ep, history[-1])
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
    ap.add_argument("--bound", type=float, default=0.5, help="|r| above which a voxel counts as out of bound")
    ap.add_argument("--no-augment", action="store_true")
    ap.add_argument("--val-every", type=int, default=0, help="validate on 8 held-out phantoms every N epochs")
    ap.add_argument("--save", default="cnn_weights.pt", help="weights + Adam state (torch.save)")
    ap.add_argument("--export", default=None, help="also write the FP32 weights as safetensors")
    ap.add_argument("--resume", default=None, help="continue from a --save file (weights and Adam state)")
    ap.add_argument("--init", default=None, help="start from these FP32 weights (safetensors), fresh Adam state")
    ap.add_argument("--seed0", type=int, default=TRAIN_SEED0, help="first training phantom seed")
    ap.add_argument("--nuclide", default="Y90", choices=["Y90", "Lu177"])
    ap.add_argument("--voxel", type=float, default=2.0, help="voxel size in mm")
    ap.add_argument("--no-photons", action="store_true", help="electrons-only Monte Carlo")
    ap.add_argument("--kernel-size", type=int, default=13, help="grid of the Monte Carlo point kernel of the plain dose")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    shape = (args.shape,) * 3
    corr = LearnedCorrector(args.nuclide, voxel_mm=args.voxel, bound=args.bound, photons=not args.no_photons,
                            kernel_size=args.kernel_size)
    if args.resume:
        corr.net.load_state_dict(torch.load(args.resume, weights_only=True))
        log(f"resumed from {args.resume} at Adam step {corr.net.step}")
    elif args.init:
        corr.net.load_weights(args.init)
        log(f"initialised from {args.init}")
    if args.seed0 + args.n_train > EVAL_SEED0 - 8:
        raise SystemExit("training seeds would overlap the validation/evaluation seeds")
    log(f"building {args.n_train} training triples {shape} from seed {args.seed0}")
    data = build_training_set(corr, args.n_train, shape, args.train_histories, log, seed0=args.seed0)
    val_log = []
    on_epoch = None
    if args.val_every > 0:
        from .pipeline import synth_phantom, voxel_errors
        val = []
        for s in range(EVAL_SEED0 - 8, EVAL_SEED0):     # disjoint from training and evaluation seeds
            act, rho = synth_phantom(s, shape, corr.device)
            val.append((act, rho, corr.mc_dose(act, rho, args.eval_histories, seed=s)))

        def on_epoch(ep, loss):
            if (ep + 1) % args.val_every:
                return
            e = [(voxel_errors(c.plain, ref), voxel_errors(c.dose, ref))
                 for c, ref in ((corr.correct(a, r), ref) for a, r, ref in val)]
            row = {"epoch": ep + 1, "loss": loss,
                   "plain_mean": sum(p.mean().item() for p, _ in e) / len(e),
                   "cnn_mean": sum(c.mean().item() for _, c in e) / len(e),
                   "plain_within": sum((p <= 0.02).float().mean().item() for p, _ in e) / len(e),
                   "cnn_within": sum((c <= 0.02).float().mean().item() for _, c in e) / len(e)}
            val_log.append(row)
            log(f"  val {row}")
    log("training")
    hist, t_train = train(corr, data, args.epochs, args.lr, log, aug=not args.no_augment, on_epoch=on_epoch)
    del data
    log("evaluating")
    ev = evaluate(corr, range(EVAL_SEED0, EVAL_SEED0 + args.n_eval), shape, args.eval_histories)
    tm = time_per_volume(corr, shape)
    torch.save(corr.net.state_dict(), args.save)
    if args.export:
        corr.net.save_weights(args.export, {"nuclide": args.nuclide, "voxel_mm": args.voxel, "shape": args.shape,
                                            "mc_photons": "0" if args.no_photons else "1",
                                            "kernel_size": args.kernel_size,
                                            "command": "python -m pyvoxeldosimetry_amd.learned.train "
                                                       + " ".join(sys.argv[1:] if argv is None else argv),
                                            "train_seeds": f"{args.seed0}..{args.seed0 + args.n_train - 1}",
                                            "init": args.init or "",
                                            "train_histories": args.train_histories, "epochs": args.epochs,
                                            "adam_steps": corr.net.step, "eval": json.dumps(ev)})
    rep = {"n_train": args.n_train, "n_eval": args.n_eval, "shape": list(shape), "epochs": args.epochs,
           "train_s": round(t_train, 2), "loss_history": hist, "validation": val_log, "eval": ev, "timing": tm,
           "params": corr.net.trainable_parameters()}
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()

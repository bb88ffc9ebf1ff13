"""Train the learned playout surrogate and measure it.

    python -m leela_amd.train_surrogate --n 200000 --seed0 10000 \\
        --playouts 256 --labels labels.npy --epochs 2 --batch 512 \\
        --lr 1e-3 --ckpt surrogate.ckpt

Training positions are generator positions from seeds seed0..seed0+n-1
(move counts uniform in 0..250); their labels are the mean outcome of
`--playouts` playouts each (playouts of position i seeded
playout_seed0 + i), read from `--labels` (a float32 .npy of n win rates)
when given, generated otherwise (GPU playouts; `--host-labels` for the
host's, which are equal). Training is Adam on minibatches of
`--batch` (learning rate ramped over `--warmup` steps), with a checkpoint
every `--ckpt-every` steps and at the end.

The evaluation labels held-out positions (seeds eval_seed0..) with a fresh
`--eval-playouts` mean and reports the mean absolute error of the
surrogate, of the constant predictor and of k-playout estimates, plus the
surrogate's inference rate at batch `--infer-batch` and the playout rate on
`--threads` host threads. Prints one JSON line.

`--policy pattern` labels with the reference's pattern playout policy
(every label, training and evaluation, from that teacher); both policies
run on the GPU (the same values as on the host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=200000)
    ap.add_argument("--seed0", type=int, default=10000)
    ap.add_argument("--playouts", type=int, default=256)
    ap.add_argument("--playout-seed0", type=int, default=1_000_000)
    ap.add_argument("--labels", default=None)
    ap.add_argument("--epochs", type=float, default=2.0)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--ckpt", default=None)
    ap.add_argument("--ckpt-every", type=int, default=1000)
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--eval-n", type=int, default=4096)
    ap.add_argument("--eval-seed0", type=int, default=300000)
    ap.add_argument("--eval-playouts", type=int, default=256)
    ap.add_argument("--eval-playout-seed0", type=int, default=2_000_000)
    ap.add_argument("--infer-batch", type=int, default=4096)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--log-every", type=int, default=200)
    ap.add_argument("--policy", choices=("light", "pattern"),
                    default="light",
                    help="playout policy of the labels (pattern: the "
                         "reference's)")
    ap.add_argument("--host-labels", action="store_true",
                    help="label with host playouts (default: GPU playouts, "
                         "the same values)")
    args = ap.parse_args(argv)

    import torch
    from . import positions
    from .surrogate import Surrogate, train

    policy = (positions.PLAYOUT_PATTERN if args.policy == "pattern"
              else positions.PLAYOUT_LIGHT)
    host = args.host_labels

    def make_labels(n, seed0, playouts, playout_seed0, planes=True):
        # GPU playouts unless asked for the host's (equal values).
        if host:
            return positions.value_labels(n, seed0, playouts, playout_seed0,
                                          threads=args.threads, planes=planes,
                                          policy=policy)
        return positions.value_labels_device(n, seed0, playouts,
                                             playout_seed0, planes=planes,
                                             policy=policy)

    t0 = time.perf_counter()
    if args.labels and os.path.exists(args.labels):
        labels = np.load(args.labels, allow_pickle=False).astype(np.float32)
        if labels.shape != (args.n,):
            raise SystemExit(f"{args.labels}: expected {args.n} labels, "
                             f"found {labels.shape}")
        planes = positions.make_batch(args.n, seed0=args.seed0,
                                      net=positions.NET_VALUE,
                                      threads=args.threads)
    else:
        planes, labels = make_labels(args.n, args.seed0, args.playouts,
                                     args.playout_seed0)
        if args.labels:
            np.save(args.labels, labels)
    t_data = time.perf_counter() - t0
    print(f"data: {args.n} positions in {t_data:.1f} s", file=sys.stderr,
          flush=True)

    sur = Surrogate(0, max_batch=max(args.batch, args.infer_batch),
                    seed=args.seed)
    if args.resume and args.ckpt and os.path.exists(args.ckpt):
        sur.load(args.ckpt)
    steps = int(args.epochs * args.n / args.batch)
    done = sur.step
    losses = []
    t0 = time.perf_counter()
    while done < steps:
        k = min(args.ckpt_every, steps - done)
        losses += [(done + i, l) for i, l in
                   train(sur, planes, labels, k, batch=args.batch, lr=args.lr,
                         seed=args.seed + done, log_every=args.log_every,
                         warmup=args.warmup)]
        done += k
        if args.ckpt:
            sur.save(args.ckpt)
        print(f"step {done}/{steps} loss {losses[-1][1]:.5f}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0

    # Held-out evaluation against a fresh playout mean.
    ev_planes, ev_y = make_labels(args.eval_n, args.eval_seed0,
                                  args.eval_playouts, args.eval_playout_seed0)
    pred = sur.predict(ev_planes)
    out = {"mae_surrogate": float(np.mean(np.abs(pred - ev_y))),
           "mae_constant": float(np.mean(np.abs(labels.mean() - ev_y)))}
    for k in (1, 4, 16, 64):
        _, est = make_labels(args.eval_n, args.eval_seed0, k,
                             args.eval_playout_seed0 + 7_000_000, planes=False)
        out[f"mae_{k}_playouts"] = float(np.mean(np.abs(est - ev_y)))
    # Fresh labels of training positions (generalisation gap).
    m = min(args.eval_n, args.n)
    _, tr_fresh = make_labels(m, args.seed0, args.eval_playouts,
                              args.eval_playout_seed0 + 9_000_000,
                              planes=False)
    out["mae_surrogate_train_positions"] = float(
        np.mean(np.abs(sur.predict(planes[:m]) - tr_fresh)))

    # Inference rate at the serving batch, inputs resident on the device.
    B = args.infer_batch
    reps = max(1, -(-B // args.eval_n))
    X = torch.from_numpy(np.concatenate([ev_planes] * reps)[:B]).cuda()
    W = torch.empty(B, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        sur.forward_device(B, X.data_ptr(), W.data_ptr(), st)
    torch.cuda.synchronize()
    iters = 20
    t0 = time.perf_counter()
    for _ in range(iters):
        sur.forward_device(B, X.data_ptr(), W.data_ptr(), st)
    torch.cuda.synchronize()
    t_inf = (time.perf_counter() - t0) / iters
    # Playout rate on the host threads (256-playout estimates).
    t0 = time.perf_counter()
    npo = max(args.threads, 64)
    positions.value_labels(npo, args.eval_seed0, 256, 123, threads=args.threads,
                           planes=False, policy=policy)
    t_po = time.perf_counter() - t0
    out.update({
        "train_positions": args.n, "label_playouts": args.playouts,
        "labels": "host playouts" if host else "GPU playouts",
        "policy": args.policy,
        "data_seconds": round(t_data, 2),
        "steps": steps, "batch": args.batch, "lr": args.lr,
        "first_loss": losses[0][1] if losses else None,
        "final_loss": losses[-1][1] if losses else None,
        "train_seconds": round(t_train, 2),
        "train_positions_per_s": round(steps * args.batch / t_train, 1),
        "eval_positions": args.eval_n, "eval_playouts": args.eval_playouts,
        "infer_batch": B, "infer_ms": round(t_inf * 1e3, 3),
        "infer_positions_per_s": round(B / t_inf, 1),
        "playout_estimates_per_s": round(npo / t_po, 2),
        "playout_threads": args.threads,
    })
    sur.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

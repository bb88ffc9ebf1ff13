#!/usr/bin/env python3
"""Benchmark: NN positions/sec of Leela's policy+value leaf evaluation on
one MI355X, batch 1024 of 19x19 positions, inputs resident in HBM.

A step is one policy forward plus one value forward over the batch (the
two nets every leaf evaluation runs). Prints one JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0}
HBM_PEAK_GBS = 8000.0


def make_planes(n: int) -> np.ndarray:
    from leela_amd import positions
    return positions.batch_planes(n, net=0, seed0=1)


def cpu_baseline(pol, val, planes, budget_s: float, threads: int):
    """Reference CPU path (oracle restatement: im2col + OpenBLAS sgemm +
    bias + ELU, FP32) on a bounded sample of the same positions."""
    import oracle
    blas = oracle.load_blas()
    pnet, vnet = oracle.Net(pol), oracle.Net(val)
    n = min(len(planes), 2 * threads)
    t0 = time.perf_counter()
    oracle.forward_f32(pnet, planes[:n], oracle.HEAD_POLICY, threads=threads)
    oracle.forward_f32(vnet, planes[:n], oracle.HEAD_VALUE, threads=threads)
    dt = time.perf_counter() - t0
    per = dt / n
    m = int(max(n, min(len(planes), budget_s / per)))
    t0 = time.perf_counter()
    oracle.forward_f32(pnet, planes[:m], oracle.HEAD_POLICY, threads=threads)
    oracle.forward_f32(vnet, planes[:m], oracle.HEAD_VALUE, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(m / dt, 2), "unit": "positions/s", "cores": threads,
            "kind": "port",
            "sample": f"{m} of the benchmark's positions, policy+value, "
                      f"{threads} worker threads x 1 BLAS thread, {blas}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--precision", default="auto",
                    choices=["auto", "fp32", "bf16", "fp8"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(leaf positions are independent; see DESIGN.md on sharding).",
              file=sys.stderr)
        return 2

    import torch
    import leela_amd
    from leela_amd import nets

    prec_name = args.precision
    if prec_name == "auto":
        prec_name = "bf16"
    prec = leela_amd.PRECISIONS[prec_name]
    B = args.batch

    pol, val = nets.weight_set("A")
    be = leela_amd.Backend(0)
    be.load_net(leela_amd.NET_POLICY, pol)
    be.load_net(leela_amd.NET_VALUE, val)
    be.finalize()

    planes = make_planes(B)
    dev = torch.device("cuda:0")
    d_planes = torch.from_numpy(planes).to(dev)
    d_praw = torch.empty((B, 361), dtype=torch.float32, device=dev)
    d_pout = torch.empty((B, 361), dtype=torch.float32, device=dev)
    d_vraw = torch.empty((B, 1), dtype=torch.float32, device=dev)
    d_vout = torch.empty((B, 1), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step(p):
        be.forward_device(leela_amd.NET_POLICY, p, B, d_planes.data_ptr(),
                          d_praw.data_ptr(), d_pout.data_ptr(), stream)
        be.forward_device(leela_amd.NET_VALUE, p, B, d_planes.data_ptr(),
                          d_vraw.data_ptr(), d_vout.data_ptr(), stream)

    def timed(p, steps, warmup):
        for _ in range(warmup):
            step(p)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(p)
        torch.cuda.synchronize(dev)
        return time.perf_counter() - t0

    # FP32 plain form on the same positions and weights (the comparison
    # the metric names), then the measured precision.
    fp32_dt = timed(leela_amd.PREC_FP32, max(2, args.steps // 4), 1)
    fp32_value = B * max(2, args.steps // 4) / fp32_dt
    step(leela_amd.PREC_FP32)
    torch.cuda.synchronize(dev)
    fp32_logits = d_praw.cpu().numpy().copy()
    fp32_win = d_vout.cpu().numpy().copy()

    dt = timed(prec, args.steps, args.warmup)
    ms_per_step = dt * 1e3 / args.steps
    value = B * args.steps / dt
    step(prec)
    torch.cuda.synchronize(dev)
    logits = d_praw.cpu().numpy()
    win = d_vout.cpu().numpy()
    top1 = float(np.mean(logits.argmax(1) == fp32_logits.argmax(1)))
    dwin = float(np.abs(win - fp32_win).max())

    # Dominant kernel, timed per launch with HIP events on its stream.
    reps = 5
    kern_ms = []
    for _ in range(reps):
        ms = be.forward_timed(leela_amd.NET_POLICY, prec, B,
                              d_planes.data_ptr(), d_praw.data_ptr(),
                              d_pout.data_ptr(), stream)
        if prec_name == "fp32":
            kern_ms += list(ms[2:12])  # launches 2..11: conv3..conv12
        else:
            kern_ms += list(ms)        # one fused launch per forward
    avg_ms = float(np.mean(kern_ms))
    if prec_name == "fp32":
        kflops = 2.0 * 128 * 1152 * 361 * B  # one trunk layer, whole batch
        kernel = "conv_f32_kernel<3> (policy trunk 3x3 128->128)"
    else:
        kflops = float(nets.flops_per_position(nets.POLICY_SHAPES)) * B
        kernel = (f"fused_net_{prec_name} policy (13 convs + softmax, "
                  f"{B} boards per launch)")
    achieved = kflops / (avg_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[prec_name]

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(pol, val, planes, args.cpu_seconds, threads)

    line = {
        "metric": "NN positions/sec, BF16 vs FP32 plain, batch 1024 at "
                  ">=99.5% top-1 agreement",
        "value": round(value, 1),
        "unit": "positions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec_name,
        "data": "synthetic",
        "config": {
            "workload": f"Leela policy (13 conv, 128-wide) + value (12 conv "
                        f"64-wide + 2 FC) forward, batch {B} of 19x19",
            "batch": B,
            "precision": prec_name,
            "weights": "He-normal seed 7 (set A), Network.cpp:82-137 shapes",
            "positions": "random legal play, seeds 1..B",
            "parallelism": "single GPU",
            "fp32_plain_positions_per_s": round(fp32_value, 1),
            "speedup_vs_fp32_plain": round(value / fp32_value, 3),
            "top1_agreement_vs_fp32_plain": round(top1, 5),
            "max_abs_value_diff_vs_fp32_plain": round(dwin, 6),
        },
        "roofline": {
            "bound": "mfma",
            "kernel": kernel,
            "achieved": round(achieved, 2),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": None,
            "avg_launch_ms": round(avg_ms, 4),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    be.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

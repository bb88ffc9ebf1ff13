#!/usr/bin/env python3
"""Benchmark: NN positions/sec of Leela's policy+value leaf evaluation on
one MI355X, batch 1024 of 19x19 positions, inputs resident in HBM.

A step is one policy forward plus one value forward over the batch (the two
nets every leaf evaluation runs). The FP32 plain form, the BF16 form, the
FP16-operand form and the FP8 form run on the same positions and weights.
The CPU-baseline leg times the reference CPU path (the oracle's restatement)
and computes the FP64 oracle's outputs for the batch, against which every
form's top-1 agreement, KL, max-abs policy error and value error are
reported. The headline `value` is the fastest 16-bit form whose policy
top-1 agrees with the FP64 oracle on >= 99.5% of the batch; the process
exits with status 3 (after printing) when no form meets that bar. Also
reported: the ragged mixed-board workload (1024 x 19x19 + 512 x 13x13 +
512 x 9x9 in one launch, autotuned, and through the leaf queue from 16
threads). Prints one JSON line.
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

PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0, "fp8": 5000.0}
AGREEMENT_BAR = 0.995

# HBM bytes per launch of the roofline kernel at batch 1024, measured with
# rocprofv3 FETCH_SIZE / WRITE_SIZE (one counter per run) and scaled by the
# counters' reading for a known-size elementwise copy (reads x2.0, writes
# x1.0). Not measured in this run: the figures, their commands and the
# build they were measured on are in TRAFFIC_SOURCE. Keyed by form.
TRAFFIC_SOURCE = "profiles/fused_net_policy_hbm_traffic_b1024.md"
TRAFFIC_B1024 = {
    "fp16": {"bytes": 75_239_424, "fetch_kib": 35266.0, "write_kib": 2944.0},
    "bf16": {"bytes": 52_370_432, "fetch_kib": 24099.5, "write_kib": 2944.0},
    "fp32": {"bytes": 579_100_000, "fetch_kib": 190213.0,
             "write_kib": 185100.0},
}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """Cores this process may run on (its CPU affinity)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def fp64_reference(pol, val, planes, vplanes, threads: int):
    """The FP64 oracle's outputs for the batch (CPU-baseline leg: the
    oracle as the checker of every form's accuracy)."""
    import oracle
    oracle.load_blas()
    t0 = time.perf_counter()
    p = oracle.forward_f64(oracle.Net(pol), planes, oracle.HEAD_POLICY,
                           threads=threads, gemm=True)
    v = oracle.forward_f64(oracle.Net(val), vplanes, oracle.HEAD_VALUE,
                           threads=threads, gemm=True)
    return p, v, time.perf_counter() - t0


def accuracy_vs_fp64(logits, probs, win, ref):
    """Top-1 agreement, mean KL(FP64 || form), max |d prob|, max |d logit|
    and max |d win rate| against the FP64 oracle."""
    (raw64, p64), (_, w64) = ref
    q = np.maximum(probs.astype(np.float64), 1e-30)
    p = np.maximum(p64, 1e-30)
    kl = np.sum(p * (np.log(p) - np.log(q)), axis=1)
    return {"top1_vs_fp64": round(float(np.mean(logits.argmax(1) ==
                                                raw64.argmax(1))), 5),
            "mean_kl_vs_fp64": float(f"{kl.mean():.4g}"),
            "max_abs_policy_err_vs_fp64": float(
                f"{np.abs(probs - p64).max():.4g}"),
            "max_abs_logit_err_vs_fp64": float(
                f"{np.abs(logits - raw64).max():.4g}"),
            "max_abs_value_err_vs_fp64": float(
                f"{np.abs(win.ravel() - w64.ravel()).max():.4g}")}


def cpu_baseline(pol, val, planes, vplanes, budget_s: float, threads: int):
    """Reference CPU path (oracle restatement: im2col + OpenBLAS sgemm +
    bias + ELU, FP32) on a bounded sample of the same positions, with
    `threads` workers x 1 BLAS thread (Network.cpp:236-250) and with one."""
    import oracle
    from leela_amd import positions
    blas = oracle.load_blas()
    pnet, vnet = oracle.Net(pol), oracle.Net(val)
    pos = [positions.Position.random(1 + i) for i in range(len(planes))]

    def rate(t, budget):
        n = min(len(planes), 2 * t)
        t0 = time.perf_counter()
        oracle.forward_f32(pnet, planes[:n], oracle.HEAD_POLICY, threads=t)
        oracle.forward_f32(vnet, vplanes[:n], oracle.HEAD_VALUE, threads=t)
        per = (time.perf_counter() - t0) / n
        m = int(max(n, min(len(planes), budget / per)))
        t0 = time.perf_counter()
        oracle.forward_f32(pnet, planes[:m], oracle.HEAD_POLICY, threads=t)
        oracle.forward_f32(vnet, vplanes[:m], oracle.HEAD_VALUE, threads=t)
        t_nn = time.perf_counter() - t0
        # The feature encoders on the same positions and threads (netbench
        # times them with the nets: each thread encodes, then evaluates).
        t0 = time.perf_counter()
        positions.encode_batch(pos[:m], positions.NET_POLICY, 0, t)
        positions.encode_batch(pos[:m], positions.NET_VALUE, 0, t)
        t_enc = time.perf_counter() - t0
        return m, m / t_nn, m / (t_nn + t_enc)

    m, v, ve = rate(threads, budget_s)
    m1, v1, ve1 = rate(1, budget_s / 3)
    return {"value": round(v, 2), "unit": "positions/s", "cores": threads,
            "kind": "port",
            "sample": f"{m} of the benchmark's positions, policy+value, "
                      f"{threads} worker threads x 1 BLAS thread (every "
                      f"core this process may use), {blas}",
            "value_1_thread": round(v1, 2), "sample_1_thread": m1,
            "with_encoder": round(ve, 2),
            "with_encoder_1_thread": round(ve1, 2),
            "host_cores_available": host_cores(),
            "host_cpus_online": os.cpu_count(),
            "cpu_model": _cpu_model()}


def ragged_leg(be, leela_amd, nets, positions, p, stream, dev, torch,
               steps):
    """The mixed-board workload (1024 x 19x19 + 512 x 13x13 + 512 x 9x9):
    one ragged launch per net after autotuning the launch shape for this
    mix, timed with HIP events on the benchmark stream; then the same
    boards through the leaf queue from 16 threads (max batch 2048)."""
    if p not in (leela_amd.PREC_BF16, leela_amd.PREC_FP16):
        return None
    sides, pol, val = positions.ragged_mix()
    n = len(sides)
    tune = be.autotune_ragged(p, sides)
    chosen = [ln for ln in tune.splitlines() if "chosen" in ln]
    pp = torch.from_numpy(np.concatenate([b.ravel() for b in pol])).to(dev)
    vp = torch.from_numpy(np.concatenate([b.ravel() for b in val])).to(dev)
    pts = sum(s * s for s in sides)
    praw = torch.empty(pts, dtype=torch.float32, device=dev)
    pout = torch.empty_like(praw)
    vraw = torch.empty(n, dtype=torch.float32, device=dev)
    vout = torch.empty_like(vraw)
    ext = torch.cuda.ExternalStream(stream)

    def rstep():
        be.forward_ragged_device(leela_amd.NET_POLICY, p, sides,
                                 pp.data_ptr(), praw.data_ptr(),
                                 pout.data_ptr(), stream)
        be.forward_ragged_device(leela_amd.NET_VALUE, p, sides,
                                 vp.data_ptr(), vraw.data_ptr(),
                                 vout.data_ptr(), stream)

    for _ in range(3):
        rstep()
    torch.cuda.synchronize(dev)
    k = max(5, steps // 2)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * k)]
    t0 = time.perf_counter()
    for i in range(k):
        ev[2 * i].record(ext)
        rstep()
        ev[2 * i + 1].record(ext)
    torch.cuda.synchronize(dev)
    wall = (time.perf_counter() - t0) / k
    dev_ms = float(np.median([ev[2 * i].elapsed_time(ev[2 * i + 1])
                              for i in range(k)]))
    conv = sum(2 * co * ci * f * f for kd, f, ci, co in
               nets.POLICY_SHAPES + nets.VALUE_SHAPES if kd == nets.CONV)
    fc = sum(2 * co * ci for kd, f, ci, co in nets.VALUE_SHAPES
             if kd == nets.FC)
    flops = conv * pts + fc * n
    tf = flops / (dev_ms * 1e-3) / 1e12

    be.queue_start(p, max_batch=2048, max_wait_us=20000)
    try:
        be.queue_benchmark(sides, pol, val, threads=16, rounds=1)
        st0 = be.queue_stats()
        q_pos_s = be.queue_benchmark(sides, pol, val, threads=16, rounds=4)
        st1 = be.queue_stats()
    finally:
        be.queue_stop()
    return {
        "workload": "1024 x 19x19 + 512 x 13x13 + 512 x 9x9 generator "
                    "positions (move counts uniform in 0..0.6 x area), "
                    "policy + value, one ragged launch per net",
        "boards": n,
        "positions_per_s": round(n / (dev_ms * 1e-3), 1),
        "positions_per_s_wall": round(n / wall, 1),
        "ms_per_step": round(dev_ms, 4),
        "tflops": round(tf, 2),
        "frac_of_peak": round(tf / 2500.0, 4),
        "flops_per_step": flops,
        "flops_rule": "conv layers x board points (P/361 of a 19x19 "
                      "board's), value inner products per board",
        "autotune": chosen,
        "queue_16_threads_positions_per_s": round(q_pos_s, 1),
        "queue_launches_per_bucket": {
            str(b): int(st1["per_bucket"][b] - st0["per_bucket"][b])
            for b in (64, 256, 1024, 2048)},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--precision", default="auto",
                    choices=["auto", "fp32", "bf16", "fp16"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-only",
                    choices=["fp32", "bf16", "fp16", "encoder", "playouts"],
                    help="profiling workload: only `steps` policy forwards "
                         "of this form, or `steps` device-encoder launches "
                         "over 4096 boards (both nets), or one launch of "
                         "2048 positions x 64 GPU playouts (plus "
                         "--copy-mib), no JSON contract")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="with --kernel-only: a fill and an elementwise copy "
                         "of this many MiB, the known-size reference for PMC "
                         "byte counts")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (--gpus 1): leaf positions are "
              "independent, so more GPUs would each run their own share of "
              "the batches; no multi-GPU path is implemented.",
              file=sys.stderr)
        return 2

    import torch
    import leela_amd
    from leela_amd import nets, positions

    B = args.batch
    pol, val = nets.weight_set("A")
    be = leela_amd.Backend(0)
    be.load_net(leela_amd.NET_POLICY, pol)
    be.load_net(leela_amd.NET_VALUE, val)
    be.finalize()

    planes = positions.batch_planes(B, net=0, seed0=1)
    vplanes = positions.batch_planes(B, net=1, seed0=1)
    dev = torch.device("cuda:0")
    d_planes = torch.from_numpy(planes).to(dev)
    d_vplanes = torch.from_numpy(vplanes).to(dev)
    d_praw = torch.empty((B, 361), dtype=torch.float32, device=dev)
    d_pout = torch.empty((B, 361), dtype=torch.float32, device=dev)
    d_vraw = torch.empty((B, 1), dtype=torch.float32, device=dev)
    d_vout = torch.empty((B, 1), dtype=torch.float32, device=dev)
    # All device work of the benchmark runs on one non-default stream (the
    # default stream's implicit synchronisation would sit inside the
    # per-launch event timing).
    bench_stream = torch.cuda.Stream(dev)
    stream = bench_stream.cuda_stream
    torch.cuda.synchronize(dev)

    def step(p, st=None):
        st = stream if st is None else st
        be.forward_device(leela_amd.NET_POLICY, p, B, d_planes.data_ptr(),
                          d_praw.data_ptr(), d_pout.data_ptr(), st)
        be.forward_device(leela_amd.NET_VALUE, p, B, d_vplanes.data_ptr(),
                          d_vraw.data_ptr(), d_vout.data_ptr(), st)

    def settle(p, seconds=0.3):
        # Back-to-back steps until the device has been busy for `seconds`:
        # after an idle gap (host-side work) the first launches run at a
        # lower clock, which would otherwise leak into the timings below.
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            for _ in range(4):
                step(p)
            torch.cuda.synchronize(dev)

    def run(name, steps, warmup):
        p = leela_amd.PRECISIONS[name]
        settle(p)
        for _ in range(warmup):
            step(p)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(p)
        torch.cuda.synchronize(dev)
        dt = time.perf_counter() - t0
        logits = d_praw.cpu().numpy().copy()
        probs = d_pout.cpu().numpy().copy()
        win = d_vout.cpu().numpy().copy()
        return {"dt": dt, "steps": steps, "pos_s": B * steps / dt,
                "logits": logits, "probs": probs, "win": win}

    if args.kernel_only:
        if args.kernel_only == "encoder":
            n = 4 * B
            enc = positions.DeviceEncoder(0, max_batch=n)
            eb = positions.PositionBatch(
                [positions.Position.random(1 + i) for i in range(n)])
            epl = torch.empty((n, 32, 361), dtype=torch.uint8, device=dev)
            evpl = torch.empty_like(epl)
            for _ in range(args.warmup + args.steps):
                enc.encode_both_into(eb, epl.data_ptr(), evpl.data_ptr(), 0,
                                     stream)
            torch.cuda.synchronize(dev)
            enc.close()
        elif args.kernel_only == "playouts":
            n = 2048
            enc = positions.DeviceEncoder(0, max_batch=n)
            pb = positions.PositionBatch(
                [positions.Position.random(10000 + i) for i in range(n)])
            w = torch.empty(n, dtype=torch.float32, device=dev)
            enc.playouts_into(pb, 64, 5, w.data_ptr(), stream)
            torch.cuda.synchronize(dev)
            enc.close()
        else:
            p = leela_amd.PRECISIONS[args.kernel_only]
            for _ in range(args.warmup + args.steps):
                be.forward_device(leela_amd.NET_POLICY, p, B,
                                  d_planes.data_ptr(), d_praw.data_ptr(),
                                  d_pout.data_ptr(), stream)
        if args.copy_mib:
            # Fill (writes N MiB), then y = x + 1 (reads N MiB, writes N
            # MiB): the known-size kernels that calibrate the byte counters.
            src = torch.ones(args.copy_mib << 18, dtype=torch.float32,
                             device=dev)
            dst = torch.empty_like(src)
            torch.add(src, 1.0, out=dst)
        torch.cuda.synchronize(dev)
        print(json.dumps({"kernel_only": args.kernel_only, "batch": B,
                          "policy_forwards": args.warmup + args.steps,
                          "copy_bytes": args.copy_mib << 20}))
        be.close()
        return 0

    forms = ["fp32", "bf16", "fp16", "fp8"]
    res = {}
    for name in forms:
        steps = args.steps if name != "fp32" else max(2, args.steps // 4)
        res[name] = run(name, steps, args.warmup)
    ref = res["fp32"]
    for name in forms:
        r = res[name]
        r["top1"] = float(np.mean(r["logits"].argmax(1) ==
                                  ref["logits"].argmax(1)))
        r["dwin"] = float(np.abs(r["win"] - ref["win"]).max())

    # CPU-baseline leg, part 1: the FP64 oracle's outputs for the batch, the
    # checker of every form's agreement (without it, agreement is judged
    # against the FP32 plain form, itself within 1e-4 of the CPU path).
    cores = host_cores()
    ref64 = None
    fp64_s = None
    if not args.no_cpu_baseline:
        p64, v64, fp64_s = fp64_reference(pol, val, planes, vplanes, cores)
        ref64 = (p64, v64)
        for name in forms:
            r = res[name]
            r["acc64"] = accuracy_vs_fp64(r["logits"], r["probs"], r["win"],
                                          ref64)
            r["agree"] = r["acc64"]["top1_vs_fp64"]
    else:
        for name in forms:
            res[name]["agree"] = res[name]["top1"]

    head = args.precision
    ok = [n for n in ("bf16", "fp16") if res[n]["agree"] >= AGREEMENT_BAR]
    if head == "auto":
        head = max(ok, key=lambda n: res[n]["pos_s"]) if ok else "bf16"
    agreement_met = head in ok
    if head != "fp32" and res[head]["steps"] != args.steps:
        res[head] = {**res[head], **run(head, args.steps, args.warmup)}
    h = res[head]
    value = h["pos_s"]
    ms_per_step = h["dt"] * 1e3 / h["steps"]

    # Host-buffer boundary (lz_forward: planes in from host memory, outputs
    # back), policy + value per step: the PCIe-inclusive rate, reported
    # beside the HBM-resident headline, never as `value`.
    p = leela_amd.PRECISIONS[head]
    be.forward(leela_amd.NET_POLICY, p, planes)
    be.forward(leela_amd.NET_VALUE, p, vplanes)
    t0 = time.perf_counter()
    host_steps = max(2, args.steps // 4)
    for _ in range(host_steps):
        be.forward(leela_amd.NET_POLICY, p, planes)
        be.forward(leela_amd.NET_VALUE, p, vplanes)
    host_pos_s = B * host_steps / (time.perf_counter() - t0)

    # The policy + value step captured once as a hipGraph and replayed:
    # per-replay HIP events, median of 50 after 10 warm-up replays.
    graph_pos_s = None
    try:
        gs = torch.cuda.Stream(dev)
        with torch.cuda.stream(gs):
            for _ in range(2):
                step(p, gs.cuda_stream)
        torch.cuda.synchronize(dev)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=gs):
            step(p, gs.cuda_stream)
        times = []
        for i in range(60):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            graph.replay()
            e1.record()
            e1.synchronize()
            if i >= 10:
                times.append(e0.elapsed_time(e1))
        graph_pos_s = B / (float(np.median(times)) * 1e-3)
        del graph
    except RuntimeError as err:  # capture unsupported: report, do not fail
        print(f"graph replay skipped: {err}", file=sys.stderr)

    # Host feature encoder (both nets' planes from the batch's positions,
    # lz_encode_batch) on the host threads.
    enc_threads = min(16, os.cpu_count() or 1)
    pos_objs = [positions.Position.random(1 + i) for i in range(B)]
    positions.encode_batch(pos_objs[:64], leela_amd.NET_POLICY, 0, enc_threads)
    t0 = time.perf_counter()
    for _ in range(3):
        positions.encode_batch(pos_objs, leela_amd.NET_POLICY, 0, enc_threads)
        positions.encode_batch(pos_objs, leela_amd.NET_VALUE, 0, enc_threads)
    enc_pos_s = 3 * B / (time.perf_counter() - t0)

    # Device encoder (lz_encode_device_both): both nets' planes computed on
    # the GPU from the positions (host side: one record per board, one
    # copy). End to end: positions in host memory -> planes -> policy +
    # value outputs on the benchmark stream, one encoder launch per E_AHEAD
    # net batches (a launch lasts as long as its slowest ladder reading,
    # which a bigger launch shares among more boards). The host packs the
    # next launch's records while the device runs the previous batches.
    # (On a second stream the encoder's waves hold CUs that the persistent
    # net kernels wait for: measured slower than in line.)
    E_AHEAD = 4
    denc = positions.DeviceEncoder(0, max_batch=E_AHEAD * B)
    pbatch = positions.PositionBatch(pos_objs)
    ebatch = positions.PositionBatch(
        [positions.Position.random(1 + i) for i in range(E_AHEAD * B)])
    d_epl = torch.empty((E_AHEAD * B,) + tuple(d_planes.shape[1:]),
                        dtype=torch.uint8, device=dev)
    d_evpl = torch.empty_like(d_epl)

    def encode_step(batch=pbatch):
        denc.encode_both_into(batch, d_epl.data_ptr(), d_evpl.data_ptr(), 0,
                              stream)

    def e2e_step(k):
        encode_step(ebatch)
        for q in range(E_AHEAD):
            pl = d_epl[q * B:(q + 1) * B]
            vpl = d_evpl[q * B:(q + 1) * B]
            be.forward_device(leela_amd.NET_POLICY, p, B, pl.data_ptr(),
                              d_praw.data_ptr(), d_pout.data_ptr(), stream)
            be.forward_device(leela_amd.NET_VALUE, p, B, vpl.data_ptr(),
                              d_vraw.data_ptr(), d_vout.data_ptr(), stream)

    encode_step()
    torch.cuda.synchronize(dev)
    denc_exact = bool(torch.equal(d_epl[:B], d_planes) and
                      torch.equal(d_evpl[:B], d_vplanes))
    reps = max(3, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(reps):
        encode_step()
    torch.cuda.synchronize(dev)
    denc_pos_s = reps * B / (time.perf_counter() - t0)
    settle(p)
    for k in range(2):
        e2e_step(k)
    torch.cuda.synchronize(dev)
    e2e_launches = max(2, args.steps // E_AHEAD)
    t0 = time.perf_counter()
    for k in range(e2e_launches):
        e2e_step(k)
    torch.cuda.synchronize(dev)
    e2e_pos_s = e2e_launches * E_AHEAD * B / (time.perf_counter() - t0)

    # Playouts of the surrogate's teacher (lz_playout_winrates_device vs
    # lz_make_value_labels on the host threads), 64 per position.
    po_n, po_k = 2048, 64
    po_pos = positions.PositionBatch(
        [positions.Position.random(10000 + i) for i in range(po_n)])
    d_po = torch.empty(po_n, dtype=torch.float32, device=dev)
    pdenc = positions.DeviceEncoder(0, max_batch=po_n)
    pdenc.playouts_into(po_pos, po_k, 5, d_po.data_ptr(), stream)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    pdenc.playouts_into(po_pos, po_k, 5, d_po.data_ptr(), stream)
    torch.cuda.synchronize(dev)
    gpu_po_s = po_n * po_k / (time.perf_counter() - t0)
    # The reference's pattern policy on the device (one playout per thread;
    # the first call uploads its tables and allocates the playout states).
    d_pw = torch.empty(po_n, dtype=torch.float32, device=dev)
    pdenc.playouts_into(po_pos, 4, 6, d_pw.data_ptr(), stream,
                        positions.PLAYOUT_PATTERN)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    pdenc.playouts_into(po_pos, po_k, 6, d_pw.data_ptr(), stream,
                        positions.PLAYOUT_PATTERN)
    torch.cuda.synchronize(dev)
    gpu_pat_s = po_n * po_k / (time.perf_counter() - t0)
    pdenc.close()
    host_n = 128
    t0 = time.perf_counter()
    positions.value_labels(host_n, 10000, po_k, 5, threads=enc_threads,
                           planes=False)
    host_po_s = host_n * po_k / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    positions.value_labels(host_n, 10000, po_k, 6, threads=enc_threads,
                           planes=False, policy=positions.PLAYOUT_PATTERN)
    host_pat_s = host_n * po_k / (time.perf_counter() - t0)
    denc.close()

    # Network::benchmark restated (the netbench command): 16 threads of
    # synchronous get_scored_moves / get_value calls (RANDOM_ROTATION, host
    # feature extraction per call), batched by the leaf queue.
    from leela_amd.network import Network
    nb_threads = 16
    be.queue_start(p, max_batch=nb_threads, max_wait_us=200)
    try:
        # One untimed run first: the queue's threads, buffers and the
        # layered kernels' first launches warm up.
        Network(be, p).benchmark(positions.Position.random(1), nb_threads)
        nb_policy, nb_value = Network(be, p).benchmark(
            positions.Position.random(1), nb_threads)
    finally:
        be.queue_stop()

    # Search-side batching (lz_search): search threads grow one tree from a
    # generator position, each new leaf expanded by the policy net and
    # valued by the value net through the synchronous facade calls, which
    # the leaf queue batches across threads; one thread without the queue
    # for comparison.
    from leela_amd import search as lz_search
    s_root = positions.Position.random(21)
    s_threads, s_sims = 128, 4096
    be.queue_start(p, max_batch=256, max_wait_us=300)
    try:
        # Warm-up search first (the queue's first launches of each bucket
        # and the threads' first allocations), then the timed one.
        lz_search.search(positions.Position.random(20), s_threads, 1024,
                         backend=be, precision=p, seed=3)
        srch = lz_search.search(s_root, s_threads, s_sims, backend=be,
                                precision=p, seed=4)
    finally:
        be.queue_stop()
    srch1 = lz_search.search(s_root, 1, 256, backend=be, precision=p, seed=4)

    # Config 3: ragged batched leaf evaluation, 1024 x 19x19 + 512 x 13x13
    # + 512 x 9x9 (random legal play, move counts uniform in 0..0.6 x board
    # area, one seed range per side) in one persistent launch per net, the
    # launch shape set by the autotuner for this mix; then the same boards
    # through the leaf queue from 16 threads (ragged graph of the 2048
    # bucket). FLOPs: the conv layers scale with the board's points
    # (P / 361); the value head's inner products see every board as 361
    # inputs (smaller boards embedded at y*19+x), so they do not.
    ragged = ragged_leg(be, leela_amd, nets, positions, p, stream, dev,
                        torch, args.steps)

    # Dominant kernel, timed per launch with HIP events on its stream.
    def kernel_ms(form, sel):
        q = leela_amd.PRECISIONS[form]
        settle(q)
        out = []
        for _ in range(20):
            ms = be.forward_timed(leela_amd.NET_POLICY, q, B,
                                  d_planes.data_ptr(), d_praw.data_ptr(),
                                  d_pout.data_ptr(), stream)
            out += list(sel(ms))
        return float(np.mean(out))

    trunk_flops = 2.0 * 128 * 1152 * 361 * B  # one trunk layer, whole batch
    avg_ms = kernel_ms(head, (lambda ms: ms[2:12]) if head == "fp32"
                       else (lambda ms: ms))
    if head == "fp32":
        kflops = trunk_flops
        kernel = "conv_f32_kernel<3> (policy trunk 3x3 128->128)"
    else:
        kflops = float(nets.flops_per_position(nets.POLICY_SHAPES)) * B
        kernel = (f"fused_net<{head}> policy (13 convs + softmax, {B} "
                  f"boards per launch)")
    achieved = kflops / (avg_ms * 1e-3) / 1e12
    w_bytes = 1664609 * (4 if head == "fp32" else 2)
    algo_bytes = B * (32 * 361 + 2 * 4 * 361) + w_bytes
    peak = PEAK_TFLOPS[head]
    # The FP32 plain form's dominant kernel (the ten 128->128 trunk convs of
    # the policy net, launches 2..11 of its forward) against the FP32
    # matrix peak.
    f32_ms = kernel_ms("fp32", lambda ms: ms[2:12])
    f32_tf = trunk_flops / (f32_ms * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(pol, val, planes, vplanes, args.cpu_seconds,
                           cores)
        cpu["fp64_oracle_seconds"] = round(fp64_s, 2)

    step_flops = float(nets.flops_per_position(nets.POLICY_SHAPES) +
                       nets.flops_per_position(nets.VALUE_SHAPES))
    forms_out = {}
    for n in forms:
        tf = res[n]["pos_s"] * step_flops / 1e12
        forms_out[n] = {
            "positions_per_s": round(res[n]["pos_s"], 1),
            "speedup_vs_fp32_plain": round(res[n]["pos_s"] / ref["pos_s"], 3),
            "step_tflops": round(tf, 2),
            "step_frac_of_peak": round(tf / PEAK_TFLOPS[n], 4),
            "top1_agreement_vs_fp32_plain": round(res[n]["top1"], 5),
            "max_abs_value_diff_vs_fp32_plain": round(res[n]["dwin"], 6),
            **res[n].get("acc64", {})}
    forms_out["fp32"]["roofline"] = {
        "kernel": "conv_f32_kernel<3> (policy trunk 3x3 128->128, ten "
                  "launches per forward)",
        "achieved": round(f32_tf, 2), "peak": PEAK_TFLOPS["fp32"],
        "unit": "TFLOP/s", "frac": round(f32_tf / PEAK_TFLOPS["fp32"], 4),
        "avg_launch_ms": round(f32_ms, 4)}
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
        "dtype": head,
        "data": "synthetic",
        "config": {
            "workload": f"Leela policy (13 conv, 128-wide) + value (12 conv "
                        f"64-wide + 2 FC) forward, batch {B} of 19x19",
            "batch": B,
            "precision": head,
            "weights": "He-normal seed 7 (set A), Network.cpp:82-137 shapes",
            "positions": "generator: random legal play, seeds 1..B, move "
                         "counts uniform in 0..250",
            "parallelism": "single GPU",
            "agreement_bar": AGREEMENT_BAR,
            "agreement_reference": ("FP64 oracle (im2col + dgemm)"
                                    if ref64 else "FP32 plain form"),
            "agreement_met": agreement_met,
            "forms": forms_out,
            "ragged": ragged,
            "host_buffer_positions_per_s": round(host_pos_s, 1),
            "graph_replay_positions_per_s": (round(graph_pos_s, 1)
                                             if graph_pos_s else None),
            "host_encoder_positions_per_s": round(enc_pos_s, 1),
            "host_encoder_threads": enc_threads,
            "device_encoder_positions_per_s": round(denc_pos_s, 1),
            "device_encoder_planes_equal_host": denc_exact,
            "end_to_end_positions_per_s": round(e2e_pos_s, 1),
            "end_to_end": "positions in host memory -> "
                          "lz_encode_device_both (one launch per 4 net "
                          "batches) -> policy + value forward per 1024, one "
                          "stream, one host thread",
            "netbench": {
                "policy_per_s": round(nb_policy, 1),
                "value_per_s": round(nb_value, 1),
                "threads": nb_threads,
                "what": "Network::benchmark (Network.cpp:147-199): 2000 "
                        "policy then 10000 value evaluations, "
                        "RANDOM_ROTATION, synchronous calls from 16 "
                        "threads through the leaf queue (max batch 16)",
            },
            "search": {
                "threads": s_threads,
                "simulations": int(srch["simulations"]),
                "simulations_per_s": round(srch["simulations_per_s"], 1),
                "mean_batch": round(srch["mean_batch"], 1),
                "collisions": int(srch["collisions"]),
                "simulations_per_s_1_thread": round(
                    srch1["simulations_per_s"], 1),
                "what": "lz_search: tree search with virtual loss, each new "
                        "leaf one policy row (async) and one value row "
                        "through the leaf queue (max batch 256); 1 thread: "
                        "direct forwards",
            },
            "playouts": {
                "device_playouts_per_s": round(gpu_po_s, 1),
                "host_playouts_per_s": round(host_po_s, 1),
                "host_threads": enc_threads,
                "sample": f"{po_n} generator positions x {po_k} light "
                          f"playouts on the device, {host_n} x {po_k} on "
                          f"the host; equal results bit for bit",
                "device_pattern_playouts_per_s": round(gpu_pat_s, 1),
                "host_pattern_playouts_per_s": round(host_pat_s, 1),
                "pattern": "the reference's playout policy "
                           "(FastState::play_random_move), same positions "
                           "and counts; device equal to host bit for bit",
            },
        },
        "roofline": {
            "bound": "mfma",
            "kernel": kernel,
            "achieved": round(achieved, 2),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": (TRAFFIC_B1024[head]["bytes"]
                        if B == 1024 and head in TRAFFIC_B1024 else None),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2.0 + "
                            "WRITE_SIZE x1.0, calibrated on a known-size "
                            "copy)",
            "traffic_source": TRAFFIC_SOURCE,
            "avg_launch_ms": round(avg_ms, 4),
            # Planes in + logits and probabilities out per board, weights
            # once per launch (16-bit), over the launch time.
            "algorithmic_bytes": algo_bytes,
            "algorithmic_gb_s": round(algo_bytes / (avg_ms * 1e-3) / 1e9, 2),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    be.close()
    if not agreement_met:
        print(f"no 16-bit form meets the {AGREEMENT_BAR} top-1 agreement "
              f"bar", file=sys.stderr)
        return 3
    return 0


if __name__ == "__main__":
    sys.exit(main())

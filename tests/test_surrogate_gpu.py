"""The surrogate's HIP kernels against the FP64 oracle (oracle/surrogate.py,
pinned by finite differences on the CPU): forward win rates, every
parameter tensor's gradient, the Adam update, checkpoint round trips,
training progress, and the gain over plain playouts.

Tolerances: against the oracle run with BF16 rounding at the points where
the kernels store 16-bit values, win rates within 2e-3 and each gradient
tensor within 2% relative L2 error; against the exact oracle, win rates
within 3e-2 and gradient cosine >= 0.98 over the whole vector."""
import os

import numpy as np
import pytest

from oracle import surrogate as orc
from leela_amd import positions
from leela_amd.surrogate import Surrogate, LeafEvaluator, train

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur():
    s = Surrogate(0, max_batch=512, seed=3)
    yield s
    s.close()


def _value_planes(n, seed0):
    return positions.make_batch(n, seed0=seed0, net=positions.NET_VALUE)


def test_forward_matches_oracle(sur):
    print("surrogate forward: 12 generator positions, seeds 500..", flush=True)
    planes = _value_planes(12, 500)
    flat = sur.params().astype(np.float64)
    win = sur.predict(planes)
    t16, _ = orc.forward(flat, planes, bf16_points=True)
    t64, _ = orc.forward(flat, planes)
    assert np.abs(win - (1 + t16) / 2).max() < 2e-3
    assert np.abs(win - (1 + t64) / 2).max() < 3e-2
    # Batches split across predict calls give the same rows.
    np.testing.assert_array_equal(sur.predict(planes[5:9]), win[5:9])


def test_gradients_match_oracle(sur):
    print("surrogate gradients: 6 positions, seeds 700..", flush=True)
    planes = _value_planes(6, 700)
    y = np.random.default_rng(1).uniform(0.1, 0.9, 6).astype(np.float32)
    flat = sur.params().astype(np.float64)
    loss, g = sur.gradients(planes, y)
    t, cache = orc.forward(flat, planes, bf16_points=True)
    lo, go = orc.backward(cache, y)
    assert abs(loss - lo) <= 2e-3 * lo + 1e-6
    for name, shape, off in orc.layout():
        n = int(np.prod(shape))
        a, b = g[off:off + n].astype(np.float64), go[off:off + n]
        nb = np.linalg.norm(b)
        if nb < 1e-9:
            assert np.linalg.norm(a) < 1e-6, name
            continue
        rel = np.linalg.norm(a - b) / nb
        assert rel < 2e-2, (name, rel)
    t0, c0 = orc.forward(flat, planes)
    _, g0 = orc.backward(c0, y)
    cos = g @ g0 / np.linalg.norm(g) / np.linalg.norm(g0)
    assert cos >= 0.98
    # Same inputs, same gradient (fixed-order reductions).
    _, g2 = sur.gradients(planes, y)
    np.testing.assert_array_equal(g, g2)


def test_adam_step_matches_oracle():
    import torch
    print("surrogate adam: 2 steps on 8 positions", flush=True)
    s = Surrogate(0, max_batch=64, seed=9)
    try:
        planes = _value_planes(8, 900)
        y = np.linspace(0.2, 0.8, 8).astype(np.float32)
        X = torch.from_numpy(planes).cuda()
        Y = torch.from_numpy(y).cuda()
        p = s.params().astype(np.float64)
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        lr = 1e-3
        for step in (1, 2):
            _, g = s.gradients(planes, y)
            loss = s.train_step_device(8, X.data_ptr(), Y.data_ptr(), lr,
                                       stream=torch.cuda.current_stream().cuda_stream)
            assert np.isfinite(loss)
            p, m, v = orc.adam(p, g.astype(np.float64), m, v, step, lr)
            got = s.params()
            assert np.abs(got - p).max() < 1e-6, step
            p = got.astype(np.float64)
        assert s.step == 2
    finally:
        s.close()


def test_checkpoint_resume_is_exact(tmp_path):
    import torch
    print("surrogate checkpoint: 3 steps, save after 2, resume", flush=True)
    planes = _value_planes(32, 1200)
    y = np.random.default_rng(2).uniform(0, 1, 32).astype(np.float32)
    X = torch.from_numpy(planes).cuda()
    Y = torch.from_numpy(y).cuda()
    st = torch.cuda.current_stream().cuda_stream
    a = Surrogate(0, max_batch=32, seed=4)
    b = Surrogate(0, max_batch=32, seed=5)
    try:
        for _ in range(2):
            a.train_step_device(32, X.data_ptr(), Y.data_ptr(), 1e-3, stream=st)
        path = tmp_path / "sg.ckpt"
        a.save(path)
        a.train_step_device(32, X.data_ptr(), Y.data_ptr(), 1e-3, stream=st)
        b.load(path)
        assert b.step == 2
        b.train_step_device(32, X.data_ptr(), Y.data_ptr(), 1e-3, stream=st)
        np.testing.assert_array_equal(a.params(), b.params())
        with open(tmp_path / "bad.ckpt", "wb") as fp:
            fp.write(b"not a checkpoint")
        with pytest.raises(Exception):
            b.load(tmp_path / "bad.ckpt")
    finally:
        a.close()
        b.close()


def test_batch_limits_rejected(sur):
    with pytest.raises(Exception):
        sur.gradients(np.zeros((0, 32, 361), np.uint8), np.zeros(0, np.float32))
    with pytest.raises(Exception):
        sur.predict(np.zeros((2, 32, 360), np.uint8))


def test_training_beats_playouts_at_equal_cost():
    """Gain over the method without a surrogate. Train on 4096 positions
    labelled with 64 playouts each; on 512 held-out positions, measure MAE
    against a fresh 256-playout mean. The surrogate must beat the constant
    predictor (the training mean) and a 4-playout estimate, and answer far
    faster than playouts do."""
    import time
    import torch
    print("surrogate training: labels 4096 x 64 playouts", flush=True)
    th = min(16, os.cpu_count() or 1)
    tr_planes, tr_y = positions.value_labels(4096, 30000, 64, 3_000_000,
                                             threads=th)
    ev_planes, ev_y = positions.value_labels(512, 40000, 256, 4_000_000,
                                             threads=th)
    s = Surrogate(0, max_batch=512, seed=7)
    try:
        print("surrogate training: 600 steps, batch 256", flush=True)
        losses = train(s, tr_planes, tr_y, steps=600, batch=256, lr=1e-3,
                       seed=1, log_every=100)
        assert losses[-1][1] < 0.7 * losses[0][1]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pred = s.predict(ev_planes)
        t_sur = time.perf_counter() - t0
    finally:
        s.close()
    mae_sur = float(np.mean(np.abs(pred - ev_y)))
    mae_const = float(np.mean(np.abs(tr_y.mean() - ev_y)))
    ps = [positions.Position.random(40000 + i, 19, 250) for i in range(512)]
    t0 = time.perf_counter()
    est4 = LeafEvaluator(None, playouts=4, seed=5_000_000).winrates(ps)
    t_po = time.perf_counter() - t0
    mae_4 = float(np.mean(np.abs(est4 - ev_y)))
    print(f"MAE surrogate {mae_sur:.4f} constant {mae_const:.4f} "
          f"4-playout {mae_4:.4f}; time surrogate {t_sur * 1e3:.1f} ms, "
          f"4 playouts {t_po * 1e3:.1f} ms", flush=True)
    assert mae_sur < mae_const
    assert mae_sur < mae_4
    assert t_sur < t_po

"""GPU checks of the FP8 form (FP16 5x5 stem, e4m3 operands for every later
conv with per-output-channel weight scales and a per-board activation scale
per layer, v_mfma_scale_f32_16x16x128_f8f6f4) on 16 generator positions
(seeds 301..316).

The FP8 emulation (oracle/fp8.py: FP64 arithmetic at the same
quantisation points, e4m3 rounding pinned to torch.float8_e4m3fn) sets the
accuracy the form should reach against the exact FP64 forward. The GPU's
FP8 MFMA sums 128 products per instruction in its own internal order and
precision, so its e4m3 rounding decisions are not the emulation's one for
one, and 12 layers of 3-bit-mantissa rounding amplify any difference: the
test compares the two statistically. Bars on 16 positions: the GPU's
top-1 agreement with FP64 within 0.15 of the emulation's, its mean KL from
FP64 at most twice the emulation's + 0.01, values within 0.1 of FP64 on
average; the figures are printed.
"""
import numpy as np
import pytest

from leela_amd.positions import make_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(weights_b):
    import leela_amd
    be = leela_amd.Backend(0)
    pol, val = weights_b
    be.load_net(leela_amd.NET_POLICY, pol)
    be.load_net(leela_amd.NET_VALUE, val)
    be.finalize()
    yield be
    be.close()


def _kl(p, q):
    p = np.maximum(p, 1e-30)
    q = np.maximum(q, 1e-30)
    return np.sum(p * (np.log(p) - np.log(q)), axis=1)


def test_fp8_policy(gpu, weights_b):
    import leela_amd
    import oracle
    from oracle import fp8
    print("fp8 policy: 16 positions, seeds 301..316", flush=True)
    pol, _ = weights_b
    planes = make_batch(16, seed0=301)
    _, p_gpu = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP8, planes)
    _, p_emu = fp8.forward_fp8(pol, planes, policy=True)
    _, p64 = oracle.forward_f64(oracle.Net(pol), planes, oracle.HEAD_POLICY)
    assert np.all(np.isfinite(p_gpu))
    np.testing.assert_allclose(p_gpu.sum(1), 1.0, atol=1e-4)
    top_gpu = np.mean(p_gpu.argmax(1) == p64.argmax(1))
    top_emu = np.mean(p_emu.argmax(1) == p64.argmax(1))
    kl_gpu = _kl(p64, p_gpu).mean()
    kl_emu = _kl(p64, p_emu).mean()
    print(f"FP8 vs FP64 over 16 positions: GPU top-1 {top_gpu:.3f}, mean KL "
          f"{kl_gpu:.4f}, max dprob {np.abs(p_gpu - p64).max():.4f}; "
          f"emulation top-1 {top_emu:.3f}, mean KL {kl_emu:.4f}; GPU vs "
          f"emulation top-1 {np.mean(p_gpu.argmax(1) == p_emu.argmax(1)):.3f}",
          flush=True)
    assert top_gpu >= top_emu - 0.15
    assert kl_gpu <= 2.0 * kl_emu + 0.01


def test_fp8_value(gpu, weights_b):
    import leela_amd
    import oracle
    from oracle import fp8
    print("fp8 value: 16 positions, seeds 301..316", flush=True)
    _, val = weights_b
    planes = make_batch(16, seed0=301, net=1)
    _, v_gpu = gpu.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP8, planes)
    _, v_emu = fp8.forward_fp8(val, planes, policy=False)
    _, v64 = oracle.forward_f64(oracle.Net(val), planes, oracle.HEAD_VALUE)
    d_gpu = np.abs(v_gpu - v64).ravel()
    d_emu = np.abs(v_emu - v64).ravel()
    print(f"FP8 value vs FP64: GPU mean {d_gpu.mean():.4f} max "
          f"{d_gpu.max():.4f}; emulation mean {d_emu.mean():.4f} max "
          f"{d_emu.max():.4f}", flush=True)
    assert np.all((v_gpu >= 0) & (v_gpu <= 1))
    assert d_gpu.mean() < 0.1 and d_gpu.mean() <= 2.0 * d_emu.mean() + 0.02


def test_fp8_batch_rows_independent(gpu):
    import leela_amd
    planes = make_batch(40, seed0=900)
    _, a = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP8, planes)
    _, b = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP8, planes[7:19])
    np.testing.assert_array_equal(a[7:19], b)


def test_fp8_ragged_rejected(gpu):
    import leela_amd
    with pytest.raises(Exception):
        gpu.forward_ragged(leela_amd.NET_POLICY, leela_amd.PREC_FP8,
                           [np.zeros((32, 81), np.uint8),
                            np.zeros((32, 169), np.uint8)])

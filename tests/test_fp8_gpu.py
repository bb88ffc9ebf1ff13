"""GPU checks of the FP8 form (FP16 5x5 stem, e4m3 operands for every later
conv with per-output-channel weight scales and a per-board activation scale
per layer, v_mfma_scale_f32_16x16x128_f8f6f4).

  * Exactness: on nets whose every quantised value is representable
    (tests/fp8_nets.py, the same nets that pin the FP64 emulation on the
    CPU), the GPU's policy logits and win rates equal the FP64 forward of
    the net with the planted roundings applied, within 1e-5 relative:
    scales, quantisation points and rounding are the form's definition.
  * Accuracy on 64 generator positions (seeds 301..364) against the FP64
    oracle, absolute bars: top-1 agreement >= 0.60, mean KL(FP64 || FP8)
    < 0.15, mean |d win| < 0.1. Three mantissa bits through twelve
    layers of this random He-normal net cost that much in the exact FP64
    emulation of the form too (its figures are printed beside the GPU's).
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


@pytest.mark.parametrize("opt", [{}, {"round_weight": True},
                                 {"round_act": True}, {"tiny_channel": True}],
                         ids=["exact", "round_weight", "round_act",
                              "tiny_channel"])
def test_fp8_exact_nets(opt):
    import leela_amd
    import oracle
    from fp8_nets import exact_net, planes
    print(f"fp8 exact nets {opt}", flush=True)
    pol, pol_exp = exact_net(True, **opt)
    val, val_exp = exact_net(False, **opt)
    be = leela_amd.Backend(0)
    try:
        be.load_net(leela_amd.NET_POLICY, pol)
        be.load_net(leela_amd.NET_VALUE, val)
        be.finalize()
        pl = planes(8, seed=6)
        raw, _ = be.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP8, pl)
        vraw, _ = be.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP8, pl)
    finally:
        be.close()
    r64, _ = oracle.forward_f64(oracle.Net(pol_exp), pl, oracle.HEAD_POLICY)
    v64, _ = oracle.forward_f64(oracle.Net(val_exp), pl, oracle.HEAD_VALUE)
    print(f"policy max|d|={np.abs(raw - r64).max():.3g} of "
          f"{np.abs(r64).max():.3g}, value max|d|={np.abs(vraw - v64).max():.3g}")
    assert np.abs(raw - r64).max() <= 1e-5 * np.abs(r64).max()
    assert np.abs(vraw - v64).max() <= 1e-5 * max(np.abs(v64).max(), 1e-3)


def test_fp8_policy_vs_fp64(gpu, weights_b):
    import leela_amd
    import oracle
    from oracle import fp8
    print("fp8 policy: 64 positions, seeds 301..364", flush=True)
    pol, _ = weights_b
    planes = make_batch(64, seed0=301)
    _, p_gpu = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP8, planes)
    _, p_emu = fp8.forward_fp8(pol, planes[:16], policy=True)
    oracle.load_blas()
    _, p64 = oracle.forward_f64(oracle.Net(pol), planes, oracle.HEAD_POLICY,
                                gemm=True)
    assert np.all(np.isfinite(p_gpu))
    np.testing.assert_allclose(p_gpu.sum(1), 1.0, atol=1e-4)
    top_gpu = np.mean(p_gpu.argmax(1) == p64.argmax(1))
    kl_gpu = _kl(p64, p_gpu).mean()
    top_emu = np.mean(p_emu.argmax(1) == p64[:16].argmax(1))
    kl_emu = _kl(p64[:16], p_emu).mean()
    print(f"FP8 vs FP64, 64 positions: GPU top-1 {top_gpu:.3f}, mean KL "
          f"{kl_gpu:.4f}, max dprob {np.abs(p_gpu - p64).max():.4f}; FP64 "
          f"emulation of the form on the first 16: top-1 {top_emu:.3f}, "
          f"mean KL {kl_emu:.4f}", flush=True)
    assert top_gpu >= 0.60
    assert kl_gpu < 0.15


def test_fp8_value_vs_fp64(gpu, weights_b):
    import leela_amd
    import oracle
    print("fp8 value: 64 positions, seeds 301..364", flush=True)
    _, val = weights_b
    planes = make_batch(64, seed0=301, net=1)
    _, v_gpu = gpu.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP8, planes)
    oracle.load_blas()
    _, v64 = oracle.forward_f64(oracle.Net(val), planes, oracle.HEAD_VALUE,
                                gemm=True)
    d = np.abs(v_gpu - v64).ravel()
    print(f"FP8 value vs FP64: mean {d.mean():.4f} max {d.max():.4f}",
          flush=True)
    assert np.all((v_gpu >= 0) & (v_gpu <= 1))
    assert d.mean() < 0.1


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

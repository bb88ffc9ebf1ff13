"""The reference's trained nets (model_4258 / model_4750: Convolution, ReLU,
BatchNorm per layer, a last 3x3 conv with no activation) and the oracle's
ReLU / batch-normalisation path that checks the GPU on them.

Pins of the oracle's additions, each against something other than its own
formula: ReLU output equals max(0, the same conv without activation) and
scipy's correlate2d gives that conv; batch normalisation with a channel's
own statistics (mean and variance of the activated outputs over the batch,
times the stored scale factor) leaves the channel with mean 0 and variance
var / (var + 1e-5) (Network.cpp:427-448); the FP32 restatement of the
reference's CPU path agrees with the FP64 one on the trained nets.
"""
from pathlib import Path

import numpy as np
import pytest
from scipy.signal import correlate2d

from tests.caffe_fixture import MODELS, load_tables, planes22

import oracle

REF = Path("/root/reference")


@pytest.fixture(scope="module")
def tables():
    return load_tables()


def test_fixture_shapes(tables):
    t = tables["model_4258"]
    assert [(L["filter"], L["cin"], L["cout"]) for L in t] == [
        (3, 22, 160), (3, 160, 96), (3, 96, 32), (3, 32, 32), (3, 32, 1)]
    t = tables["model_4750"]
    assert [(L["filter"], L["cin"], L["cout"]) for L in t] == [
        (5, 22, 96), (3, 96, 160), (3, 160, 128), (3, 128, 128),
        (3, 128, 96), (3, 96, 96), (3, 96, 96), (3, 96, 64), (3, 64, 64),
        (3, 64, 1)]
    for m in MODELS:
        t = tables[m]
        assert all(L["act"] == oracle.ACT_RELU for L in t[:-1])
        assert t[-1]["act"] == oracle.ACT_NONE
        assert all(L["bn"] is not None for L in t[:-1]) and t[-1]["bn"] is None


@pytest.mark.skipif(not (REF / "model_4258.caffemodel").exists(),
                    reason="reference files not present")
def test_fixture_equals_the_caffemodels(tables):
    from leela_amd import caffe_weights as cw
    for m in MODELS:
        t = cw.layer_table(cw.read_caffemodel(str(REF / f"{m}.caffemodel")))
        assert len(t) == len(tables[m])
        for a, b in zip(t, tables[m]):
            np.testing.assert_array_equal(a["w"], b["w"])
            np.testing.assert_array_equal(a["b"], b["b"])
            assert a["act"] == b["act"]
            if a["bn"] is not None:
                for x, y in zip(a["bn"], b["bn"]):
                    np.testing.assert_array_equal(x, y)


def _probe_net(w1, b1, act, bn, k):
    """conv (w1, b1, act, bn) then a head that copies channel k's centre
    tap, so the policy logits are that channel's map."""
    cout, cin = w1.shape[:2]
    wh = np.zeros((1, cout, 3, 3), np.float32)
    wh[0, k, 1, 1] = 1.0
    return oracle.Net([(0, w1.shape[2], cin, cout, w1, b1, act, bn),
                       (0, 3, cout, 1, wh, np.zeros(1, np.float32),
                        oracle.ACT_NONE, None)])


def _maps(net, planes):
    raw, _ = oracle.forward_f64(net, planes, oracle.HEAD_POLICY)
    return raw


def test_relu_is_max_of_the_plain_conv_and_matches_scipy():
    rng = np.random.default_rng(5)
    cin, cout = 5, 3
    w1 = rng.normal(0, 0.5, (cout, cin, 3, 3)).astype(np.float32)
    b1 = rng.normal(0, 0.5, cout).astype(np.float32)
    planes = (rng.random((4, cin, 361)) < 0.4).astype(np.uint8)
    for k in range(cout):
        none = _maps(_probe_net(w1, b1, oracle.ACT_NONE, None, k), planes)
        relu = _maps(_probe_net(w1, b1, oracle.ACT_RELU, None, k), planes)
        np.testing.assert_array_equal(relu, np.maximum(none, 0.0))
        assert (none < 0).any() and (none > 0).any()
        for i in range(planes.shape[0]):
            x = planes[i].reshape(cin, 19, 19).astype(np.float64)
            want = sum(correlate2d(x[c], w1[k, c].astype(np.float64),
                                   mode="same") for c in range(cin))
            np.testing.assert_allclose(none[i], want.ravel() + b1[k],
                                       rtol=0, atol=1e-12)


def test_batchnorm_with_own_statistics_standardises():
    rng = np.random.default_rng(6)
    cin, cout, scale = 6, 4, 999.98236
    w1 = rng.normal(0, 0.7, (cout, cin, 3, 3)).astype(np.float32)
    b1 = rng.normal(0.3, 0.2, cout).astype(np.float32)
    planes = (rng.random((8, cin, 361)) < 0.5).astype(np.uint8)
    pre = np.stack([_maps(_probe_net(w1, b1, oracle.ACT_RELU, None, k),
                          planes) for k in range(cout)])
    mean = pre.reshape(cout, -1).mean(axis=1)
    var = pre.reshape(cout, -1).var(axis=1)
    # The stored blobs are the statistics times the scale factor.
    bn = ((mean * scale).astype(np.float32), (var * scale).astype(np.float32),
          scale)
    s32 = float(np.float32(scale))  # the scale factor is an FP32 blob
    m32 = bn[0].astype(np.float64) / s32
    v32 = bn[1].astype(np.float64) / s32
    eps = float(np.float32(1e-5))
    for k in range(cout):
        y = _maps(_probe_net(w1, b1, oracle.ACT_RELU, bn, k), planes)
        # The FP32 blobs carry the statistics to ~1e-7 relative.
        assert abs(y.mean() - (mean[k] - m32[k]) / np.sqrt(v32[k] + eps)) \
            < 1e-12
        assert abs(y.mean()) < 1e-5 * (1 + abs(mean[k]) / np.sqrt(var[k]))
        want_var = var[k] / (v32[k] + eps)
        assert y.var() == pytest.approx(want_var, rel=1e-6)
        assert y.var() == pytest.approx(1.0, rel=1e-3)


@pytest.mark.parametrize("model", MODELS)
def test_reference_cpu_path_agrees_with_fp64_on_trained_nets(tables, model):
    net = oracle.Net.from_table(tables[model])
    planes = planes22(16, seed0=1)
    oracle.load_blas()
    raw32, p32 = oracle.forward_f32(net, planes, oracle.HEAD_POLICY,
                                    temperature=1.0)
    raw64, p64 = oracle.forward_f64(net, planes, oracle.HEAD_POLICY,
                                    temperature=1.0, gemm=True)
    # Logits reach |3,400| on these inputs (model_4258): FP32 rounding is
    # relative to that scale, and the softmax (T = 1) follows it.
    scale = np.abs(raw64).max()
    assert np.abs(raw32 - raw64).max() < 1e-5 * (1 + scale)
    assert np.abs(p32 - p64).max() < 1e-4
    assert (p32.argmax(1) == p64.argmax(1)).all()
    # Trained nets: confident, varied move choices.
    assert p64.max(axis=1).mean() > 0.1
    assert len(set(p64.argmax(1).tolist())) > 4

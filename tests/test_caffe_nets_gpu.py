"""The reference's trained nets on the GPU: model_4258 and model_4750
(22-plane Caffe policy nets: Convolution, ReLU, BatchNorm per layer, a last
3x3 conv with no activation, softmax at T = 1; read from the shipped
.caffemodel files into tests/golden/caffe_models.npz), pushed through
lz_push_convolve / lz_push_activation / lz_push_batchnorm and run by the
FP32 form, against the FP64 oracle on the same planes. Inputs: the first 22
of the 32 policy planes of generator positions (the 22-plane feature set is
not in the reference).

Bars: logits within 1e-5 of the logit scale (these nets reach |logit| ~
3,400 on these inputs), probabilities within 1e-4, the same top-1 move on
every board.
"""
import numpy as np
import pytest

from tests.caffe_fixture import MODELS, load_tables, planes22

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables():
    return load_tables()


def _backend(table):
    import leela_amd
    from leela_amd import caffe_weights as cw
    from leela_amd import nets
    be = leela_amd.Backend(0)
    cw.push_layers(be, leela_amd.NET_POLICY, table)
    # A value net in the reference's shapes fills the context's other slot.
    be.load_net(leela_amd.NET_VALUE, nets.weight_set("A")[1])
    be.finalize()
    be.set_temperature(1.0)
    return be


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("n", [1, 64])
def test_trained_net_fp32_matches_fp64_oracle(tables, model, n):
    import leela_amd
    import oracle
    print(f"case: {model} n={n}", flush=True)
    planes = planes22(n, seed0=300 + n)
    be = _backend(tables[model])
    try:
        raw, prob = be.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32,
                               planes)
    finally:
        be.close()
    oracle.load_blas()
    raw64, p64 = oracle.forward_f64(oracle.Net.from_table(tables[model]),
                                    planes, oracle.HEAD_POLICY,
                                    temperature=1.0, gemm=True)
    scale = np.abs(raw64).max()
    assert np.abs(raw - raw64).max() < 1e-5 * (1 + scale)
    assert np.abs(prob - p64).max() < 1e-4
    assert (prob.argmax(1) == p64.argmax(1)).all()
    np.testing.assert_allclose(prob.sum(1), 1.0, atol=1e-5)


def test_general_nets_refuse_the_16bit_forms_and_the_queue(tables):
    import leela_amd
    from leela_amd.backend import LeelaError
    be = _backend(tables["model_4258"])
    try:
        planes = planes22(4)
        for prec in (leela_amd.PREC_BF16, leela_amd.PREC_FP16,
                     leela_amd.PREC_FP8):
            with pytest.raises(LeelaError) as e:
                be.forward(leela_amd.NET_POLICY, prec, planes)
            assert e.value.code == -5  # LZ_ERR_UNSUPPORTED
        with pytest.raises(LeelaError) as e:
            be.queue_start(leela_amd.PREC_FP32, 16, 100)
        assert e.value.code == -5
        # The value net (reference shapes, 32 planes) still runs in 16 bit.
        vp = np.zeros((2, 32, 361), np.uint8)
        be.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP16, vp)
    finally:
        be.close()

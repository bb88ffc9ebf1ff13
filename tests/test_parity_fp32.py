"""GPU parity of the FP32 plain form against the reference CPU path
(im2col + sgemm + bias + ELU, restated in the oracle): logits and value
within 1e-4 absolute, identical top-1 moves."""
import numpy as np
import pytest

from leela_amd.positions import make_batch, synthetic_planes

pytestmark = pytest.mark.gpu

TOL = 1e-4  # absolute, on policy logits, policy probabilities and value


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


@pytest.fixture(scope="module")
def cpu_nets(weights_b):
    import oracle
    oracle.load_blas()
    pol, val = weights_b
    return oracle.Net(pol), oracle.Net(val)


@pytest.mark.parametrize("n", [1, 7, 64])
def test_policy_fp32_matches_cpu_path(gpu, cpu_nets, n):
    import oracle
    import leela_amd
    planes = synthetic_planes(n, seed=100 + n)
    raw_ref, prob_ref = oracle.forward_f32(cpu_nets[0], planes,
                                           oracle.HEAD_POLICY)
    raw, prob = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32, planes)
    print(f"policy n={n}: max|dlogit|={np.abs(raw - raw_ref).max():.3g} "
          f"max|dprob|={np.abs(prob - prob_ref).max():.3g}")
    assert np.abs(raw - raw_ref).max() < TOL
    assert np.abs(prob - prob_ref).max() < TOL
    assert (raw.argmax(1) == raw_ref.argmax(1)).all()
    np.testing.assert_allclose(prob.sum(1), 1.0, atol=1e-5)


@pytest.mark.parametrize("n", [1, 5, 64])
def test_value_fp32_matches_cpu_path(gpu, cpu_nets, n):
    import oracle
    import leela_amd
    planes = synthetic_planes(n, seed=200 + n)
    raw_ref, win_ref = oracle.forward_f32(cpu_nets[1], planes,
                                          oracle.HEAD_VALUE)
    raw, win = gpu.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP32, planes)
    print(f"value n={n}: max|draw|={np.abs(raw - raw_ref).max():.3g}")
    assert np.abs(raw - raw_ref).max() < TOL
    assert np.abs(win - win_ref).max() < TOL


def test_fp32_rows_independent_of_batch(gpu):
    """A position's output does not depend on what else is in the batch:
    bitwise equal in the FP32 path."""
    import leela_amd
    planes = synthetic_planes(33, seed=5)
    raw_all, _ = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32, planes)
    raw_sub, _ = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32,
                             planes[[20, 3, 31]])
    assert np.array_equal(raw_all[[20, 3, 31]], raw_sub)
    v_all, _ = gpu.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP32, planes)
    v_sub, _ = gpu.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP32,
                           planes[[7]])
    assert np.array_equal(v_all[[7]], v_sub)


def test_fp32_config1_batch256_vs_cpu(gpu, cpu_nets):
    """Batch 256 of generator positions (seeds 1..256, up to 250 moves): the
    plain form against the CPU path on every row."""
    import oracle
    import leela_amd
    planes = make_batch(256, seed0=1)
    raw_ref, _ = oracle.forward_f32(cpu_nets[0], planes, oracle.HEAD_POLICY)
    raw, _ = gpu.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32, planes)
    assert np.abs(raw - raw_ref).max() < TOL
    assert (raw.argmax(1) == raw_ref.argmax(1)).all()
    v_ref, w_ref = oracle.forward_f32(cpu_nets[1], planes, oracle.HEAD_VALUE)
    v, w = gpu.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP32, planes)
    assert np.abs(v - v_ref).max() < TOL
    assert np.abs(w - w_ref).max() < TOL

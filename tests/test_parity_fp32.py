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


@pytest.fixture(scope="module")
def gpu_a(weights_a):
    import leela_amd
    be = leela_amd.Backend(0)
    pol, val = weights_a
    be.load_net(leela_amd.NET_POLICY, pol)
    be.load_net(leela_amd.NET_VALUE, val)
    be.finalize()
    yield be
    be.close()


def test_fp32_config0_seed1_position(gpu_a, weights_a):
    """The single-position workload: the generator position of seed 1 with
    a move count uniform in 0..200, its policy planes from the library's
    encoder, batch 1, weight set A (He-normal seed 7; the shipped values
    are not available). 361 probabilities and the top-1 move against the
    reference CPU path."""
    import oracle
    import leela_amd
    from leela_amd.positions import Position
    oracle.load_blas()
    pos = Position.random(1, max_moves=200)
    planes = pos.planes(leela_amd.NET_POLICY, 0)[None]
    assert np.array_equal(planes, make_batch(1, seed0=1, max_moves=200))
    raw_ref, prob_ref = oracle.forward_f32(oracle.Net(weights_a[0]), planes,
                                           oracle.HEAD_POLICY)
    raw, prob = gpu_a.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32,
                              planes)
    top = int(raw_ref[0].argmax())
    print(f"config 0: moves played {pos.info()['movenum']}, top-1 "
          f"{top} (x={top % 19}, y={top // 19}) p={prob_ref[0, top]:.4f}, "
          f"max|dlogit|={np.abs(raw - raw_ref).max():.3g}")
    assert prob.shape == (1, 361)
    assert np.abs(raw - raw_ref).max() < TOL
    assert np.abs(prob - prob_ref).max() < TOL
    assert int(raw[0].argmax()) == top
    np.testing.assert_allclose(prob.sum(), 1.0, atol=1e-5)


@pytest.mark.parametrize("wset", ["A", "B"])
def test_fp32_config1_batch256_vs_cpu(gpu, gpu_a, weights_a, weights_b,
                                      wset):
    """Batch 256 of generator positions (seeds 1..256, up to 250 moves): the
    plain form against the CPU path on every row, for weight set A
    (He-normal seed 7, zero biases) and set B (plus biases)."""
    import oracle
    import leela_amd
    oracle.load_blas()
    be, (pol, val) = (gpu_a, weights_a) if wset == "A" else (gpu, weights_b)
    planes = make_batch(256, seed0=1)
    vplanes = make_batch(256, seed0=1, net=leela_amd.NET_VALUE)
    raw_ref, p_ref = oracle.forward_f32(oracle.Net(pol), planes,
                                        oracle.HEAD_POLICY)
    raw, p = be.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32, planes)
    assert np.abs(raw - raw_ref).max() < TOL
    assert np.abs(p - p_ref).max() < TOL
    assert (raw.argmax(1) == raw_ref.argmax(1)).all()
    v_ref, w_ref = oracle.forward_f32(oracle.Net(val), vplanes,
                                      oracle.HEAD_VALUE)
    v, w = be.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP32, vplanes)
    assert np.abs(v - v_ref).max() < TOL
    assert np.abs(w - w_ref).max() < TOL

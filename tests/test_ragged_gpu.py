"""GPU tests of ragged batches (mixed 9x9 / 13x13 / 19x19 boards in one
persistent launch of the fused 16-bit kernels) against the FP64 oracle run
per board size, with the value head's build-defined embedding of smaller
boards into the 361-input inner product (y*19+x, zeros elsewhere).

Tolerances (FP16 operands): max |d prob| < 0.005, max |d win| < 0.005,
top-1 agreement >= 95% over the mixed batch."""
import numpy as np
import pytest

from leela_amd.positions import NET_POLICY, NET_VALUE, make_batch

pytestmark = pytest.mark.gpu

SIDES = [19, 13, 9, 9, 13, 9, 19, 9, 9, 13, 13, 9, 19, 9, 13, 9, 9]
MAXM = {19: 216, 13: 101, 9: 48}  # 0.6 * board area


@pytest.fixture(scope="module")
def be(weights_b):
    import leela_amd
    b = leela_amd.Backend(0)
    pol, val = weights_b
    b.load_net(leela_amd.NET_POLICY, pol)
    b.load_net(leela_amd.NET_VALUE, val)
    b.finalize()
    yield b
    b.close()


def boards_for(net):
    return [make_batch(1, seed0=9000 + i, side=s, max_moves=MAXM[s],
                       net=net)[0] for i, s in enumerate(SIDES)]


def test_ragged_policy_and_value_vs_fp64(be, weights_b):
    import leela_amd
    import oracle
    pol, val = weights_b
    pb = boards_for(NET_POLICY)
    vb = boards_for(NET_VALUE)
    raw, prob = be.forward_ragged(leela_amd.NET_POLICY, leela_amd.PREC_FP16, pb)
    vraw, win = be.forward_ragged(leela_amd.NET_VALUE, leela_amd.PREC_FP16, vb)
    agree = []
    for i, s in enumerate(SIDES):
        r64, p64 = oracle.forward_f64(oracle.Net(pol), pb[i][None],
                                      oracle.HEAD_POLICY, S=s)
        _, w64 = oracle.forward_f64(oracle.Net(val), vb[i][None],
                                    oracle.HEAD_VALUE, S=s)
        assert prob[i].shape == (s * s,)
        assert np.abs(prob[i] - p64[0]).max() < 0.005, (i, s)
        assert abs(win[i][0] - w64[0, 0]) < 0.005, (i, s)
        agree.append(raw[i].argmax() == r64[0].argmax())
        np.testing.assert_allclose(prob[i].sum(), 1.0, atol=1e-5)
    assert np.mean(agree) >= 0.95


def test_ragged_19_matches_dense_path(be):
    """19x19 boards in a ragged batch give exactly the dense path's
    numbers, and a board's result does not depend on its slot mates."""
    import leela_amd
    pb = boards_for(NET_POLICY)
    for prec in (leela_amd.PREC_BF16, leela_amd.PREC_FP16):
        raw, _ = be.forward_ragged(leela_amd.NET_POLICY, prec, pb)
        idx19 = [i for i, s in enumerate(SIDES) if s == 19]
        dense, _ = be.forward(leela_amd.NET_POLICY, prec,
                              np.stack([pb[i] for i in idx19]))
        for k, i in enumerate(idx19):
            assert np.array_equal(raw[i], dense[k])
        nine = [i for i, s in enumerate(SIDES) if s == 9]
        alone, _ = be.forward_ragged(leela_amd.NET_POLICY, prec,
                                     [pb[nine[-1]]])
        assert np.array_equal(alone[0], raw[nine[-1]])
        vb = boards_for(NET_VALUE)
        v_all, _ = be.forward_ragged(leela_amd.NET_VALUE, prec, vb)
        v_one, _ = be.forward_ragged(leela_amd.NET_VALUE, prec, [vb[2]])
        assert np.array_equal(v_all[2], v_one[0])


def test_ragged_rejects_fp32_and_bad_sides(be):
    import leela_amd
    from leela_amd.backend import LeelaError
    pb = boards_for(NET_POLICY)[:3]
    with pytest.raises(LeelaError):
        be.forward_ragged(leela_amd.NET_POLICY, leela_amd.PREC_FP32, pb)
    with pytest.raises(LeelaError):
        be.forward_ragged(leela_amd.NET_POLICY, leela_amd.PREC_BF16,
                          [np.zeros((32, 4 * 4), np.uint8)])


def test_small_board_value_embedding_closed_form():
    """The GPU's value head for boards below 19x19 against the closed form
    of its definition (head plane at y*19+x of the 361 inputs): a value net
    whose trunk copies input plane 0 to the head plane (tests/fp8_nets.py),
    boards of 9, 13 and 19 in one ragged batch, FP16 and BF16 (the copied
    plane is 0/1, exact in both)."""
    import leela_amd
    from fp8_nets import plane_copy_value, plane_copy_value_net
    from leela_amd import nets
    layers = plane_copy_value_net(seed=4)
    b = leela_amd.Backend(0)
    try:
        b.load_net(leela_amd.NET_POLICY, nets.weight_set("A")[0])
        b.load_net(leela_amd.NET_VALUE, layers)
        b.finalize()
        rng = np.random.default_rng(12)
        boards = [(rng.random((32, s * s)) < 0.5).astype(np.uint8)
                  for s in (9, 13, 19, 9, 13, 9)]
        for prec in (leela_amd.PREC_FP16, leela_amd.PREC_BF16):
            raw, win = b.forward_ragged(leela_amd.NET_VALUE, prec, boards)
            for i, pl in enumerate(boards):
                s = int(round(np.sqrt(pl.shape[1])))
                want = plane_copy_value(layers, pl[0], s)
                assert abs(raw[i][0] - want) < 1e-4, (prec, i, s)
                assert abs(win[i][0] - (1 + np.tanh(want)) / 2) < 1e-4
    finally:
        b.close()

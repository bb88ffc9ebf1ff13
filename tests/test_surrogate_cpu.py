"""The surrogate's FP64 oracle pinned on the CPU: its convolution against
scipy's correlate2d, its backward pass against central finite differences,
its Adam against the closed form of Algorithm 1, and its BF16 rounding
against torch's. Also the host-side fallback of the leaf evaluator."""
import numpy as np
import pytest
from scipy.signal import correlate2d

from oracle import surrogate as orc
from leela_amd import positions
from leela_amd.backend import LeelaError


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    flat = np.zeros(orc.num_params())
    for name, shape, off in orc.layout():
        n = int(np.prod(shape))
        if name.startswith("W"):
            std = np.sqrt(2.0 / (shape[1] * 9))
        elif name == "F1":
            std = np.sqrt(2.0 / orc.P)
        else:
            std = 0.1
        flat[off:off + n] = rng.normal(0, std * scale, n)
    return flat


def _planes(n, seed):
    return positions.synthetic_planes(n, seed)


def test_layout_and_abi_param_count():
    assert orc.num_params() == 484930
    from leela_amd import surrogate
    assert surrogate.num_params() == orc.num_params()


def test_conv_is_same_padded_cross_correlation():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 5, 19, 19))
    W = rng.normal(size=(4, 5, 3, 3))
    got = np.einsum("ok,nkp->nop", W.reshape(4, -1), orc._cols(x))
    for n in range(2):
        for o in range(4):
            ref = sum(correlate2d(x[n, c], W[o, c], mode="same")
                      for c in range(5))
            np.testing.assert_allclose(got[n, o].reshape(19, 19), ref,
                                       atol=1e-12)
    # _col2im is the adjoint of _cols: <cols(x), d> == <x, col2im(d)>.
    d = rng.normal(size=(2, 45, 361))
    lhs = np.sum(orc._cols(x) * d)
    rhs = np.sum(x * orc._col2im(d, 5))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_backward_matches_finite_differences():
    flat = _params(11)
    planes = _planes(2, 5)
    y = np.array([0.2, 0.7])
    t, cache = orc.forward(flat, planes)
    loss, grad = orc.backward(cache, y)
    assert abs(loss - orc.loss_of(t, y)) < 1e-15
    rng = np.random.default_rng(0)
    h = 1e-6
    checked = 0
    for name, shape, off in orc.layout():
        n = int(np.prod(shape))
        for idx in rng.choice(n, size=min(2, n), replace=False):
            i = off + idx
            fp, fm = flat.copy(), flat.copy()
            fp[i] += h
            fm[i] -= h
            num = (orc.loss_of(orc.forward(fp, planes)[0], y) -
                   orc.loss_of(orc.forward(fm, planes)[0], y)) / (2 * h)
            assert abs(num - grad[i]) <= 1e-6 * max(1e-3, abs(num)) + 1e-9, \
                (name, idx, num, grad[i])
            checked += 1
    assert checked == 2 * len(orc.layout()) - 2   # bh and c2 have one entry


def test_skip_connections_carry_gradient():
    # Zero every conv weight of the blocks: the block convs then pass only
    # their bias, and the gradient at the first conv must still flow through
    # the six skips (a dropped skip term zeroes it).
    flat = _params(2)
    p = orc.unpack(flat)
    for layer in range(1, orc.NCONV):
        p[f"W{layer}"][...] = 0.0
    flat2 = np.concatenate([p[n].ravel() for n, _, _ in orc.layout()])
    t, cache = orc.forward(flat2, _planes(2, 9))
    _, g = orc.backward(cache, np.array([0.1, 0.9]))
    off = orc.layout()[0][2]
    assert np.abs(g[off:off + 18432]).max() > 1e-6


def test_adam_closed_form():
    # With a constant gradient g, m_hat = g and v_hat = g^2 at every step,
    # so each step moves by lr * g / (|g| + eps).
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p = np.array([1.0, 1.0, 1.0, 1.0])
    m = np.zeros(4)
    v = np.zeros(4)
    lr, eps = 1e-2, 1e-8
    for step in range(1, 6):
        p, m, v = orc.adam(p, g, m, v, step, lr, eps=eps)
        np.testing.assert_allclose(
            p, 1.0 - step * lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-15)
    # After the gradient stops, the moments decay: the step is
    # lr * (b1^k m-part) ... strictly smaller than lr in size.
    p2, _, _ = orc.adam(p, np.zeros(4), m, v, 6, lr)
    assert np.all(np.abs(p2 - p) < lr)


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.normal(size=10000) * 10.0 ** rng.integers(-6, 6, 10000),
                        [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.5 - 2 ** -8, 0.0]])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(orc.bf16(x).astype(np.float32), ref)


def test_bf16_oracle_is_close_to_exact():
    flat = _params(5)
    planes = _planes(4, 2)
    y = np.array([0.3, 0.6, 0.5, 0.8])
    t0, c0 = orc.forward(flat, planes)
    t1, c1 = orc.forward(flat, planes, bf16_points=True)
    assert np.abs(t0 - t1).max() < 0.05
    _, g0 = orc.backward(c0, y)
    _, g1 = orc.backward(c1, y)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.98


class _Broken:
    def predict(self, planes):
        raise LeelaError(-3, "unavailable")


def test_leaf_evaluator_falls_back_to_playouts():
    from leela_amd.surrogate import LeafEvaluator
    ps = [positions.Position.random(s, 19, 60) for s in (1, 2, 3)]
    plain = LeafEvaluator(None, playouts=8, seed=40).winrates(ps)
    ev = LeafEvaluator(_Broken(), playouts=8, seed=40)
    got = ev.winrates(ps)
    assert ev.fallbacks == 1
    np.testing.assert_array_equal(got, plain)
    for i, p in enumerate(ps):
        assert plain[i] == p.playout_winrate(8, 40 + i)

"""Pins of the CPU oracle (oracle/oracle.c) before it is trusted as the
checker of the HIP kernels.

The reference ships no CNN golden vectors (the weight blobs NN128.cpp and
NNValue.cpp are absent and the only known-answer test needs them), so the
oracle is pinned by mathematics: a pure-Python direct convolution on tiny
shapes, single-tap shift identities that fail for a flipped kernel or a
transposed weight layout, the ELU / softmax / tanh closed forms, the
symmetry group, and two independent whole-net restatements (im2col + sgemm
in FP32, direct convolution in FP64) that must agree.
"""
import math

import numpy as np
import pytest

import oracle


def py_conv(x, w, b, filt, elu=True):
    """Same-padded cross-correlation in plain Python loops."""
    cin, S, _ = x.shape
    cout = w.shape[0]
    pad = filt // 2
    out = np.zeros((cout, S, S))
    for o in range(cout):
        for yy in range(S):
            for xx in range(S):
                acc = float(b[o])
                for c in range(cin):
                    for ky in range(filt):
                        for kx in range(filt):
                            iy, ix = yy + ky - pad, xx + kx - pad
                            if 0 <= iy < S and 0 <= ix < S:
                                acc += float(w[o, c, ky, kx]) * float(x[c, iy, ix])
                if elu and acc <= 0:
                    acc = math.expm1(acc)
                out[o, yy, xx] = acc
    return out


@pytest.mark.parametrize("filt", [3, 5])
def test_conv_matches_python_loops(filt):
    rng = np.random.default_rng(filt)
    x = rng.standard_normal((3, 5, 5)).astype(np.float32)
    w = rng.standard_normal((4, 3, filt, filt)).astype(np.float32)
    b = rng.standard_normal(4).astype(np.float32)
    want = py_conv(x, w, b, filt)
    got32 = oracle.convolve_f32(x, w, b, filt)
    got64 = oracle.conv_direct_f64(x.astype(np.float64), w, b, filt)
    np.testing.assert_allclose(got64, want, rtol=0, atol=1e-12)
    np.testing.assert_allclose(got32, want, rtol=0, atol=2e-5)


def test_single_tap_orientation():
    """w[o][c][0][0] = 1 multiplies in[y-1][x-1] (cross-correlation, OIHW):
    a flipped kernel or swapped in/out channel axes moves the output."""
    S = 6
    x = np.arange(2 * S * S, dtype=np.float32).reshape(2, S, S) / 100.0
    w = np.zeros((3, 2, 3, 3), np.float32)
    w[2, 1, 0, 0] = 1.0
    b = np.zeros(3, np.float32)
    out = oracle.conv_direct_f64(x.astype(np.float64), w, b, 3, elu=False)
    want = np.zeros((3, S, S))
    want[2, 1:, 1:] = x[1, :-1, :-1]
    np.testing.assert_array_equal(out, want)
    out32 = oracle.convolve_f32(x, w, b, 3)
    np.testing.assert_allclose(out32, np.where(want > 0, want, 0.0), atol=1e-6)


def test_im2col_layout():
    """Im2Col.h:9-50: rows (c, kr, kc), columns output pixels."""
    S, filt = 4, 3
    x = np.arange(2 * S * S, dtype=np.float32).reshape(2, S, S) + 1
    col = oracle.im2col(x, filt)
    for c in range(2):
        for kr in range(filt):
            for kc in range(filt):
                row = col[(c * filt + kr) * filt + kc]
                for r in range(S):
                    for q in range(S):
                        iy, ix = r + kr - 1, q + kc - 1
                        v = x[c, iy, ix] if 0 <= iy < S and 0 <= ix < S else 0
                        assert row[r * S + q] == v


def test_elu_and_fc_head_rules():
    """ELU on convs and multi-output FCs; no ELU on a 1-output FC
    (Network.cpp:395-423)."""
    x = np.array([1.0, -2.0, 0.5], np.float32)
    w = np.eye(3, dtype=np.float32)
    b = np.zeros(3, np.float32)
    y = oracle.innerproduct_f32(x, w, b)
    np.testing.assert_allclose(y, [1.0, math.expm1(-2.0), 0.5], rtol=1e-6)
    y1 = oracle.innerproduct_f32(x, np.array([[0.0, 1.0, 0.0]], np.float32),
                                 np.zeros(1, np.float32))
    assert y1[0] == -2.0


def test_softmax_closed_forms():
    from scipy.special import softmax
    rng = np.random.default_rng(3)
    x = rng.standard_normal(361).astype(np.float32) * 3
    for T in (0.75, 1.0, 0.094):
        got = oracle.softmax_f32(x, T)
        np.testing.assert_allclose(got, softmax(x.astype(np.float64) / T),
                                   rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(oracle.softmax_f32(np.full(7, 2.5, np.float32),
                                                  0.75), np.full(7, 1 / 7),
                               rtol=1e-6)


def test_symmetry_group():
    inv = [0, 1, 2, 3, 4, 6, 5, 7]
    for s in range(8):
        m = [oracle.rotate_nn_idx(v, s) for v in range(361)]
        assert sorted(m) == list(range(361))
        for v in range(0, 361, 7):
            assert oracle.rev_rotate_nn_idx(oracle.rotate_nn_idx(v, s), s) == v
            assert oracle.rotate_nn_idx(oracle.rotate_nn_idx(v, s), inv[s]) == v
    # corner (x=0, y=0): flip-y -> (0, 18), flip-x -> (18, 0), swap fixes it
    assert oracle.rotate_nn_idx(0, 1) == 18 * 19
    assert oracle.rotate_nn_idx(0, 2) == 18
    assert oracle.rotate_nn_idx(1, 4) == 19  # (x=1,y=0) -> (x=0,y=1)


def test_rng_matches_reference_build():
    import oracle.ref_board as rb
    if not rb.available():
        pytest.skip("reference board checker not built")
    for seed in (1, 6, 7, 5489, 123456):
        s = oracle.rng_stream(seed, 64)
        assert np.array_equal((s >> np.uint64(32)).astype(np.uint32),
                              rb.rng_u32(seed, 64))
        f = (s >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
        assert np.array_equal(f, rb.rng_flt(seed, 64))


@pytest.fixture(scope="module")
def small_nets():
    """A 3-conv policy net and a 3-conv + 2-FC value net on 7x7 boards,
    random weights with biases."""
    rng = np.random.default_rng(5)

    def conv(f, cin, cout):
        return (0, f, cin, cout,
                (rng.standard_normal((cout, cin, f, f))
                 * np.sqrt(2 / (cin * f * f))).astype(np.float32),
                (rng.standard_normal(cout) * 0.1).astype(np.float32))

    pol = [conv(5, 4, 6), conv(3, 6, 6), conv(3, 6, 1)]
    val = [conv(5, 4, 6), conv(3, 6, 6), conv(3, 6, 1),
           (1, 0, 49, 8, (rng.standard_normal((8, 49)) * .2).astype(np.float32),
            (rng.standard_normal(8) * .1).astype(np.float32)),
           (1, 0, 8, 1, (rng.standard_normal((1, 8)) * .3).astype(np.float32),
            (rng.standard_normal(1) * .1).astype(np.float32))]
    return pol, val


def py_forward(layers, planes, S, head, T=0.75):
    x = planes.reshape(-1, S, S).astype(np.float64)
    flat = None
    for kind, f, cin, cout, w, b in layers:
        if kind == 0:
            x = py_conv(x, w, b, f)
        else:
            v = x.ravel() if flat is None else flat
            flat = w.astype(np.float64) @ v + b
            if cout > 1:
                flat = np.where(flat > 0, flat, np.expm1(flat))
    if head == 0:
        z = x.ravel()
        e = np.exp((z - z.max()) / T)
        return z, e / e.sum()
    return flat, (1 + np.tanh(flat)) / 2


def test_whole_net_three_ways(small_nets):
    """Python loops, FP64 direct and FP32 im2col+sgemm forwards agree on
    small nets of both head types."""
    pol, val = small_nets
    rng = np.random.default_rng(9)
    planes = (rng.random((3, 4, 49)) < 0.3).astype(np.uint8)
    for layers, head in ((pol, 0), (val, 1)):
        net = oracle.Net(layers)
        r64, o64 = oracle.forward_f64(net, planes, head, S=7)
        r32, o32 = oracle.forward_f32(net, planes, head, S=7)
        for i in range(3):
            pr, po = py_forward(layers, planes[i], 7, head)
            np.testing.assert_allclose(r64[i], pr, atol=1e-12)
            np.testing.assert_allclose(o64[i], po, atol=1e-12)
        np.testing.assert_allclose(r32, r64, atol=2e-5)
        np.testing.assert_allclose(o32, o64, atol=2e-5)


def test_reference_shapes_f32_vs_f64(weights_b):
    """The 19x19 reference shapes: the FP32 CPU path stays within 1e-4 of
    the FP64 oracle (the bar the HIP FP32 form is held to)."""
    from leela_amd.positions import make_batch
    pol, val = weights_b
    oracle.load_blas()
    planes = make_batch(3, seed0=21)
    for layers, head in ((pol, 0), (val, 1)):
        net = oracle.Net(layers)
        r32, o32 = oracle.forward_f32(net, planes, head)
        r64, o64 = oracle.forward_f64(net, planes, head)
        assert np.abs(r32 - r64).max() < 1e-4
        assert np.abs(o32 - o64).max() < 1e-5


def test_f64_gemm_equals_direct(small_nets, weights_a):
    """The im2col + dgemm FP64 forward equals the FP64 direct convolution
    (pinned above by Python loops) up to summation order, on small nets
    with both heads and on the reference shapes at 19x19 and 9x9 (the
    value head's 361-input embedding included)."""
    from leela_amd.positions import make_batch
    oracle.load_blas()
    pol, val = small_nets
    rng = np.random.default_rng(3)
    planes = (rng.random((4, 4, 49)) < 0.3).astype(np.uint8)
    for layers, head in ((pol, 0), (val, 1)):
        net = oracle.Net(layers)
        a = oracle.forward_f64(net, planes, head, S=7)
        b = oracle.forward_f64(net, planes, head, S=7, gemm=True)
        np.testing.assert_allclose(b[0], a[0], atol=1e-12)
        np.testing.assert_allclose(b[1], a[1], atol=1e-12)
    pa, va = weights_a
    for side in (19, 9):
        for layers, head, net_id in ((pa, 0, 0), (va, 1, 1)):
            planes = make_batch(2, seed0=77, side=side, max_moves=40,
                                net=net_id)
            net = oracle.Net(layers)
            a = oracle.forward_f64(net, planes, head, S=side)
            b = oracle.forward_f64(net, planes, head, S=side, gemm=True)
            np.testing.assert_allclose(b[0], a[0], atol=1e-10)
            np.testing.assert_allclose(b[1], a[1], atol=1e-10)


def test_fp8_e4m3_rounding_matches_torch():
    """The FP8 emulation's e4m3 rounding is torch's float8_e4m3fn (OCP
    E4M3FN, round to nearest even) on every finite input range, with
    saturation at 448."""
    import torch
    from oracle import fp8
    rng = np.random.default_rng(8)
    x = np.concatenate([
        rng.normal(size=20000) * 2.0 ** rng.integers(-12, 9, 20000),
        np.arange(-448, 449, 0.25),
        [2.0 ** -9, 1.5 * 2.0 ** -10, 2.0 ** -10, 3 * 2.0 ** -10, 0.0,
         2.0 ** -6 - 2.0 ** -10, 15.5 * 2.0 ** -4, 464.0, -464.0]])
    x = x[np.abs(x) <= 448].astype(np.float32)
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).float().numpy()
    np.testing.assert_array_equal(fp8.e4m3(x).astype(np.float32), ref)
    assert fp8.e4m3(1000.0) == 448.0 and fp8.e4m3(-1e9) == -448.0


@pytest.mark.parametrize("side", [9, 13, 19])
def test_small_board_value_embedding_closed_form(side):
    """The value head of a board smaller than 19x19 sees its head plane at
    index y*19+x of the 361-input inner product (zeros elsewhere): on a net
    whose trunk copies plane 0 to the head plane, the FP64 oracle equals
    the closed form of that definition (a contiguous S*S embedding, or a
    transposed one, would not)."""
    from fp8_nets import plane_copy_value, plane_copy_value_net
    layers = plane_copy_value_net(seed=side)
    rng = np.random.default_rng(side)
    planes = (rng.random((3, 32, side * side)) < 0.5).astype(np.uint8)
    raw, win = oracle.forward_f64(oracle.Net(layers), planes,
                                  oracle.HEAD_VALUE, S=side)
    for i in range(3):
        want = plane_copy_value(layers, planes[i, 0], side)
        assert abs(raw[i, 0] - want) < 1e-12
        assert abs(win[i, 0] - (1 + np.tanh(want)) / 2) < 1e-12
    if side < 19:
        # A contiguous embedding differs: the check can tell them apart.
        e_bad = np.zeros(361)
        e_bad[:side * side] = planes[0, 0]
        (_, _, _, _, w1, b1), (_, _, _, _, w2, b2) = layers[-2:]
        h = w1.astype(np.float64) @ e_bad + b1
        h = np.where(h > 0, h, np.expm1(np.minimum(h, 0)))
        assert abs(float(w2[0] @ h + b2[0]) - raw[0, 0]) > 1e-3

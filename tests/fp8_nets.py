"""Nets of the reference shapes whose FP8 form is exact, for pinning the FP8
emulation (oracle/fp8.py) and the GPU's FP8 kernels against FP64.

The stem copies one input plane per output channel at one tap, with a
weight that is an e4m3 value times 2^-7 (the largest 448 * 2^-7 = 3.5);
every later conv copies one input channel per output channel at one tap
with weight 1 or 1/2 (the head conv sums a few channels). All activations
stay >= 0 (ELU is the identity), each board's maximum stays 3.5 times a
power of two, so its activation scale (max / 448) is a power of two, and
every activation is an e4m3 value times that scale; every weight divided
by its channel's scale (max |W[o]| / 448) is 448 or another e4m3 value.
Quantisation then changes nothing and the form must equal the FP64
forward, up to the FP32 rounding of the weight scales. Options plant one
value each, with the value the form must turn it into:
  round_weight: a head weight of 1.0625 beside a 2.0 in the same output
      channel: 1.0625 / (2 / 448) = 238 rounds to 240 -> 240 * 2 / 448;
  round_act:    a stem channel on the path into the head of 1.0625 * 2^-7,
      on boards whose activation scale is 2^-7: 1.0625 rounds (tie to
      even) to 1.0 -> 2^-7;
  tiny_channel: a trunk channel on the path into the head whose only
      weight is 1.25 * 2^-2 (exact under its own channel's scale; under
      one scale for the whole layer, 0.3125 * 448 = 140 would round to 144).
"""
from __future__ import annotations

import numpy as np

E4M3_VALUES = np.array([448, 384, 320, 256, 224, 192, 160, 128, 112, 96,
                        80, 64, 56, 48, 40, 32], np.float64)
HEAD_TAPS = ((3, 1, 1, 2.0), (17, 0, 1, 0.5), (40, 2, 2, 1.0))
TINY_LAYER = 3


def _conv(cout, cin, f):
    return np.zeros((cout, cin, f, f), np.float32)


def exact_net(policy: bool, round_weight=False, round_act=False,
              tiny_channel=False, seed=0):
    """(layers, expected_layers): the net, and the same net with every
    planted value replaced by what the FP8 form rounds it to."""
    rng = np.random.default_rng(seed)
    c = 128 if policy else 64
    m1 = 96 if policy else 64
    nmid = 11 if policy else 10
    # Stem 5x5 32 -> m1: channel o copies plane o % 32 at tap (dy, dx).
    stem_val = np.array([E4M3_VALUES[o % len(E4M3_VALUES)] / 128.0
                         for o in range(m1)])
    stem_tap = [(o % 32, *rng.integers(1, 4, 2)) for o in range(m1)]
    # Middle 3x3 convs: output channel o copies channel src[l][o].
    cins = [m1] + [c] * (nmid - 1)
    srcs = [[(o + 7 * l) % cins[l] for o in range(c)] for l in range(nmid)]
    tiny = None
    if tiny_channel:
        # Head input channel 40 followed back to layer TINY_LAYER and to the
        # stem, whose value becomes a power of two (2.0) so that 1.25 times
        # every value on the path is an e4m3 value.
        ch = 40
        for l in range(nmid - 1, TINY_LAYER, -1):
            ch = srcs[l][ch]
        tiny = ch
        for l in range(TINY_LAYER, -1, -1):
            ch = srcs[l][ch]
        stem_val[ch] = 2.0
    act_ch = None
    if round_act:
        # The stem channel at the start of head input channel 17's path.
        ch = 17
        for l in range(nmid - 1, -1, -1):
            ch = srcs[l][ch]
        act_ch = ch
        stem_val[ch] = 1.0625 / 128.0
    stem = _conv(m1, 32, 5)
    for o, (plane, dy, dx) in enumerate(stem_tap):
        stem[o, plane, dy, dx] = stem_val[o]
    layers = [[0, 5, 32, m1, stem, np.zeros(m1, np.float32)]]
    # Weights 1 or 1/2; a channel that copies one of the layer's largest
    # channels keeps weight 1 and the centre tap, so every board's maximum
    # stays 3.5 and its activation scale 2^-7.
    val = stem_val.copy()
    for l in range(nmid):
        w = _conv(c, cins[l], 3)
        top = val.max()
        nval = np.zeros(c)
        for o in range(c):
            sv = val[srcs[l][o]]
            if sv == top:
                wv, dy, dx = 1.0, 1, 1
            else:
                wv = 1.0 if o % 4 else 0.5
                dy, dx = rng.integers(0, 3, 2)
            if l == TINY_LAYER and o == tiny:
                wv = 1.25 * 2.0 ** -2
            w[o, srcs[l][o], dy, dx] = wv
            nval[o] = sv * wv
        val = nval
        layers.append([0, 3, cins[l], c, w, np.zeros(c, np.float32)])
    # Head 3x3 c -> 1: a few channels.
    head = _conv(1, c, 3)
    for ch, dy, dx, v in HEAD_TAPS:
        head[0, ch, dy, dx] = v
    if round_weight:
        head[0, 61, 1, 0] = 1.0625
    layers.append([0, 3, c, 1, head, np.zeros(1, np.float32)])
    if not policy:
        fw1 = (rng.standard_normal((256, 361)) * 0.05).astype(np.float32)
        fb1 = (rng.standard_normal(256) * 0.1).astype(np.float32)
        fw2 = (rng.standard_normal((1, 256)) * 0.05).astype(np.float32)
        layers.append([1, 0, 361, 256, fw1, fb1])
        layers.append([1, 0, 256, 1, fw2, np.zeros(1, np.float32)])
    expected = [[k, f, ci, co, w.copy(), b.copy()]
                for k, f, ci, co, w, b in layers]
    if round_weight:
        expected[nmid + 1][4][0, 61, 1, 0] = np.float32(240.0 * 2.0 / 448.0)
    if round_act:
        plane, dy, dx = stem_tap[act_ch]
        expected[0][4][act_ch, plane, dy, dx] = 1.0 / 128.0
    return [tuple(x) for x in layers], [tuple(x) for x in expected]


def planes(n: int, seed: int = 1) -> np.ndarray:
    """Random binary planes in which every plane has set points (so every
    stem channel reaches its value somewhere on every board)."""
    rng = np.random.default_rng(seed)
    p = (rng.random((n, 32, 361)) < 0.4).astype(np.uint8)
    p[:, :, 180] = 1
    return p


def plane_copy_value_net(seed=0):
    """A value net of the reference shapes whose conv trunk copies input
    plane 0 through to the head plane (one weight 1 per layer at the centre
    tap, every other conv weight and bias 0; the plane is 0/1, so each ELU
    is the identity), with random inner products. Its value on a board of
    side S is then, in closed form, fc2 . ELU(fc1 e + b1) + b2 where e is
    plane 0 placed at y*19+x of a 361-vector (zeros elsewhere)."""
    rng = np.random.default_rng(seed)
    layers = []
    w = _conv(64, 32, 5)
    w[0, 0, 2, 2] = 1.0
    layers.append((0, 5, 32, 64, w, np.zeros(64, np.float32)))
    for _ in range(10):
        w = _conv(64, 64, 3)
        w[0, 0, 1, 1] = 1.0
        layers.append((0, 3, 64, 64, w, np.zeros(64, np.float32)))
    w = _conv(1, 64, 3)
    w[0, 0, 1, 1] = 1.0
    layers.append((0, 3, 64, 1, w, np.zeros(1, np.float32)))
    fw1 = (rng.standard_normal((256, 361)) * 0.1).astype(np.float32)
    fb1 = (rng.standard_normal(256) * 0.1).astype(np.float32)
    fw2 = (rng.standard_normal((1, 256)) * 0.1).astype(np.float32)
    fb2 = (rng.standard_normal(1) * 0.1).astype(np.float32)
    layers.append((1, 0, 361, 256, fw1, fb1))
    layers.append((1, 0, 256, 1, fw2, fb2))
    return layers


def plane_copy_value(layers, plane0, side):
    """The closed form of plane_copy_value_net's output (raw value) for a
    board of `side` whose plane 0 is plane0 [side*side]."""
    e = np.zeros(361)
    for y in range(side):
        for x in range(side):
            e[y * 19 + x] = plane0[y * side + x]
    (_, _, _, _, w1, b1), (_, _, _, _, w2, b2) = layers[-2:]
    h = w1.astype(np.float64) @ e + b1
    h = np.where(h > 0, h, np.expm1(np.minimum(h, 0)))
    return float(w2.astype(np.float64)[0] @ h + b2[0])

"""Layer tables of Leela's nets and their synthetic weight sets.

Shapes are the ones Network.cpp declares for the BLAS build (Network.cpp:
82-137): the 128-wide policy net and the 64-wide value net. The shipped
weight values (NN128.cpp, NNValue.cpp) are not available, so weights are
drawn He-normal, w ~ N(0, 2 / fan_in), from the reference RNG (Box-Muller,
lz_normal_fill):
  set "A": weights only, biases zero;
  set "B": as A plus biases N(0, 0.1^2), exercising the bias path.
Layer l of net k (0 policy, 1 value) draws its weights from stream
seed * 1000 + 100 * k + l and its biases from bias_seed * 1000 + 100 * k + l.
"""
from __future__ import annotations

import math

import numpy as np

from .backend import normal_fill

CONV, FC = 0, 1

# (kind, filter, cin, cout) in push order (Network.cpp:206-233).
POLICY_SHAPES = ([(CONV, 5, 32, 96), (CONV, 3, 96, 128)]
                 + [(CONV, 3, 128, 128)] * 10
                 + [(CONV, 3, 128, 1)])
VALUE_SHAPES = ([(CONV, 5, 32, 64)]
                + [(CONV, 3, 64, 64)] * 10
                + [(CONV, 3, 64, 1), (FC, 0, 361, 256), (FC, 0, 256, 1)])

NET_SHAPES = (POLICY_SHAPES, VALUE_SHAPES)


def param_count(shapes) -> int:
    return sum((cin * max(f, 1) ** 2 + 1) * cout for _, f, cin, cout in shapes)


def flops_per_position(shapes, side: int = 19) -> int:
    """Multiply-adds x 2 of one forward (conv over side^2 pixels)."""
    P = side * side
    tot = 0
    for kind, f, cin, cout in shapes:
        tot += 2 * cout * cin * f * f * P if kind == CONV else 2 * cout * cin
    return tot


def he_normal(net: int, seed: int = 7, bias_seed: int | None = None,
              bias_std: float = 0.1):
    """Synthetic weights for net 0 (policy) or 1 (value)."""
    layers = []
    for l, (kind, f, cin, cout) in enumerate(NET_SHAPES[net]):
        fan_in = cin * f * f if kind == CONV else cin
        w = normal_fill(seed * 1000 + 100 * net + l, cout * fan_in,
                        math.sqrt(2.0 / fan_in))
        if bias_seed is None:
            b = np.zeros(cout, np.float32)
        else:
            b = normal_fill(bias_seed * 1000 + 100 * net + l, cout, bias_std)
        if kind == CONV:
            w = w.reshape(cout, cin, f, f)
        else:
            w = w.reshape(cout, cin)
        layers.append((kind, f, cin, cout, w, b))
    return layers


def weight_set(name: str):
    """(policy_layers, value_layers) for set "A" or "B"."""
    if name == "A":
        return he_normal(0, 7), he_normal(1, 7)
    if name == "B":
        return he_normal(0, 7, 8), he_normal(1, 7, 8)
    raise ValueError(name)

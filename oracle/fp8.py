"""FP8 e4m3 emulation of the library's FP8 form — TEST INFRASTRUCTURE ONLY.

The FP8 form is this library's own variant (the reference computes in FP32),
so this restates the form's definition, in FP64 arithmetic:

  * the 5x5 stem: weights rounded to FP16, inputs the binary planes;
  * every later conv l: weights quantised per output channel,
    scale_o = max_k |W[o][k]| / 448 (FP32), Wq = e4m3(W / scale_o);
    activations quantised per board, s = max |a| / 448 over the board's
    previous layer output, aq = e4m3(a / s);
    output = ELU(sum Wq aq * (scale_o * s) + b);
  * the head conv's ELU output feeds softmax(T) (policy) or the FP32
    inner-product head (value), as in the reference (Network.cpp:676-832).

e4m3 here is the OCP E4M3FN format (bias 7, 3 mantissa bits, subnormals of
2^-9, maximum 448, no infinities) with round-to-nearest-even and saturation
at 448; tests pin it against torch.float8_e4m3fn.
"""
from __future__ import annotations

import numpy as np

from . import conv_direct_f64, innerproduct_f32, softmax_f32

E4M3_MAX = 448.0


def e4m3(x):
    """Round to the nearest e4m3 value (ties to even), saturating at 448."""
    x = np.asarray(x, np.float64)
    a = np.abs(x)
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -40)))
    step = 2.0 ** (np.maximum(e, -6.0) - 3.0)
    q = np.minimum(np.round(a / step) * step, E4M3_MAX)
    return np.copysign(q, x)


def _elu(v):
    return np.where(v > 0.0, v, np.expm1(np.minimum(v, 0.0)))


def forward_fp8(layers, planes, policy: bool, temperature: float = 0.75):
    """layers: (kind, filter, cin, cout, w, b) in push order; planes uint8
    [n][32][361]. Returns (raw, out) float64 like the library's outputs."""
    convs = [l for l in layers if l[0] == 0]
    fcs = [l for l in layers if l[0] == 1]
    n = planes.shape[0]
    raw = np.zeros((n, 361 if policy else 1))
    out = np.zeros_like(raw)
    for i in range(n):
        x = planes[i].reshape(32, 19, 19).astype(np.float64)
        s_in = 1.0
        for li, (_, f, cin, cout, w, b) in enumerate(convs):
            w = np.asarray(w, np.float32).reshape(cout, cin, f, f)
            zb = np.zeros(cout, np.float32)
            if li == 0:
                acc = conv_direct_f64(x, w.astype(np.float16).astype(np.float32),
                                      zb, f, elu=False)
                v = _elu(acc + np.asarray(b, np.float64)[:, None, None])
            else:
                mx = np.abs(w.reshape(cout, -1)).max(axis=1)
                scale = np.where(mx > 0, np.float32(mx) / np.float32(E4M3_MAX),
                                 np.float32(1.0)).astype(np.float32)
                wq = e4m3(w / scale[:, None, None, None]).astype(np.float32)
                acc = conv_direct_f64(x, wq, zb, f, elu=False)
                sc = (scale.astype(np.float64) * s_in)[:, None, None]
                v = _elu(acc * sc + np.asarray(b, np.float64)[:, None, None])
            if li == len(convs) - 1:
                break
            m = np.abs(v).max()
            s_in = m / E4M3_MAX if m > 0 else 1.0
            x = e4m3(np.clip(v / s_in, -E4M3_MAX, E4M3_MAX))
        head = v.reshape(-1)
        if policy:
            raw[i] = head
            out[i] = softmax_f32(head.astype(np.float32), temperature)
        else:
            (_, _, _, _, w1, b1), (_, _, _, _, w2, b2) = fcs
            # The oracle's inner product applies ELU when outputs > 1
            # (Network.cpp:395-423).
            h = innerproduct_f32(head.astype(np.float32), w1, b1)
            z = innerproduct_f32(h, w2, b2)[0]
            raw[i, 0] = z
            out[i, 0] = (1.0 + np.tanh(z)) / 2.0
    return raw, out

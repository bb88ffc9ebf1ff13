"""FP64 CPU oracle of the learned playout surrogate — TEST INFRASTRUCTURE ONLY.

The surrogate is this library's own network (the reference has none; its
teacher is Playout::mc_owner, Playout.cpp:174-217), so the oracle restates
the network's definition, not a reference routine:

    a0  = elu(conv3x3(planes; W0) + b0)                       32 -> 64
    block k = 1..6:
        u_k = elu(conv3x3(h_{k-1}; W_{2k-1}) + b_{2k-1})
        h_k = elu(conv3x3(u_k; W_{2k}) + b_{2k} + h_{k-1})    (h_0 = a0)
    s   = elu(sum_c wh[c] h_6[c, p] + bh)                     1x1, 64 -> 1
    f   = elu(F1 s + c1)                                      361 -> 64
    t   = tanh(F2 . f + c2)                                   64 -> 1
    win = (1 + t) / 2
    loss = mean_n (t_n - (2 y_n - 1))^2

Convolutions are zero-padded "same" cross-correlations with OIHW weights
(as the reference's convolve, Network.cpp:344-393). The flat parameter
vector is, per conv layer, W[64][Cin][3][3] then b[64], followed by wh[64],
bh, F1[64][361], c1[64], F2[64], c2.

`bf16=True` rounds at the points where the HIP path stores 16-bit values
(conv weights, every conv activation and every conv-gradient map) with
round-to-nearest-even, so that the GPU's gradients can be compared at
tight tolerance; `bf16=False` is the exact math that the finite-difference
tests pin. Adam is Kingma & Ba (2015), Algorithm 1.
"""
from __future__ import annotations

import numpy as np

S = 19
P = S * S
CH = 64
NCONV = 13
HID = 64


def layout():
    """[(name, shape, offset)] of the flat parameter vector."""
    out, off = [], 0
    for layer in range(NCONV):
        cin = 32 if layer == 0 else CH
        for name, shape in ((f"W{layer}", (CH, cin, 3, 3)), (f"b{layer}", (CH,))):
            out.append((name, shape, off))
            off += int(np.prod(shape))
    for name, shape in (("wh", (CH,)), ("bh", (1,)), ("F1", (HID, P)),
                        ("c1", (HID,)), ("F2", (HID,)), ("c2", (1,))):
        out.append((name, shape, off))
        off += int(np.prod(shape))
    return out


def num_params() -> int:
    name, shape, off = layout()[-1]
    return off + int(np.prod(shape))


def unpack(flat):
    return {n: flat[o:o + int(np.prod(s))].reshape(s) for n, s, o in layout()}


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _cols(x):
    """x [n][C][19][19] -> im2col [n][C*9][361], index c*9 + kh*3 + kw."""
    n, c = x.shape[:2]
    xp = np.zeros((n, c, S + 2, S + 2))
    xp[:, :, 1:-1, 1:-1] = x
    cols = np.empty((n, c, 9, S, S))
    for kh in range(3):
        for kw in range(3):
            cols[:, :, kh * 3 + kw] = xp[:, :, kh:kh + S, kw:kw + S]
    return cols.reshape(n, c * 9, P)


def _col2im(dcols, c):
    """Adjoint of _cols: [n][C*9][361] -> [n][C][19][19]."""
    n = dcols.shape[0]
    d = dcols.reshape(n, c, 9, S, S)
    xp = np.zeros((n, c, S + 2, S + 2))
    for kh in range(3):
        for kw in range(3):
            xp[:, :, kh:kh + S, kw:kw + S] += d[:, :, kh * 3 + kw]
    return xp[:, :, 1:-1, 1:-1]


def conv_pre(x, W, b, res=None):
    """Pre-activation of one conv: x [n][C][19][19], W [64][C][3][3]
    (already rounded when emulating), b [64], res [n][64][361] or None.
    Returns (z [n][64][361], im2col of x)."""
    c = _cols(x)
    z = np.einsum("ok,nkp->nop", W.reshape(W.shape[0], -1), c,
                  optimize=True) + b[None, :, None]
    if res is not None:
        z = z + res
    return z, c


def conv_wgrad(e, cols, cin):
    """dW [64][cin][3][3] and db [64] from the gradient e [n][64][361] at a
    conv's pre-activation and the im2col of its input."""
    dW = np.einsum("nop,nkp->ok", e, cols, optimize=True)
    return dW.reshape(e.shape[1], cin, 3, 3), e.sum((0, 2))


def conv_bwd_data(e, W):
    """Gradient at a conv's input [n][cin][361] from e [n][64][361]."""
    n, cin = e.shape[0], W.shape[1]
    d = np.einsum("ok,nop->nkp", W.reshape(W.shape[0], -1), e, optimize=True)
    return _col2im(d, cin).reshape(n, cin, P)


def _elu(v):
    return np.where(v > 0.0, v, np.expm1(np.minimum(v, 0.0)))


def _elu_grad(a):
    """d ELU / d v through the output a = ELU(v)."""
    return np.where(a > 0.0, 1.0, a + 1.0)


def head_fwd(h6, p):
    """s [n][361], f [n][64], t [n] from the last conv's output [n][64][361]."""
    s = _elu(np.einsum("c,ncp->np", p["wh"], h6, optimize=True) + p["bh"][0])
    f = _elu(s @ p["F1"].T + p["c1"])
    t = np.tanh(f @ p["F2"] + p["c2"][0])
    return s, f, t


def head_bwd(h6, s, f, t, y, p):
    """Head gradients {wh, bh, F1, c1, F2, c2} and ds [n][361] of the loss
    mean_n (t - (2y - 1))^2 over the n rows given."""
    n = t.shape[0]
    y = np.asarray(y, np.float64)
    dz = 2.0 * (t - (2.0 * y - 1.0)) / n * (1.0 - t * t)
    g = {"F2": dz @ f, "c2": np.array([dz.sum()])}
    df = np.outer(dz, p["F2"]) * _elu_grad(f)
    g["F1"] = df.T @ s
    g["c1"] = df.sum(0)
    ds = (df @ p["F1"]) * _elu_grad(s)
    g["wh"] = np.einsum("np,ncp->c", ds, h6, optimize=True)
    g["bh"] = np.array([ds.sum()])
    return g, ds


def forward(flat, planes, bf16_points=False):
    """planes uint8 [n][32][361] -> (t [n], cache)."""
    p = unpack(np.asarray(flat, np.float64))
    rnd = bf16 if bf16_points else (lambda v: v)
    n = planes.shape[0]
    x = planes.reshape(n, 32, S, S).astype(np.float64)
    acts, cols = [], []
    h = None
    for layer in range(NCONV):
        res = h.reshape(n, CH, P) if (layer >= 2 and layer % 2 == 0) else None
        z, c = conv_pre(x, rnd(p[f"W{layer}"]), p[f"b{layer}"], res)
        cols.append(c)
        a = rnd(_elu(z)).reshape(n, CH, S, S)
        acts.append(a)
        if layer == 0 or layer % 2 == 0:
            h = a
        x = a
    s, f, t = head_fwd(acts[-1].reshape(n, CH, P), p)
    cache = dict(p=p, planes=planes, acts=acts, cols=cols, s=s, f=f, t=t,
                 bf16=bf16_points)
    return t, cache


def loss_of(t, y):
    return float(np.mean((t - (2.0 * np.asarray(y, np.float64) - 1.0)) ** 2))


def backward(cache, y):
    """Loss and flat gradient of loss_of(t, y) for the cached forward."""
    p, acts, cols = cache["p"], cache["acts"], cache["cols"]
    rnd = bf16 if cache["bf16"] else (lambda v: v)
    s, f, t = cache["s"], cache["f"], cache["t"]
    n = t.shape[0]
    h6 = acts[-1].reshape(n, CH, P)
    g, ds = head_bwd(h6, s, f, t, y, p)
    # Map l: gradient at the pre-activation of conv l.
    maps = [None] * NCONV
    maps[-1] = rnd(ds[:, None, :] * p["wh"][None, :, None] * _elu_grad(h6))
    for layer in range(NCONV - 1, -1, -1):
        cin = 32 if layer == 0 else CH
        e = maps[layer]
        g[f"W{layer}"], g[f"b{layer}"] = conv_wgrad(e, cols[layer], cin)
        if layer == 0:
            break
        dx = conv_bwd_data(e, rnd(p[f"W{layer}"]))
        if layer % 2 == 1:
            # First conv of a block: add the skip's gradient (map l + 1).
            dx = dx + maps[layer + 1]
        below = acts[layer - 1].reshape(n, CH, P)
        maps[layer - 1] = rnd(dx * _elu_grad(below))
    cache["maps"] = maps
    flat = np.concatenate([g[name].ravel() for name, _, _ in layout()])
    return loss_of(t, y), flat


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); step counts from 1.
    Returns new (params, m, v)."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mhat = m / (1.0 - b1 ** step)
    vhat = v / (1.0 - b2 ** step)
    return params - lr * mhat / (np.sqrt(vhat) + eps), m, v

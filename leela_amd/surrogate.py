"""Learned playout surrogate: a leaf win-rate estimator trained on playout
outcomes (the labels of Playout::mc_owner, Playout.cpp:174-217), with plain
playouts as the fallback.

``Surrogate`` wraps the lz_surrogate_* C ABI (HIP kernels: BF16 MFMA convs,
FP32 master weights, Adam). ``train`` runs minibatch training on planes and
labels resident in device memory. ``LeafEvaluator`` answers win-rate
queries with the surrogate when one is loaded and with playouts otherwise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .backend import LZ_OK, LeelaError, load_library
from . import positions as _positions

P = 361
PLANES = 32


def _lib():
    L = load_library()
    if not getattr(L, "_surrogate_bound", False):
        vp, f32p = C.c_void_p, C.POINTER(C.c_float)
        u8p = C.POINTER(C.c_uint8)
        L.lz_surrogate_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int,
                                          C.c_uint32]
        L.lz_surrogate_destroy.argtypes = [vp]
        L.lz_surrogate_destroy.restype = None
        L.lz_surrogate_num_params.argtypes = []
        L.lz_surrogate_last_error.argtypes = [vp]
        L.lz_surrogate_last_error.restype = C.c_char_p
        L.lz_surrogate_get_params.argtypes = [vp, f32p]
        L.lz_surrogate_set_params.argtypes = [vp, f32p]
        L.lz_surrogate_set_adam.argtypes = [vp, C.c_float, C.c_float, C.c_float]
        L.lz_surrogate_step.argtypes = [vp]
        L.lz_surrogate_step.restype = C.c_longlong
        L.lz_surrogate_forward_device.argtypes = [vp, C.c_int, vp, vp, vp]
        L.lz_surrogate_predict.argtypes = [vp, C.c_int, u8p, f32p]
        L.lz_surrogate_gradients.argtypes = [vp, C.c_int, u8p, f32p, f32p, f32p]
        L.lz_surrogate_train_step_device.argtypes = [vp, C.c_int, vp, vp,
                                                     C.c_float, f32p, vp]
        L.lz_surrogate_keep_maps.argtypes = [vp, C.c_int]
        L.lz_surrogate_read_map.argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_uint16)]
        L.lz_surrogate_read_head.argtypes = [vp, C.c_int, f32p, f32p, f32p]
        L.lz_surrogate_save.argtypes = [vp, C.c_char_p]
        L.lz_surrogate_load.argtypes = [vp, C.c_char_p]
        L._surrogate_bound = True
    return L


def num_params() -> int:
    return int(_lib().lz_surrogate_num_params())


class Surrogate:
    """One lz_surrogate: parameters, Adam state and buffers for batches of
    up to max_batch 19x19 positions on one GPU."""

    def __init__(self, device: int = 0, max_batch: int = 512, seed: int = 1):
        self._L = _lib()
        h = C.c_void_p()
        rc = self._L.lz_surrogate_create(C.byref(h), device, max_batch, seed)
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_surrogate_create")
        self._h = h
        self.max_batch = max_batch
        self.n_params = num_params()

    def _chk(self, rc, what):
        if rc != LZ_OK:
            msg = self._L.lz_surrogate_last_error(self._h)
            raise LeelaError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.lz_surrogate_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def step(self) -> int:
        return int(self._L.lz_surrogate_step(self._h))

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        self._chk(self._L.lz_surrogate_get_params(
            self._h, out.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return out

    def set_params(self, flat) -> None:
        a = np.ascontiguousarray(flat, np.float32)
        if a.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters")
        self._chk(self._L.lz_surrogate_set_params(
            self._h, a.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    def set_adam(self, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        self._chk(self._L.lz_surrogate_set_adam(self._h, beta1, beta2, eps),
                  "set_adam")

    @staticmethod
    def _planes(planes) -> np.ndarray:
        a = np.ascontiguousarray(planes, np.uint8)
        if a.ndim != 3 or a.shape[1:] != (PLANES, P):
            raise ValueError("planes must be [n][32][361] uint8")
        return a

    def predict(self, planes) -> np.ndarray:
        """Win rates for the side to move of value planes [n][32][361]."""
        a = self._planes(planes)
        out = np.empty(a.shape[0], np.float32)
        for i in range(0, a.shape[0], self.max_batch):
            blk = np.ascontiguousarray(a[i:i + self.max_batch])
            o = np.empty(blk.shape[0], np.float32)
            self._chk(self._L.lz_surrogate_predict(
                self._h, blk.shape[0], blk.ctypes.data_as(C.POINTER(C.c_uint8)),
                o.ctypes.data_as(C.POINTER(C.c_float))), "predict")
            out[i:i + blk.shape[0]] = o
        return out

    def forward_device(self, n: int, planes_ptr: int, out_ptr: int,
                       stream: int = 0) -> None:
        self._chk(self._L.lz_surrogate_forward_device(
            self._h, n, planes_ptr, out_ptr, stream or None), "forward")

    def gradients(self, planes, targets):
        """(loss, flat gradient) of mean (t - (2y - 1))^2; no update."""
        a = self._planes(planes)
        y = np.ascontiguousarray(targets, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        self._chk(self._L.lz_surrogate_gradients(
            self._h, a.shape[0], a.ctypes.data_as(C.POINTER(C.c_uint8)),
            y.ctypes.data_as(C.POINTER(C.c_float)),
            g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)),
            "gradients")
        return loss.value, g

    def train_step_device(self, n: int, planes_ptr: int, targets_ptr: int,
                          lr: float, want_loss: bool = True,
                          stream: int = 0):
        loss = C.c_float()
        self._chk(self._L.lz_surrogate_train_step_device(
            self._h, n, planes_ptr, targets_ptr, lr,
            C.byref(loss) if want_loss else None, stream or None),
            "train_step")
        return loss.value if want_loss else None

    def keep_maps(self, on: bool = True) -> None:
        """Keep all 13 gradient maps of the next backward pass."""
        self._chk(self._L.lz_surrogate_keep_maps(self._h, int(on)), "keep_maps")

    def read_map(self, kind: int, layer: int, n: int) -> np.ndarray:
        """Raw BF16 bits [n][361][64] of conv `layer`'s output (kind 0) or
        of the gradient at its pre-activation (kind 1), from the last pass."""
        out = np.empty((n, P, 64), np.uint16)
        self._chk(self._L.lz_surrogate_read_map(
            self._h, kind, layer, n, out.ctypes.data_as(C.POINTER(C.c_uint16))),
            "read_map")
        return out

    def read_head(self, n: int):
        """The head's saved (s [n][361], f [n][64], t [n])."""
        sv = np.empty((n, P), np.float32)
        fv = np.empty((n, 64), np.float32)
        tv = np.empty(n, np.float32)
        f32p = C.POINTER(C.c_float)
        self._chk(self._L.lz_surrogate_read_head(
            self._h, n, sv.ctypes.data_as(f32p), fv.ctypes.data_as(f32p),
            tv.ctypes.data_as(f32p)), "read_head")
        return sv, fv, tv

    def save(self, path) -> None:
        self._chk(self._L.lz_surrogate_save(self._h, os.fsencode(path)),
                  "save")

    def load(self, path) -> None:
        self._chk(self._L.lz_surrogate_load(self._h, os.fsencode(path)),
                  "load")


def train(sur: Surrogate, planes, labels, steps: int, batch: int = 512,
          lr: float = 1e-3, seed: int = 0, log_every: int = 0,
          warmup: int = 200, device: str = "cuda:0"):
    """Minibatch Adam training on planes [N][32][361] and win-rate labels
    [N], both copied once to device memory. Minibatches are drawn without
    replacement within each pass over the data (shuffled per pass from
    `seed`); the learning rate ramps linearly to `lr` over the surrogate's
    first `warmup` steps. Returns [(step, loss)] every `log_every` steps."""
    import torch
    dev = torch.device(device)
    X = torch.from_numpy(np.ascontiguousarray(planes, np.uint8)).to(dev)
    Y = torch.from_numpy(np.ascontiguousarray(labels, np.float32)).to(dev)
    N = X.shape[0]
    if batch > sur.max_batch or batch > N:
        raise ValueError("batch larger than max_batch or the data")
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    perm = torch.randperm(N, generator=gen).to(dev)
    pos = 0
    stream = torch.cuda.current_stream(dev).cuda_stream
    losses = []
    xb = torch.empty((batch, PLANES, P), dtype=torch.uint8, device=dev)
    yb = torch.empty((batch,), dtype=torch.float32, device=dev)
    for it in range(steps):
        if pos + batch > N:
            perm = torch.randperm(N, generator=gen).to(dev)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        torch.index_select(X, 0, idx, out=xb)
        torch.index_select(Y, 0, idx, out=yb)
        want = bool(log_every) and (it % log_every == 0 or it == steps - 1)
        k = sur.step + 1
        lr_k = lr * min(1.0, k / warmup) if warmup > 0 else lr
        loss = sur.train_step_device(batch, xb.data_ptr(), yb.data_ptr(), lr_k,
                                     want_loss=want, stream=stream)
        if want:
            losses.append((it, loss))
    torch.cuda.synchronize(dev)
    return losses


class LeafEvaluator:
    """Win rates of leaf positions for the side to move: the surrogate when
    one is given, plain playouts (Playout::mc_owner with the light policy)
    when it is None or fails."""

    def __init__(self, surrogate: Surrogate | None = None,
                 playouts: int = 64, seed: int = 0):
        self.surrogate = surrogate
        self.playouts = playouts
        self.seed = seed
        self.fallbacks = 0

    def winrates(self, positions) -> np.ndarray:
        if self.surrogate is not None:
            try:
                planes = np.stack([p.planes(_positions.NET_VALUE)
                                   for p in positions])
                return self.surrogate.predict(planes)
            except LeelaError:
                self.fallbacks += 1
        return np.array([p.playout_winrate(self.playouts, self.seed + i)
                         for i, p in enumerate(positions)], np.float32)

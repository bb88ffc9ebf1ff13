"""CPU oracle of Leela's CNN leaf evaluation — TEST INFRASTRUCTURE ONLY.

ctypes wrapper over oracle/liboracle.so (oracle.c). Only the tests, the
smoke check and the CPU-baseline leg of bench.py may import this package;
the shipped library never calls it.
"""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

CONV, FC = 0, 1
HEAD_POLICY, HEAD_VALUE = 0, 1


class _Layer(C.Structure):
    _fields_ = [("kind", C.c_int), ("filter", C.c_int), ("cin", C.c_int),
                ("cout", C.c_int), ("w", C.POINTER(C.c_float)),
                ("b", C.POINTER(C.c_float)), ("act", C.c_int),
                ("bn_means", C.POINTER(C.c_float)),
                ("bn_vars", C.POINTER(C.c_float)), ("bn_scale", C.c_float)]


ACT_ELU, ACT_RELU, ACT_NONE = 0, 1, 2


def build() -> Path:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        f32p = C.POINTER(C.c_float)
        f64p = C.POINTER(C.c_double)
        u8p = C.POINTER(C.c_uint8)
        L.orc_forward_f32.argtypes = [C.c_int, C.POINTER(_Layer), C.c_int,
                                      C.c_int, C.c_float, C.c_int, u8p, f32p,
                                      f32p, C.c_int]
        L.orc_forward_f64.argtypes = [C.c_int, C.POINTER(_Layer), C.c_int,
                                      C.c_int, C.c_double, C.c_int, u8p, f64p,
                                      f64p, C.c_int]
        L.orc_forward_f64_gemm.argtypes = L.orc_forward_f64.argtypes
        L.orc_im2col.argtypes = [C.c_int, C.c_int, C.c_int, f32p, f32p]
        L.orc_softmax_f32.argtypes = [C.c_int, f32p, f32p, C.c_float]
        L.orc_convolve_f32.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                       f32p, f32p, f32p, f32p, f32p]
        L.orc_conv_direct_f64.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                          f64p, f32p, f32p, f64p, C.c_int]
        L.orc_innerproduct_f32.argtypes = [C.c_int, C.c_int, f32p, f32p,
                                           f32p, f32p]
        L.orc_rng_stream.argtypes = [C.c_uint32, C.c_int,
                                     C.POINTER(C.c_uint64)]
        L.orc_rotate_nn_idx.argtypes = [C.c_int, C.c_int]
        L.orc_rev_rotate_nn_idx.argtypes = [C.c_int, C.c_int]
        L.orc_load_blas.argtypes = [C.c_char_p]
        L.orc_blas_name.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def find_openblas() -> str | None:
    """Path of an OpenBLAS shared library on this host (numpy's bundled
    ILP64 build in this image), or None."""
    pats = []
    try:
        import numpy
        base = Path(numpy.__file__).resolve().parent.parent / "numpy.libs"
        pats.append(str(base / "libscipy_openblas64_*.so"))
        pats.append(str(base / "libopenblas*.so"))
    except Exception:
        pass
    pats += ["/usr/lib/x86_64-linux-gnu/libopenblas.so*",
             "/opt/OpenBLAS/lib/libopenblas.so*"]
    for p in pats:
        hits = sorted(glob.glob(p))
        if hits:
            return hits[0]
    return None


_blas_loaded = None


def load_blas() -> str:
    """dlopen OpenBLAS once; returns the description of the sgemm in use."""
    global _blas_loaded
    if _blas_loaded is None:
        path = find_openblas()
        ok = bool(path) and lib().orc_load_blas(path.encode()) == 1
        _blas_loaded = ok
    return lib().orc_blas_name().decode()


class Net:
    """Layer list in the reference's push order: (kind, filter, cin, cout,
    w, b) with w OIHW for convs and [out][in] for FCs, each conv followed
    by ELU; or (kind, filter, cin, cout, w, b, act, bn) with act ACT_* and
    bn None or (means, variances, scale factor) for a Caffe net's conv."""

    def __init__(self, layers):
        self.layers = []
        for layer in layers:
            kind, filt, cin, cout, w, b = layer[:6]
            act = layer[6] if len(layer) > 6 else ACT_ELU
            bn = layer[7] if len(layer) > 7 else None
            w = np.ascontiguousarray(w, dtype=np.float32).ravel()
            b = np.ascontiguousarray(b, dtype=np.float32).ravel()
            if bn is not None:
                bn = (np.ascontiguousarray(bn[0], dtype=np.float32).ravel(),
                      np.ascontiguousarray(bn[1], dtype=np.float32).ravel(),
                      float(bn[2]))
            self.layers.append((kind, filt, cin, cout, w, b, act, bn))
        arr = (_Layer * len(self.layers))()
        null = C.POINTER(C.c_float)()
        for i, (kind, filt, cin, cout, w, b, act, bn) in enumerate(
                self.layers):
            arr[i] = _Layer(kind, filt, cin, cout, _p(w, C.c_float),
                            _p(b, C.c_float), act,
                            _p(bn[0], C.c_float) if bn else null,
                            _p(bn[1], C.c_float) if bn else null,
                            bn[2] if bn else 0.0)
        self._arr = arr

    @classmethod
    def from_table(cls, table):
        """From a layer table of leela_amd.caffe_weights.layer_table (or
        the same dicts read back from a fixture)."""
        return cls([(0 if L["kind"] == "conv" else 1, L["filter"], L["cin"],
                     L["cout"], L["w"], L["b"], L["act"], L["bn"])
                    for L in table])


def forward_f32(net: Net, planes: np.ndarray, head: int, S: int = 19,
                temperature: float = 0.75, threads: int = 0):
    """Reference CPU path (im2col + sgemm + bias + ELU) over n positions.
    Returns (raw, out) as float32 arrays."""
    planes = np.ascontiguousarray(planes, dtype=np.uint8)
    n = planes.shape[0]
    per = S * S if head == HEAD_POLICY else 1
    raw = np.zeros((n, per), np.float32)
    out = np.zeros((n, per), np.float32)
    rc = lib().orc_forward_f32(len(net.layers), net._arr, S, head,
                               temperature, n, _p(planes, C.c_uint8),
                               _p(raw, C.c_float), _p(out, C.c_float), threads)
    if rc != 0:
        raise ValueError(f"orc_forward_f32 failed: {rc}")
    return raw, out


def forward_f64(net: Net, planes: np.ndarray, head: int, S: int = 19,
                temperature: float = 0.75, threads: int = 0,
                gemm: bool = False):
    """FP64 forward, direct convolution (or im2col + dgemm when gemm is
    true: the same values up to FP64 summation order, much faster with
    OpenBLAS loaded). Returns (raw, out) float64."""
    planes = np.ascontiguousarray(planes, dtype=np.uint8)
    n = planes.shape[0]
    per = S * S if head == HEAD_POLICY else 1
    raw = np.zeros((n, per), np.float64)
    out = np.zeros((n, per), np.float64)
    fn = lib().orc_forward_f64_gemm if gemm else lib().orc_forward_f64
    rc = fn(len(net.layers), net._arr, S, head, temperature, n,
            _p(planes, C.c_uint8), _p(raw, C.c_double), _p(out, C.c_double),
            threads)
    if rc != 0:
        raise ValueError(f"orc_forward_f64 failed: {rc}")
    return raw, out


def im2col(x: np.ndarray, filt: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    c, S, _ = x.shape
    col = np.zeros((c * filt * filt, S * S), np.float32)
    lib().orc_im2col(c, filt, S, _p(x, C.c_float), _p(col, C.c_float))
    return col


def convolve_f32(x, w, b, filt):
    x = np.ascontiguousarray(x, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    cin, S, _ = x.shape
    cout = b.shape[0]
    out = np.zeros((cout, S * S), np.float32)
    col = np.zeros((cin * filt * filt * S * S,), np.float32)
    lib().orc_convolve_f32(filt, cin, cout, S, _p(x, C.c_float),
                           _p(w, C.c_float), _p(b, C.c_float),
                           _p(out, C.c_float), _p(col, C.c_float))
    return out.reshape(cout, S, S)


def conv_direct_f64(x, w, b, filt, elu=True):
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    cin, S, _ = x.shape
    cout = b.shape[0]
    out = np.zeros((cout, S * S), np.float64)
    lib().orc_conv_direct_f64(filt, cin, cout, S, _p(x, C.c_double),
                              _p(w, C.c_float), _p(b, C.c_float),
                              _p(out, C.c_double), int(elu))
    return out.reshape(cout, S, S)


def innerproduct_f32(x, w, b):
    x = np.ascontiguousarray(x, dtype=np.float32)
    w = np.ascontiguousarray(w, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros(b.shape[0], np.float32)
    lib().orc_innerproduct_f32(x.shape[0], b.shape[0], _p(x, C.c_float),
                               _p(w, C.c_float), _p(b, C.c_float),
                               _p(out, C.c_float))
    return out


def softmax_f32(x, temperature):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    lib().orc_softmax_f32(x.shape[0], _p(x, C.c_float), _p(out, C.c_float),
                          temperature)
    return out


def rng_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().orc_rng_stream(seed, n, _p(out, C.c_uint64))
    return out


def rotate_nn_idx(v: int, s: int) -> int:
    return lib().orc_rotate_nn_idx(v, s)


def rev_rotate_nn_idx(v: int, s: int) -> int:
    return lib().orc_rev_rotate_nn_idx(v, s)

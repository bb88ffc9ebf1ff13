"""Python binding of the Network facade (include/leela_amd.h,
lz_network_*): get_scored_moves / get_value / async_scored_moves with the
reference's Netresult semantics (Network.cpp:536-832)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .backend import LZ_OK, LeelaError, load_library

DIRECT, RANDOM_ROTATION, AVERAGE_ALL = 0, 1, 2

MOVES_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_float),
                       C.POINTER(C.c_int32))


def _lib():
    L = load_library()
    if not getattr(L, "_network_bound", False):
        vp = C.c_void_p
        L.lz_network_scored_moves.argtypes = [
            vp, C.c_int, vp, C.c_int, C.c_int, C.POINTER(C.c_float),
            C.POINTER(C.c_int32), C.POINTER(C.c_int)]
        L.lz_network_value.argtypes = [vp, C.c_int, vp, C.c_int,
                                       C.POINTER(C.c_float)]
        L.lz_network_async_scored_moves.argtypes = [vp, vp, C.c_int,
                                                    MOVES_FN, vp]
        L.lz_network_heatmap.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int,
                                         C.c_int, C.c_char_p, C.c_size_t]
        L.lz_network_benchmark.argtypes = [vp, C.c_int, vp, C.c_int,
                                           C.POINTER(C.c_float),
                                           C.POINTER(C.c_float)]
        L._network_bound = True
    return L


class Network:
    """Network facade on a finalized Backend."""

    def __init__(self, backend, precision: int):
        self._be = backend
        self.precision = precision

    def get_scored_moves(self, pos, ensemble: int = DIRECT,
                         rotation: int = 0):
        """[(probability, vertex)] for the empty points, row-major."""
        probs = np.zeros(361, np.float32)
        verts = np.zeros(361, np.int32)
        cnt = C.c_int(0)
        rc = _lib().lz_network_scored_moves(
            self._be._ctx, self.precision, pos._p, ensemble, rotation,
            probs.ctypes.data_as(C.POINTER(C.c_float)),
            verts.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(cnt))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_network_scored_moves")
        return probs[:cnt.value].copy(), verts[:cnt.value].copy()

    def get_value(self, pos, ensemble: int = DIRECT) -> float:
        v = C.c_float(0)
        rc = _lib().lz_network_value(self._be._ctx, self.precision, pos._p,
                                     ensemble, C.byref(v))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_network_value")
        return float(v.value)

    def show_heatmap(self, pos, ensemble: int = DIRECT, rotation: int = 0,
                     topmoves: bool = True) -> str:
        """Network::show_heatmap's text for get_scored_moves(pos)."""
        buf = C.create_string_buffer(8192)
        rc = _lib().lz_network_heatmap(self._be._ctx, self.precision, pos._p,
                                       ensemble, rotation, int(topmoves), buf,
                                       len(buf))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_network_heatmap")
        return buf.value.decode()

    def benchmark(self, pos, threads: int):
        """Network::benchmark (netbench): (policy p/s, value p/s)."""
        p, v = C.c_float(0), C.c_float(0)
        rc = _lib().lz_network_benchmark(self._be._ctx, self.precision,
                                         pos._p, threads, C.byref(p),
                                         C.byref(v))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_network_benchmark")
        return float(p.value), float(v.value)

    def async_scored_moves(self, pos, rotation: int, done, user: int = 0):
        """done: a MOVES_FN-wrapped callable (user, count, probs, vertices)
        run on the queue's dispatcher thread."""
        rc = _lib().lz_network_async_scored_moves(self._be._ctx, pos._p,
                                                  rotation, done,
                                                  C.c_void_p(user))
        if rc != LZ_OK:
            raise LeelaError(rc, "lz_network_async_scored_moves")

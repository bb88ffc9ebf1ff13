"""TEST INFRASTRUCTURE ONLY: the reference's board engine compiled from its
own sources (oracle/ref_board/Makefile -> oracle/_ref/libref_board.so).

Only the tests and tests/golden/make_board_golden.py use it, as the checker
of the library's board engine, generator and encoders.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_ref" / "libref_board.so"
REF = Path(os.environ.get("LEELA_REFERENCE", "/root/reference"))

_lib = None


def build() -> bool:
    """Compile the checker when the reference sources are present."""
    if not (REF / "FastBoard.cpp").exists():
        return LIB.exists()
    subprocess.run(["make", "-s", "-j8", "-C", str(HERE / "ref_board"),
                    f"REF={REF}"], check=True)
    return LIB.exists()


def available() -> bool:
    return LIB.exists()


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise FileNotFoundError(f"{LIB} not built")
        L = C.CDLL(str(LIB), mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        i32p = C.POINTER(C.c_int32)
        u8p = C.POINTER(C.c_uint8)
        L.ref_random_position.argtypes = [C.c_int, C.c_uint32, C.c_int, i32p,
                                          u8p, i32p, i32p, i32p]
        L.ref_replay.argtypes = [C.c_int, C.c_float, C.c_int, i32p, i32p,
                                 u8p, i32p]
        L.ref_rng_u32.argtypes = [C.c_uint32, C.c_int, C.c_uint32,
                                  C.POINTER(C.c_uint32)]
        L.ref_rng_u32.restype = None
        L.ref_rng_flt.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_float)]
        L.ref_final_score.argtypes = [C.c_int, C.c_float, C.c_int, i32p]
        L.ref_final_score.restype = C.c_float
        L.ref_rng_flt.restype = None
        L.ref_pattern_playouts.argtypes = [C.c_int, C.c_float, C.c_int, i32p,
                                           C.c_int, C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_float)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def random_position(side: int, seed: int, max_moves: int):
    """(info[7], squares, queries[P][12], moves) of the generator run on
    the reference engine."""
    P = side * side
    info = np.zeros(7, np.int32)
    sq = np.zeros(P, np.uint8)
    q = np.zeros((P, 12), np.int32)
    moves = np.zeros(2 * P + 600, np.int32)
    nm = C.c_int32(0)
    lib().ref_random_position(side, seed, max_moves, _p(info, C.c_int32),
                              _p(sq, C.c_uint8), _p(q, C.c_int32),
                              _p(moves, C.c_int32), C.byref(nm))
    return info, sq, q, moves[:nm.value].copy()


def replay(side: int, komi: float, moves):
    P = side * side
    moves = np.ascontiguousarray(moves, np.int32)
    info = np.zeros(7, np.int32)
    sq = np.zeros(P, np.uint8)
    q = np.zeros((P, 12), np.int32)
    lib().ref_replay(side, komi, len(moves), _p(moves, C.c_int32),
                     _p(info, C.c_int32), _p(sq, C.c_uint8), _p(q, C.c_int32))
    return info, sq, q


def final_score(side: int, komi: float, moves) -> float:
    """FastBoard::final_mc_score after replaying `moves`."""
    moves = np.ascontiguousarray(moves, np.int32)
    return float(lib().ref_final_score(side, komi, len(moves),
                                       _p(moves, C.c_int32)))


def rng_u32(seed: int, n: int, max_: int = 0) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().ref_rng_u32(seed, n, max_, _p(out, C.c_uint32))
    return out


def rng_flt(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().ref_rng_flt(seed, n, _p(out, C.c_float))
    return out


def pattern_playouts(side: int, komi: float, moves, seeds) -> np.ndarray:
    """final_mc_score of one reference pattern playout per seed from the
    position after `moves` (rebuilt in raster order; see driver.cpp)."""
    moves = np.ascontiguousarray(moves, np.int32)
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.zeros(len(seeds), np.float32)
    lib().ref_pattern_playouts(side, komi, len(moves), _p(moves, C.c_int32),
                               len(seeds), _p(seeds, C.c_uint32),
                               _p(out, C.c_float))
    return out

"""The device feature encoder's kernel source (leela_amd/csrc/encoder.hip)
and the playout kernel's (playout.hip) run on the CPU, before any GPU run,
against the reference's own board queries, the host encoder and the host
playouts.

The kernel is compiled by g++ against tests/wave_emul/hip/hip_runtime.h,
which runs each workgroup as 64 host threads (one wavefront): __shared__
data shared by them, __syncthreads a barrier, __ballot a gather of the 64
predicates. The same source, unchanged, is what hipcc builds for gfx950;
the GPU tests compare that build with the same expectations.

Pins: the planes restated from tests/golden/board_golden.npz (queries of
the reference's FastBoard, including crafted ladder-rich positions), the
host encoder on generator positions of mixed sides, and the symmetry
identity planes(s)[i] = planes(0)[rotate(i, s)].
"""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from leela_amd.positions import NET_POLICY, NET_VALUE

from test_board_cpu import expected_planes, rot_closed_form

ROOT = Path(__file__).resolve().parents[1]
EMUL = Path(__file__).parent / "wave_emul"
CSRC = ROOT / "leela_amd" / "csrc"
OUT = ROOT / "_build" / "wave_emul" / "libencoder_emul.so"
GOLD = Path(__file__).parent / "golden" / "board_golden.npz"
SOURCES = [CSRC / "encoder.hip", CSRC / "playout.hip", EMUL / "driver.cpp",
           CSRC / "go_board.cpp", CSRC / "input_planes.cpp",
           CSRC / "lz_enc_record.cpp", CSRC / "playout.cpp"]
HIP_SOURCES = 2


def _build():
    deps = SOURCES + [EMUL / "hip" / "hip_runtime.h"] + list(CSRC.glob("*.h"))
    if OUT.exists() and OUT.stat().st_mtime > max(p.stat().st_mtime
                                                   for p in deps):
        return OUT
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    OUT.parent.mkdir(parents=True, exist_ok=True)
    cmd = [cxx, "-std=c++20", "-O2", "-pthread", "-fPIC", "-shared",
           "-Wno-unknown-pragmas", f"-I{EMUL}", f"-I{CSRC}",
           f"-I{ROOT / 'include'}", "-x", "c++",
           *map(str, SOURCES[:HIP_SOURCES]), "-x", "none",
           *map(str, SOURCES[HIP_SOURCES:]), "-o", str(OUT)]
    subprocess.run(cmd, check=True, capture_output=True)
    return OUT


@pytest.fixture(scope="module")
def emul():
    L = C.CDLL(str(_build()))
    i32p = C.POINTER(C.c_int32)
    u8p = C.POINTER(C.c_uint8)
    L.emul_encode_moves.argtypes = [C.c_int, C.c_int, i32p, C.c_int, C.c_int,
                                    u8p]
    L.emul_encode_moves_both.argtypes = [C.c_int, C.c_int, i32p, C.c_int,
                                         u8p, u8p]
    f32p = C.POINTER(C.c_float)
    L.emul_playouts.argtypes = [C.c_int, i32p, C.POINTER(C.c_uint32), i32p,
                                C.c_int, C.c_uint32, f32p, f32p]
    L.emul_encode_generated.argtypes = [C.c_int, i32p,
                                        C.POINTER(C.c_uint32), i32p, C.c_int,
                                        C.c_int, u8p, u8p]
    return L


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def encode_moves(L, side, moves, net, sym=0):
    out = np.full((32, side * side), 0xAB, np.uint8)
    mv = np.ascontiguousarray(moves, np.int32)
    L.emul_encode_moves(side, len(mv), _p(mv, C.c_int32), net, sym,
                        _p(out, C.c_uint8))
    return out


@pytest.mark.parametrize("net", [NET_POLICY, NET_VALUE])
def test_kernel_matches_reference_queries(emul, gold, net):
    """Every golden board (random games, ko and ladder prefixes, crafted
    ladder-rich positions): the kernel's planes equal the layout restated
    from the reference's queries."""
    for i, (side, *_rest) in enumerate(gold["cases"]):
        got = encode_moves(emul, int(side), gold[f"moves_{i}"], net)
        want = expected_planes(net, int(side), gold[f"info_{i}"],
                               gold[f"squares_{i}"], gold[f"queries_{i}"])
        assert np.array_equal(got, want), (i, np.argwhere(got != want)[:5])


def test_kernel_symmetries(emul, gold):
    cases = gold["cases"]
    for i in (0, 80, 100, len(cases) - 1):
        side = int(cases[i][0])
        base = encode_moves(emul, side, gold[f"moves_{i}"], NET_POLICY)
        for s in range(1, 8):
            got = encode_moves(emul, side, gold[f"moves_{i}"], NET_POLICY, s)
            idx = [rot_closed_form(k, s, side) for k in range(side * side)]
            assert np.array_equal(got, base[:, idx]), (i, s)


def test_kernel_both_layouts_in_one_pass(emul, gold):
    """The one-pass form writes each net's planes exactly as the single-net
    form does."""
    for i in (3, 90, 130, 160, 200):
        side = int(gold["cases"][i][0])
        mv = np.ascontiguousarray(gold[f"moves_{i}"], np.int32)
        pol = np.zeros((32, side * side), np.uint8)
        val = np.zeros((32, side * side), np.uint8)
        emul.emul_encode_moves_both(side, len(mv), _p(mv, C.c_int32), 2,
                                    _p(pol, C.c_uint8), _p(val, C.c_uint8))
        assert np.array_equal(pol, encode_moves(emul, side, mv, NET_POLICY, 2))
        assert np.array_equal(val, encode_moves(emul, side, mv, NET_VALUE, 2))


@pytest.mark.parametrize("net,sym", [(NET_POLICY, 0), (NET_POLICY, 5),
                                     (NET_VALUE, 3)])
def test_kernel_matches_host_encoder_ragged(emul, net, sym):
    """A ragged batch of generator positions (19/13/9, move caps 250 / 101
    / 48): one launch, planes packed board after board, equal to the host
    encoder board by board."""
    n = 150
    sides = np.array([(19, 13, 9)[i % 3] for i in range(n)], np.int32)
    seeds = np.arange(5000, 5000 + n, dtype=np.uint32)
    mm = np.array([250 if s == 19 else int(0.6 * s * s) for s in sides],
                  np.int32)
    tot = int(sum(32 * s * s for s in sides))
    out = np.full(tot, 0xAB, np.uint8)
    ref = np.zeros(tot, np.uint8)
    emul.emul_encode_generated(n, _p(sides, C.c_int32), _p(seeds, C.c_uint32),
                               _p(mm, C.c_int32), net, sym,
                               _p(out, C.c_uint8), _p(ref, C.c_uint8))
    off = 0
    for i, s in enumerate(sides):
        sz = 32 * int(s) * int(s)
        assert np.array_equal(out[off:off + sz], ref[off:off + sz]), (i, s)
        off += sz


def test_playout_kernel_matches_host_playouts(emul):
    """playout.hip's kernel source against the host's mc_winrate, bit for
    bit: generator positions of three sides (early, middle and end of
    game), 8 playouts each."""
    n = 6
    sides = np.array([(19, 13, 9)[i % 3] for i in range(n)], np.int32)
    seeds = np.arange(40, 40 + n, dtype=np.uint32)
    mm = np.array([(250, 101, 48)[i % 3] for i in range(n)], np.int32)
    out = np.full(n, -1.0, np.float32)
    ref = np.zeros(n, np.float32)
    emul.emul_playouts(n, _p(sides, C.c_int32), _p(seeds, C.c_uint32),
                       _p(mm, C.c_int32), 8, 777, _p(out, C.c_float),
                       _p(ref, C.c_float))
    assert np.array_equal(out, ref), (out, ref)
    assert len(set(ref.tolist())) > 3  # the outcomes differ by position


def lattice_moves(shift, extra=()):
    """Moves of a 19x19 board whose stones are all single-stone strings
    (black at x % 2 == shift, y even; white at the other parity, y odd:
    180-181 strings), then `extra` moves; white passes when it runs out."""
    v = lambda x, y: (y + 1) * 21 + (x + 1)  # noqa: E731
    blk = [v(x, y) for y in range(0, 19, 2) for x in range(19)
           if x % 2 == shift]
    wht = [v(x, y) for y in range(1, 19, 2) for x in range(19)
           if x % 2 != shift]
    moves = []
    for i, b in enumerate(blk):
        moves += [b, wht[i] if i < len(wht) else -1]
    return moves + list(extra)


@pytest.mark.parametrize("net", [NET_POLICY, NET_VALUE])
def test_kernel_many_strings(emul, net):
    """Boards with more strings than the kernel keeps in LDS (128): the
    rest go through its global overflow area; planes equal the host
    encoder's."""
    from leela_amd.positions import Position
    v = lambda x, y: (y + 1) * 21 + (x + 1)  # noqa: E731
    cases = [lattice_moves(0), lattice_moves(1),
             lattice_moves(0, [v(1, 0), v(3, 0), v(0, 1), v(18, 17)]),
             lattice_moves(1, [v(0, 0), v(1, 1), v(17, 17), v(8, 8)])]
    for k, moves in enumerate(cases):
        pos = Position(19)
        for m in moves:
            pos.play(m)
        for sym in (0, 5):
            got = encode_moves(emul, 19, moves, net, sym)
            want = pos.planes(net, sym).reshape(32, 361)
            assert np.array_equal(got, want), (k, sym,
                                               np.argwhere(got != want)[:5])

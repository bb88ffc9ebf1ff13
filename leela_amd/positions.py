"""Positions and input planes through the library's host-side C ABI.

Mirrors the reference's FastState / Network feature interface:
``Position.play`` is FastState::play_move, ``Position.planes`` is
Network::gather_features_policy / gather_features_value followed by the
symmetry rotation of get_scored_moves_internal (Network.cpp:765-773).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .backend import LZ_OK, LeelaError, load_library

NET_POLICY, NET_VALUE = 0, 1
PLAYOUT_LIGHT, PLAYOUT_PATTERN = 0, 1
PASS = -1
QUERY_FIELDS = 12
BLACK, WHITE, EMPTY = 0, 1, 2

# Generator move caps: 0.6 * board area for the ragged mix, 250 / 200 for
# the 19x19 batches.
MAX_MOVES_19 = 250


def _lib():
    L = load_library()
    if not getattr(L, "_positions_bound", False):
        vp = C.c_void_p
        u8p = C.POINTER(C.c_uint8)
        i32p = C.POINTER(C.c_int32)
        L.lz_position_new.restype = vp
        L.lz_position_new.argtypes = [C.c_int, C.c_float]
        L.lz_position_free.argtypes = [vp]
        L.lz_position_free.restype = None
        L.lz_position_play.argtypes = [vp, C.c_int]
        L.lz_position_random.argtypes = [vp, C.c_uint32, C.c_int]
        L.lz_position_info.argtypes = [vp, i32p]
        L.lz_position_squares.argtypes = [vp, u8p]
        L.lz_encode_planes.argtypes = [vp, C.c_int, C.c_int, u8p]
        L.lz_board_queries.argtypes = [vp, i32p]
        L.lz_make_batch.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_int,
                                    C.c_int, C.c_int, u8p]
        f32p = C.POINTER(C.c_float)
        L.lz_encode_batch.argtypes = [C.POINTER(vp), C.c_int, C.c_int,
                                      C.c_int, C.c_int, u8p]
        L.lz_position_score.argtypes = [vp, f32p]
        L.lz_position_set_handicap.argtypes = [vp, C.c_int]
        L.lz_playout_winrate.argtypes = [vp, C.c_int, C.c_uint32, f32p]
        L.lz_make_value_labels.argtypes = [C.c_int, C.c_int, C.c_uint32,
                                           C.c_int, C.c_int, C.c_uint32,
                                           C.c_int, u8p, f32p]
        L.lz_make_value_labels_policy.argtypes = [
            C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int,
            C.c_uint32, C.c_int, u8p, f32p]
        L.lz_playout_winrate_policy.argtypes = [vp, C.c_int, C.c_int,
                                                C.c_uint32, f32p]
        L.lz_mc_policy_load.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
        L.lz_playout_scores.argtypes = [vp, C.c_int, C.c_int, C.c_uint32,
                                        f32p]
        L.lz_encoder_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int]
        L.lz_encoder_destroy.argtypes = [vp]
        L.lz_encoder_destroy.restype = None
        L.lz_encoder_last_error.argtypes = [vp]
        L.lz_encoder_last_error.restype = C.c_char_p
        L.lz_encode_device.argtypes = [vp, C.POINTER(vp), C.c_int, C.c_int,
                                       C.c_int, vp, vp]
        L.lz_encode_device_both.argtypes = [vp, C.POINTER(vp), C.c_int,
                                            C.c_int, vp, vp, vp]
        L.lz_playout_winrates_device.argtypes = [vp, C.POINTER(vp), C.c_int,
                                                 C.c_int, C.c_uint32, vp, vp]
        L.lz_playout_winrates_device_policy.argtypes = [
            vp, C.POINTER(vp), C.c_int, C.c_int, C.c_uint32, C.c_int, vp, vp]
        L._positions_bound = True
    return L


def _chk(rc, what):
    if rc != LZ_OK:
        raise LeelaError(rc, what)


class Position:
    """A game position (FastState counterpart)."""

    def __init__(self, side: int = 19, komi: float = 7.5):
        self._L = _lib()
        self._p = self._L.lz_position_new(side, komi)
        if not self._p:
            raise ValueError(f"bad board side {side}")
        self.side = side

    def __del__(self):
        try:
            if self._p:
                self._L.lz_position_free(self._p)
        except Exception:
            pass

    @classmethod
    def random(cls, seed: int, side: int = 19, max_moves: int = MAX_MOVES_19):
        p = cls(side)
        _chk(p._L.lz_position_random(p._p, seed, max_moves), "random")
        return p

    def vertex(self, x: int, y: int) -> int:
        return (y + 1) * (self.side + 2) + (x + 1)

    def play(self, vertex: int):
        _chk(self._L.lz_position_play(self._p, vertex), f"play {vertex}")

    def info(self) -> dict:
        a = np.zeros(7, np.int32)
        _chk(self._L.lz_position_info(self._p, a.ctypes.data_as(
            C.POINTER(C.c_int32))), "info")
        keys = ("side", "to_move", "komove", "last_move", "prev_last",
                "movenum", "passes")
        return dict(zip(keys, map(int, a)))

    def squares(self) -> np.ndarray:
        a = np.zeros(self.side * self.side, np.uint8)
        _chk(self._L.lz_position_squares(self._p, a.ctypes.data_as(
            C.POINTER(C.c_uint8))), "squares")
        return a

    def planes(self, net: int = NET_POLICY, symmetry: int = 0) -> np.ndarray:
        a = np.zeros((32, self.side * self.side), np.uint8)
        _chk(self._L.lz_encode_planes(self._p, net, symmetry,
                                      a.ctypes.data_as(C.POINTER(C.c_uint8))),
             "encode")
        return a

    def queries(self) -> np.ndarray:
        a = np.zeros((self.side * self.side, QUERY_FIELDS), np.int32)
        _chk(self._L.lz_board_queries(self._p, a.ctypes.data_as(
            C.POINTER(C.c_int32))), "queries")
        return a

    def set_handicap(self, handicap: int):
        """FastState::set_handicap: added to komi in every score."""
        _chk(self._L.lz_position_set_handicap(self._p, handicap), "handicap")

    def score(self) -> float:
        """FastBoard::final_mc_score with the position's komi."""
        f = C.c_float()
        _chk(self._L.lz_position_score(self._p, C.byref(f)), "score")
        return f.value

    def playout_scores(self, n: int, seed: int,
                       policy: int = PLAYOUT_LIGHT) -> np.ndarray:
        """final_mc_score (black's view) of each of the n playouts behind
        playout_winrate(n, seed, policy)."""
        out = np.zeros(n, np.float32)
        _chk(self._L.lz_playout_scores(self._p, policy, n, seed,
                                       out.ctypes.data_as(
                                           C.POINTER(C.c_float))),
             "playout scores")
        return out

    def playout_winrate(self, n: int, seed: int,
                        policy: int = PLAYOUT_LIGHT) -> float:
        """Playout::mc_owner: mean outcome of n playouts for the side to
        move (tie 0.5), light (uniform) or with the reference's pattern
        policy."""
        f = C.c_float()
        _chk(self._L.lz_playout_winrate_policy(self._p, policy, n, seed,
                                               C.byref(f)), "playout")
        return f.value


def make_batch(n: int, seed0: int = 1, side: int = 19,
               max_moves: int = MAX_MOVES_19, net: int = NET_POLICY,
               threads: int = 0) -> np.ndarray:
    """Planes [n][32][side*side] of positions from seeds seed0..seed0+n-1."""
    out = np.zeros((n, 32, side * side), np.uint8)
    _chk(_lib().lz_make_batch(n, side, seed0, max_moves, net, threads,
                              out.ctypes.data_as(C.POINTER(C.c_uint8))),
         "make_batch")
    return out


def encode_batch(positions, net: int = NET_POLICY, symmetry: int = 0,
                 threads: int = 0) -> np.ndarray:
    """Planes [n][32][S*S] of Position objects of one side, encoded on
    `threads` host threads (lz_encode_batch)."""
    n = len(positions)
    side = positions[0].side if n else 19
    out = np.zeros((n, 32, side * side), np.uint8)
    arr = (C.c_void_p * max(n, 1))(*[p._p for p in positions])
    _chk(_lib().lz_encode_batch(arr, n, net, symmetry, threads,
                                out.ctypes.data_as(C.POINTER(C.c_uint8))),
         "encode_batch")
    return out


class PositionBatch:
    """Positions with their C pointer array built once, for callers that
    encode the same batch repeatedly (the array is what the ABI takes)."""

    def __init__(self, positions):
        self.positions = list(positions)
        self.array = (C.c_void_p * max(len(self.positions), 1))(
            *[p._p for p in self.positions])

    def __len__(self):
        return len(self.positions)


def _as_array(positions):
    if isinstance(positions, PositionBatch):
        return positions.array, len(positions)
    n = len(positions)
    return (C.c_void_p * max(n, 1))(*[p._p for p in positions]), n


class DeviceEncoder:
    """The feature encoder on the GPU (lz_encode_device): planes of a batch
    of Position objects written into device memory, one wavefront per
    board. Sides may differ within a batch (planes packed board after
    board)."""

    def __init__(self, device: int = 0, max_batch: int = 1024):
        self._L = _lib()
        h = C.c_void_p()
        _chk(self._L.lz_encoder_create(C.byref(h), device, max_batch),
             "lz_encoder_create")
        self._h = h
        self.device = device
        self.max_batch = max_batch

    def close(self):
        if getattr(self, "_h", None):
            self._L.lz_encoder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def encode_into(self, positions, d_planes: int, net: int = NET_POLICY,
                    symmetry: int = 0, stream: int = 0):
        """Enqueue the planes of `positions` (a list of Position or a
        PositionBatch) into device address d_planes on `stream` (0: the
        HIP default stream)."""
        arr, n = _as_array(positions)
        rc = self._L.lz_encode_device(self._h, arr, n, net, symmetry,
                                      C.c_void_p(d_planes),
                                      C.c_void_p(stream))
        if rc != LZ_OK:
            raise LeelaError(rc, self._L.lz_encoder_last_error(self._h)
                             .decode())

    def encode_both_into(self, positions, d_policy: int, d_value: int,
                         symmetry: int = 0, stream: int = 0):
        """Both nets' planes from one pass (lz_encode_device_both)."""
        arr, n = _as_array(positions)
        rc = self._L.lz_encode_device_both(self._h, arr, n, symmetry,
                                           C.c_void_p(d_policy),
                                           C.c_void_p(d_value),
                                           C.c_void_p(stream))
        if rc != LZ_OK:
            raise LeelaError(rc, self._L.lz_encoder_last_error(self._h)
                             .decode())

    def playouts_into(self, positions, playouts: int, seed0: int,
                      d_winrates: int, stream: int = 0,
                      policy: int = PLAYOUT_LIGHT):
        """Enqueue the positions' playout win rates (lz_playout_winrate_policy
        on the GPU, playouts of position i seeded seed0 + i) into device
        address d_winrates (float per position; NaN where a pattern
        playout was abandoned)."""
        arr, n = _as_array(positions)
        rc = self._L.lz_playout_winrates_device_policy(
            self._h, arr, n, playouts, seed0 & 0xFFFFFFFF, policy,
            C.c_void_p(d_winrates), C.c_void_p(stream))
        if rc != LZ_OK:
            raise LeelaError(rc, self._L.lz_encoder_last_error(self._h)
                             .decode())

    def playout_winrates(self, positions, playouts: int, seed0: int,
                         policy: int = PLAYOUT_LIGHT) -> np.ndarray:
        """Win rates as a numpy array (synchronous); raises if a pattern
        playout outgrew the device's candidate list."""
        import torch
        dev = torch.device("cuda", self.device)
        out = torch.empty(max(len(positions), 1), dtype=torch.float32,
                          device=dev)
        self.playouts_into(positions, playouts, seed0, out.data_ptr(),
                           torch.cuda.current_stream(dev).cuda_stream, policy)
        w = out[:len(positions)].cpu().numpy()
        if np.isnan(w).any():
            raise LeelaError(-5, f"{int(np.isnan(w).sum())} positions with a "
                             "playout beyond the device candidate list")
        return w

    def encode(self, positions, net: int = NET_POLICY, symmetry: int = 0):
        """Planes as a uint8 CUDA tensor: [n, 32, S*S] for one side, flat
        for mixed sides."""
        import torch
        sides = {p.side for p in positions}
        total = sum(32 * p.side * p.side for p in positions)
        dev = torch.device("cuda", self.device)
        out = torch.empty(max(total, 1), dtype=torch.uint8, device=dev)
        self.encode_into(positions, out.data_ptr(), net, symmetry,
                         torch.cuda.current_stream(dev).cuda_stream)
        out = out[:total]
        if len(sides) == 1:
            side = sides.pop()
            out = out.view(len(positions), 32, side * side)
        return out


def load_mc_policy(path: str) -> None:
    """lz_mc_policy_load: the pattern policy's tables from `path` (a file
    written by leela_amd/data/make_mc_policy.py); raises with the loader's
    message when the file is missing or malformed."""
    err = C.create_string_buffer(256)
    rc = _lib().lz_mc_policy_load(str(path).encode(), err, len(err))
    if rc != 0:
        raise LeelaError(rc, f"mc policy: {err.value.decode()}")


def value_labels(n: int, seed0: int, playouts: int, playout_seed: int,
                 side: int = 19, max_moves: int = MAX_MOVES_19,
                 threads: int = 0, planes: bool = True,
                 policy: int = PLAYOUT_LIGHT):
    """Surrogate training data: value planes [n][32][side*side] (or None)
    of generator positions from seeds seed0.., and their `playouts`-playout
    win rates for the side to move, playouts of position i seeded
    playout_seed + i, with the light or the pattern playout policy."""
    pl = np.zeros((n, 32, side * side), np.uint8) if planes else None
    w = np.zeros(n, np.float32)
    _chk(_lib().lz_make_value_labels_policy(
        policy, n, side, seed0, max_moves, playouts, playout_seed, threads,
        pl.ctypes.data_as(C.POINTER(C.c_uint8)) if planes else None,
        w.ctypes.data_as(C.POINTER(C.c_float))), "value_labels")
    return pl, w


def value_labels_device(n: int, seed0: int, playouts: int, playout_seed: int,
                        side: int = 19, max_moves: int = MAX_MOVES_19,
                        planes: bool = True, device: int = 0,
                        batch: int = 4096, policy: int = PLAYOUT_LIGHT):
    """value_labels with the playouts (and value planes) computed on the
    GPU: lz_playout_winrates_device_policy and lz_encode_device over
    batches of generator positions. Equal to value_labels bit for bit."""
    import torch
    enc = DeviceEncoder(device, max_batch=min(batch, max(n, 1)))
    dev = torch.device("cuda", device)
    pl = np.zeros((n, 32, side * side), np.uint8) if planes else None
    w = np.zeros(n, np.float32)
    try:
        for a in range(0, n, enc.max_batch):
            b = min(n, a + enc.max_batch)
            pos = PositionBatch([Position.random(seed0 + i, side, max_moves)
                                 for i in range(a, b)])
            d_w = torch.empty(b - a, dtype=torch.float32, device=dev)
            st = torch.cuda.current_stream(dev).cuda_stream
            enc.playouts_into(pos, playouts, playout_seed + a, d_w.data_ptr(),
                              st, policy)
            if planes:
                d_pl = torch.empty((b - a, 32, side * side), dtype=torch.uint8,
                                   device=dev)
                enc.encode_into(pos, d_pl.data_ptr(), NET_VALUE, 0, st)
                pl[a:b] = d_pl.cpu().numpy()
            w[a:b] = d_w.cpu().numpy()
    finally:
        enc.close()
    if np.isnan(w).any():
        raise LeelaError(-5, f"{int(np.isnan(w).sum())} positions with a "
                         "playout beyond the device candidate list")
    return pl, w


def batch_planes(n: int, net: int = 0, seed0: int = 1) -> np.ndarray:
    """Planes of the benchmark batch: random legal play from seeds
    seed0..seed0+n-1, move counts uniform in 0..250, 19x19."""
    return make_batch(n, seed0=seed0, side=19, max_moves=MAX_MOVES_19,
                      net=net)


# The mixed-board workload: (side, boards, move cap = 0.6 x area, first
# seed); seed ranges are disjoint.
RAGGED_MIX = ((19, 1024, 216, 30001), (13, 512, 101, 40001),
              (9, 512, 48, 50001))


def ragged_mix(mix=RAGGED_MIX, pattern=(19, 13, 19, 9)):
    """Boards of mixed sides from the generator, interleaved by `pattern`
    (19, 13, 19, 9, ...): (sides, policy planes, value planes), the planes
    as lists of uint8 [32][S*S]."""
    per_side = {}
    for side, n, mmax, seed0 in mix:
        per_side[side] = (make_batch(n, seed0=seed0, side=side,
                                     max_moves=mmax, net=NET_POLICY),
                          make_batch(n, seed0=seed0, side=side,
                                     max_moves=mmax, net=NET_VALUE))
    total = sum(n for _, n, _, _ in mix)
    idx = {side: 0 for side, _, _, _ in mix}
    order = []
    while len(order) < total:
        for s in pattern:
            if idx[s] < len(per_side[s][0]):
                order.append((s, idx[s]))
                idx[s] += 1
    sides = [s for s, _ in order]
    pol = [per_side[s][0][i] for s, i in order]
    val = [per_side[s][1][i] for s, i in order]
    return sides, pol, val


def synthetic_planes(n: int, seed: int, side: int = 19) -> np.ndarray:
    """uint8 [n][32][side*side] stacks shaped like the encoder's output but
    drawn independently per plane (no board engine): planes 0/1/2 partition
    the board, liberty planes mark stones, after-move planes mark empties,
    the rest are sparse. For kernel tests that need arbitrary inputs."""
    rng = np.random.default_rng(seed)
    P = side * side
    out = np.zeros((n, 32, P), np.uint8)
    xs = np.arange(P) % side
    ys = np.arange(P) // side
    line3 = (xs == 2) | (xs == side - 3) | (ys == 2) | (ys == side - 3)
    for i in range(n):
        fill = rng.uniform(0.0, 0.6)
        occ = rng.random(P)
        own = occ < fill / 2
        opp = (occ >= fill / 2) & (occ < fill)
        emp = ~(own | opp)
        out[i, 0], out[i, 1], out[i, 2] = emp, own, opp
        libs = rng.integers(0, 5, P)
        for k in range(5):
            out[i, 3 + k] = own & (libs == k)
            out[i, 8 + k] = opp & (libs == k)
        al = rng.integers(0, 6, P)
        at = rng.integers(0, 6, P)
        for k in range(6):
            out[i, 13 + k] = emp & (al == k)
            out[i, 19 + k] = emp & (at == k)
        out[i, 25:30] = (rng.random((5, P)) < 0.01) & emp
        out[i, 30] = opp
        out[i, 31] = line3
    return out

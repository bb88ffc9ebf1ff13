"""CPU parity of the board engine, position generator and feature encoders
against the reference's own board engine.

Pins, in order of strength:
  * tests/golden/board_golden.npz — moves, FastState fields, squares and
    per-point board queries produced by the reference sources compiled by
    oracle/ref_board/Makefile (see tests/golden/make_board_golden.py);
  * the live reference build (oracle/_ref), when present, on more seeds;
  * closed forms: liberties by flood fill, the plane layout of
    Network.cpp:883-1201 restated here from the queries, symmetry group
    identities.
"""
from collections import deque
from pathlib import Path

import numpy as np
import pytest

from leela_amd import positions
from leela_amd.positions import NET_POLICY, NET_VALUE, Position

GOLD = Path(__file__).parent / "golden" / "board_golden.npz"
FIELDS = ("side", "to_move", "komove", "last_move", "prev_last", "movenum",
          "passes")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def replay_mine(side, moves):
    p = Position(side)
    for mv in moves:
        p.play(int(mv))
    return p


def info_vec(p):
    i = p.info()
    return np.array([i[k] for k in FIELDS], np.int32)


def test_golden_cover_the_hard_branches(gold):
    """The fixture exercises ko, losing ladders and winning ladders."""
    n = len(gold["cases"])
    ko = sum(gold[f"info_{i}"][2] > 0 for i in range(n))
    ll = sum((gold[f"queries_{i}"][:, 7] == 1).sum() for i in range(n))
    wl = sum((gold[f"queries_{i}"][:, 8] == 1).sum() for i in range(n))
    assert ko >= 5 and ll >= 20 and wl >= 50


def test_replayed_positions_match_reference(gold):
    cases = gold["cases"]
    for i, (side, seed, mm, prefix) in enumerate(cases):
        p = replay_mine(side, gold[f"moves_{i}"])
        assert np.array_equal(info_vec(p), gold[f"info_{i}"]), (i, seed)
        assert np.array_equal(p.squares(), gold[f"squares_{i}"]), (i, seed)
        assert np.array_equal(p.queries(), gold[f"queries_{i}"]), (i, seed)


def test_generator_reproduces_reference_games(gold):
    cases = gold["cases"]
    for i, (side, seed, mm, prefix) in enumerate(cases):
        if prefix != -1:  # prefixes and crafted positions
            continue
        p = Position.random(int(seed), int(side), int(mm))
        assert np.array_equal(info_vec(p), gold[f"info_{i}"]), (i, seed)
        assert np.array_equal(p.squares(), gold[f"squares_{i}"]), (i, seed)


def test_move_counts_cover_range():
    counts = [Position.random(s, 19, 250).info()["movenum"]
              for s in range(1, 201)]
    assert min(counts) < 30 and max(counts) > 220


def test_live_reference_more_seeds():
    import oracle.ref_board as rb
    if not rb.available():
        pytest.skip("reference board checker not built")
    for side, mm, seeds in ((19, 250, range(2000, 2060)),
                            (13, 101, range(50, 80)), (9, 48, range(50, 80))):
        for seed in seeds:
            info, sq, q, moves = rb.random_position(side, seed, mm)
            p = Position.random(seed, side, mm)
            assert np.array_equal(info_vec(p), info), (side, seed)
            assert np.array_equal(p.squares(), sq), (side, seed)
            assert np.array_equal(p.queries(), q), (side, seed)


def test_reference_rng_values(gold):
    import oracle
    for seed in (1, 7, 8, 5489):
        s = oracle.rng_stream(seed, 16)
        assert np.array_equal((s >> np.uint64(32)).astype(np.uint32),
                              gold[f"rng_u32_{seed}"])
        top = (s >> np.uint64(32)).astype(np.uint64)
        assert np.array_equal(((top * np.uint64(361)) >> np.uint64(32))
                              .astype(np.uint32), gold[f"rng_u32max_{seed}"])


# ---- closed forms --------------------------------------------------------


def flood_liberties(sq, side):
    """Liberties of every stone's string by breadth-first flood fill."""
    libs = np.full(side * side, -1)
    seen = np.zeros(side * side, bool)
    for s in range(side * side):
        if sq[s] == 2 or seen[s]:
            continue
        col, group, lib = sq[s], [], set()
        dq = deque([s])
        seen[s] = True
        while dq:
            v = dq.popleft()
            group.append(v)
            y, x = divmod(v, side)
            for yy, xx in ((y - 1, x), (y + 1, x), (y, x - 1), (y, x + 1)):
                if 0 <= yy < side and 0 <= xx < side:
                    u = yy * side + xx
                    if sq[u] == 2:
                        lib.add(u)
                    elif sq[u] == col and not seen[u]:
                        seen[u] = True
                        dq.append(u)
        for v in group:
            libs[v] = len(lib)
    return libs


def test_liberties_equal_flood_fill(gold):
    for i, (side, *_rest) in enumerate(gold["cases"]):
        sq = gold[f"squares_{i}"]
        q = gold[f"queries_{i}"]
        fl = flood_liberties(sq, side)
        stones = sq != 2
        assert np.array_equal(q[stones, 1], fl[stones])


def expected_planes(net, side, info, sq, q, komi=7.5):
    """Network.cpp:883-1201 restated from the board queries."""
    P = side * side
    out = np.zeros((32, P), np.uint8)
    tomove = info[1]
    top = 5 if net == NET_POLICY else 6
    own, opp = (3, 8) if net == NET_POLICY else (3, 9)
    aft_own, aft_opp = (13, 19) if net == NET_POLICY else (15, 21)
    lad, ladw = (25, 26) if net == NET_POLICY else (27, 28)
    komi_p, line_p, ko_p = (30, 31, 27) if net == NET_POLICY else (29, 30, 31)
    for idx in range(P):
        y, x = divmod(idx, side)
        if x in (2, side - 3) or y in (2, side - 3):
            out[line_p, idx] = 1
        c = sq[idx]
        if c != 2:
            if c == 1 and abs(komi) > 0.75:
                out[komi_p, idx] = 1
            out[1 if c == tomove else 2, idx] = 1
            libs = min(q[idx, 1], top)
            out[(own if c == tomove else opp) + libs - 1, idx] = 1
        else:
            out[0, idx] = 1
            for val, base in ((q[idx, 3], aft_own), (q[idx, 4], aft_opp)):
                if val >= 1:
                    out[base + min(val, 6) - 1, idx] = 1
            if q[idx, 7] == 1:
                out[lad, idx] = 1
            if q[idx, 8] == 1:
                out[ladw, idx] = 1
    W = side + 2

    def to_idx(v):
        return (v // W - 1) * side + (v % W - 1)

    if info[2] > 0:
        out[ko_p, to_idx(info[2])] = 1
    if net == NET_POLICY and info[3] > 0:
        out[28, to_idx(info[3])] = 1
        if info[4] > 0:
            out[29, to_idx(info[4])] = 1
    return out


@pytest.mark.parametrize("net", [NET_POLICY, NET_VALUE])
def test_planes_match_restated_layout(gold, net):
    for i, (side, *_rest) in enumerate(gold["cases"]):
        p = replay_mine(side, gold[f"moves_{i}"])
        want = expected_planes(net, side, gold[f"info_{i}"],
                               gold[f"squares_{i}"], gold[f"queries_{i}"])
        got = p.planes(net, 0)
        assert np.array_equal(got, want), (i, np.argwhere(got != want)[:5])


def test_planes_partition_and_history():
    for seed in range(1, 30):
        p = Position.random(seed, 19, 250)
        pl = p.planes(NET_POLICY)
        assert np.all(pl[0] + pl[1] + pl[2] == 1)
        # each stone in exactly one liberty bucket of its colour
        assert np.array_equal(pl[3:8].sum(0), pl[1])
        assert np.array_equal(pl[8:13].sum(0), pl[2])
        assert pl[28].sum() == (1 if p.info()["last_move"] > 0 else 0)


def rot_closed_form(idx, s, side=19):
    x, y = idx % side, idx // side
    if s >= 4:
        x, y = y, x
    if s % 4 in (1, 3):
        y = side - 1 - y
    if s % 4 in (2, 3):
        x = side - 1 - x
    return y * side + x


def test_symmetry_planes_rotate():
    p = Position.random(11, 19, 250)
    base = p.planes(NET_POLICY, 0)
    inverse = [0, 1, 2, 3, 4, 6, 5, 7]
    for s in range(8):
        rot = p.planes(NET_POLICY, s)
        idx = np.array([rot_closed_form(i, s) for i in range(361)])
        assert np.array_equal(rot, base[:, idx])
        # the inverse table of rev_rotate_nn_idx undoes the map
        inv = np.array([rot_closed_form(i, inverse[s]) for i in range(361)])
        assert np.array_equal(idx[inv], np.arange(361))
        back = np.empty_like(base)
        back[:, idx] = rot
        assert np.array_equal(back, base)


@pytest.mark.parametrize("side", [5, 9, 13])
def test_symmetry_planes_rotate_small_boards(side):
    p = Position.random(12, side, 60)
    for net in (NET_POLICY, NET_VALUE):
        base = p.planes(net, 0)
        for s in range(8):
            idx = np.array([rot_closed_form(i, s, side)
                            for i in range(side * side)])
            assert np.array_equal(p.planes(net, s), base[:, idx])


def test_make_batch_matches_single_positions():
    b = positions.make_batch(6, seed0=40, net=NET_VALUE, threads=3)
    for i in range(6):
        assert np.array_equal(b[i], Position.random(40 + i, 19, 250)
                              .planes(NET_VALUE))


def test_encode_batch_matches_single_encoder():
    """lz_encode_batch (threads, symmetries) gives the per-position encoder's
    planes; the bitboard after-liberties inside it is pinned through the
    plane tests above, which restate the planes from the reference's
    per-point queries."""
    ps = [positions.Position.random(s, 19, 250) for s in range(700, 760)]
    for net in (positions.NET_POLICY, positions.NET_VALUE):
        for sym in (0, 3, 6):
            got = positions.encode_batch(ps, net, sym, threads=4)
            ref = np.stack([p.planes(net, sym) for p in ps])
            assert np.array_equal(got, ref), (net, sym)
    small = [positions.Position.random(s, 9, 40) for s in range(5)]
    assert positions.encode_batch(small, 0, 0, 2).shape == (5, 32, 81)
    with pytest.raises(Exception):
        positions.encode_batch(ps[:2] + small[:1], 0, 0, 1)

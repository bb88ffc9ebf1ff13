"""Light playouts and the Monte Carlo score: the surrogate's teacher.

FastBoard::final_mc_score is pinned three ways: hand-counted boards,
the reference engine (oracle/_ref, when built) on generator positions
played out to the end (only eyes left), and the invariant that a finished
position's playout outcome is the sign of its score. Playout::mc_owner's
tie (0.5) and side-to-move flip are pinned on finished positions, where no
random move is left to play. The light playout policy itself has no
reference counterpart (the reference draws from its pattern policy), so
its win rates are parity unpinned; the tests check determinism, range and
agreement between the batch and single-position entry points.
"""
import numpy as np
import pytest

from leela_amd import positions
from leela_amd.positions import Position, PASS

try:
    from oracle import ref_board
    HAVE_REF = ref_board.available()
except Exception:  # pragma: no cover - the checker is optional
    HAVE_REF = False


def _pos(side, komi, moves):
    p = Position(side, komi)
    for m in moves:
        p.play(m)
    return p


def test_empty_board_score_is_minus_komi():
    for side, komi in ((5, 7.5), (9, 0.5), (19, 6.5)):
        assert _pos(side, komi, []).score() == -komi


def test_hand_counted_scores():
    p = Position(5, 0.5)
    v = p.vertex
    # Black stones at (1,0) and (0,1) make (0,0) a black eye: 2 + 1.
    for m in (v(1, 0), PASS, v(0, 1)):
        p.play(m)
    assert p.score() == 3 - 0.5
    # A white stone in the far corner and a white eye next to it.
    for m in (v(4, 3), PASS, v(3, 4)):
        p.play(m)
    # White: 2 stones + eye (4,4); black: 2 stones + eye (0,0).
    assert p.score() == (3 - 3) - 0.5
    # A point with three black neighbours and one white one is nobody's.
    q = Position(5, 0.0)
    w = q.vertex
    for m in (w(1, 2), w(2, 1), w(3, 2), w(2, 3)):
        q.play(m)
    # Black (1,2), white (2,1), black (3,2), white (2,3): (2,2) mixed.
    assert q.score() == 0.0


@pytest.mark.skipif(not HAVE_REF, reason="reference engine not built")
def test_score_matches_reference_on_played_out_boards():
    checked = 0
    for side in (5, 9, 13):
        for seed in range(1, 41):
            info, sq, q, moves = ref_board.random_position(side, seed, 4 * side * side)
            mine = Position.random(seed, side, 4 * side * side)
            assert np.array_equal(mine.squares(), sq)
            ref = ref_board.final_score(side, 7.5, moves)
            assert mine.score() == ref, (side, seed)
            checked += 1
    for seed in range(1, 41):
        info, sq, q, moves = ref_board.random_position(19, seed, 250)
        assert Position.random(seed, 19, 250).score() == \
            ref_board.final_score(19, 7.5, moves)
    assert checked == 120


def test_finished_position_outcome_and_flip():
    # Four passes end the game before any playout move: the outcome is the
    # score's sign, for the side to move.
    p = _pos(5, 0.5, [PASS] * 4)            # black to move, W wins by 0.5
    assert p.info()["passes"] == 4 and p.info()["to_move"] == 0
    assert p.playout_winrate(16, 1) == 0.0
    p = _pos(5, 0.5, [PASS] * 5)            # white to move
    assert p.playout_winrate(16, 1) == 1.0
    p = _pos(5, 0.0, [PASS] * 4)            # jigo counts half
    assert p.playout_winrate(16, 1) == 0.5
    p = _pos(5, -3.5, [PASS] * 4)           # negative komi: black wins
    assert p.playout_winrate(16, 1) == 1.0


def test_playouts_deterministic_and_bounded():
    p = Position.random(3, 9, 40)
    a = p.playout_winrate(64, 11)
    assert a == p.playout_winrate(64, 11)
    assert 0.0 <= a <= 1.0
    assert a * 64 * 2 == int(a * 64 * 2)     # a multiple of 1/128
    # The position itself is not changed by the playouts.
    before = p.squares().copy()
    p.playout_winrate(8, 5)
    assert np.array_equal(before, p.squares())


def test_playouts_see_a_won_position():
    # Black owns most of a 9x9 board (a wall on column 6, nothing else):
    # black wins most playouts under komi 7.5 (0.876 over 4096 playouts;
    # random play breaks the wall now and then).
    p = Position(9, 7.5)
    for y in range(9):
        p.play(p.vertex(6, y))
        p.play(PASS)
    w = p.playout_winrate(64, 3)             # black to move
    assert w > 0.75


def test_value_labels_match_single_entry_points():
    pl, w = positions.value_labels(6, seed0=100, playouts=16,
                                   playout_seed=900, side=9, max_moves=48,
                                   threads=3)
    ref_planes = positions.make_batch(6, seed0=100, side=9, max_moves=48,
                                      net=positions.NET_VALUE)
    assert np.array_equal(pl, ref_planes)
    for i in range(6):
        p = Position.random(100 + i, 9, 48)
        assert w[i] == p.playout_winrate(16, 900 + i)
    _, w2 = positions.value_labels(6, 100, 16, 900, side=9, max_moves=48,
                                   threads=1, planes=False)
    assert np.array_equal(w, w2)


def test_handicap_adds_to_komi_in_every_score():
    """FastState::calculate_mc_score scores with komi + handicap
    (FastState.cpp:395-401): a position with handicap h scores, and plays
    out, exactly as the same position with komi raised by h."""
    from leela_amd.positions import PLAYOUT_LIGHT, PLAYOUT_PATTERN
    a = Position.random(17, 9, 40)
    # The same generator game with handicap 0 and with handicap 3.
    base = Position.random(17, 9, 40)
    sq = base.squares()
    for h in (0, 3):
        p = Position.random(17, 9, 40)
        p.set_handicap(h)
        q = Position.random(17, 9, 40)
        np.testing.assert_array_equal(p.squares(), sq)
        if h == 0:
            s0 = p.score()
        else:
            assert p.score() == s0 - 3.0
            for pol in (PLAYOUT_LIGHT, PLAYOUT_PATTERN):
                ws = p.playout_scores(16, 5, pol)
                w0 = q.playout_scores(16, 5, pol)
                # The same moves: only the final subtraction differs.
                np.testing.assert_array_equal(ws, w0 - 3.0)
    with pytest.raises(Exception):
        a.set_handicap(-1)

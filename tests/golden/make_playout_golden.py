#!/usr/bin/env python3
"""Writes tests/golden/playout_golden.npz from the reference's own playout
policy (FastState::play_random_move with the Matcher pattern tables and the
MCPolicy feature weights, compiled from the reference sources by
oracle/ref_board/Makefile).

Each case is a position (the moves that reach it from the empty board) and
a playout seed s: playout j is seeded playout_seed(s, j) (the library's
per-playout seed, restated below), two outputs discarded, run to four
passes or move 2 * side^2, and stored as its final_mc_score with komi 7.5.
Positions are generator games on 9x9, 13x13 and 19x19 at several lengths,
plus states right after a capture (the policy's capture branch) and
simple-ko states (the ko square barred from the draw).

Run: python tests/golden/make_playout_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle.ref_board as rb  # noqa: E402

PLAYOUTS = 16
MAX_LEN = 800


def playout_seed(seed: int, j: int) -> int:
    """The library's per-playout seed (lz::go::playout_seed)."""
    m = 0xFFFFFFFF
    h = (seed ^ ((j * 0x9E3779B9) & m)) & m
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & m
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & m
    h ^= h >> 16
    return h


def stones(sq):
    return int((sq == 0).sum() + (sq == 1).sum())


def special_prefixes(side, seeds, max_moves, want_capture, want_ko):
    """Prefixes that end right after a capture, and ones in a ko state."""
    cap, ko = [], []
    for seed in seeds:
        info, sq, q, moves = rb.random_position(side, seed, max_moves)
        prev = 0
        for k in range(1, len(moves) + 1):
            i2, s2, _ = rb.replay(side, 7.5, moves[:k])
            n = stones(s2)
            if moves[k - 1] > 0 and n < prev + 1 and len(cap) < want_capture:
                cap.append(moves[:k].copy())
            if i2[2] > 0 and len(ko) < want_ko:
                ko.append(moves[:k].copy())
            prev = n
            if len(cap) >= want_capture and len(ko) >= want_ko:
                return cap, ko
    return cap, ko


def main():
    cases = []
    for side, lens in ((9, (0, 30, 80, 200)), (13, (60, 150, 400)),
                       (19, (0, 120, 300, 700))):
        for i, mm in enumerate(lens):
            for s in range(3):
                _, _, _, moves = rb.random_position(side, 100 * side + 10 * i
                                                    + s, mm)
                cases.append((side, moves))
    for side in (9, 19):
        cap, ko = special_prefixes(side, range(1, 60), 3 * side * side, 4, 4)
        assert len(cap) == 4 and len(ko) == 4, (side, len(cap), len(ko))
        cases += [(side, m) for m in cap + ko]
    n = len(cases)
    sides = np.zeros(n, np.int32)
    nmoves = np.zeros(n, np.int32)
    moves = np.zeros((n, MAX_LEN), np.int16)
    seeds = np.zeros(n, np.uint32)
    scores = np.zeros((n, PLAYOUTS), np.float32)
    for c, (side, mv) in enumerate(cases):
        assert len(mv) <= MAX_LEN
        sides[c] = side
        nmoves[c] = len(mv)
        moves[c, :len(mv)] = mv
        seeds[c] = 1000 + 7 * c
        js = [playout_seed(int(seeds[c]), j) for j in range(PLAYOUTS)]
        scores[c] = rb.pattern_playouts(side, 7.5, mv, js)
    out = Path(__file__).resolve().parent / "playout_golden.npz"
    np.savez_compressed(out, sides=sides, nmoves=nmoves, moves=moves,
                        seeds=seeds, scores=scores)
    print(f"wrote {out}: {n} positions x {PLAYOUTS} playouts")


if __name__ == "__main__":
    main()

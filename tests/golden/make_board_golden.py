#!/usr/bin/env python3
"""Writes tests/golden/board_golden.npz from the reference's own board
engine (FastBoard / FullBoard / FastState compiled from the reference
sources by oracle/ref_board/Makefile).

Cases are (side, seed, max_moves, prefix): the synthetic-position generator
is run on the reference engine from `seed`; prefix -1 keeps the whole game,
otherwise the position after the first `prefix` moves of that game;
prefix -2 marks a crafted ladder-rich position (crafted_moves). Full
games cover 19x19, 13x13 and 9x9; prefixes are picked to include simple-ko
states and losing-ladder escapes, which random-play end positions rarely
contain. Each case stores the moves, the FastState fields, the squares and
the per-point board queries the feature planes are built from. The first
outputs of the reference RNG for a few seeds are stored too.

Run: python tests/golden/make_board_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle.ref_board as rb  # noqa: E402
from tests.crafted import crafted_moves  # noqa: E402


def pick_prefixes(n_ko=12, n_ladder=24, seeds=range(1, 400)):
    ko, lad = [], []
    for seed in seeds:
        info, sq, q, moves = rb.random_position(19, seed, 250)
        got_l = False
        for k in range(1, len(moves) + 1):
            i2, s2, q2 = rb.replay(19, 7.5, moves[:k])
            if i2[2] > 0 and len(ko) < n_ko:
                ko.append((19, seed, 250, k))
            if not got_l and (q2[:, 7] == 1).any() and len(lad) < n_ladder \
                    and k > 40:
                lad.append((19, seed, 250, k))
                got_l = True
        if len(ko) >= n_ko and len(lad) >= n_ladder:
            break
    return ko + lad


def main():
    if not rb.build():
        raise SystemExit("reference sources not found; cannot build checker")
    cases = ([(19, s, 250, -1) for s in range(1, 65)]
             + [(19, s, 200, -1) for s in range(1000, 1008)]
             + [(13, s, 101, -1) for s in range(1, 25)]
             + [(9, s, 48, -1) for s in range(1, 25)]
             + pick_prefixes()
             + [(19, s, 0, -2) for s in range(1, 61)])
    out = {"cases": np.array(cases, np.int32)}
    for i, (side, seed, mm, prefix) in enumerate(cases):
        if prefix == -2:
            moves = crafted_moves(seed, side)
            info, sq, q = rb.replay(side, 7.5, moves)
            out[f"info_{i}"] = info
            out[f"squares_{i}"] = sq
            out[f"queries_{i}"] = q
            out[f"moves_{i}"] = moves
            continue
        info, sq, q, moves = rb.random_position(side, seed, mm)
        if prefix >= 0:
            moves = moves[:prefix]
            info, sq, q = rb.replay(side, 7.5, moves)
        out[f"info_{i}"] = info
        out[f"squares_{i}"] = sq
        out[f"queries_{i}"] = q
        out[f"moves_{i}"] = moves
    for seed in (1, 7, 8, 5489):
        out[f"rng_u32_{seed}"] = rb.rng_u32(seed, 16)
        out[f"rng_u32max_{seed}"] = rb.rng_u32(seed, 16, 361)
    path = Path(__file__).with_name("board_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({path.stat().st_size} bytes, {len(cases)} cases)")


if __name__ == "__main__":
    main()

"""Input batches for the benchmark and tests.

`synthetic_planes` builds plausible encoder-shaped plane stacks without a
board engine."""
from __future__ import annotations

import numpy as np


def synthetic_planes(n: int, seed: int, side: int = 19) -> np.ndarray:
    """uint8 [n][32][side*side] stacks shaped like the encoder's output:
    planes 0/1/2 partition the board into empty / own / opponent points, the
    liberty planes mark stones, the after-move planes mark empties, and the
    remaining planes are sparse. Used where the values only need to be
    plausible inputs (kernel parity), not positions of a real game."""
    rng = np.random.default_rng(seed)
    P = side * side
    out = np.zeros((n, 32, P), np.uint8)
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
        xs = np.arange(P) % side
        ys = np.arange(P) // side
        out[i, 31] = (xs == 2) | (xs == side - 3) | (ys == 2) | (ys == side - 3)
    return out


def batch_planes(n: int, net: int = 0, seed0: int = 1) -> np.ndarray:
    """Planes of the benchmark batch."""
    return synthetic_planes(n, seed0)

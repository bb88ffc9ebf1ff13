"""The reference's pattern playout policy on the GPU (mc_playout.hip: one
playout per thread, the host's source, mc_core.h). Pinned to the host
playouts bit for bit (same positions, same seeds: equal win rates), and to
the reference itself through the golden playouts: the device win rate of a
golden position equals the mean outcome of the reference's own playouts
stored in tests/golden/playout_golden.npz."""
import time
from pathlib import Path

import numpy as np
import pytest

from leela_amd import positions
from leela_amd.positions import (PLAYOUT_LIGHT, PLAYOUT_PATTERN,
                                 DeviceEncoder, Position, PositionBatch)

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "playout_golden.npz"


@pytest.fixture(scope="module")
def enc():
    e = DeviceEncoder(0, max_batch=512)
    yield e
    e.close()


def _replay(side, moves):
    p = Position(side)
    for m in moves:
        p.play(int(m))
    return p


def test_device_pattern_playouts_equal_host(enc):
    pos = [Position.random(100 + i, s, s * s * 2 // 3)
           for i, s in enumerate([9] * 8 + [13] * 8 + [19] * 8)]
    n, seed0 = 32, 777
    print(f"case: {len(pos)} positions x {n} pattern playouts", flush=True)
    dev = enc.playout_winrates(PositionBatch(pos), n, seed0, PLAYOUT_PATTERN)
    host = np.array([p.playout_winrate(n, seed0 + i, PLAYOUT_PATTERN)
                     for i, p in enumerate(pos)], np.float32)
    np.testing.assert_array_equal(dev, host)
    # The light policy through the same entry point is unchanged.
    dl = enc.playout_winrates(PositionBatch(pos), n, seed0, PLAYOUT_LIGHT)
    hl = np.array([p.playout_winrate(n, seed0 + i, PLAYOUT_LIGHT)
                   for i, p in enumerate(pos)], np.float32)
    np.testing.assert_array_equal(dl, hl)
    assert (dev != dl).any()


def test_device_pattern_playouts_match_reference_golden(enc):
    g = np.load(GOLDEN)
    scores = g["scores"]
    n_cases, n_play = scores.shape
    for c in range(n_cases):
        side = int(g["sides"][c])
        p = _replay(side, g["moves"][c, :g["nmoves"][c]])
        got = enc.playout_winrates(PositionBatch([p]), n_play,
                                   int(g["seeds"][c]), PLAYOUT_PATTERN)[0]
        s = scores[c]
        black = np.float32(np.where(s > 0, 1.0, np.where(s < 0, 0.0, 0.5))
                           .astype(np.float32).sum()) / np.float32(n_play)
        want = black if p.info()["to_move"] == 0 else np.float32(1) - black
        assert got == want, (c, side, got, want)


def test_device_pattern_labels_and_rate(enc):
    n, playouts, pseed = 256, 64, 4242
    t0 = time.perf_counter()
    _, wd = positions.value_labels_device(n, 9000, playouts, pseed,
                                          planes=False, batch=256,
                                          policy=PLAYOUT_PATTERN)
    dt = time.perf_counter() - t0
    _, wh = positions.value_labels(n, 9000, playouts, pseed, planes=False,
                                   threads=16, policy=PLAYOUT_PATTERN)
    np.testing.assert_array_equal(wd, wh)
    print(f"pattern playouts on the device: {n * playouts / dt:.0f}/s "
          f"(19x19 generator positions, first call incl. setup)")

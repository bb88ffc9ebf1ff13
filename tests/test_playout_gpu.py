"""GPU playouts (lz_playout_winrates_device, playout.hip) against the host
playouts, bit for bit: the same light policy, the same per-playout random
streams, the same score. The host playouts are pinned to the reference's
final score by the CPU tests; the kernel source itself is also run on the
CPU under the wavefront emulation (test_encoder_emul_cpu.py)."""
import numpy as np
import pytest

from leela_amd import positions
from leela_amd.positions import PASS, DeviceEncoder, Position

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def enc():
    import torch
    assert torch.cuda.is_available()
    e = DeviceEncoder(0, max_batch=512)
    yield e
    e.close()


def _pos(side, komi, moves):
    p = Position(side, komi)
    for m in moves:
        p.play(m)
    return p


def test_gpu_playouts_equal_host(enc):
    pos = []
    for i in range(60):
        side, mm = ((19, 250), (13, 101), (9, 48))[i % 3]
        pos.append(Position.random(600 + i, side, mm if i % 2 else mm // 3))
    got = enc.playout_winrates(pos, 16, 4242)
    want = np.array([p.playout_winrate(16, 4242 + i)
                     for i, p in enumerate(pos)], np.float32)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert len(set(want.tolist())) > 10


def test_gpu_playouts_finished_positions_and_flip(enc):
    pos = [_pos(5, 0.5, [PASS] * 4),      # black to move, white wins by 0.5
           _pos(5, 0.5, [PASS] * 5),      # white to move
           _pos(5, 0.0, [PASS] * 4),      # jigo counts half
           _pos(5, -3.5, [PASS] * 4)]     # negative komi: black wins
    assert enc.playout_winrates(pos, 16, 1).tolist() == [0.0, 1.0, 0.5, 1.0]


def test_gpu_playouts_see_a_won_position(enc):
    p = Position(9, 7.5)
    for y in range(9):
        p.play(p.vertex(6, y))
        p.play(PASS)
    w = enc.playout_winrates([p], 64, 3)
    assert w[0] == p.playout_winrate(64, 3) and w[0] > 0.75


def test_value_labels_device_equal_host():
    pl_d, w_d = positions.value_labels_device(40, 300, 8, 900, side=13,
                                              max_moves=101, batch=16)
    pl_h, w_h = positions.value_labels(40, 300, 8, 900, side=13,
                                       max_moves=101, threads=4)
    assert np.array_equal(pl_d, pl_h)
    assert np.array_equal(w_d, w_h)

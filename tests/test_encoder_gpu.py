"""The device feature encoder (lz_encode_device, encoder.hip built for
gfx950) against the reference's board queries and the host encoder,
bit-exact.

Expected planes: the layout restated from tests/golden/board_golden.npz
(queries of the reference's FastBoard over random games, ko and ladder
prefixes and crafted ladder-rich positions), and the host encoder
(pinned to the same fixture by the CPU tests) on generator positions of
mixed sides, all eight symmetries and a full 1024-board batch.
"""
from pathlib import Path

import numpy as np
import pytest

from leela_amd.backend import LeelaError
from leela_amd.positions import (NET_POLICY, NET_VALUE, DeviceEncoder,
                                 Position, encode_batch)

from test_board_cpu import expected_planes

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).parent / "golden" / "board_golden.npz"


@pytest.fixture(scope="module")
def enc():
    import torch
    assert torch.cuda.is_available()
    e = DeviceEncoder(0, max_batch=1024)
    yield e
    e.close()


def replay(side, moves):
    p = Position(side)
    for m in moves:
        p.play(int(m))
    return p


@pytest.mark.parametrize("net", [NET_POLICY, NET_VALUE])
def test_golden_boards_in_one_ragged_batch(enc, net):
    gold = np.load(GOLD)
    cases = gold["cases"]
    pos = [replay(int(c[0]), gold[f"moves_{i}"]) for i, c in enumerate(cases)]
    flat = enc.encode(pos, net).cpu().numpy()
    off = 0
    for i, c in enumerate(cases):
        side = int(c[0])
        sz = 32 * side * side
        got = flat[off:off + sz].reshape(32, side * side)
        want = expected_planes(net, side, gold[f"info_{i}"],
                               gold[f"squares_{i}"], gold[f"queries_{i}"])
        assert np.array_equal(got, want), (i, np.argwhere(got != want)[:5])
        off += sz
    assert off == flat.size


@pytest.mark.parametrize("side,max_moves", [(19, 250), (13, 101), (9, 48)])
def test_all_symmetries_match_host(enc, side, max_moves):
    pos = [Position.random(7000 + i, side, max_moves) for i in range(96)]
    for net in (NET_POLICY, NET_VALUE):
        for sym in range(8):
            got = enc.encode(pos, net, sym).cpu().numpy()
            want = encode_batch(pos, net, sym)
            assert np.array_equal(got, want), (net, sym)


def test_full_batch_matches_host(enc):
    pos = [Position.random(1 + i) for i in range(1024)]
    for net in (NET_POLICY, NET_VALUE):
        got = enc.encode(pos, net).cpu().numpy()
        assert np.array_equal(got, encode_batch(pos, net)), net


def test_both_nets_in_one_pass(enc):
    import torch
    pos = [Position.random(300 + i) for i in range(256)]
    pol = torch.empty((256, 32, 361), dtype=torch.uint8, device="cuda")
    val = torch.empty_like(pol)
    for sym in (0, 6):
        enc.encode_both_into(pos, pol.data_ptr(), val.data_ptr(), sym,
                             torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(pol.cpu().numpy(),
                              encode_batch(pos, NET_POLICY, sym))
        assert np.array_equal(val.cpu().numpy(),
                              encode_batch(pos, NET_VALUE, sym))


def test_back_to_back_calls_reuse_slots(enc):
    """Three calls in flight on one stream (the record slots alternate);
    each output equals its own host encoding."""
    import torch
    batches = [[Position.random(9000 + 100 * k + i) for i in range(64)]
               for k in range(3)]
    outs = [torch.empty((64, 32, 361), dtype=torch.uint8, device="cuda")
            for _ in batches]
    st = torch.cuda.current_stream().cuda_stream
    for b, o in zip(batches, outs):
        enc.encode_into(b, o.data_ptr(), NET_POLICY, 0, st)
    torch.cuda.synchronize()
    for b, o in zip(batches, outs):
        assert np.array_equal(o.cpu().numpy(), encode_batch(b, NET_POLICY))


def test_bad_arguments_rejected(enc):
    import torch
    pos = [Position.random(1)]
    buf = torch.empty(32 * 361, dtype=torch.uint8, device="cuda")
    with pytest.raises(LeelaError):
        enc.encode_into(pos, buf.data_ptr(), 2, 0)
    with pytest.raises(LeelaError):
        enc.encode_into(pos, buf.data_ptr(), NET_POLICY, 8)
    with pytest.raises(LeelaError):
        enc.encode_into([Position.random(1)] * 1025, buf.data_ptr())
    enc.encode_into([], buf.data_ptr())  # empty batch: nothing to do


def test_crafted_ladder_positions_match_host(enc):
    """1,024 crafted ladder-rich positions (ladder starters and random
    breakers, as in the golden fixture but on fresh seeds): the device
    encoder equals the host encoder on both nets."""
    from crafted import crafted_moves
    pos = [replay(19, crafted_moves(20000 + s)) for s in range(1024)]
    for net in (NET_POLICY, NET_VALUE):
        got = enc.encode(pos, net).cpu().numpy()
        want = encode_batch(pos, net)
        assert np.array_equal(got, want), (net, np.argwhere(got != want)[:5])
    # These positions do read ladders: both ladder planes are set somewhere.
    pl = encode_batch(pos, NET_POLICY)
    assert pl[:, 25].sum() > 20 and pl[:, 26].sum() > 100


def test_many_strings_overflow_matches_host(enc):
    """Boards with more strings (180-181) than the kernel keeps in LDS
    (128), mixed in one batch with generator positions: the strings past
    the cap live in the encoder's global overflow area; planes equal the
    host encoder's."""
    from test_encoder_emul_cpu import lattice_moves
    v = lambda x, y: (y + 1) * 21 + (x + 1)  # noqa: E731
    many = [replay(19, lattice_moves(0)), replay(19, lattice_moves(1)),
            replay(19, lattice_moves(0, [v(1, 0), v(3, 0), v(0, 1)])),
            replay(19, lattice_moves(1, [v(0, 0), v(1, 1), v(8, 8)]))]
    pos = []
    for i in range(64):
        pos.append(many[i % 4] if i % 3 == 0 else Position.random(500 + i))
    for net in (NET_POLICY, NET_VALUE):
        got = enc.encode(pos, net).cpu().numpy()
        want = encode_batch(pos, net, 0, 8)
        assert np.array_equal(got.reshape(want.shape), want)

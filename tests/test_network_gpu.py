"""GPU tests of the Network facade: get_scored_moves / get_value /
async_scored_moves against the reference semantics (Network.cpp:536-832)
restated here from the raw batched forward, and the FP32 drop-in against
the oracle's reference CPU path end to end."""
import ctypes as C
import threading

import numpy as np
import pytest

from leela_amd.positions import NET_POLICY, NET_VALUE, Position

pytestmark = pytest.mark.gpu

INVERSE = [0, 1, 2, 3, 4, 6, 5, 7]


def rot(idx, s):
    x, y = idx % 19, idx // 19
    if s >= 4:
        x, y = y, x
    if s % 4 in (1, 3):
        y = 18 - y
    if s % 4 in (2, 3):
        x = 18 - x
    return y * 19 + x


def expected_moves(pos, probs_by_rot, rots):
    """Network.cpp:820-829 + 645-667: reverse-rotate each symmetry's output,
    keep empty points in row-major order, average, zero ladder escapes."""
    sq = pos.squares()
    ladder = pos.planes(NET_POLICY, 0)[25]
    empt = [i for i in range(361) if sq[i] == 2]
    acc = None
    for probs, r in zip(probs_by_rot, rots):
        vals = np.array([probs[rot(i, INVERSE[r])] for i in empt], np.float32)
        acc = vals if acc is None else (acc + vals).astype(np.float32)
    if len(rots) == 8:
        acc = (acc / np.float32(8)).astype(np.float32)
    acc = np.where(ladder[empt] == 1, np.float32(0), acc).astype(np.float32)
    verts = np.array([(i // 19 + 1) * 21 + i % 19 + 1 for i in empt], np.int32)
    return acc, verts


@pytest.fixture(scope="module")
def be(weights_b):
    import leela_amd
    b = leela_amd.Backend(0)
    pol, val = weights_b
    b.load_net(leela_amd.NET_POLICY, pol)
    b.load_net(leela_amd.NET_VALUE, val)
    b.finalize()
    yield b
    b.close()


def ladder_positions():
    """Generator positions that contain a losing-ladder escape point."""
    found = []
    for seed in range(1, 400):
        p = Position.random(seed, 19, 250)
        if p.planes(NET_POLICY)[25].any():
            found.append(p)
        if len(found) == 3:
            break
    return found


@pytest.mark.parametrize("prec_name", ["fp32", "fp16"])
def test_scored_moves_direct_and_average(be, prec_name):
    import leela_amd
    from leela_amd.network import AVERAGE_ALL, DIRECT, Network
    prec = leela_amd.PRECISIONS[prec_name]
    net = Network(be, prec)
    poss = [Position.random(s, 19, 250) for s in (3, 17, 40)] + ladder_positions()
    assert any(p.planes(NET_POLICY)[25].any() for p in poss)
    for pos in poss:
        rows = np.stack([pos.planes(NET_POLICY, r) for r in range(8)])
        _, probs8 = be.forward(leela_amd.NET_POLICY, prec, rows)
        for r in (0, 5):
            got_p, got_v = net.get_scored_moves(pos, DIRECT, r)
            want_p, want_v = expected_moves(pos, [probs8[r]], [r])
            assert np.array_equal(got_v, want_v)
            assert np.array_equal(got_p, want_p)
        got_p, got_v = net.get_scored_moves(pos, AVERAGE_ALL)
        want_p, want_v = expected_moves(pos, list(probs8), list(range(8)))
        assert np.array_equal(got_v, want_v)
        np.testing.assert_allclose(got_p, want_p, rtol=0, atol=1e-7)
        assert got_p.sum() <= 1.0 + 1e-5
        assert np.all(np.diff(got_v) > 0)


def test_fp32_facade_matches_cpu_path(be, weights_b):
    """The drop-in end to end: FP32 get_scored_moves / get_value against the
    oracle's reference CPU path on the same encoded planes."""
    import leela_amd
    import oracle
    from leela_amd.network import DIRECT, Network
    pol, val = weights_b
    oracle.load_blas()
    net = Network(be, leela_amd.PREC_FP32)
    for seed in (8, 21, 99):
        pos = Position.random(seed, 19, 250)
        for r in (0, 6):
            pl = pos.planes(NET_POLICY, r)[None]
            _, cpu = oracle.forward_f32(oracle.Net(pol), pl, oracle.HEAD_POLICY)
            want_p, want_v = expected_moves(pos, [cpu[0]], [r])
            got_p, got_v = net.get_scored_moves(pos, DIRECT, r)
            assert np.array_equal(got_v, want_v)
            assert np.abs(got_p - want_p).max() < 1e-4
            if want_p.size:
                assert got_v[got_p.argmax()] == want_v[want_p.argmax()]
        vpl = pos.planes(NET_VALUE, 0)[None]
        _, vw = oracle.forward_f32(oracle.Net(val), vpl, oracle.HEAD_VALUE)
        assert abs(net.get_value(pos, DIRECT) - vw[0, 0]) < 1e-4


def test_value_ensembles(be):
    import leela_amd
    from leela_amd.network import AVERAGE_ALL, DIRECT, RANDOM_ROTATION, Network
    net = Network(be, leela_amd.PREC_FP32)
    pos = Position.random(77, 19, 250)
    rows = np.stack([pos.planes(NET_VALUE, r) for r in range(8)])
    _, w = be.forward(leela_amd.NET_VALUE, leela_amd.PREC_FP32, rows)
    assert net.get_value(pos, DIRECT) == w[0, 0]
    assert abs(net.get_value(pos, AVERAGE_ALL) - w[:, 0].mean()) < 1e-6
    # One random rotation per call: each draw is one of the 8 rows exactly.
    for _ in range(8):
        v = net.get_value(pos, RANDOM_ROTATION)
        assert any(v == w[r, 0] for r in range(8)), v


def test_non_19_boards(be):
    import leela_amd
    from leela_amd.network import DIRECT, Network
    net = Network(be, leela_amd.PREC_FP32)
    pos = Position.random(5, 13, 101)
    p, v = net.get_scored_moves(pos, DIRECT, 0)
    assert p.size == 0 and v.size == 0
    assert net.get_value(pos, DIRECT) == 0.5


def test_async_scored_moves_through_queue(be):
    import leela_amd
    from leela_amd.network import DIRECT, MOVES_FN, Network
    net = Network(be, leela_amd.PREC_FP32)
    poss = [Position.random(s, 19, 250) for s in range(600, 664)]
    want = [net.get_scored_moves(p, DIRECT, i % 8) for i, p in enumerate(poss)]
    got = [None] * len(poss)

    def done(user, count, probs, verts):
        i = int(user or 0)
        got[i] = (np.ctypeslib.as_array(probs, (count,)).copy() if count else
                  np.zeros(0, np.float32),
                  np.ctypeslib.as_array(verts, (count,)).copy() if count else
                  np.zeros(0, np.int32))
    cb = MOVES_FN(done)
    be.queue_start(leela_amd.PREC_FP32, max_batch=64, max_wait_us=500)
    try:
        def producer(t):
            for i in range(t, len(poss), 16):
                net.async_scored_moves(poss[i], i % 8, cb, i)
        ths = [threading.Thread(target=producer, args=(t,)) for t in range(16)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        be.join()
    finally:
        be.queue_stop()
    for i in range(len(poss)):
        assert np.array_equal(got[i][1], want[i][1])
        assert np.array_equal(got[i][0], want[i][0])


def test_sync_calls_batch_through_queue(be):
    """With the leaf queue running, the synchronous facade calls of many
    threads are batched together and give the direct results bit for bit
    (FP32: rows are independent of their batch)."""
    import leela_amd
    from leela_amd.network import DIRECT, Network
    net = Network(be, leela_amd.PREC_FP32)
    poss = [Position.random(s, 19, 250) for s in range(700, 748)]
    want_p = [net.get_scored_moves(p, DIRECT, i % 8) for i, p in enumerate(poss)]
    want_v = [net.get_value(p, DIRECT) for p in poss]
    got_p = [None] * len(poss)
    got_v = [None] * len(poss)
    be.queue_start(leela_amd.PREC_FP32, max_batch=64, max_wait_us=2000)
    try:
        def worker(t):
            for i in range(t, len(poss), 16):
                got_p[i] = net.get_scored_moves(poss[i], DIRECT, i % 8)
                got_v[i] = net.get_value(poss[i], DIRECT)
        ths = [threading.Thread(target=worker, args=(t,)) for t in range(16)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        stats = be.queue_stats()
    finally:
        be.queue_stop()
    for i in range(len(poss)):
        assert np.array_equal(got_p[i][0], want_p[i][0])
        assert np.array_equal(got_p[i][1], want_p[i][1])
        assert got_v[i] == want_v[i]
    # 96 calls arrived in fewer launches than calls: they were batched.
    assert stats["positions"] == 2 * len(poss)
    assert stats["batches"] < stats["positions"]


def test_netbench(be):
    """Network::benchmark restated: both rates positive, through the queue
    and through direct forwards."""
    import leela_amd
    from leela_amd.network import Network
    net = Network(be, leela_amd.PREC_FP16)
    pos = Position.random(3, 19, 250)
    direct = net.benchmark(pos, 4)
    be.queue_start(leela_amd.PREC_FP16, max_batch=64, max_wait_us=200)
    try:
        queued = net.benchmark(pos, 16)
    finally:
        be.queue_stop()
    assert min(direct) > 0 and min(queued) > 0


def test_first_layer_plane_count():
    """A net whose first conv takes another plane count than the
    reference's 32 (a Caffe net's own inputs) runs in FP32 on planes
    [n][C][361] and matches the oracle; the paths that lay out the
    reference's 32 planes (the leaf queue) and the 16-bit forms refuse
    it."""
    import leela_amd
    import oracle
    from leela_amd import nets
    from leela_amd.backend import LeelaError
    rng = np.random.default_rng(0)
    w1 = (rng.standard_normal((32, 16, 5, 5)) * 0.1).astype(np.float32)
    w2 = (rng.standard_normal((1, 32, 3, 3)) * 0.1).astype(np.float32)
    b1, b2 = np.zeros(32, np.float32), np.zeros(1, np.float32)
    be = leela_amd.Backend(0)
    try:
        be.push_convolve(leela_amd.NET_POLICY, 5, 16, 32, w1, b1)
        be.push_convolve(leela_amd.NET_POLICY, 3, 32, 1, w2, b2)
        be.load_net(leela_amd.NET_VALUE, nets.weight_set("A")[1])
        be.finalize()
        planes = (rng.random((3, 16, 361)) < 0.3).astype(np.uint8)
        raw, prob = be.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP32,
                               planes)
        raw64, _ = oracle.forward_f64(
            oracle.Net([(0, 5, 16, 32, w1, b1), (0, 3, 32, 1, w2, b2)]),
            planes, oracle.HEAD_POLICY)
        assert np.abs(raw - raw64).max() < 1e-5
        with pytest.raises(LeelaError) as e:
            be.forward(leela_amd.NET_POLICY, leela_amd.PREC_FP16, planes)
        assert e.value.code == leela_amd.backend.LZ_ERR_UNSUPPORTED
        with pytest.raises(LeelaError) as e:
            be.queue_start(leela_amd.PREC_FP32, 16, 100)
        assert "32 input planes" in str(e.value)
    finally:
        be.close()


def test_show_heatmap_text(be):
    """Network::show_heatmap's layout (Network.cpp:834-881): 19 rows of
    int(p x 1000) as "%3d ", top row first, 0 on occupied points; then the
    top moves in descending order, as "%1.3f (D4)", until 0.85 is covered
    or a move is below 0.01."""
    import leela_amd
    from leela_amd.network import Network
    pos = Position.random(41)
    net = Network(be, leela_amd.PREC_FP32)
    probs, verts = net.get_scored_moves(pos)
    text = net.show_heatmap(pos)
    lines = text.splitlines()
    grid = [[int(t) for t in ln.split()] for ln in lines[:19]]
    assert all(len(row) == 19 for row in grid)
    want = np.zeros((19, 19), int)
    for p, v in zip(probs, verts):
        x, y = v % 21 - 1, v // 21 - 1
        want[18 - y, x] = int(np.float32(p) * np.float32(1000))
    assert np.array_equal(np.array(grid), want)
    sq = pos.squares().reshape(19, 19)
    assert all(want[18 - y, x] == 0 for y in range(19) for x in range(19)
               if sq[y, x] != 2)
    top = lines[19:]
    order = sorted(zip(probs, verts), reverse=True)
    cum = 0.0
    for i, ln in enumerate(top):
        p, v = order[i]
        x, y = v % 21 - 1, v // 21 - 1
        col = chr(ord("A") + x + (1 if x >= 8 else 0))
        assert ln == f"{p:1.3f} ({col}{y + 1})"
        cum += p
    assert len(top) >= 1


def test_network_get_is_the_reference_net_with_set_a(weights_a):
    """Network::get_Network (lz_network_get): one process-wide context,
    initialised once with the reference's layer tables and weight set A;
    its forwards equal a context loaded with set A from Python, bit for
    bit, and closing a handle to it leaves it alive."""
    import leela_amd
    from leela_amd.positions import make_batch
    d1 = leela_amd.Backend.default()
    d2 = leela_amd.Backend.default()
    assert d1._ctx.value == d2._ctx.value
    ref = leela_amd.Backend(0)
    pol, val = weights_a
    ref.load_net(leela_amd.NET_POLICY, pol)
    ref.load_net(leela_amd.NET_VALUE, val)
    ref.finalize()
    try:
        for net in (NET_POLICY, NET_VALUE):
            planes = make_batch(9, seed0=4242 + net, net=net)
            for prec in (leela_amd.PREC_FP32, leela_amd.PREC_FP16):
                got = d1.forward(net, prec, planes)
                want = ref.forward(net, prec, planes)
                np.testing.assert_array_equal(got[0], want[0])
                np.testing.assert_array_equal(got[1], want[1])
        d1.close()
        raw, _ = d2.forward(NET_POLICY, leela_amd.PREC_FP16,
                            make_batch(2, seed0=1))
        assert np.isfinite(raw).all()
    finally:
        ref.close()

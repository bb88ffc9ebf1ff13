"""GPU tests of ragged batched leaf evaluation through the leaf queue, at
the mixed-board workload: 1024 boards of 19x19, 512 of 13x13 and 512 of
9x9 (random legal play, move counts uniform in 0..0.6 x board area, one
seed range per side) submitted from 16 threads into a queue of max batch
2048, so both nets run one ragged launch of the 2048-board bucket's
captured graph.

Checks:
  * every board's policy and value come back equal, bit for bit, to one
    direct lz_forward_ragged call over the same 2048 boards (a board's
    result does not depend on its slot or batch), for FP16 and BF16;
  * a sample of 8 boards per side against the FP64 oracle run at that side
    (FP16: max |d prob| < 0.005, |d win| < 0.005, as the direct ragged
    test);
  * part-filled buckets (a ragged batch of 5 and one of 300, 19x19-only
    batches between them) evaluate exactly their rows, equal to direct
    calls;
  * an FP32 queue refuses boards other than 19x19.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def be(weights_a):
    import leela_amd
    b = leela_amd.Backend(0)
    pol, val = weights_a
    b.load_net(leela_amd.NET_POLICY, pol)
    b.load_net(leela_amd.NET_VALUE, val)
    b.finalize()
    yield b
    b.close()


@pytest.fixture(scope="module")
def mix():
    """Boards interleaved 19, 13, 19, 9 ...: (sides, policy planes, value
    planes) as lists of [32][S*S] arrays."""
    from leela_amd.positions import ragged_mix
    sides, pol, val = ragged_mix()
    assert len(sides) == 2048 and sides.count(19) == 1024
    assert sides.count(13) == 512 and sides.count(9) == 512
    return sides, pol, val


class Collector:
    def __init__(self, sides, policy):
        self.sides = sides
        self.raw = [None] * len(sides)
        self.out = [None] * len(sides)
        self.status = np.full(len(sides), -99, np.int32)
        from leela_amd.backend import DONE_FN

        def done(user, status, raw, out):
            i = int(user or 0)
            self.status[i] = status
            if status == 0:
                w = sides[i] * sides[i] if policy else 1
                self.raw[i] = np.ctypeslib.as_array(raw, (w,)).copy()
                self.out[i] = np.ctypeslib.as_array(out, (w,)).copy()
        self.cb = DONE_FN(done)


def submit_all(be, net, sides, planes, col, idx, threads=16):
    def producer(t):
        for k in range(t, len(idx), threads):
            i = idx[k]
            be.submit_side(net, sides[i], planes[i], col.cb, int(i))
    ths = [threading.Thread(target=producer, args=(t,))
           for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()


@pytest.mark.parametrize("prec_name", ["fp16", "bf16"])
def test_mixed_2048_from_16_threads_equals_direct(be, mix, prec_name):
    import leela_amd
    prec = leela_amd.PRECISIONS[prec_name]
    sides, pol, val = mix
    print(f"ragged queue {prec_name}: direct forwards", flush=True)
    want_p = be.forward_ragged(leela_amd.NET_POLICY, prec, pol)
    want_v = be.forward_ragged(leela_amd.NET_VALUE, prec, val)
    be.queue_start(prec, max_batch=2048, max_wait_us=5_000_000)
    try:
        pc = Collector(sides, True)
        vc = Collector(sides, False)
        idx = list(range(2048))
        print("ragged queue: 16 producers", flush=True)
        submit_all(be, leela_amd.NET_POLICY, sides, pol, pc, idx)
        submit_all(be, leela_amd.NET_VALUE, sides, val, vc, idx)
        be.join()
        st = be.queue_stats()
        assert st["positions"] == 4096
        assert st["per_bucket"][2048] == 2, st
        assert (pc.status == 0).all() and (vc.status == 0).all()
        for i in range(2048):
            assert np.array_equal(pc.raw[i], want_p[0][i]), i
            assert np.array_equal(pc.out[i], want_p[1][i]), i
            assert np.array_equal(vc.raw[i], want_v[0][i]), i
            assert np.array_equal(vc.out[i], want_v[1][i]), i

        # Part-filled buckets: the graphs read the row count from device
        # memory, so rows past the batch are neither computed nor needed.
        rng = np.random.default_rng(5)
        for m in (5, 300, 64, 77):
            sel = rng.choice(2048, size=m, replace=False)
            if m in (64, 77):  # 19x19-only batches use the dense graphs
                sel = np.array([i for i in range(2048)
                                if sides[i] == 19][:m])
            c = Collector(sides, True)
            submit_all(be, leela_amd.NET_POLICY, sides, pol, c, list(sel),
                       threads=4)
            be.join()
            for i in sel:
                assert c.status[i] == 0
                assert np.array_equal(c.raw[i], want_p[0][i]), (m, i)
                assert np.array_equal(c.out[i], want_p[1][i]), (m, i)
        st = be.queue_stats()
        assert st["positions"] == 4096 + 5 + 300 + 64 + 77
    finally:
        be.queue_stop()


def test_mixed_sample_vs_fp64(be, mix, weights_a):
    import leela_amd
    import oracle
    pol_w, val_w = weights_a
    sides, pol, val = mix
    be.queue_start(leela_amd.PREC_FP16, max_batch=2048, max_wait_us=200)
    try:
        sample = []
        for s in (19, 13, 9):
            sample += [i for i in range(2048) if sides[i] == s][:8]
        pc = Collector(sides, True)
        vc = Collector(sides, False)
        submit_all(be, leela_amd.NET_POLICY, sides, pol, pc, sample)
        submit_all(be, leela_amd.NET_VALUE, sides, val, vc, sample)
        be.join()
    finally:
        be.queue_stop()
    for i in sample:
        s = sides[i]
        _, p64 = oracle.forward_f64(oracle.Net(pol_w), pol[i][None],
                                    oracle.HEAD_POLICY, S=s)
        _, w64 = oracle.forward_f64(oracle.Net(val_w), val[i][None],
                                    oracle.HEAD_VALUE, S=s)
        assert np.abs(pc.out[i] - p64[0]).max() < 0.005, (i, s)
        assert abs(vc.out[i][0] - w64[0, 0]) < 0.005, (i, s)
        np.testing.assert_allclose(pc.out[i].sum(), 1.0, atol=1e-5)


def test_fp32_queue_takes_19x19_only(be, mix):
    import leela_amd
    from leela_amd.backend import LeelaError, DONE_FN
    sides, pol, _ = mix
    cb = DONE_FN(lambda *a: None)
    be.queue_start(leela_amd.PREC_FP32, max_batch=64, max_wait_us=100)
    try:
        i13 = sides.index(13)
        with pytest.raises(LeelaError) as e:
            be.submit_side(leela_amd.NET_POLICY, 13, pol[i13], cb, 0)
        assert e.value.code == leela_amd.backend.LZ_ERR_UNSUPPORTED
        with pytest.raises(LeelaError) as e:
            be.submit_side(leela_amd.NET_POLICY, 4, np.zeros(32 * 16,
                                                             np.uint8), cb, 0)
        assert e.value.code == leela_amd.backend.LZ_ERR_ARG
    finally:
        be.queue_stop()

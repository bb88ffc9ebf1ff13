"""GPU tests of the batched leaf queue: positions submitted from 16
threads come back through their callbacks with exactly the values a direct
batched forward gives (bitwise: a row's result does not depend on the batch
it was in), across graph replays of every bucket with eager forwards and
synchronisations in between."""
import ctypes as C
import threading

import numpy as np
import pytest

from leela_amd.positions import make_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend(weights_a):
    import leela_amd
    be = leela_amd.Backend(0)
    pol, val = weights_a
    be.load_net(leela_amd.NET_POLICY, pol)
    be.load_net(leela_amd.NET_VALUE, val)
    be.finalize()
    yield be
    be.close()


class Collector:
    def __init__(self, n, width):
        self.raw = np.full((n, width), np.nan, np.float32)
        self.out = np.full((n, width), np.nan, np.float32)
        self.status = np.full(n, -99, np.int32)
        self.width = width
        from leela_amd.backend import DONE_FN

        def done(user, status, raw, out):
            i = int(user or 0)
            self.status[i] = status
            if status == 0:
                self.raw[i] = np.ctypeslib.as_array(raw, (self.width,))
                self.out[i] = np.ctypeslib.as_array(out, (self.width,))
        self.cb = DONE_FN(done)


@pytest.mark.parametrize("prec_name", ["fp32", "bf16"])
def test_queue_matches_direct_forward(backend, prec_name):
    import leela_amd
    prec = leela_amd.PRECISIONS[prec_name]
    n = 1500
    pol_planes = make_batch(n, seed0=500)
    val_planes = make_batch(n, seed0=500, net=leela_amd.NET_VALUE)
    want_p, want_pp = backend.forward(leela_amd.NET_POLICY, prec, pol_planes)
    want_v, want_vv = backend.forward(leela_amd.NET_VALUE, prec, val_planes)

    backend.queue_start(prec, max_batch=1024, max_wait_us=300)
    try:
        pc = Collector(n, 361)
        vc = Collector(n, 1)

        def producer(t):
            for i in range(t, n, 16):
                backend.submit(leela_amd.NET_POLICY, pol_planes[i], pc.cb, i)
                backend.submit(leela_amd.NET_VALUE, val_planes[i], vc.cb, i)

        ths = [threading.Thread(target=producer, args=(t,)) for t in range(16)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        backend.join()
        assert backend.queue_depth() == 0
        assert (pc.status == 0).all() and (vc.status == 0).all()
        assert np.array_equal(pc.raw, want_p)
        assert np.array_equal(pc.out, want_pp)
        assert np.array_equal(vc.raw, want_v)
        assert np.array_equal(vc.out, want_vv)

        # Graph replays of small buckets with eager forwards in between.
        for rnd, m in enumerate((5, 64, 70, 300)):
            eager, _ = backend.forward(leela_amd.NET_POLICY, prec,
                                       pol_planes[:m])
            c = Collector(m, 361)
            for i in range(m):
                backend.submit(leela_amd.NET_POLICY, pol_planes[i], c.cb, i)
            backend.join()
            assert np.array_equal(c.raw, eager), rnd
        st = backend.queue_stats()
        assert st["positions"] == 2 * n + 5 + 64 + 70 + 300
        assert st["per_bucket"][64] >= 1
    finally:
        backend.queue_stop()


def test_queue_rejects_without_start(backend):
    import leela_amd
    from leela_amd.backend import LeelaError, DONE_FN
    cb = DONE_FN(lambda *a: None)
    with pytest.raises(LeelaError):
        backend.submit(leela_amd.NET_POLICY,
                       np.zeros((32, 361), np.uint8), cb, 0)

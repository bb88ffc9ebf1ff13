"""The layered form of the 16-bit kernels (fused_forward_layered: one launch
per conv, each conv of a board split over workgroups of output channels)
against the persistent fused kernel: the same outputs bit for bit for both
nets, both 16-bit forms and batch sizes 1..64, directly and through the
leaf queue's smallest bucket; and its latency for small batches."""
import time

import numpy as np
import pytest

from leela_amd.positions import make_batch

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


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("n", [1, 7, 64])
def test_layered_equals_fused(be, prec, n):
    import leela_amd
    p = leela_amd.PRECISIONS[prec]
    print(f"case: {prec} n={n}", flush=True)
    for net, seed in ((leela_amd.NET_POLICY, 11), (leela_amd.NET_VALUE, 12)):
        planes = make_batch(n, seed0=seed * 1000 + n, net=net)
        be.set_layered_max(64)
        raw_l, out_l = be.forward(net, p, planes)
        be.set_layered_max(0)
        raw_f, out_f = be.forward(net, p, planes)
        be.set_layered_max(64)
        np.testing.assert_array_equal(raw_l, raw_f)
        np.testing.assert_array_equal(out_l, out_f)


def test_layered_queue_bucket_equals_direct(be):
    """16 positions through the queue's 64-board bucket (layered graph,
    device-side count) equal a direct fused forward of the same boards."""
    import leela_amd
    p = leela_amd.PREC_FP16
    planes = make_batch(16, seed0=777)
    be.set_layered_max(0)
    want_raw, want = be.forward(leela_amd.NET_POLICY, p, planes)
    be.set_layered_max(64)
    from leela_amd.backend import DONE_FN
    got = np.full((16, 361), np.nan, np.float32)
    status = np.full(16, -99, np.int32)

    def done(user, st, raw, out):
        i = int(user or 0)
        status[i] = st
        if st == 0:
            got[i] = np.ctypeslib.as_array(out, (361,))
    cb = DONE_FN(done)
    be.queue_start(p, 64, 2000)
    try:
        for rep in range(3):  # graph replays with eager work between
            got[:] = np.nan
            for i in range(16):
                be.submit(leela_amd.NET_POLICY, planes[i], cb, i)
            be.join()
            assert (status == 0).all()
            np.testing.assert_array_equal(got, want)
            be.forward(leela_amd.NET_VALUE, p, make_batch(3, seed0=rep + 1,
                                                         net=1))
    finally:
        be.queue_stop()


def test_layered_latency(be):
    """Small-batch latency of the two forms (printed; the layered form must
    not be slower at one board)."""
    import torch
    import leela_amd
    dev = torch.device("cuda", 0)
    p = leela_amd.PREC_FP16
    res = {}
    for n in (1, 16, 64):
        planes = torch.from_numpy(make_batch(n, seed0=5)).to(dev)
        raw = torch.empty((n, 361), device=dev)
        out = torch.empty_like(raw)
        st = torch.cuda.current_stream(dev).cuda_stream
        for lm in (0, 64):
            be.set_layered_max(lm)
            for _ in range(3):
                be.forward_device(0, p, n, planes.data_ptr(), raw.data_ptr(),
                                  out.data_ptr(), st)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 20
            for _ in range(reps):
                be.forward_device(0, p, n, planes.data_ptr(), raw.data_ptr(),
                                  out.data_ptr(), st)
            torch.cuda.synchronize()
            res[(n, lm)] = (time.perf_counter() - t0) / reps * 1e6
    be.set_layered_max(64)
    print("policy fp16 latency (us): " + ", ".join(
        f"n={n} fused {res[(n, 0)]:.0f} layered {res[(n, 64)]:.0f}"
        for n in (1, 16, 64)), flush=True)
    assert res[(1, 64)] < res[(1, 0)]


@pytest.mark.parametrize("n,max_batch", [(6, 64), (100, 256)])
def test_queue_rows_that_are_not_binary_go_in_as_bytes(be, n, max_batch):
    """The queue packs the rows of small batches one bit per plane for the
    layered stem; a batch with a plane byte other than 0 / 1 goes in as
    bytes, so it still equals a direct forward of the same bytes, as does a
    batch too large for the layered kernels (a graph replay of bytes)."""
    import leela_amd
    from leela_amd.backend import DONE_FN
    p = leela_amd.PREC_FP16
    print(f"case: n={n} max_batch={max_batch}", flush=True)
    planes = make_batch(n, seed0=991)
    planes[2, 7, 100] = 2  # not a binary plane value
    be.set_layered_max(64)
    want_raw, want = be.forward(leela_amd.NET_POLICY, p, planes)
    got = np.full((n, 361), np.nan, np.float32)
    status = np.full(n, -99, np.int32)

    def done(user, st, raw, out):
        i = int(user or 0)
        status[i] = st
        if st == 0:
            got[i] = np.ctypeslib.as_array(out, (361,))
    cb = DONE_FN(done)
    be.queue_start(p, max_batch, 2000)
    try:
        for i in range(n):
            be.submit(leela_amd.NET_POLICY, planes[i], cb, i)
        be.join()
    finally:
        be.queue_stop()
    assert (status == 0).all()
    np.testing.assert_array_equal(got, want)
    # and the 2 matters: the same row with a 1 there differs
    planes[2, 7, 100] = 1
    _, other = be.forward(leela_amd.NET_POLICY, p, planes)
    assert not np.array_equal(other[2], want[2])

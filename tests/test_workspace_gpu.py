"""Caller-owned workspaces (lz_workspace_create / lz_forward_device_ws):
forwards of different batches on two streams with a workspace each,
enqueued back to back so that they overlap, give exactly the results of
the same forwards run one after the other (fused and layered forms, both
nets)."""
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


@pytest.mark.parametrize("na,nb", [(300, 257), (16, 40)])
def test_two_streams_two_workspaces(be, na, nb):
    import torch
    import leela_amd
    p = leela_amd.PREC_FP16
    dev = torch.device("cuda", 0)
    print(f"case: n={na},{nb}", flush=True)
    for net, width in ((leela_amd.NET_POLICY, 361), (leela_amd.NET_VALUE, 1)):
        pa = make_batch(na, seed0=101, net=net)
        pb = make_batch(nb, seed0=909, net=net)
        want_a = be.forward(net, p, pa)
        want_b = be.forward(net, p, pb)
        da, db = torch.from_numpy(pa).to(dev), torch.from_numpy(pb).to(dev)
        ra = torch.empty((na, width), device=dev)
        oa = torch.empty_like(ra)
        rb = torch.empty((nb, width), device=dev)
        ob = torch.empty_like(rb)
        s1, s2 = torch.cuda.Stream(dev), torch.cuda.Stream(dev)
        w1, w2 = be.workspace(na), be.workspace(8)  # w2 grows on first use
        try:
            torch.cuda.synchronize()
            for _ in range(3):
                be.forward_device_ws(w1, net, p, na, da.data_ptr(),
                                     ra.data_ptr(), oa.data_ptr(),
                                     s1.cuda_stream)
                be.forward_device_ws(w2, net, p, nb, db.data_ptr(),
                                     rb.data_ptr(), ob.data_ptr(),
                                     s2.cuda_stream)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(ra.cpu().numpy(), want_a[0])
            np.testing.assert_array_equal(oa.cpu().numpy(), want_a[1])
            np.testing.assert_array_equal(rb.cpu().numpy(), want_b[0])
            np.testing.assert_array_equal(ob.cpu().numpy(), want_b[1])
        finally:
            be.free_workspace(w1)
            be.free_workspace(w2)

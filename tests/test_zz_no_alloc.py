"""lz_forward_device_noalloc against lz_forward_device: the same fused
kernels with the value head's plane in the library's static scratch array
instead of the context's workspace, so results must be equal bit for bit.
Two calls in a row per form, the second on a smaller batch (the array keeps
the first call's rows beyond the second's n), through the C ABI."""
import numpy as np
import pytest

from leela_amd.positions import make_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(weights_b):
    import leela_amd
    be = leela_amd.Backend(0)
    pol, val = weights_b
    be.load_net(leela_amd.NET_POLICY, pol)
    be.load_net(leela_amd.NET_VALUE, val)
    be.finalize()
    yield be
    be.close()


def _run(gpu, fn, net, prec, d_planes, n):
    import torch
    per = 361 if net == 0 else 1
    raw = torch.full((n, per), float("nan"), device="cuda:0")
    out = torch.full((n, per), float("nan"), device="cuda:0")
    torch.cuda.synchronize()
    fn(net, prec, n, d_planes.data_ptr(), raw.data_ptr(), out.data_ptr(),
       None)
    torch.cuda.synchronize()
    return raw.cpu().numpy(), out.cpu().numpy()


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp8"])
def test_noalloc_equals_workspace_forward(gpu, prec):
    import torch
    import leela_amd
    p = leela_amd.PRECISIONS[prec]
    planes = torch.from_numpy(make_batch(300, seed0=71)).to("cuda:0")
    for n in (300, 77):
        for net in (leela_amd.NET_POLICY, leela_amd.NET_VALUE):
            print(f"noalloc {prec} net {net} n {n}", flush=True)
            ref = _run(gpu, gpu.forward_device, net, p, planes, n)
            got = _run(gpu, gpu.forward_device_noalloc, net, p, planes, n)
            assert np.isfinite(ref[0]).all() and np.isfinite(ref[1]).all()
            assert np.array_equal(ref[0], got[0])
            assert np.array_equal(ref[1], got[1])


def test_noalloc_rejects_what_it_cannot_hold(gpu):
    import torch
    import leela_amd
    from leela_amd.backend import LeelaError
    planes = torch.zeros((1, 32, 361), dtype=torch.uint8, device="cuda:0")
    out = torch.zeros((1, 361), device="cuda:0")
    with pytest.raises(LeelaError) as e:
        gpu.forward_device_noalloc(leela_amd.NET_VALUE, leela_amd.PREC_BF16,
                                   leela_amd.backend.NOALLOC_MAX_BOARDS + 1,
                                   planes.data_ptr(), out.data_ptr(),
                                   out.data_ptr())
    assert e.value.code == leela_amd.backend.LZ_ERR_ARG
    with pytest.raises(LeelaError) as e:
        gpu.forward_device_noalloc(leela_amd.NET_POLICY, leela_amd.PREC_FP32,
                                   1, planes.data_ptr(), out.data_ptr(),
                                   out.data_ptr())
    assert e.value.code == leela_amd.backend.LZ_ERR_UNSUPPORTED

"""CPU checks of the C ABI library: it loads, exports every function the
header declares, and its host-only helpers behave. No GPU compute."""
import ctypes as C

import numpy as np
import pytest

import leela_amd
from leela_amd import backend, nets


def test_library_exports_every_header_function():
    lib = leela_amd.load_library()
    names = backend.header_functions()
    assert "lz_forward" in names and "lz_init" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_init_without_device_fails_cleanly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    lib = leela_amd.load_library()
    ctx = C.c_void_p()
    rc = lib.lz_init(C.byref(ctx), 0)
    assert rc in (backend.LZ_ERR_HIP, backend.LZ_ERR_ARG)
    assert not ctx.value
    assert lib.lz_last_error(None)


def test_null_context_is_rejected():
    lib = leela_amd.load_library()
    assert lib.lz_finalize(None) == backend.LZ_ERR_ARG
    assert lib.lz_forward(None, 0, 0, 1, None, None, None) == backend.LZ_ERR_ARG
    lib.lz_destroy(None)


def test_normal_fill_deterministic_and_moments():
    a = backend.normal_fill(7, 200000, 2.0)
    b = backend.normal_fill(7, 200000, 2.0)
    c = backend.normal_fill(8, 1000, 2.0)
    assert np.array_equal(a, b)
    assert not np.array_equal(a[:1000], c)
    assert abs(a.mean()) < 0.02
    assert abs(a.std() - 2.0) < 0.02
    # Gaussian tails: P(|z| > 3 sigma) = 0.0027
    assert 0.0015 < np.mean(np.abs(a) > 6.0) < 0.004


def test_net_shapes_match_reference_declarations():
    # Array sizes of Network.cpp:82-137 (BLAS build + value net).
    sizes = [cout * cin * f * f if k == nets.CONV else cout * cin
             for k, f, cin, cout in nets.POLICY_SHAPES]
    assert sizes == [76800, 110592] + [147456] * 10 + [1152]
    vsizes = [cout * cin * f * f if k == nets.CONV else cout * cin
              for k, f, cin, cout in nets.VALUE_SHAPES]
    assert vsizes == [51200] + [36864] * 10 + [576, 92416, 256]
    assert nets.flops_per_position(nets.POLICY_SHAPES) == 1200761088
    assert nets.flops_per_position(nets.VALUE_SHAPES) == 303725696
    assert nets.param_count(nets.POLICY_SHAPES) == 1664609
    assert nets.param_count(nets.VALUE_SHAPES) == 514050


def test_he_normal_scale():
    pol, _ = nets.weight_set("A")
    for kind, f, cin, cout, w, b in pol:
        fan_in = cin * f * f
        if w.size > 10000:
            assert abs(w.std() / np.sqrt(2.0 / fan_in) - 1) < 0.02
        assert not b.any()

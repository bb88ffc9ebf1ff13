"""Pins of the FP8 form's emulation (oracle/fp8.py:forward_fp8) that do not
go through the emulation's own formulas: nets built so that every value the
form quantises is exactly representable (tests/fp8_nets.py), where the form
must equal the FP64 forward, and planted values whose rounding is worked
out by hand, where it must equal the FP64 forward of the net with the
rounded value in place. A dropped or doubled scale, quantisation at the
wrong point (before the bias, after the ELU of the wrong layer), one weight
scale per layer instead of per output channel, or rounding other than to
nearest even moves the result far outside the 1e-5 relative tolerance."""
import numpy as np
import pytest

import oracle
from oracle import fp8

from fp8_nets import exact_net, planes

REL = 1e-5  # FP32 rounding of the per-channel weight scales (max|W| / 448)
OPTIONS = [{}, {"round_weight": True}, {"round_act": True},
           {"tiny_channel": True}]


def _f64(layers, pl, policy):
    head = oracle.HEAD_POLICY if policy else oracle.HEAD_VALUE
    return oracle.forward_f64(oracle.Net(layers), pl, head)


@pytest.mark.parametrize("policy", [True, False])
@pytest.mark.parametrize("opt", OPTIONS, ids=["exact", "round_weight",
                                               "round_act", "tiny_channel"])
def test_emulation_equals_fp64_where_exact(policy, opt):
    pl = planes(3, seed=4)
    net, expected = exact_net(policy, **opt)
    raw8, out8 = fp8.forward_fp8(net, pl, policy=policy)
    raw64, out64 = _f64(expected, pl, policy)
    scale = np.abs(raw64).max()
    assert scale > 0.05
    assert np.abs(raw8 - raw64).max() <= REL * scale
    assert np.abs(out8 - out64).max() <= REL * max(np.abs(out64).max(), 1e-3)
    if opt.get("round_weight") or opt.get("round_act"):
        # The planted value's rounding is visible in the output: without it
        # the result would be a different one.
        raw_un, _ = _f64(net, pl, policy)
        assert np.abs(raw8 - raw_un).max() > 10 * REL * scale


def test_weight_scales_are_per_output_channel():
    """The tiny channel is exact only under its own channel's scale: one
    scale for the whole layer (max |W| = 1) would quantise its weight
    0.3125 to 144 / 448, and that net's FP64 forward is far from the
    emulation's."""
    from fp8_nets import TINY_LAYER
    pl = planes(3, seed=4)
    net, _ = exact_net(True, tiny_channel=True)
    raw8, _ = fp8.forward_fp8(net, pl, policy=True)
    per_layer = [list(x) for x in net]
    w = per_layer[1 + TINY_LAYER][4].copy()
    w[np.isclose(w, 0.3125)] = np.float32(144.0 / 448.0)
    assert np.isclose(w, 144.0 / 448.0).sum() == 1
    per_layer[1 + TINY_LAYER][4] = w
    raw_pl, _ = _f64([tuple(x) for x in per_layer], pl, True)
    assert np.abs(raw8 - raw_pl).max() > 10 * REL * np.abs(raw8).max()

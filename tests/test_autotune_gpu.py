"""GPU tests of the launch-shape autotuner: it reports every candidate,
keeps one per net, leaves results bitwise unchanged (every shape runs the
same k-step order per output), and its choice survives a save/load round
trip."""
import numpy as np
import pytest

from leela_amd.positions import make_batch

pytestmark = pytest.mark.gpu


def test_autotune_choice_and_results(weights_a, tmp_path):
    import leela_amd
    be = leela_amd.Backend(0)
    pol, val = weights_a
    be.load_net(leela_amd.NET_POLICY, pol)
    be.load_net(leela_amd.NET_VALUE, val)
    be.finalize()
    try:
        planes = make_batch(300, seed0=40)
        before_p, _ = be.forward(leela_amd.NET_POLICY, leela_amd.PREC_BF16,
                                 planes)
        before_v, _ = be.forward(leela_amd.NET_VALUE, leela_amd.PREC_BF16,
                                 planes)
        report = be.autotune(leela_amd.PREC_BF16, 300)
        print(report)
        lines = report.strip().splitlines()
        assert sum("chosen" in ln for ln in lines) == 2
        assert sum("grid_mult=" in ln and "chosen" not in ln
                   for ln in lines) == 12
        after_p, _ = be.forward(leela_amd.NET_POLICY, leela_amd.PREC_BF16,
                                planes)
        after_v, _ = be.forward(leela_amd.NET_VALUE, leela_amd.PREC_BF16,
                                planes)
        assert np.array_equal(before_p, after_p)
        assert np.array_equal(before_v, after_v)
        path = str(tmp_path / "tune.txt")
        be.autotune_save(path)
        text = open(path).read()
        be.autotune_load(path)
        be.autotune_save(path)
        assert open(path).read() == text
    finally:
        be.close()

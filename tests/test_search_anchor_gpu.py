"""Anchors of the search-side batching: the search's decisions do not
depend on the leaf queue, and match a search whose nets are the oracle's
restatement of the reference CPU path.

  * One search thread with the leaf queue running (policy rows submitted
    asynchronously, value rows through the synchronous facade, batched
    launches) builds the same tree as with direct forwards: every tree
    statistic equal, root value bit for bit (FP32 and FP16; a row's result
    does not depend on its batch).
  * The FP32 form's search and lz_search_host with the oracle's FP32 CPU
    path (im2col + sgemm, within 1e-4 of the GPU) pick the same best move
    on fixed generator positions, one thread, 96 simulations each.
"""
import pytest

from leela_amd import search
from leela_amd.positions import Position

pytestmark = pytest.mark.gpu

TREE_KEYS = ("simulations", "nodes", "policy_evals", "value_evals",
             "root_visits", "best_move", "best_visits", "root_children",
             "root_value", "visit_sum_mismatches", "virtual_loss_left")


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


@pytest.mark.parametrize("prec_name", ["fp32", "fp16"])
def test_queue_does_not_change_the_tree(be, prec_name):
    import leela_amd
    prec = leela_amd.PRECISIONS[prec_name]
    for seed in (21, 22):
        root = Position.random(seed)
        print(f"search anchor {prec_name} position {seed}", flush=True)
        direct = search.search(root, 1, 160, backend=be, precision=prec,
                               seed=7)
        be.queue_start(prec, max_batch=64, max_wait_us=100)
        try:
            queued = search.search(root, 1, 160, backend=be, precision=prec,
                                   seed=7)
        finally:
            be.queue_stop()
        assert queued["queue_positions"] > 0
        for k in TREE_KEYS:
            assert queued[k] == direct[k], (seed, k, queued[k], direct[k])


def test_best_move_matches_reference_cpu_path(be, weights_a):
    import leela_amd
    import oracle
    oracle.load_blas()
    nets_ = (oracle.Net(weights_a[0]), oracle.Net(weights_a[1]))

    def forward(net, planes):
        head = oracle.HEAD_POLICY if net == 0 else oracle.HEAD_VALUE
        _, out = oracle.forward_f32(nets_[net], planes, head, threads=8)
        return out
    same_tree = 0
    for seed in (31, 32, 33, 34):
        root = Position.random(seed)
        print(f"search vs oracle position {seed}", flush=True)
        gpu = search.search(root, 1, 96, backend=be,
                            precision=leela_amd.PREC_FP32, seed=9)
        cpu = search.search_host(root, forward, 1, 96, seed=9)
        assert gpu["best_move"] == cpu["best_move"], (seed, gpu, cpu)
        assert abs(gpu["root_value"] - cpu["root_value"]) < 1e-3
        same_tree += all(gpu[k] == cpu[k] for k in
                         ("nodes", "best_visits", "root_children"))
    print(f"identical tree statistics on {same_tree} of 4 positions")

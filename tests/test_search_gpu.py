"""GPU tests of lz_search with the nets: the search threads' leaf expansions
(policy) and evaluations (value) go through the synchronous facade calls,
which the leaf queue batches across threads (UCTNode::netscore_children /
run_value_net, UCTNode.cpp:74-121, 316-344)."""
import sys

import pytest

from leela_amd import search
from leela_amd.positions import Position

pytestmark = pytest.mark.gpu


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


def check_tree(st, simulations):
    assert st["simulations"] == simulations
    assert st["root_visits"] == simulations
    assert st["virtual_loss_left"] == 0
    assert st["visit_sum_mismatches"] == 0
    assert st["policy_evals"] == st["value_evals"]
    assert st["root_children"] > 0
    assert 0.0 < st["root_value"] < 1.0


def test_search_batched_through_queue(be):
    import leela_amd
    root = Position.random(21)
    be.queue_start(leela_amd.PREC_FP16, max_batch=256, max_wait_us=300)
    try:
        st = search.search(root, 64, 2000, backend=be,
                           precision=leela_amd.PREC_FP16, seed=4)
    finally:
        be.queue_stop()
    print("search 64 threads:", {k: st[k] for k in (
        "simulations_per_s", "mean_batch", "collisions", "nodes")},
        flush=True, file=sys.stderr)
    check_tree(st, 2000)
    # Each expansion is one policy row and one value row, all queued; the
    # root's policy is the eight-symmetry average (seven rows more).
    assert st["queue_positions"] == st["policy_evals"] + st["value_evals"] + 7
    assert st["mean_batch"] > 4.0


def test_search_direct_forwards(be):
    """No queue: each facade call is its own lz_forward; same tree
    bookkeeping, no queue launches."""
    import leela_amd
    st = search.search(Position.random(22), 4, 200, backend=be,
                       precision=leela_amd.PREC_FP16, seed=5)
    check_tree(st, 200)
    assert st["queue_batches"] == 0


def test_search_single_thread_repeatable(be):
    """One thread: the tree is a function of the nets' outputs, which the
    fused kernels compute deterministically."""
    import leela_amd
    root = Position.random(23)
    runs = [search.search(root, 1, 150, backend=be,
                          precision=leela_amd.PREC_FP16, seed=6)
            for _ in range(2)]
    for k in ("best_move", "best_visits", "nodes", "root_value"):
        assert runs[0][k] == runs[1][k], k

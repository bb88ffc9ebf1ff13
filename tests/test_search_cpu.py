"""lz_search on the host (playout evaluator): the tree's bookkeeping under
several threads, determinism with one thread, and a capture it must find.
The GPU test (test_search_gpu.py) drives the same search with the nets
through the leaf queue."""
import pytest

from leela_amd import positions, search


def vertex(x, y, side=9):
    return (y + 1) * (side + 2) + (x + 1)


def capture_position():
    """9x9, black to move: five white stones on row 4 (x = 2..6) with a
    single liberty at (7, 4); capturing there is the obvious move."""
    p = positions.Position(9)
    black = [(x, 3) for x in range(2, 7)] + [(x, 5) for x in range(2, 7)] + \
        [(1, 4)]
    white = [(x, 4) for x in range(2, 7)]
    for i, b in enumerate(black):
        p.play(vertex(*b))
        p.play(vertex(*white[i]) if i < len(white) else -1)
    return p


def test_capture_position_setup():
    p = capture_position()
    sq = p.squares().reshape(9, 9)
    assert p.info()["to_move"] == 0  # black to move
    assert all(sq[4, x] == 1 for x in range(2, 7))
    assert sq[4, 7] == 2


@pytest.mark.parametrize("threads", [1, 4])
def test_search_bookkeeping(threads):
    p = positions.Position.random(5, side=9, max_moves=20)
    st = search.search(p, threads, 300, evaluator=search.EVAL_PLAYOUTS,
                       playouts=4, seed=3)
    assert st["simulations"] == 300
    assert st["root_visits"] == 300
    assert st["virtual_loss_left"] == 0
    assert st["visit_sum_mismatches"] == 0
    assert st["policy_evals"] == st["value_evals"]
    # Every expansion adds its children; the root is node 1.
    assert st["nodes"] >= 1 + st["root_children"]
    assert 0.0 <= st["root_value"] <= 1.0
    assert st["queue_batches"] == 0  # no GPU context


def test_search_deterministic_single_thread():
    p = positions.Position.random(11, side=9, max_moves=30)
    a = search.search(p, 1, 200, evaluator=search.EVAL_PLAYOUTS, playouts=4,
                      seed=9)
    b = search.search(p, 1, 200, evaluator=search.EVAL_PLAYOUTS, playouts=4,
                      seed=9)
    for k in ("best_move", "best_visits", "root_value", "nodes",
              "root_children"):
        assert a[k] == b[k], k


def test_search_finds_capture():
    # Uniform priors over ~65 candidates: a larger exploration constant
    # than with net priors, so that the capture gets visits early.
    st = search.search(capture_position(), 4, 1200,
                       evaluator=search.EVAL_PLAYOUTS, playouts=16,
                       c_puct=3.0, seed=1)
    assert st["best_move"] == vertex(7, 4)
    assert st["root_value"] > 0.5


def test_search_rejects_bad_arguments():
    p = positions.Position(9)
    with pytest.raises(Exception):
        search.search(p, 0, 10, evaluator=search.EVAL_PLAYOUTS)
    with pytest.raises(Exception):
        search.search(p, 1, 10, evaluator=search.EVAL_NETS)  # no context


def _oracle_forward(weights):
    import oracle
    oracle.load_blas()
    nets_ = (oracle.Net(weights[0]), oracle.Net(weights[1]))

    def forward(net, planes):
        head = oracle.HEAD_POLICY if net == 0 else oracle.HEAD_VALUE
        _, out = oracle.forward_f32(nets_[net], planes, head, threads=1)
        return out
    return forward


def test_search_host_with_the_reference_cpu_path(weights_a):
    """lz_search_host with the oracle's restatement of the reference CPU
    path as the nets: a consistent tree, the same tree on a second run (one
    thread), and the AVERAGE_ALL root expansion passing eight rows."""
    seen = []
    fwd = _oracle_forward(weights_a)

    def forward(net, planes):
        seen.append((net, planes.shape[0]))
        return fwd(net, planes)
    p = positions.Position.random(21)
    a = search.search_host(p, forward, 1, 24, seed=5)
    b = search.search_host(p, fwd, 1, 24, seed=5)
    assert a["simulations"] == 24 and a["root_visits"] == 24
    assert a["virtual_loss_left"] == 0 and a["visit_sum_mismatches"] == 0
    assert a["policy_evals"] == a["value_evals"]
    assert seen[0] == (0, 8)
    assert all(n == 1 for _, n in seen[1:])
    for k in ("nodes", "best_move", "best_visits", "root_children",
              "root_value", "policy_evals"):
        assert a[k] == b[k], k

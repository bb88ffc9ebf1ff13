"""Search-side batching (lz_search): threads of a tree search whose leaf
expansions and evaluations go through the network facade, and so through
the leaf queue when one runs (UCTNode::netscore_children / run_value_net,
UCTNode.cpp:74-121, 316-344; UCTSearch.cpp:79-149)."""
from __future__ import annotations

import ctypes as C

from .backend import LZ_OK, LeelaError, load_library

EVAL_NETS, EVAL_PLAYOUTS = 0, 1

# int (*lz_host_forward_fn)(void *user, int net, int n,
#                           const uint8_t *planes, float *out)
HOST_FORWARD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int,
                              C.POINTER(C.c_uint8), C.POINTER(C.c_float))


class SearchStats(C.Structure):
    """lz_search_stats (include/leela_amd.h)."""
    _fields_ = [
        ("seconds", C.c_double),
        ("simulations", C.c_int64),
        ("policy_evals", C.c_int64),
        ("value_evals", C.c_int64),
        ("collisions", C.c_int64),
        ("nodes", C.c_int64),
        ("queue_batches", C.c_int64),
        ("queue_positions", C.c_int64),
        ("root_visits", C.c_int64),
        ("virtual_loss_left", C.c_int64),
        ("visit_sum_mismatches", C.c_int64),
        ("best_move", C.c_int32),
        ("best_visits", C.c_int32),
        ("root_children", C.c_int32),
        ("root_value", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def _lib():
    L = load_library()
    if not getattr(L, "_search_bound", False):
        vp = C.c_void_p
        L.lz_search.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_float, C.c_uint32,
                                C.POINTER(SearchStats)]
        L.lz_search_host.argtypes = [HOST_FORWARD_FN, vp, vp, C.c_int,
                                     C.c_int, C.c_float, C.c_uint32,
                                     C.POINTER(SearchStats)]
        L._search_bound = True
    return L


def search(pos, threads: int, simulations: int, *, backend=None,
           precision: int = 0, evaluator: int = EVAL_NETS,
           playouts: int = 16, c_puct: float = 1.5, seed: int = 1) -> dict:
    """Runs lz_search from `pos` and returns its statistics as a dict.

    evaluator EVAL_NETS needs a finalized Backend (its leaf queue, if
    started at `precision`, batches the threads' calls); EVAL_PLAYOUTS runs
    on the host only."""
    st = SearchStats()
    ctx = backend._ctx if backend is not None else None
    rc = _lib().lz_search(ctx, precision, pos._p, evaluator, playouts,
                          threads, simulations, c_puct, seed & 0xFFFFFFFF,
                          C.byref(st))
    if rc != LZ_OK:
        raise LeelaError(rc, "lz_search")
    d = st.as_dict()
    d["simulations_per_s"] = (d["simulations"] / d["seconds"]
                              if d["seconds"] > 0 else 0.0)
    d["mean_batch"] = (d["queue_positions"] / d["queue_batches"]
                       if d["queue_batches"] else 0.0)
    return d


def search_host(pos, forward, threads: int, simulations: int, *,
                c_puct: float = 1.5, seed: int = 1) -> dict:
    """lz_search_host: the net-evaluated search with the nets computed by
    `forward(net, planes) -> out` (planes uint8 [n][32][361]; out softmax
    probabilities [n][361] for net 0, win rates [n] for net 1)."""
    import numpy as np

    def cb(user, net, n, planes, out):
        try:
            pl = np.ctypeslib.as_array(planes, (n, 32, 361)).copy()
            res = np.asarray(forward(net, pl), np.float32).reshape(-1)
            np.ctypeslib.as_array(out, (res.size,))[:] = res
            return LZ_OK
        except Exception:  # reported to the search as a failed forward
            return -2
    fn = HOST_FORWARD_FN(cb)
    st = SearchStats()
    rc = _lib().lz_search_host(fn, None, pos._p, threads, simulations,
                               c_puct, seed & 0xFFFFFFFF, C.byref(st))
    if rc != LZ_OK:
        raise LeelaError(rc, "lz_search_host")
    return st.as_dict()

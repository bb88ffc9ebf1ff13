"""Search-side batching (lz_search): threads of a tree search whose leaf
expansions and evaluations go through the network facade, and so through
the leaf queue when one runs (UCTNode::netscore_children / run_value_net,
UCTNode.cpp:74-121, 316-344; UCTSearch.cpp:79-149)."""
from __future__ import annotations

import ctypes as C

from .backend import LZ_OK, LeelaError, load_library

EVAL_NETS, EVAL_PLAYOUTS = 0, 1


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

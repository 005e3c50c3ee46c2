"""ctypes binding of librinnd_hip.so (the C ABI declared in include/rinnd_hip.h).

The library is loaded from this package directory.  There is no fallback: if the HIP library is
missing or no GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RINND_HIP_LIB") or os.path.join(HERE, "librinnd_hip.so")

OK = 0
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "HIP_ERROR", 4: "UNSUPPORTED", 5: "INTERNAL"}

DIST = {
    "sqeuclidean": 0, "cosine": 1, "inner_product": 2, "alt_dot": 3, "direct_cosine": 4,
    "alt_cosine": 5, "alt_inner_product": 6,
}
CORR = {"none": 0, "sqrt": 1, "alt_cosine": 2, "alt_inner_product": 3}
RNG = {"reference": 0, "counter": 1}
JOIN = {"fp32": 0, "bf16": 1}

# every symbol include/rinnd_hip.h declares
EXPORTS = [
    "rinnd_ctx_create", "rinnd_ctx_destroy", "rinnd_last_error", "rinnd_version", "rinnd_rng_seed",
    "rinnd_rng_jump", "rinnd_pair_distances", "rinnd_nn_descent", "rinnd_nn_descent_batch", "rinnd_build_forest",
    "rinnd_forest_n_trees", "rinnd_forest_tree_info", "rinnd_forest_tree_copy", "rinnd_forest_free",
    "rinnd_sort_rows", "rinnd_search_graph_build", "rinnd_index_build", "rinnd_index_free",
    "rinnd_index_query", "rinnd_index_query_noalloc", "rinnd_index_query_one", "rinnd_index_query_quantized",
    "rinnd_index_quantize_trees", "rinnd_index_encode_i8", "rinnd_index_query_i8", "rinnd_index_release_data",
    "rinnd_index_neighbor_graph",
    "rinnd_index_search_graph_info", "rinnd_index_search_graph", "rinnd_index_vertex_order",
    "rinnd_index_trees", "rinnd_server_create", "rinnd_server_submit", "rinnd_server_submit_quantized", "rinnd_server_poll", "rinnd_server_wait",
    "rinnd_server_get_stats", "rinnd_server_last_error", "rinnd_server_destroy",
]


class Params(C.Structure):
    """rinnd_params (mirrors NNDescentParams, core/nndescent/mod.rs:21-40)."""
    _fields_ = [
        ("n_neighbors", C.c_int64), ("n_trees", C.c_int64), ("leaf_size", C.c_int64),
        ("max_candidates", C.c_int64), ("n_iters", C.c_int64), ("delta", C.c_float),
        ("angular", C.c_int32), ("max_depth", C.c_int64), ("verbose", C.c_int32),
        ("n_threads", C.c_int32), ("rng_mode", C.c_int32), ("join_precision", C.c_int32),
    ]


class KernelTiming(C.Structure):
    _fields_ = [("ms", C.c_double), ("launches", C.c_int64), ("bytes", C.c_double)]


# rinnd_timed_kernel order, with the kernel each entry times
TIMED_KERNELS = ("join", "margins", "leaf_dist", "leaf_push")


class BatchStats(C.Structure):
    """rinnd_batch_stats."""
    _fields_ = [
        ("passes", C.c_int32), ("graph_replays", C.c_int32), ("graph_memsets_replaced", C.c_int32),
        ("join_threads_per_row", C.c_int32), ("join_rows_per_group", C.c_int32),
        ("changes", C.c_int64 * 64), ("join_pairs", C.c_int64 * 64),
        ("ms_forest", C.c_double), ("ms_init", C.c_double), ("ms_candidates_join", C.c_double),
        ("ms_apply", C.c_double), ("ms_total", C.c_double),
    ]

    def to_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("changes", "join_pairs")}
        p = self.passes
        d["changes"] = list(self.changes[:min(p, 64)])
        d["join_pairs"] = list(self.join_pairs[:min(p, 64)])
        return d


BATCH_NO_GRAPH = 1


class BuildStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("converged", C.c_int32), ("changes", C.c_int64 * 64),
        ("join_pairs", C.c_int64 * 64), ("join_rescored", C.c_int64 * 64), ("join_updates", C.c_int64 * 64),
        ("ms_forest", C.c_double), ("ms_init", C.c_double), ("ms_candidates", C.c_double),
        ("ms_join", C.c_double), ("ms_apply", C.c_double), ("ms_total", C.c_double),
        ("timed", KernelTiming * 4),
    ]

    def to_dict(self):
        it = self.iterations
        return {
            "iterations": it, "converged": bool(self.converged),
            "changes": list(self.changes[: min(it, 64)]),
            "join_pairs": list(self.join_pairs[: min(it, 64)]),
            "join_rescored": list(self.join_rescored[: min(it, 64)]),
            "join_updates": list(self.join_updates[: min(it, 64)]),
            "ms_forest": self.ms_forest, "ms_init": self.ms_init, "ms_candidates": self.ms_candidates,
            "ms_join": self.ms_join, "ms_apply": self.ms_apply, "ms_total": self.ms_total,
            "timed": {name: {"ms": self.timed[i].ms, "launches": self.timed[i].launches, "bytes": self.timed[i].bytes}
                      for i, name in enumerate(TIMED_KERNELS)},
        }


class IndexParams(C.Structure):
    """rinnd_index_params (mirrors NNDescentBuilder, core/index.rs:449-583)."""
    _fields_ = [
        ("n_neighbors", C.c_int64), ("n_trees", C.c_int64), ("n_search_trees", C.c_int64),
        ("search_tree_leaf_budget", C.c_int64), ("leaf_size", C.c_int64), ("max_candidates", C.c_int64),
        ("n_iters", C.c_int64), ("delta", C.c_float), ("diversify_prob", C.c_float),
        ("pruning_degree_multiplier", C.c_float), ("angular", C.c_int32), ("random_seed", C.c_uint64),
        ("verbose", C.c_int32), ("n_threads", C.c_int32), ("rng_mode", C.c_int32),
        ("join_precision", C.c_int32),
    ]


class ServerParams(C.Structure):
    """rinnd_server_params (include/rinnd_hip.h)."""
    _fields_ = [
        ("max_batch", C.c_int64), ("max_inflight", C.c_int32), ("max_k", C.c_int64), ("visited_bytes", C.c_int64),
        ("quantized", C.c_int32), ("candidate_width", C.c_int64), ("rerank_width", C.c_int64),
        ("target_batch", C.c_int64), ("min_width", C.c_int64),
    ]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_int64), ("queries", C.c_int64), ("launches", C.c_int64),
                ("max_launch_queries", C.c_int64), ("gpu_ms", C.c_double)]


_lib = None
_lock = threading.Lock()
_p = C.c_void_p
_i64 = C.c_int64


def lib():
    """Load librinnd_hip.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is missing: run `python -m rinnd_amd.build` first")
            L = C.CDLL(LIB_PATH)
            sig = {
                "rinnd_ctx_create": (C.c_int, [C.c_int, C.POINTER(_p)]),
                "rinnd_ctx_destroy": (None, [_p]),
                "rinnd_last_error": (C.c_char_p, [_p]),
                "rinnd_version": (C.c_char_p, []),
                "rinnd_rng_seed": (None, [C.c_uint64, _p]),
                "rinnd_rng_jump": (None, [_p, C.c_uint64]),
                "rinnd_pair_distances": (C.c_int, [_p, C.c_int, _p, _i64, _i64, _p, _p, _i64, _p]),
                "rinnd_nn_descent": (C.c_int, [_p, _p, _i64, _i64, C.c_int, C.POINTER(Params), _p, _p, _p,
                                               _p, C.POINTER(_p), C.POINTER(BuildStats)]),
                "rinnd_nn_descent_batch": (C.c_int, [_p, _p, _p, _i64, _i64, C.c_int, C.POINTER(Params), _p, _p,
                                                     _p, _p, _p, _p, _p, _p, C.c_int32, C.POINTER(BatchStats)]),
                "rinnd_build_forest": (C.c_int, [_p, _p, _i64, _i64, _i64, _i64, C.c_int32, _i64, C.c_int32,
                                                 _p, C.POINTER(_p)]),
                "rinnd_forest_n_trees": (_i64, [_p]),
                "rinnd_forest_tree_info": (C.c_int, [_p, _i64, _p, _p]),
                "rinnd_forest_tree_copy": (C.c_int, [_p, _i64, _p, _p, _p, _p]),
                "rinnd_forest_free": (None, [_p]),
                "rinnd_sort_rows": (C.c_int, [_p, _p, _p, _p, _i64, _i64]),
                "rinnd_search_graph_build": (C.c_int, [_p, _p, _p, _p, _i64, _i64, _i64, C.c_int, C.c_float,
                                                       C.c_float, _p, _p, _p, _p]),
                "rinnd_index_build": (C.c_int, [_p, _p, _i64, _i64, C.c_int, C.c_int, C.POINTER(IndexParams),
                                                C.POINTER(_p), C.POINTER(BuildStats)]),
                "rinnd_index_free": (None, [_p]),
                "rinnd_index_query": (C.c_int, [_p, _p, _i64, _i64, C.c_float, _p, _p]),
                "rinnd_index_query_noalloc": (C.c_int, [_p, _p, _i64, _i64, C.c_float, _p, _p]),
                "rinnd_index_query_one": (C.c_int, [_p, _p, _i64, C.c_float, _p, _p]),
                "rinnd_index_query_quantized": (C.c_int, [_p, _p, _i64, _i64, C.c_float, _i64, _i64, _p, _p]),
                "rinnd_index_quantize_trees": (C.c_int, [_p]),
                "rinnd_index_encode_i8": (C.c_int, [_p, C.c_int]),
                "rinnd_index_query_i8": (C.c_int, [_p, _p, _i64, _i64, C.c_float, _i64, _i64, _p, _p]),
                "rinnd_index_release_data": (C.c_int, [_p]),
                "rinnd_index_neighbor_graph": (C.c_int, [_p, _p, _p]),
                "rinnd_index_search_graph_info": (C.c_int, [_p, _p, _p, _p]),
                "rinnd_index_search_graph": (C.c_int, [_p, _p, _p]),
                "rinnd_index_vertex_order": (C.c_int, [_p, _p]),
                "rinnd_index_trees": (_p, [_p]),
                "rinnd_server_create": (C.c_int, [_p, C.POINTER(ServerParams), C.POINTER(_p)]),
                "rinnd_server_submit": (C.c_int, [_p, _p, _i64, _i64, C.c_float, _p, _p, C.POINTER(C.c_uint64)]),
                "rinnd_server_submit_quantized": (C.c_int, [_p, _p, _i64, _i64, C.c_float, _i64, _i64, _p, _p,
                                                            C.POINTER(C.c_uint64)]),
                "rinnd_server_poll": (C.c_int, [_p, C.c_uint64, C.POINTER(C.c_int32)]),
                "rinnd_server_wait": (C.c_int, [_p, C.c_uint64]),
                "rinnd_server_get_stats": (C.c_int, [_p, C.POINTER(ServerStats)]),
                "rinnd_server_last_error": (C.c_char_p, [_p]),
                "rinnd_server_destroy": (None, [_p]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def ptr(a):
    """Raw pointer of a numpy array (host) or a torch tensor (device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())


class RinndError(RuntimeError):
    pass


class Context:
    """One rinnd_ctx (device + HIP stream)."""

    _default = {}

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().rinnd_ctx_create(device, C.byref(h))
        if st != OK:
            raise RinndError(f"rinnd_ctx_create({device}) failed: {STATUS.get(st, st)} (is a GPU visible?)")
        self.h = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def check(self, status: int, what: str):
        if status != OK:
            msg = lib().rinnd_last_error(self.h)
            msg = msg.decode() if msg else ""
            if status == 1:
                raise ValueError(f"{what}: {msg}")
            raise RinndError(f"{what} failed ({STATUS.get(status, status)}): {msg}")

    def __del__(self):
        h = getattr(self, "h", None)
        if h and lib is not None and _lib is not None:
            _lib.rinnd_ctx_destroy(h)
            self.h = None


def rng_seed(seed: int) -> np.ndarray:
    st = np.zeros(4, np.uint64)
    lib().rinnd_rng_seed(C.c_uint64(seed), ptr(st))
    return st


def rng_jump(state: np.ndarray, steps: int) -> np.ndarray:
    st = np.ascontiguousarray(state, np.uint64).copy()
    lib().rinnd_rng_jump(ptr(st), C.c_uint64(steps))
    return st


def make_params(n_neighbors, n_trees, leaf_size, max_candidates, n_iters, delta=0.001, angular=False,
                max_depth=200, verbose=False, n_threads=1, rng_mode="reference", join_precision="fp32"):
    return Params(int(n_neighbors), int(n_trees), int(leaf_size), int(max_candidates), int(n_iters),
                  float(delta), int(bool(angular)), int(max_depth), int(bool(verbose)), int(n_threads),
                  RNG[rng_mode], JOIN[join_precision])


def nn_descent_batch(ctx, data, offsets, dist, params, seeds, n_trees=None, n_iters=None, graph=True):
    """rinnd_nn_descent_batch over host arrays: data (n_total x dim float32), offsets (n_indexes + 1), one FastRng
    seed per index.  Returns (indices, distances, flags, iterations, converged, rng_states, stats dict); the heap
    rows of index s are rows offsets[s]..offsets[s+1], ids in its own numbering."""
    X = np.ascontiguousarray(data, np.float32)
    off = np.ascontiguousarray(offsets, np.int64)
    m = len(off) - 1
    n, d = X.shape
    k = int(params.n_neighbors)
    states = np.stack([rng_seed(int(sd)) for sd in seeds]) if m else np.zeros((0, 4), np.uint64)
    states = np.ascontiguousarray(states, np.uint64)
    nt = None if n_trees is None else np.ascontiguousarray(n_trees, np.int64)
    ni = None if n_iters is None else np.ascontiguousarray(n_iters, np.int64)
    idx = np.zeros((n, k), np.int32)
    dst = np.zeros((n, k), np.float32)
    flg = np.zeros((n, k), np.uint8)
    its = np.zeros(m, np.int32)
    conv = np.zeros(m, np.int32)
    st = BatchStats()
    rc = lib().rinnd_nn_descent_batch(ctx.h, ptr(X), ptr(off), m, d, DIST[dist] if isinstance(dist, str) else dist,
                                      C.byref(params), ptr(nt) if nt is not None else None,
                                      ptr(ni) if ni is not None else None, ptr(states), ptr(idx), ptr(dst), ptr(flg),
                                      ptr(its), ptr(conv), 0 if graph else BATCH_NO_GRAPH, C.byref(st))
    ctx.check(rc, "rinnd_nn_descent_batch")
    return idx, dst, flg, its, conv, states, st.to_dict()


class Forest:
    """rinnd_forest handle; trees are copied to host on demand."""

    def __init__(self, h, owned=True):
        self.h = h
        self.owned = owned

    def __del__(self):
        if getattr(self, "owned", False) and getattr(self, "h", None) and _lib is not None:
            _lib.rinnd_forest_free(self.h)
            self.h = None

    @property
    def n_trees(self) -> int:
        return int(lib().rinnd_forest_n_trees(self.h))

    def n_nodes(self, t: int) -> int:
        nn = np.zeros(1, np.int64)
        dim = np.zeros(1, np.int64)
        if lib().rinnd_forest_tree_info(self.h, t, ptr(nn), ptr(dim)) != OK:
            raise RinndError("rinnd_forest_tree_info failed")
        return int(nn[0])

    def tree(self, t: int, n: int):
        nn = np.zeros(1, np.int64)
        dim = np.zeros(1, np.int64)
        st = lib().rinnd_forest_tree_info(self.h, t, ptr(nn), ptr(dim))
        if st != OK:
            raise RinndError("rinnd_forest_tree_info failed")
        nn, dim = int(nn[0]), int(dim[0])
        hp = np.zeros((nn, dim), np.float32)
        off = np.zeros(nn, np.float32)
        ch = np.zeros((nn, 2), np.int32)
        idx = np.zeros(n, np.int32)
        st = lib().rinnd_forest_tree_copy(self.h, t, ptr(hp), ptr(off), ptr(ch), ptr(idx))
        if st != OK:
            raise RinndError("rinnd_forest_tree_copy failed")
        return {"hyperplanes": hp, "offsets": off, "children": ch, "indices": idx, "n_nodes": nn}


class Server:
    """rinnd_server: asynchronous batched queries over a resident index (handle of rinnd_index_build).
    quantized: False (FP32), True or "fp8" (FP8 e4m3 candidates) or "int8" (the index's int8 codes)."""

    def __init__(self, index_handle, max_batch=4096, max_inflight=3, max_k=128, visited_bytes=8 << 30,
                 quantized=False, candidate_width=0, rerank_width=0, target_batch=0, min_width=0):
        q = {"fp8": 1, "int8": 2}.get(quantized, None) if isinstance(quantized, str) else int(bool(quantized))
        if q is None:
            raise ValueError(f"unknown quantized mode {quantized!r}")
        p = ServerParams(int(max_batch), int(max_inflight), int(max_k), int(visited_bytes), q,
                         int(candidate_width), int(rerank_width), int(target_batch), int(min_width))
        h = C.c_void_p()
        st = lib().rinnd_server_create(index_handle, C.byref(p), C.byref(h))
        if st != OK:
            raise RinndError(f"rinnd_server_create failed ({STATUS.get(st, st)})")
        self.h = h

    def _check(self, st, what):
        if st != OK:
            msg = lib().rinnd_server_last_error(self.h)
            msg = msg.decode() if msg else ""
            if st == 1:
                raise ValueError(f"{what}: {msg}")
            raise RinndError(f"{what} failed ({STATUS.get(st, st)}): {msg}")

    def submit(self, queries, nq, k, epsilon, out_idx, out_dist, candidate_width=None, rerank_width=None):
        """Queues a request; queries / out_idx / out_dist are arrays, tensors or raw addresses (int) and must stay
        valid until the request completes.  Widths (quantized servers) default to the server's."""
        t = C.c_uint64()
        addr = lambda a: C.c_void_p(a) if isinstance(a, int) else ptr(a)  # noqa: E731
        if candidate_width is None and rerank_width is None:
            st = lib().rinnd_server_submit(self.h, addr(queries), int(nq), int(k), C.c_float(epsilon), addr(out_idx),
                                           addr(out_dist), C.byref(t))
        else:
            st = lib().rinnd_server_submit_quantized(self.h, addr(queries), int(nq), int(k), C.c_float(epsilon),
                                                     int(candidate_width or 0), int(rerank_width or 0), addr(out_idx),
                                                     addr(out_dist), C.byref(t))
        self._check(st, "rinnd_server_submit")
        return t.value

    def poll(self, ticket):
        d = C.c_int32()
        self._check(lib().rinnd_server_poll(self.h, C.c_uint64(ticket), C.byref(d)), "rinnd_server_poll")
        return bool(d.value)

    def wait(self, ticket):
        self._check(lib().rinnd_server_wait(self.h, C.c_uint64(ticket)), "rinnd_server_wait")

    def stats(self):
        s = ServerStats()
        lib().rinnd_server_get_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.rinnd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

"""ctypes front end of the oracle (CPU restatement of lmcinnes/rinnd's build and search path).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker.  The product package (rinnd_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

METRICS = {
    "sqeuclidean": 0, "cosine": 1, "inner_product": 2, "alt_dot": 3, "direct_cosine": 4,
    "alt_cosine": 5, "alt_inner_product": 6,
}
CORR = {"none": 0, "sqrt": 1, "alt_cosine": 2, "alt_inner_product": 3}

_p = C.c_void_p
_i64 = C.c_int64
_f = C.c_float


def build():
    subprocess.run(["bash", os.path.join(_HERE, "build.sh")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_rng_seed": (None, [C.c_uint64, _p]),
            "orc_rng_u64": (None, [_p, _i64, _p]),
            "orc_rng_float": (None, [_p, _i64, _p]),
            "orc_rng_bool": (None, [_p, _i64, _p]),
            "orc_rng_index": (None, [_p, C.c_uint64, _i64, _p]),
            "orc_distance": (_f, [C.c_int, _p, _p, _i64]),
            "orc_pair_distances": (None, [C.c_int, _p, _i64, _p, _p, _i64, _p]),
            "orc_correct": (_f, [C.c_int, _f]),
            "orc_heap_checked_flagged_push": (C.c_int, [_p, _p, _p, _i64, C.c_int32, _f, C.c_int]),
            "orc_heap_simple_push": (C.c_int, [_p, _p, _p, _i64, C.c_int32, _f]),
            "orc_heap_unchecked_flagged_push": (None, [_p, _p, _p, _i64, _i64, _i64, C.c_int32, _f, C.c_int]),
            "orc_cand_push": (None, [_p, _p, _i64, _f, C.c_int32]),
            "orc_sort_all": (None, [_p, _p, _p, _i64, _i64]),
            "orc_build_forest": (_p, [_p, _i64, _i64, _i64, _i64, _p, C.c_int, _i64, C.c_int]),
            "orc_build_tree": (_p, [_p, _i64, _i64, _i64, _p, C.c_int, _i64]),
            "orc_forest_from_arrays": (_p, [_i64, _i64, _p, _p, _p, _i64, _p]),
            "orc_forest_free": (None, [_p]),
            "orc_forest_n_trees": (_i64, [_p]),
            "orc_forest_n_nodes": (_i64, [_p, _i64]),
            "orc_forest_n_indices": (_i64, [_p, _i64]),
            "orc_forest_copy": (None, [_p, _i64, _p, _p, _p, _p]),
            "orc_forest_leaves": (_i64, [_p, _i64, _p, _i64]),
            "orc_tree_search_leaves": (_i64, [_p, _i64, _p, _p, _i64, _p, _i64, _p]),
            "orc_init_from_leaves": (None, [_p, _p, _p, _i64, _i64, _p, _p, _i64, C.c_int, C.c_int]),
            "orc_build_candidates": (None, [_p, _p, _p, _i64, _i64, _i64, _p, C.c_int, _p, _p]),
            "orc_update_iteration": (_i64, [_p, _p, _p, _i64, _i64, _p, _p, _i64, _p, _i64, C.c_int, C.c_int]),
            "orc_nn_descent": (_p, [_p, _i64, _i64, C.c_int, _i64, _i64, _i64, _i64, _i64, _f, C.c_int,
                                    _i64, _p, C.c_int, _p, _p, _p, _p, _p, _p]),
            "orc_search_graph": (_i64, [_p, _p, _p, _i64, _i64, _i64, C.c_int, _f, _f, C.c_int, _p, _p,
                                        _i64, _p]),
            "orc_weighted_union": (_i64, [_i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _i64]),
            "orc_bounded_heap": (_i64, [_i64, _p, _p, _i64, _p, _p, _p]),
            "orc_index_build": (_p, [_p, _i64, _i64, C.c_int, C.c_int, _i64, _i64, _i64, _i64, _i64, _i64,
                                     _i64, _f, _f, _f, C.c_int, C.c_uint64, C.c_int]),
            "orc_index_free": (None, [_p]),
            "orc_index_query": (None, [_p, _p, _i64, _i64, _f, C.c_int, _p, _p]),
            "orc_index_query_one": (None, [_p, _p, _i64, _f, _p, _p]),
            "orc_index_neighbor_graph": (None, [_p, _p, _p]),
            "orc_index_search_graph_nnz": (_i64, [_p]),
            "orc_index_min_distance": (_f, [_p]),
            "orc_index_search_graph": (None, [_p, _p, _p]),
            "orc_index_vertex_order": (None, [_p, _p]),
            "orc_index_n_trees": (_i64, [_p]),
            "orc_index_forest": (_p, [_p]),
            "orc_index_trace": (None, [_p, _p, _p, _p]),
            "orc_diversify": (_i64, [_p, _p, _p, _i64, C.c_int, _p, _i64, C.c_int, _i64, _p, _p]),
            "orc_greedy": (_i64, [_p, _p, _i64, _i64, C.c_int, _p, _p, _p, _i64, _i64, _f, _f, _p, _p, _p]),
            "orc_index_quantize_trees": (None, [_p]),
            "orc_quantized_margin": (_f, [_p, _f, _i64, _p, _f, _p, _p]),
            "orc_quantize_query_i8": (_f, [_p, _i64, _p]),
            "orc_tree_search_leaves_q": (_i64, [_p, _i64, _p, _f, _p, _i64, _p]),
            "orc_index_encode_i8": (_p, [_p, C.c_int]),
            "orc_i8_free": (None, [_p]),
            "orc_i8_codes": (None, [_p, _p, _p, _p]),
            "orc_encode_i8_rows": (_p, [_p, _i64, _i64, C.c_int]),
            "orc_quantize_query_codes": (_f, [_p, _p, _p]),
            "orc_i8_cosine_distance": (_f, [_p, _f, _p, _f, _i64]),
            "orc_index_query_i8": (None, [_p, _p, _p, _i64, _i64, _i64, _i64, _f, C.c_int, _p, _p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---- RNG ------------------------------------------------------------------------------------------
def rng_state(seed):
    st = np.zeros(4, np.uint64)
    lib().orc_rng_seed(C.c_uint64(seed), ptr(st))
    return st


def rng_u64(st, count):
    out = np.zeros(count, np.uint64)
    lib().orc_rng_u64(ptr(st), count, ptr(out))
    return out


def rng_float(st, count):
    out = np.zeros(count, np.float32)
    lib().orc_rng_float(ptr(st), count, ptr(out))
    return out


def rng_bool(st, count):
    out = np.zeros(count, np.uint8)
    lib().orc_rng_bool(ptr(st), count, ptr(out))
    return out.astype(bool)


def rng_index(st, n, count):
    out = np.zeros(count, np.uint64)
    lib().orc_rng_index(ptr(st), C.c_uint64(n), count, ptr(out))
    return out


# ---- distances -------------------------------------------------------------------------------------
def distance(metric, a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return float(lib().orc_distance(METRICS[metric], ptr(a), ptr(b), a.shape[0]))


def pair_distances(metric, data, a, b):
    data = np.ascontiguousarray(data, np.float32)
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    out = np.zeros(a.shape[0], np.float32)
    lib().orc_pair_distances(METRICS[metric], ptr(data), data.shape[1], ptr(a), ptr(b), a.shape[0], ptr(out))
    return out


def correct(corr, d):
    return float(lib().orc_correct(CORR[corr], C.c_float(d)))


# ---- heaps -----------------------------------------------------------------------------------------
class Heap:
    """Single-row NeighborHeap for unit tests."""

    def __init__(self, k):
        self.idx = np.full(k, -1, np.int32)
        self.dist = np.full(k, np.inf, np.float32)
        self.flags = np.zeros(k, np.uint8)
        self.k = k

    def checked_flagged_push(self, nb, d, is_new):
        return bool(lib().orc_heap_checked_flagged_push(ptr(self.idx), ptr(self.dist), ptr(self.flags),
                                                        self.k, nb, C.c_float(d), int(is_new)))

    def simple_push(self, nb, d):
        return bool(lib().orc_heap_simple_push(ptr(self.idx), ptr(self.dist), ptr(self.flags), self.k, nb,
                                               C.c_float(d)))


def cand_push(pri, idx, priority, index):
    lib().orc_cand_push(ptr(pri), ptr(idx), pri.shape[0], C.c_float(priority), int(index))


def sort_all(idx, dist, flags):
    n, k = idx.shape
    lib().orc_sort_all(ptr(idx), ptr(dist), ptr(flags), n, k)


def bounded_heap(k, ds, ids):
    ds = np.ascontiguousarray(ds, np.float32)
    ids = np.ascontiguousarray(ids, np.int32)
    ins = np.zeros(len(ds), np.uint8)
    oi = np.zeros(max(k, 1), np.int32)
    od = np.zeros(max(k, 1), np.float32)
    m = lib().orc_bounded_heap(k, ptr(ds), ptr(ids), len(ds), ptr(ins), ptr(oi), ptr(od))
    return oi[:m], od[:m], ins.astype(bool)


# ---- forest ----------------------------------------------------------------------------------------
class Forest:
    def __init__(self, handle):
        self.h = handle

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_forest_free(self.h)
            self.h = None

    @property
    def n_trees(self):
        return int(lib().orc_forest_n_trees(self.h))

    def tree(self, t):
        L = lib()
        nn = int(L.orc_forest_n_nodes(self.h, t))
        ni = int(L.orc_forest_n_indices(self.h, t))
        dim_hp = None
        hp_len = nn  # dim unknown here; recover from hyperplane length below
        off = np.zeros(nn, np.float32)
        ch = np.zeros((nn, 2), np.int32)
        idx = np.zeros(ni, np.int32)
        L.orc_forest_copy(self.h, t, None, ptr(off), ptr(ch), ptr(idx))
        return {"offsets": off, "children": ch, "indices": idx, "n_nodes": nn}

    def hyperplanes(self, t, dim):
        nn = int(lib().orc_forest_n_nodes(self.h, t))
        hp = np.zeros((nn, dim), np.float32)
        lib().orc_forest_copy(self.h, t, ptr(hp), None, None, None)
        return hp

    def leaves(self, t):
        cap = int(lib().orc_forest_n_nodes(self.h, t)) + 1
        out = np.zeros((cap, 2), np.int64)
        m = int(lib().orc_forest_leaves(self.h, t, ptr(out), cap))
        return [tuple(map(int, r)) for r in out[:m]]

    def search_leaves(self, t, point, st, budget):
        point = np.ascontiguousarray(point, np.float32)
        out = np.zeros((max(budget, 1), 2), np.int64)
        ev = np.zeros(1, np.int64)
        m = int(lib().orc_tree_search_leaves(self.h, t, ptr(point), ptr(st), budget, ptr(out), budget, ptr(ev)))
        return [tuple(map(int, r)) for r in out[:m]], int(ev[0])

    def search_leaves_quantized(self, t, codes, query_scale, st, budget):
        """QuantizedFlatTree::search_leaves of tree t's int8 copy (flat_tree.rs:103-164)."""
        codes = np.ascontiguousarray(codes, np.int8)
        out = np.zeros((max(budget, 1), 2), np.int64)
        m = int(lib().orc_tree_search_leaves_q(self.h, t, ptr(codes), C.c_float(query_scale), ptr(st), budget,
                                               ptr(out)))
        return [tuple(map(int, r)) for r in out[:m]]


def build_forest(data, n_trees, leaf_size, st, angular=False, max_depth=200, threads=1):
    data = np.ascontiguousarray(data, np.float32)
    h = lib().orc_build_forest(ptr(data), data.shape[0], data.shape[1], n_trees, leaf_size, ptr(st),
                               int(angular), max_depth, threads)
    return Forest(h)


def build_tree(data, leaf_size, st, angular=False, max_depth=200):
    data = np.ascontiguousarray(data, np.float32)
    h = lib().orc_build_tree(ptr(data), data.shape[0], data.shape[1], leaf_size, ptr(st), int(angular), max_depth)
    return Forest(h)


def forest_from_arrays(hyperplanes, offsets, children, indices):
    hp = np.ascontiguousarray(hyperplanes, np.float32)
    off = np.ascontiguousarray(offsets, np.float32)
    ch = np.ascontiguousarray(children, np.int32)
    idx = np.ascontiguousarray(indices, np.int32)
    h = lib().orc_forest_from_arrays(hp.shape[1], hp.shape[0], ptr(hp), ptr(off), ptr(ch), idx.shape[0], ptr(idx))
    return Forest(h)


# ---- NN-Descent ------------------------------------------------------------------------------------
def empty_heap(n, k):
    return (np.full((n, k), -1, np.int32), np.full((n, k), np.inf, np.float32), np.zeros((n, k), np.uint8))


def init_from_leaves(heap, forest, data, metric="sqeuclidean", threads=1):
    idx, dist, flags = heap
    n, k = idx.shape
    data = np.ascontiguousarray(data, np.float32)
    lib().orc_init_from_leaves(ptr(idx), ptr(dist), ptr(flags), n, k, forest.h, ptr(data), data.shape[1],
                               METRICS[metric], threads)


def build_candidates(heap, C_, st, threads=1):
    idx, dist, flags = heap
    n, k = idx.shape
    new = np.zeros((n, C_), np.int32)
    old = np.zeros((n, C_), np.int32)
    lib().orc_build_candidates(ptr(idx), ptr(dist), ptr(flags), n, k, C_, ptr(st), threads, ptr(new), ptr(old))
    return new, old


def update_iteration(heap, new, old, data, metric="sqeuclidean", threads=1):
    idx, dist, flags = heap
    n, k = idx.shape
    data = np.ascontiguousarray(data, np.float32)
    new = np.ascontiguousarray(new, np.int32)
    old = np.ascontiguousarray(old, np.int32)
    return int(lib().orc_update_iteration(ptr(idx), ptr(dist), ptr(flags), n, k, ptr(new), ptr(old), new.shape[1],
                                          ptr(data), data.shape[1], METRICS[metric], threads))


def nn_descent(data, k, n_trees, leaf_size, max_candidates, n_iters, delta=0.001, metric="sqeuclidean",
               angular=False, max_depth=200, seed=42, st=None, threads=1):
    data = np.ascontiguousarray(data, np.float32)
    n = data.shape[0]
    if st is None:
        st = rng_state(seed)
    idx, dist, flags = empty_heap(n, k)
    changes = np.zeros(64, np.int64)
    iters = np.zeros(1, np.int64)
    ms = np.zeros(4, np.float64)
    h = lib().orc_nn_descent(ptr(data), n, data.shape[1], METRICS[metric], k, n_trees, leaf_size, max_candidates,
                             n_iters, C.c_float(delta), int(angular), max_depth, ptr(st), threads, ptr(idx), ptr(dist),
                             ptr(flags), ptr(changes), ptr(iters), ptr(ms))
    return {"indices": idx, "distances": dist, "flags": flags, "forest": Forest(h),
            "changes": changes[: int(iters[0])].copy(), "ms": ms, "rng_state": st}


def search_graph(sorted_idx, sorted_dist, data, metric="sqeuclidean", prob=1.0, mult=1.5, threads=1):
    data = np.ascontiguousarray(data, np.float32)
    n, k = sorted_idx.shape
    maxdeg = int(np.round(np.float32(mult) * np.float32(k)))
    indptr = np.zeros(n + 1, np.int32)
    cap = n * max(maxdeg, 1)
    indices = np.zeros(cap, np.int32)
    md = np.zeros(1, np.float32)
    nnz = lib().orc_search_graph(ptr(np.ascontiguousarray(sorted_idx, np.int32)),
                                 ptr(np.ascontiguousarray(sorted_dist, np.float32)), ptr(data), n, k, data.shape[1],
                                 METRICS[metric], C.c_float(prob), C.c_float(mult), threads, ptr(indptr), ptr(indices),
                                 cap, ptr(md))
    return indptr, indices[:nnz].copy(), float(md[0])


def diversify(row_idx, row_dist, data, max_degree, order=None, stop_at_neg=True, metric="sqeuclidean"):
    """diversify_row (stop_at_neg, positions in row order) or diversify_sorted_csr (positions in `order`,
    -1 entries skipped) on one row; returns the retained (ids, distances)."""
    ri = np.ascontiguousarray(row_idx, np.int32)
    rd = np.ascontiguousarray(row_dist, np.float32)
    data = np.ascontiguousarray(data, np.float32)
    order = np.arange(len(ri), dtype=np.int64) if order is None else np.ascontiguousarray(order, np.int64)
    oi = np.zeros(max(len(ri), 1), np.int32)
    od = np.zeros(max(len(ri), 1), np.float32)
    m = int(lib().orc_diversify(ptr(ri), ptr(rd), ptr(order), len(order), int(stop_at_neg), ptr(data), data.shape[1],
                                METRICS[metric], max_degree, ptr(oi), ptr(od)))
    return oi[:m].copy(), od[:m].copy()


def greedy(q, data, indptr, indices, seeds, k, epsilon, min_distance, st, metric="sqeuclidean"):
    """Greedy search from explicit seed ids over a CSR graph; `st` (FastRng state) is advanced in place.
    Returns (ids, distances) padded to k with (-1, +inf)."""
    q = np.ascontiguousarray(q, np.float32)
    data = np.ascontiguousarray(data, np.float32)
    indptr = np.ascontiguousarray(indptr, np.int32)
    indices = np.ascontiguousarray(indices if len(indices) else [0], np.int32)
    seeds = np.ascontiguousarray(seeds if len(seeds) else [-1], np.int32)
    oi = np.zeros(k, np.int32)
    od = np.zeros(k, np.float32)
    lib().orc_greedy(ptr(q), ptr(data), data.shape[0], data.shape[1], METRICS[metric], ptr(indptr), ptr(indices),
                     ptr(seeds), len(seeds), k, C.c_float(epsilon), C.c_float(min_distance), ptr(st), ptr(oi),
                     ptr(od))
    return oi, od


def weighted_union(forward, reverse, max_degree):
    """forward/reverse: lists (one per vertex) of (idx, dist) pairs."""
    n = len(forward)

    def csr(rows):
        ptr_ = np.zeros(n + 1, np.int64)
        for v, r in enumerate(rows):
            ptr_[v + 1] = ptr_[v] + len(r)
        ii = np.array([e[0] for r in rows for e in r] or [0], np.int32)
        dd = np.array([e[1] for r in rows for e in r] or [0], np.float32)
        return ptr_, ii, dd

    fp, fi, fd = csr(forward)
    rp, ri, rd = csr(reverse)
    indptr = np.zeros(n + 1, np.int32)
    cap = n * max_degree + 1
    indices = np.zeros(cap, np.int32)
    nnz = lib().orc_weighted_union(n, max_degree, ptr(fp), ptr(fi), ptr(fd), ptr(rp), ptr(ri), ptr(rd),
                                   ptr(indptr), ptr(indices), cap)
    return indptr, indices[:nnz].copy()


class Index:
    """NNDescentBuilder::build_with_distance + NNDescentIndex::query (core/index.rs)."""

    def __init__(self, data, metric="sqeuclidean", correction="none", n_neighbors=30, n_trees=8,
                 n_search_trees=1, search_tree_leaf_budget=1, leaf_size=None, max_candidates=None, n_iters=None,
                 delta=0.001, diversify_prob=1.0, pruning_degree_multiplier=1.5, angular=False, seed=42, threads=1):
        self.data = np.ascontiguousarray(data, np.float32)
        self.n, self.dim = self.data.shape
        self.k = n_neighbors
        self.threads = threads
        self.h = lib().orc_index_build(ptr(self.data), self.n, self.dim, METRICS[metric], CORR[correction],
                                       n_neighbors, n_trees, n_search_trees, search_tree_leaf_budget,
                                       -1 if leaf_size is None else leaf_size,
                                       -1 if max_candidates is None else max_candidates,
                                       -1 if n_iters is None else n_iters, C.c_float(delta),
                                       C.c_float(diversify_prob), C.c_float(pruning_degree_multiplier), int(angular),
                                       C.c_uint64(seed), threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_index_free(self.h)
            self.h = None

    def quantize_trees(self):
        """tree_quantization = "int8": the search descends int8 copies of the search trees (flat_tree.rs:35-183)."""
        lib().orc_index_quantize_trees(self.h)

    def query(self, q, k=10, epsilon=0.1, threads=None):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        oi = np.zeros((nq, k), np.int32)
        od = np.zeros((nq, k), np.float32)
        lib().orc_index_query(self.h, ptr(q), nq, k, C.c_float(epsilon), threads or self.threads, ptr(oi), ptr(od))
        return oi, od

    def encode_i8(self, encoding):
        """The quantized index's codes of this index's rows (crates/rinnd/src/lib.rs:158-224); encoding "int8",
        "sq8u" or "sq8p"."""
        return I8Codes(lib().orc_index_encode_i8(self.h, I8_ENCODINGS[encoding]), self.n, self.dim)

    def query_i8(self, codes, q, k=10, epsilon=0.1, candidate_width=0, rerank_width=0, threads=None):
        """query_one_quantized_i8_with_workspace (core/index.rs:283-330) for query ids 0..nq-1."""
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        oi = np.zeros((nq, k), np.int32)
        od = np.zeros((nq, k), np.float32)
        lib().orc_index_query_i8(self.h, codes.h, ptr(q), nq, k, candidate_width, rerank_width, C.c_float(epsilon),
                                 threads or self.threads, ptr(oi), ptr(od))
        return oi, od

    def query_one(self, q, k=10, epsilon=0.1):
        q = np.ascontiguousarray(q, np.float32)
        oi = np.zeros(k, np.int32)
        od = np.zeros(k, np.float32)
        lib().orc_index_query_one(self.h, ptr(q), k, C.c_float(epsilon), ptr(oi), ptr(od))
        return oi, od

    @property
    def neighbor_graph(self):
        oi = np.zeros((self.n, self.k), np.int32)
        od = np.zeros((self.n, self.k), np.float32)
        lib().orc_index_neighbor_graph(self.h, ptr(oi), ptr(od))
        return oi, od

    def internal_search_graph(self):
        nnz = int(lib().orc_index_search_graph_nnz(self.h))
        indptr = np.zeros(self.n + 1, np.int32)
        indices = np.zeros(max(nnz, 1), np.int32)
        lib().orc_index_search_graph(self.h, ptr(indptr), ptr(indices))
        return indptr, indices[:nnz], float(lib().orc_index_min_distance(self.h))

    @property
    def vertex_order(self):
        out = np.zeros(self.n, np.int64)
        lib().orc_index_vertex_order(self.h, ptr(out))
        return out

    def forest(self):
        return Forest(lib().orc_index_forest(self.h))

    def trace(self):
        ch = np.zeros(64, np.int64)
        it = np.zeros(1, np.int64)
        ms = np.zeros(4, np.float64)
        lib().orc_index_trace(self.h, ptr(ch), ptr(it), ptr(ms))
        return ch[: int(it[0])].copy(), ms


# ---- int8 routing trees (flat_tree.rs:35-183) -------------------------------------------------------
def quantized_margin(hyperplane, offset, query_codes, query_scale):
    """One node's int8 plane and scale (QuantizedFlatTree::from_flat_tree) and its reconstructed margin for an int8
    query: (codes, scale, margin)."""
    h = np.ascontiguousarray(hyperplane, np.float32)
    qc = np.ascontiguousarray(query_codes, np.int8)
    codes = np.zeros(len(h), np.int8)
    sc = np.zeros(1, np.float32)
    m = lib().orc_quantized_margin(ptr(h), C.c_float(offset), len(h), ptr(qc), C.c_float(query_scale), ptr(codes),
                                   ptr(sc))
    return codes, float(sc[0]), float(m)


I8_ENCODINGS = {"int8": 0, "sq8u": 1, "sq8p": 2}


class I8Codes:
    """Codes of the quantized index (QuantizedI8Index, crates/rinnd/src/lib.rs:133-260)."""

    def __init__(self, h, n, dim):
        self.h, self.n, self.dim = h, n, dim

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_i8_free(self.h)
            self.h = None

    def arrays(self):
        codes = np.zeros((self.n, self.dim), np.int8)
        inv = np.zeros(self.n, np.float32)
        g = np.zeros(2, np.float32)
        lib().orc_i8_codes(self.h, ptr(codes), ptr(inv), ptr(g))
        return codes, inv, float(g[0]), float(g[1])

    def quantize_query(self, q):
        """quantize_query (lib.rs:325-358): the query's codes and its inverse norm / dequantisation scale."""
        q = np.ascontiguousarray(q, np.float32)
        codes = np.zeros(self.dim, np.int8)
        s = lib().orc_quantize_query_codes(self.h, ptr(q), ptr(codes))
        return codes, float(np.float32(s))


def encode_i8_rows(X, encoding):
    X = np.ascontiguousarray(X, np.float32)
    return I8Codes(lib().orc_encode_i8_rows(ptr(X), X.shape[0], X.shape[1], I8_ENCODINGS[encoding]), X.shape[0],
                   X.shape[1])


def i8_cosine_distance(q_codes, q_inv, x_codes, x_inv):
    """greedy.rs:381-387."""
    a = np.ascontiguousarray(q_codes, np.int8)
    b = np.ascontiguousarray(x_codes, np.int8)
    return float(np.float32(lib().orc_i8_cosine_distance(ptr(a), C.c_float(q_inv), ptr(b), C.c_float(x_inv),
                                                         len(a))))


def quantize_query_i8(q):
    """The per-vector symmetric int8 query (crates/rinnd/src/lib.rs:303-347): (codes, scale)."""
    x = np.ascontiguousarray(q, np.float32)
    codes = np.zeros(len(x), np.int8)
    s = lib().orc_quantize_query_i8(ptr(x), len(x), ptr(codes))
    return codes, float(s)

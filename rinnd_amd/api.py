"""Python mirror of the reference's `RINND` class (crates/rinnd/src/lib.rs:615-1128).

Same constructor signature, defaults, metric mapping and error texts as the PyO3 class; the build
and the queries run in librinnd_hip.so on the GPU.  Inputs may be numpy arrays (copied to the
device) or torch tensors already resident on the device.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time

import numpy as np

from . import _lib as L

# Metric::from_str (core/distance/mod.rs:211-246): every name the reference parses
_METRIC_NAMES = {
    "euclidean": "euclidean", "l2": "euclidean", "sqeuclidean": "sqeuclidean",
    "squared_euclidean": "sqeuclidean", "manhattan": "manhattan", "taxicab": "manhattan",
    "l1": "manhattan", "chebyshev": "chebyshev", "linfinity": "chebyshev", "linfty": "chebyshev",
    "linf": "chebyshev", "canberra": "canberra", "braycurtis": "braycurtis", "cosine": "cosine",
    "inner_product": "inner_product", "ip": "inner_product", "dot": "dot", "correlation": "correlation",
    "true_angular": "true_angular", "tsss": "tsss", "hamming": "hamming", "jaccard": "jaccard",
    "dice": "dice", "matching": "matching", "kulsinski": "kulsinski", "rogerstanimoto": "rogerstanimoto",
    "russellrao": "russellrao", "sokalsneath": "sokalsneath", "sokalmichener": "sokalmichener",
    "yule": "yule", "hellinger": "hellinger", "jensen_shannon": "jensen_shannon",
    "jensen-shannon": "jensen_shannon", "symmetric_kl": "symmetric_kl", "symmetric-kl": "symmetric_kl",
    "symmetric_kullback_liebler": "symmetric_kl",
}
# metrics with angular RP trees (core/index.rs:606-614)
_ANGULAR = {"cosine", "inner_product", "dot", "correlation", "true_angular", "tsss"}
# QuantizedI8Encoding (crates/rinnd/src/lib.rs:151-156) and its storage_info names (lib.rs:475-479)
_I8_ENCODINGS = {"int8": 0, "sq8u": 1, "sq8p": 2}
_I8_ENCODING_NAMES = {"int8": "int8_per_vector_cosine", "sq8u": "sq8u_global_symmetric",
                      "sq8p": "sq8p_per_vector_symmetric"}


def version() -> str:
    """Library version (py: `version()`)."""
    return L.lib().rinnd_version().decode()


def simd_info() -> str:
    """Describes the compute backend (py: `simd_info()`, crates/rinnd/src/lib.rs:1307-1336)."""
    return "MI355X gfx950 HIP kernels: wave64 FP32 reference-order distances, BF16 MFMA join"


def normalize_rows(x: np.ndarray) -> np.ndarray:
    """normalize_rows_inplace (crates/rinnd/src/lib.rs:1131-1146) in float32, sequential sums."""
    x = np.array(x, dtype=np.float32, copy=True, order="C")
    if x.shape[1] == 0:
        return x
    norm_sq = np.cumsum(x * x, axis=1, dtype=np.float32)[:, -1]
    ok = norm_sq > np.float32(1e-12)
    inv = np.float32(1.0) / np.sqrt(norm_sq[ok])
    x[ok] *= inv[:, None]
    return x


def _is_device_tensor(a) -> bool:
    return hasattr(a, "is_cuda") and bool(getattr(a, "is_cuda"))


def _as_input(a, name: str):
    """2-D float32 C-contiguous host array or device tensor."""
    if _is_device_tensor(a):
        import torch
        if a.dim() != 2:
            raise ValueError(f"{name} must be a 2-D array")
        return a.to(torch.float32).contiguous()
    arr = np.ascontiguousarray(a, dtype=np.float32)
    if arr.ndim != 2:
        raise ValueError(f"{name} must be a 2-D array")
    return arr


def _round_half_away(x: float) -> int:
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def default_threads() -> int:
    """rayon::current_num_threads(): RAYON_NUM_THREADS or the number of logical CPUs."""
    v = os.environ.get("RAYON_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return int(v)
    return os.cpu_count() or 1


class RINND:
    """NNDescent index for approximate nearest neighbour search (py:615-755).

    MI355X-specific keywords (not in the reference signature): ``n_threads`` (the Rayon thread count
    whose candidate-sampling streams are reproduced), ``rng_mode`` ("reference" or "counter"),
    ``join_precision`` ("fp32" or "bf16") and ``device``.
    """

    def __init__(self, data, metric="euclidean", n_neighbors=30, n_trees=None, n_search_trees=1,
                 search_tree_leaf_budget=1, leaf_size=None, max_candidates=None, n_iters=None, delta=0.001,
                 random_state=None, diversify_prob=1.0, pruning_degree_multiplier=1.5, verbose=False,
                 normalize=False, cosine_distance_mode="log", quantized_candidate_width=0,
                 quantized_rerank_width=0, tree_quantization="none", retain_fp32_data=None, *,
                 n_threads=None, rng_mode="reference", join_precision="fp32", device=0):
        if quantized_candidate_width > 0 and quantized_rerank_width > quantized_candidate_width:
            raise ValueError("quantized_rerank_width must not exceed quantized_candidate_width when candidate "
                             "width is explicit")
        x = _as_input(data, "data")
        n_points, dim = int(x.shape[0]), int(x.shape[1])
        key = str(metric).lower()
        if key not in _METRIC_NAMES:
            raise ValueError(f"Unknown metric: {metric}")
        m = _METRIC_NAMES[key]
        normalize_cosine = bool(normalize) and m == "cosine"
        if cosine_distance_mode == "log":
            direct, quantized = False, None
        elif cosine_distance_mode == "direct":
            direct, quantized = True, None
        elif cosine_distance_mode in ("int8", "sq8u", "sq8p"):
            direct, quantized = True, cosine_distance_mode
        else:
            raise ValueError(f"Unknown cosine_distance_mode: {cosine_distance_mode}; expected 'log', 'direct', "
                             "'int8', 'sq8u', or 'sq8p'")
        if tree_quantization not in ("none", "int8"):
            raise ValueError(f"Unknown tree_quantization: {tree_quantization}; expected 'none' or 'int8'")
        if tree_quantization == "int8" and quantized is None:
            raise ValueError("tree_quantization='int8' is available only with a quantized cosine_distance_mode")
        if tree_quantization == "int8" and quantized == "sq8u":
            raise ValueError("tree_quantization='int8' is incompatible with cosine_distance_mode='sq8u'; use "
                             "tree_quantization='none'")
        if direct and not normalize_cosine:
            raise ValueError("direct and quantized cosine modes require metric='cosine' and normalize=True")
        # metric -> (distance functor, correction) dispatch (py:1239-1275)
        if m == "euclidean":
            dist, corr = "sqeuclidean", "sqrt"
        elif m == "sqeuclidean":
            dist, corr = "sqeuclidean", "none"
        elif m == "cosine":
            if normalize_cosine:
                dist, corr = ("direct_cosine", "none") if direct else ("alt_dot", "alt_cosine")
            else:
                dist, corr = "alt_cosine", "alt_cosine"
        elif m == "dot":
            dist, corr = "alt_dot", "alt_cosine"
        elif m == "inner_product":
            dist, corr = "alt_inner_product", "alt_inner_product"
        else:
            raise NotImplementedError(f"metric {metric!r} is not available on the MI355X path")
        if normalize_cosine:
            if _is_device_tensor(x):
                x = x.cpu().numpy()
            x = normalize_rows(x)
        if n_trees is None:  # py:1200-1204
            n_trees = min(12, max(3, _round_half_away(2.0 * math.log10(max(n_points, 1)))))
        self.n_points, self.dim = n_points, dim
        # quantized cosine modes (py:714-731): the candidate search runs on the int8 codes of the rows
        # (int8 / sq8u / sq8p) with these default widths; FP32 rows are kept for the exact re-ranking unless
        # released
        self._quantized = quantized
        self._qcw, self._qrw = int(quantized_candidate_width), int(quantized_rerank_width)
        self.n_neighbors = int(n_neighbors)
        self.normalize = normalize_cosine
        self.metric = m
        self._dist, self._corr = dist, corr
        self._ctx = L.Context.default(device)
        p = L.IndexParams(
            int(n_neighbors), int(n_trees), int(n_search_trees), max(1, int(search_tree_leaf_budget)),
            -1 if leaf_size is None else int(leaf_size), -1 if max_candidates is None else int(max_candidates),
            -1 if n_iters is None else int(n_iters), float(delta), float(diversify_prob),
            float(pruning_degree_multiplier), int(m in _ANGULAR), int(42 if random_state is None else random_state),
            int(bool(verbose)), int(n_threads or default_threads()), L.RNG[rng_mode], L.JOIN[join_precision])
        self.build_stats = L.BuildStats()
        h = C.c_void_p()
        st = L.lib().rinnd_index_build(self._ctx.h, L.ptr(x), n_points, dim, L.DIST[dist], L.CORR[corr],
                                       C.byref(p), C.byref(h), C.byref(self.build_stats))
        self._ctx.check(st, "rinnd_index_build")
        self._h = h
        # search-tree node counts: the storage accounting of the trees (lib.rs:498-541, 228-240)
        forest = L.Forest(L.lib().rinnd_index_trees(h), owned=False)
        self._tree_nodes = [forest.n_nodes(t) for t in range(forest.n_trees)]
        self._tree_i8 = tree_quantization == "int8"
        self._released = False
        self._encode_seconds = 0.0
        if quantized is not None:  # QuantizedI8Index::new (lib.rs:158-276)
            t0 = time.perf_counter()
            self._ctx.check(L.lib().rinnd_index_encode_i8(h, _I8_ENCODINGS[quantized]), "rinnd_index_encode_i8")
            if self._tree_i8:  # lib.rs:242-248: int8 routing trees replace the FP32 search trees
                self._ctx.check(L.lib().rinnd_index_quantize_trees(h), "rinnd_index_quantize_trees")
            # lib.rs:714-718: by default FP32 rows go only with int8 trees and no constructor rerank width
            retain = (not self._tree_i8 or self._qrw > 0) if retain_fp32_data is None else bool(retain_fp32_data)
            if not retain:
                self._ctx.check(L.lib().rinnd_index_release_data(h), "rinnd_index_release_data")
                self._released = True
            self._encode_seconds = time.perf_counter() - t0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and L._lib is not None:
            L._lib.rinnd_index_free(h)
            self._h = None

    # ---- queries (py:775-821, 877-951) ----
    def _prep_queries(self, q, name="query_data"):
        q = _as_input(q, name)
        if int(q.shape[1]) != self.dim:
            raise ValueError(f"Query dimension {int(q.shape[1])} does not match data dimension {self.dim}")
        if self.normalize:
            if _is_device_tensor(q):
                q = q.cpu().numpy()
            q = normalize_rows(q)
        return q

    def _outputs(self, q, nq, k):
        if _is_device_tensor(q):
            import torch
            return (torch.empty((nq, k), dtype=torch.int32, device=q.device),
                    torch.empty((nq, k), dtype=torch.float32, device=q.device))
        return np.empty((nq, k), np.int32), np.empty((nq, k), np.float32)

    def _query_i8(self, q, k, epsilon, candidate_width, rerank_width):
        nq = int(q.shape[0])
        oi, od = self._outputs(q, nq, k)
        st = L.lib().rinnd_index_query_i8(self._h, L.ptr(q), nq, int(k), C.c_float(epsilon),
                                          int(candidate_width), int(rerank_width), L.ptr(oi), L.ptr(od))
        self._ctx.check(st, "rinnd_index_query_i8")
        return oi, od

    def query(self, query_data, k=10, epsilon=0.1):
        q = self._prep_queries(query_data)
        if self._quantized is not None:  # query_one_quantized with the constructor's widths
            return self._query_i8(q, k, epsilon, self._qcw, self._qrw)
        nq = int(q.shape[0])
        oi, od = self._outputs(q, nq, k)
        st = L.lib().rinnd_index_query(self._h, L.ptr(q), nq, int(k), C.c_float(epsilon), L.ptr(oi), L.ptr(od))
        self._ctx.check(st, "rinnd_index_query")
        return oi, od

    @staticmethod
    def _validate_widths(k, candidate_width, rerank_width):
        """validate_quantized_widths (py:1148-1172)."""
        if candidate_width != 0:
            if candidate_width < k:
                raise ValueError(f"candidate_width {candidate_width} must be at least k ({k})")
            if rerank_width > candidate_width:
                raise ValueError(f"rerank_width {rerank_width} must not exceed candidate_width {candidate_width}")
            if 0 < rerank_width < k:
                raise ValueError(f"rerank_width {rerank_width} must be zero or at least k ({k})")

    def query_quantized(self, query_data, k=10, epsilon=0.1, candidate_width=0, rerank_width=0):
        """Quantized query with explicit widths (py:825-870)."""
        self._validate_widths(k, candidate_width, rerank_width)
        q = self._prep_queries(query_data)
        if self._quantized is None:
            raise ValueError("query-time quantized widths require a quantized cosine_distance_mode")
        return self._query_i8(q, k, epsilon, candidate_width, rerank_width)

    def query_one(self, query, k=10, epsilon=0.1):
        q = np.ascontiguousarray(query, np.float32)
        if q.ndim != 1:
            raise ValueError("query must be a 1-D array")
        if q.shape[0] != self.dim:
            raise ValueError(f"Query dimension {q.shape[0]} does not match data dimension {self.dim}")
        q2 = self._prep_queries(q[None, :], "query")
        if self._quantized is not None:
            oi, od = self._query_i8(q2, k, epsilon, self._qcw, self._qrw)
            return oi[0], od[0]
        oi = np.empty(k, np.int32)
        od = np.empty(k, np.float32)
        st = L.lib().rinnd_index_query_one(self._h, L.ptr(q2), int(k), C.c_float(epsilon), L.ptr(oi), L.ptr(od))
        self._ctx.check(st, "rinnd_index_query_one")
        return oi, od

    def query_one_indices(self, query, k=10, epsilon=0.1):
        return self.query_one(query, k, epsilon)[0]

    def query_one_indices_quantized(self, query, k=10, epsilon=0.1, candidate_width=0, rerank_width=0):
        """py:960-1011: the indices of one quantized query with explicit widths."""
        self._validate_widths(k, candidate_width, rerank_width)
        q = np.ascontiguousarray(query, np.float32)
        if q.ndim != 1:
            raise ValueError("query must be a 1-D array")
        if q.shape[0] != self.dim:
            raise ValueError(f"Query dimension {q.shape[0]} does not match data dimension {self.dim}")
        if self._quantized is None:
            raise ValueError("query-time quantized widths require a quantized cosine_distance_mode")
        q2 = self._prep_queries(q[None, :], "query")
        return self._query_i8(q2, k, epsilon, candidate_width, rerank_width)[0][0]

    # ---- graph / tree exports (py:1013-1025, 1087-1128) ----
    @property
    def neighbor_graph(self):
        oi = np.empty((self.n_points, self.n_neighbors), np.int32)
        od = np.empty((self.n_points, self.n_neighbors), np.float32)
        self._ctx.check(L.lib().rinnd_index_neighbor_graph(self._h, L.ptr(oi), L.ptr(od)),
                        "rinnd_index_neighbor_graph")
        return oi, od

    @property
    def search_graph(self):
        n = np.zeros(1, np.int64)
        nnz = np.zeros(1, np.int64)
        md = np.zeros(1, np.float32)
        L.lib().rinnd_index_search_graph_info(self._h, L.ptr(n), L.ptr(nnz), L.ptr(md))
        indptr = np.empty(int(n[0]) + 1, np.int32)
        indices = np.empty(max(int(nnz[0]), 1), np.int32)
        self._ctx.check(L.lib().rinnd_index_search_graph(self._h, L.ptr(indptr), L.ptr(indices)),
                        "rinnd_index_search_graph")
        return indptr, indices[: int(nnz[0])], float(md[0])

    @property
    def vertex_order(self):
        out = np.empty(self.n_points, np.int64)
        self._ctx.check(L.lib().rinnd_index_vertex_order(self._h, L.ptr(out)), "rinnd_index_vertex_order")
        return out

    def export_search_tree(self):
        f = L.lib().rinnd_index_trees(self._h)
        forest = L.Forest(f, owned=False)
        if not f or forest.n_trees == 0:
            raise ValueError("no retained search tree is available")
        t = forest.tree(0, self.n_points)
        order = self.vertex_order
        return t["hyperplanes"], t["offsets"], t["children"], order[t["indices"]].astype(np.int32)

    def storage_info(self):
        """py:1028-1077: the FP32 representation, or the int8 codes of a quantized index (one byte a value,
        one FP32 inverse norm or scale a row), its retained FP32 arrays and its search trees."""
        data_bytes = self.n_points * self.dim * 4
        if self._quantized is not None:
            code_bytes = self.n_points * self.dim
            meta = self.n_points * 4
            # search trees: children 8 bytes and offsets 4 bytes a node, hyperplanes dim values a node, leaf
            # indices 4 bytes a point (lib.rs:498-541, 228-240)
            nodes = sum(self._tree_nodes)
            topology = nodes * 8 + len(self._tree_nodes) * self.n_points * 4
            # retained_fp32_bytes (core/index.rs:92-101): rows, neighbour distances and FP32 tree planes and offsets
            retained = ((0 if self._released else data_bytes) + self.n_points * self.n_neighbors * 4
                        + (0 if self._tree_i8 else (nodes * self.dim + nodes) * 4))
            if self._tree_i8:
                trees = {"tree_code_bytes": nodes * self.dim, "tree_scale_bytes": nodes * 4,
                         "tree_offset_bytes": nodes * 4, "tree_topology_bytes": topology,
                         "hypothetical_releasable_fp32_tree_bytes": 0,
                         "released_fp32_tree_bytes": nodes * self.dim * 4 + nodes * 4 + topology,
                         "retained_fp32_tree_topology_bytes": 0}
            else:
                trees = {"tree_code_bytes": 0, "tree_scale_bytes": 0, "tree_offset_bytes": 0,
                         "tree_topology_bytes": 0, "hypothetical_releasable_fp32_tree_bytes": 0,
                         "released_fp32_tree_bytes": 0, "retained_fp32_tree_topology_bytes": topology}
            allocated = (code_bytes + meta + retained + trees["retained_fp32_tree_topology_bytes"]
                         + trees["tree_code_bytes"] + trees["tree_scale_bytes"] + trees["tree_offset_bytes"]
                         + trees["tree_topology_bytes"])
            return {
                "encoding": _I8_ENCODING_NAMES[self._quantized], "code_bytes": code_bytes, "metadata_bytes": meta,
                "retained_fp32_bytes": retained, "encode_seconds": self._encode_seconds, **trees,
                "released_fp32_data_bytes": data_bytes if self._released else 0,
                "allocated_bytes": allocated, "bytes_per_vector": float(allocated) / max(self.n_points, 1),
            }
        return {
            "encoding": "fp32", "code_bytes": data_bytes, "metadata_bytes": 0, "retained_fp32_bytes": 0,
            "encode_seconds": 0.0, "tree_code_bytes": 0, "tree_scale_bytes": 0, "tree_offset_bytes": 0,
            "tree_topology_bytes": 0, "hypothetical_releasable_fp32_tree_bytes": 0, "released_fp32_tree_bytes": 0,
            "released_fp32_data_bytes": 0, "retained_fp32_tree_topology_bytes": 0, "allocated_bytes": data_bytes,
            "bytes_per_vector": float(data_bytes) / max(self.n_points, 1),
        }

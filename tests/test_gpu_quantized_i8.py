"""GPU parity of the quantized cosine index (cosine_distance_mode "int8" / "sq8u" / "sq8p": QuantizedI8Index,
crates/rinnd/src/lib.rs:133-398; core/index.rs:55-126, 283-330; core/search/greedy.rs:101-146, 365-408) against
the oracle.

The int8 codes, the query's codes, the candidate distances (exact integer dot products, then the reference's
three float operations), the search and the FP32 re-ranking are compared bit-exactly with the oracle's Index on
the same normalized rows, for every encoding and width rule, with FP32 and with int8 routing trees.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import RINND
from rinnd_amd import _lib as L
from rinnd_amd.api import normalize_rows

pytestmark = pytest.mark.gpu

KW = dict(n_neighbors=12, n_trees=4, n_search_trees=2, n_iters=5, random_state=17)


def data(n, d, seed):
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((24, d)).astype(np.float32)
    X = centres[rng.integers(0, 24, n)] + 0.6 * rng.standard_normal((n, d)).astype(np.float32)
    return X.astype(np.float32)


def pair(X, mode, threads=4, **kw):
    p = dict(KW)
    p.update(kw)
    g = RINND(X, metric="cosine", normalize=True, cosine_distance_mode=mode, n_threads=threads, **p)
    r = O.Index(normalize_rows(X), metric="direct_cosine", correction="none", n_neighbors=p["n_neighbors"],
                n_trees=p["n_trees"], n_iters=p["n_iters"], angular=True, seed=p["random_state"], threads=threads,
                n_search_trees=p["n_search_trees"], search_tree_leaf_budget=p.get("search_tree_leaf_budget", 1))
    gi, gd = g.neighbor_graph
    ri, rd = r.neighbor_graph
    np.testing.assert_array_equal(gi, ri)  # the same index underneath
    return g, r


def same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@pytest.fixture(scope="module", params=["int8", "sq8u", "sq8p"])
def indexed(request):
    X = data(5000, 40, 3)
    g, r = pair(X, request.param)
    codes = r.encode_i8(request.param)
    Q = data(300, 40, 4)  # the index normalizes queries itself (lib.rs:775-821); the oracle takes them normalized
    return request.param, g, r, codes, Q


@pytest.mark.parametrize("k,eps,cw,rw", [(10, 0.1, 0, 0), (10, 0.1, 10, 0), (10, 0.1, 30, 0), (10, 0.2, 30, 20),
                                         (10, 0.1, 0, 16), (6, 0.0, 0, 6), (15, 0.05, 40, 40), (1, 0.1, 8, 8)])
def test_query_i8_matches_oracle(indexed, k, eps, cw, rw):
    mode, g, r, codes, Q = indexed
    same(g.query_quantized(Q, k=k, epsilon=eps, candidate_width=cw, rerank_width=rw),
         r.query_i8(codes, normalize_rows(Q), k=k, epsilon=eps, candidate_width=cw, rerank_width=rw))


def test_constructor_widths_and_single_queries(indexed):
    """query() takes the constructor widths (0, 0 here); query_one is the same search (lib.rs:1774-1793)."""
    mode, g, r, codes, Q = indexed
    want = r.query_i8(codes, normalize_rows(Q), k=10, epsilon=0.1)
    same(g.query(Q, k=10, epsilon=0.1), want)
    for j in (0, 7, 299):
        i, d = g.query_one(Q[j], k=10, epsilon=0.1)
        np.testing.assert_array_equal(i, want[0][j])
        np.testing.assert_array_equal(d.view(np.uint32), want[1][j].view(np.uint32))


def test_query_one_indices_quantized(indexed):
    """py:960-1011: the indices of query_one_quantized_widths for one query; width errors as query_quantized."""
    mode, g, r, codes, Q = indexed
    want = r.query_i8(codes, normalize_rows(Q[:3]), k=6, epsilon=0.1, candidate_width=18, rerank_width=12)[0]
    for j in range(3):
        np.testing.assert_array_equal(g.query_one_indices_quantized(Q[j], k=6, epsilon=0.1, candidate_width=18,
                                                                    rerank_width=12), want[j])
    with pytest.raises(ValueError, match="must be at least k"):
        g.query_one_indices_quantized(Q[0], k=6, candidate_width=3)
    with pytest.raises(ValueError, match="does not match data dimension"):
        g.query_one_indices_quantized(Q[0, :5], k=6)
    plain = RINND(data(500, 40, 1), n_neighbors=8, n_trees=2, n_iters=2)
    with pytest.raises(ValueError, match="quantized cosine_distance_mode"):
        plain.query_one_indices_quantized(Q[0], k=5)


@pytest.mark.parametrize("mode", ["int8", "sq8p"])
def test_int8_trees_with_int8_codes(mode):
    X = data(4000, 32, 5)
    g, r = pair(X, mode, tree_quantization="int8", retain_fp32_data=True, search_tree_leaf_budget=3)
    r.quantize_trees()
    codes = r.encode_i8(mode)
    Q = data(200, 32, 6)
    for cw, rw in ((0, 0), (24, 12)):
        same(g.query_quantized(Q, k=8, epsilon=0.1, candidate_width=cw, rerank_width=rw),
             r.query_i8(codes, normalize_rows(Q), k=8, epsilon=0.1, candidate_width=cw, rerank_width=rw))


def test_released_rows_leave_the_int8_search_alone():
    """lib.rs:714-718 and 1873-1898: int8 trees with no rerank width release the FP32 rows by default; the
    int8 search still runs, exact re-ranking is refused."""
    X = data(3000, 24, 7)
    g, r = pair(X, "int8", tree_quantization="int8")
    r.quantize_trees()
    codes = r.encode_i8("int8")
    Q = data(100, 24, 8)
    Qn = normalize_rows(Q)
    same(g.query(Q, k=6, epsilon=0.08), r.query_i8(codes, Qn, k=6, epsilon=0.08))
    same(g.query_quantized(Q, k=6, epsilon=0.08, candidate_width=18),
         r.query_i8(codes, Qn, k=6, epsilon=0.08, candidate_width=18))
    with pytest.raises(ValueError, match="FP32 vector data was released"):
        g.query_quantized(Q, k=6, epsilon=0.08, candidate_width=18, rerank_width=12)
    oi = np.zeros((2, 6), np.int32)
    od = np.zeros((2, 6), np.float32)
    q2 = np.ascontiguousarray(Qn[:2])
    assert L.lib().rinnd_index_query(g._h, L.ptr(q2), 2, 6, C.c_float(0.1), L.ptr(oi), L.ptr(od)) == 1
    info = g.storage_info()
    assert info["released_fp32_data_bytes"] == 3000 * 24 * 4
    assert info["retained_fp32_bytes"] == 3000 * KW["n_neighbors"] * 4  # neighbour distances only
    kept = RINND(X, metric="cosine", normalize=True, cosine_distance_mode="int8", tree_quantization="int8",
                 quantized_rerank_width=12, **KW)
    assert kept.storage_info()["released_fp32_data_bytes"] == 0


def test_storage_info_names_and_sizes():
    X = data(1000, 20, 9)
    for mode, name in (("int8", "int8_per_vector_cosine"), ("sq8u", "sq8u_global_symmetric"),
                       ("sq8p", "sq8p_per_vector_symmetric")):
        g = RINND(X, metric="cosine", normalize=True, cosine_distance_mode=mode, **KW)
        s = g.storage_info()
        assert s["encoding"] == name
        assert s["code_bytes"] == 1000 * 20 and s["metadata_bytes"] == 1000 * 4
        assert s["encode_seconds"] > 0.0
        nodes = sum(g._tree_nodes)
        assert s["retained_fp32_bytes"] == (1000 * 20 + 1000 * KW["n_neighbors"] + nodes * 20 + nodes) * 4


def test_int8_recall_at_default_widths():
    """The int8 candidate distances keep 7 bits a coordinate: at the default widths (a width-k search, no
    re-ranking) recall@10 stays close to the FP32 search's on clustered data."""
    rng = np.random.default_rng(3)
    centres = rng.random((40, 64)).astype(np.float32)
    X = (centres[rng.integers(0, 40, 6000)] + 0.08 * rng.standard_normal((6000, 64))).astype(np.float32)
    Q = (centres[rng.integers(0, 40, 300)] + 0.08 * rng.standard_normal((300, 64))).astype(np.float32)
    kw = dict(metric="cosine", normalize=True, n_neighbors=15, random_state=7)
    q8 = RINND(X, cosine_distance_mode="int8", **kw)
    f32 = RINND(X, cosine_distance_mode="direct", **kw)
    Xn, Qn = normalize_rows(X), normalize_rows(Q)
    truth = np.argsort(-(Qn @ Xn.T), axis=1)[:, :10]
    rec = lambda I: np.mean([len(set(I[i]) & set(truth[i])) / 10 for i in range(len(Q))])
    r8, r32 = rec(q8.query(Q, k=10, epsilon=0.1)[0]), rec(f32.query(Q, k=10, epsilon=0.1)[0])
    assert r8 >= r32 - 0.1 and r8 >= 0.8, (r8, r32)

"""GPU parity of build_with_distance + query (core/index.rs) against the oracle, through RINND.

The k-NN graph, the diversified search graph, the vertex order and the query results are compared
bit-exactly on the FP32 squared-Euclidean path (cfg-1 style data, euclidean => sqrt correction).
"""
import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import RINND

pytestmark = pytest.mark.gpu


def uniform(n, d, seed=0):
    return np.random.default_rng(seed).random((n, d), dtype=np.float32)


def make_pair(X, metric="euclidean", threads=4, **kw):
    g = RINND(X, metric=metric, n_threads=threads, **kw)
    dist, corr, ang = {
        "euclidean": ("sqeuclidean", "sqrt", False), "sqeuclidean": ("sqeuclidean", "none", False),
        "inner_product": ("alt_inner_product", "alt_inner_product", True),
    }[metric]
    n_trees = kw.get("n_trees")
    if n_trees is None:
        import math
        n_trees = min(12, max(3, int(math.floor(2 * math.log10(len(X)) + 0.5))))
    r = O.Index(X, metric=dist, correction=corr, n_neighbors=kw.get("n_neighbors", 30), n_trees=n_trees,
                leaf_size=kw.get("leaf_size"), max_candidates=kw.get("max_candidates"), n_iters=kw.get("n_iters"),
                delta=kw.get("delta", 0.001), angular=ang, seed=kw.get("random_state", 42), threads=threads,
                n_search_trees=kw.get("n_search_trees", 1),
                search_tree_leaf_budget=kw.get("search_tree_leaf_budget", 1))
    return g, r


@pytest.mark.parametrize("threads", [1, 8])
def test_index_matches_reference(threads):
    X = uniform(3000, 32, seed=11)
    g, r = make_pair(X, n_neighbors=10, n_iters=6, threads=threads)
    gi, gd = g.neighbor_graph
    ri, rd = r.neighbor_graph
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_array_equal(gd.view(np.uint32), rd.view(np.uint32))
    np.testing.assert_array_equal(g.vertex_order, r.vertex_order)
    # search graph in original order == oracle's internal graph mapped back
    indptr, indices, md = g.search_graph
    rp, rc, rmd = r.internal_search_graph()
    assert md == rmd
    vo = r.vertex_order
    inv = np.empty_like(vo)
    inv[vo] = np.arange(len(vo))
    # search_graph_original_order (core/index.rs:137-141): every row, in the reference's entry order
    want_ptr = np.zeros(len(X) + 1, np.int64)
    want = []
    for o in range(len(X)):
        ir = inv[o]
        row = vo[rc[rp[ir]:rp[ir + 1]]]
        want.append(row)
        want_ptr[o + 1] = want_ptr[o] + len(row)
    np.testing.assert_array_equal(indptr, want_ptr)
    np.testing.assert_array_equal(indices, np.concatenate(want))
    # export_search_tree (core/index.rs:143-154): the first retained tree with leaf ids in input order
    hp, off, ch, leaf_ids = g.export_search_tree()
    t = r.forest().tree(0)
    np.testing.assert_array_equal(hp.view(np.uint32), r.forest().hyperplanes(0, X.shape[1]).view(np.uint32))
    np.testing.assert_array_equal(off.view(np.uint32), t["offsets"].view(np.uint32))
    np.testing.assert_array_equal(ch, t["children"])
    np.testing.assert_array_equal(leaf_ids, vo[t["indices"]])
    assert sorted(leaf_ids.tolist()) == list(range(len(X)))
    Q = uniform(200, 32, seed=12)
    qi, qd = g.query(Q, k=10, epsilon=0.1)
    oi, od = r.query(Q, k=10, epsilon=0.1)
    np.testing.assert_array_equal(qi, oi)
    np.testing.assert_array_equal(qd.view(np.uint32), od.view(np.uint32))


def test_cfg1_queries_bit_exact():
    """10,000 x 128 U[0,1), k=15, leaf 60, <=10 iterations; 100 queries k=10 eps=0.1."""
    X = uniform(10000, 128, seed=0)
    Q = uniform(100, 128, seed=1)
    g, r = make_pair(X, n_neighbors=15, leaf_size=60, n_iters=10, n_trees=8, threads=8)
    gi, gd = g.neighbor_graph
    ri, rd = r.neighbor_graph
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_array_equal(gd.view(np.uint32), rd.view(np.uint32))
    qi, qd = g.query(Q, k=10, epsilon=0.1)
    oi, od = r.query(Q, k=10, epsilon=0.1)
    np.testing.assert_array_equal(qi, oi)
    np.testing.assert_array_equal(qd.view(np.uint32), od.view(np.uint32))


def test_single_query_matches_batch():
    """test_rinnd.py:27-36."""
    X = np.arange(32 * 4, dtype=np.float32).reshape(32, 4)
    g = RINND(X, n_neighbors=6, n_trees=2, n_iters=3, random_state=42)
    bi, bd = g.query(X[0][None, :], k=4)
    si, sd = g.query_one(X[0], k=4)
    np.testing.assert_array_equal(si, bi[0])
    np.testing.assert_allclose(sd, bd[0])


def test_shapes_dtypes_and_errors():
    """test_rinnd.py:11-24, 58-65."""
    X = np.arange(32 * 4, dtype=np.float32).reshape(32, 4)
    g = RINND(X, n_neighbors=6, n_trees=2, n_iters=3, random_state=42)
    i, d = g.neighbor_graph
    assert i.shape == (32, 6) and d.shape == (32, 6)
    assert i.dtype == np.int32 and d.dtype == np.float32
    qi, qd = g.query(X[:3], k=4)
    assert qi.shape == (3, 4) and np.all(np.isfinite(qd))
    with pytest.raises(ValueError, match="dimension"):
        g.query(np.zeros((1, 5), np.float32), k=2)


def test_cosine_direct_mode_range():
    """test_rinnd.py:39-55."""
    X = np.arange(32 * 4, dtype=np.float32).reshape(32, 4)
    g = RINND(X, metric="cosine", normalize=True, cosine_distance_mode="direct", n_neighbors=6, n_trees=2,
              n_iters=3, random_state=42)
    i, d = g.query(X[:2], k=3)
    assert i.shape == (2, 3)
    assert np.all((d >= 0.0) & (d <= 2.0))


def test_query_finds_itself():
    """index.rs:844-847, 870-873."""
    X = uniform(500, 10, seed=3)
    g = RINND(X, n_neighbors=10, n_trees=3, n_search_trees=2, n_iters=3)
    for q in (37, 101, 499):
        i, _ = g.query_one(X[q], k=10)
        assert q in i


def test_inner_product_and_budget_paths():
    X = uniform(2000, 24, seed=9) - np.float32(0.4)
    g, r = make_pair(X, metric="inner_product", n_neighbors=8, n_iters=4, threads=4, n_search_trees=2,
                     search_tree_leaf_budget=3)
    gi, _ = g.neighbor_graph
    ri, _ = r.neighbor_graph
    np.testing.assert_array_equal(gi, ri)
    Q = uniform(50, 24, seed=10) - np.float32(0.4)
    qi, _ = g.query(Q, k=8, epsilon=0.2)
    oi, _ = r.query(Q, k=8, epsilon=0.2)
    np.testing.assert_array_equal(qi, oi)


def test_wide_queries_with_spilled_frontier():
    """A large epsilon and k widen the search until the frontier (the seed min-heap) outgrows the 2,048 entries a
    wave keeps in LDS and moves to the per-slot memory region; results still equal the oracle's bit for bit."""
    X = uniform(12000, 16, seed=5)
    Q = uniform(40, 16, seed=6)
    g, r = make_pair(X, n_neighbors=20, n_iters=6, n_trees=4, threads=4)
    for k, eps in ((100, 3.0), (40, 8.0)):
        qi, qd = g.query(Q, k=k, epsilon=eps)
        oi, od = r.query(Q, k=k, epsilon=eps)
        np.testing.assert_array_equal(qi, oi)
        np.testing.assert_array_equal(qd.view(np.uint32), od.view(np.uint32))

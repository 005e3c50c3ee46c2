"""GPU parity of the int8 routing trees (tree_quantization="int8", crates/rinnd/src/lib.rs:228-249, 303-347;
QuantizedFlatTree, core/tree/flat_tree.rs:35-183) against the oracle.

The descent through int8 trees is compared bit-exactly with the oracle's Index after quantize_trees(): the same
per-node codes and scales, the same per-query codes, the same reconstructed margins and therefore the same
leaves, seeds and query results.  The quantized cosine index is checked through the public API: storage
accounting (lib.rs:498-541, 1029-1075), the released FP32 trees (lib.rs:1117-1120) and recall.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import RINND
from rinnd_amd import _lib as L

pytestmark = pytest.mark.gpu


def quantized_pair(X, metric, **kw):
    """An FP32 index whose search is routed by int8 trees, with its oracle counterpart."""
    from test_gpu_index import make_pair
    g, r = make_pair(X, metric=metric, **kw)
    assert L.lib().rinnd_index_quantize_trees(g._h) == L.OK
    r.quantize_trees()
    return g, r


@pytest.mark.parametrize("metric,budget,trees", [("euclidean", 1, 1), ("euclidean", 4, 3),
                                                 ("inner_product", 1, 2), ("inner_product", 3, 1)])
def test_int8_routing_matches_oracle(metric, budget, trees):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((4000, 48)).astype(np.float32)
    g, r = quantized_pair(X, metric, n_neighbors=10, n_iters=5, n_search_trees=trees,
                          search_tree_leaf_budget=budget)
    Q = rng.standard_normal((300, 48)).astype(np.float32)
    for k, eps in ((10, 0.1), (5, 0.0)):
        gi, gd = g.query(Q, k=k, epsilon=eps)
        oi, od = r.query(Q, k=k, epsilon=eps)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_int8_routing_differs_from_fp32_routing_only_in_the_seeds():
    """The int8 margins change some routing decisions (else the test above would pin nothing new) while the
    search results stay those of a search: the oracle with FP32 trees and with int8 trees disagree somewhere,
    and the GPU follows the int8 one."""
    rng = np.random.default_rng(8)
    X = rng.standard_normal((3000, 64)).astype(np.float32)
    g, r = quantized_pair(X, "euclidean", n_neighbors=8, n_iters=4, search_tree_leaf_budget=2)
    from test_gpu_index import make_pair
    _, r32 = make_pair(X, metric="euclidean", n_neighbors=8, n_iters=4, search_tree_leaf_budget=2)
    Q = rng.standard_normal((400, 64)).astype(np.float32)
    gi, _ = g.query(Q, k=8, epsilon=0.0)
    oi, _ = r.query(Q, k=8, epsilon=0.0)
    fi, _ = r32.query(Q, k=8, epsilon=0.0)
    np.testing.assert_array_equal(gi, oi)
    assert (oi != fi).any(), "int8 routing changed no result: the workload does not exercise it"


def test_int8_trees_release_fp32_trees_and_repeat_is_harmless():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((1500, 32)).astype(np.float32)
    g, r = quantized_pair(X, "euclidean", n_neighbors=8, n_iters=4)
    assert not L.lib().rinnd_index_trees(g._h)
    with pytest.raises(ValueError, match="no retained search tree"):
        g.export_search_tree()
    Q = rng.standard_normal((50, 32)).astype(np.float32)
    a = g.query(Q, k=8)
    assert L.lib().rinnd_index_quantize_trees(g._h) == L.OK
    b = g.query(Q, k=8)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    oi, od = r.query(Q, k=8)
    np.testing.assert_array_equal(a[0], oi)


def test_quantized_cosine_index_with_int8_trees():
    rng = np.random.default_rng(3)
    centres = rng.random((40, 64)).astype(np.float32)
    X = (centres[rng.integers(0, 40, 6000)] + 0.08 * rng.standard_normal((6000, 64))).astype(np.float32)
    Q = (centres[rng.integers(0, 40, 300)] + 0.08 * rng.standard_normal((300, 64))).astype(np.float32)
    kw = dict(metric="cosine", normalize=True, cosine_distance_mode="int8", n_neighbors=15, random_state=7)
    a = RINND(X, tree_quantization="int8", retain_fp32_data=True, **kw)
    b = RINND(X, **kw)
    sa, sb = a.storage_info(), b.storage_info()
    nodes = sum(b._tree_nodes)
    assert nodes > 0 and a._tree_nodes == b._tree_nodes
    topo = nodes * 8 + len(b._tree_nodes) * len(X) * 4
    assert sa["tree_code_bytes"] == nodes * 64
    assert sa["tree_scale_bytes"] == sa["tree_offset_bytes"] == nodes * 4
    assert sa["tree_topology_bytes"] == topo and sa["retained_fp32_tree_topology_bytes"] == 0
    assert sa["released_fp32_tree_bytes"] == nodes * 64 * 4 + nodes * 4 + topo
    assert sb["retained_fp32_tree_topology_bytes"] == topo and sb["tree_code_bytes"] == 0
    for s in (sa, sb):
        parts = (s["code_bytes"] + s["metadata_bytes"] + s["retained_fp32_bytes"]
                 + s["retained_fp32_tree_topology_bytes"] + s["tree_code_bytes"] + s["tree_scale_bytes"]
                 + s["tree_offset_bytes"] + s["tree_topology_bytes"])
        assert s["allocated_bytes"] == parts
        assert s["bytes_per_vector"] == pytest.approx(parts / len(X))
    with pytest.raises(ValueError, match="no retained search tree"):
        a.export_search_tree()
    b.export_search_tree()
    # the same graph, searched from int8-routed seeds: recall against brute force stays that of FP32 routing
    ia, _ = a.query_quantized(Q, k=10, epsilon=0.1, candidate_width=128, rerank_width=128)
    ib, _ = b.query_quantized(Q, k=10, epsilon=0.1, candidate_width=128, rerank_width=128)
    Xn = X / np.linalg.norm(X, axis=1, keepdims=True)
    Qn = Q / np.linalg.norm(Q, axis=1, keepdims=True)
    truth = np.argsort(-(Qn @ Xn.T), axis=1)[:, :10]
    rec = lambda I: np.mean([len(set(I[i]) & set(truth[i])) / 10 for i in range(len(Q))])
    ra, rb = rec(ia), rec(ib)
    assert ra >= 0.95 and ra >= rb - 0.02, (ra, rb)
    np.testing.assert_array_equal(a.neighbor_graph[0], b.neighbor_graph[0])

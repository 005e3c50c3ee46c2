"""The oracle's int8 routing trees (QuantizedFlatTree, core/tree/flat_tree.rs:35-183; the int8 query of
crates/rinnd/src/lib.rs:303-347), pinned by the reference's own tests (flat_tree.rs:480-570) and by hand-computed
codes that tell rounding half away from zero (Rust's f32::round) from rounding half to even."""
import numpy as np
import pytest

from oracle import oracle as O


def f32(x):
    return np.float32(x)


def test_margin_reconstruction_sign_and_odd_dimension():
    """flat_tree.rs:480-495: margin = i8 dot as f32 * scale * query scale + offset, positive here."""
    h = [0.25, -0.5, 0.75, -1.0, 0.125]
    q = np.array([64, -32, 16, -8, 4], np.int8)
    codes, scale, m = O.quantized_margin(h, 0.2, q, 0.01)
    np.testing.assert_array_equal(codes, [32, -64, 95, -127, 16])  # h / (1 / 127), rounded
    assert scale == f32(1.0) / f32(127.0)
    dot = int(np.dot(codes.astype(np.int32), q.astype(np.int32)))
    assert dot == 6696
    want = f32(dot) * f32(scale) * f32(0.01) + f32(0.2)  # left to right, no fused multiply-add
    assert np.float32(m) == want and m > 0.0


def test_plane_codes_round_half_away_from_zero():
    """max |h| = 127 gives scale 1, so h / scale are the exact halves 2.5, -0.5, 6.5: rounded away from zero (3, -1, 7),
    where rounding to even would give 2, 0, 6."""
    codes, scale, _ = O.quantized_margin([2.5, -0.5, 127.0, 6.5], 0.0, np.zeros(4, np.int8), 1.0)
    assert scale == 1.0
    np.testing.assert_array_equal(codes, [3, -1, 127, 7])


def test_zero_plane_has_zero_scale_and_codes():
    codes, scale, m = O.quantized_margin([0.0, 0.0, 0.0], 0.5, np.array([1, 2, 3], np.int8), 0.1)
    assert scale == 0.0 and not codes.any() and m == np.float32(0.5)


def test_query_codes_round_half_away_and_clamp():
    """lib.rs:320-347: codes round(x * 127 / max|x|), clamped; scale max|x| / 127 (tree_query_scale)."""
    codes, scale = O.quantize_query_i8([127.0, 2.5, -0.5, 6.5, -127.0])
    assert scale == 1.0
    np.testing.assert_array_equal(codes, [127, 3, -1, 7, -127])
    codes, scale = O.quantize_query_i8([0.0, 0.0])
    assert scale == 0.0 and not codes.any()
    codes, scale = O.quantize_query_i8([0.5, -0.25])
    assert scale == f32(0.5) / f32(127.0)
    np.testing.assert_array_equal(codes, [127, -64])  # -0.25 * 254 = -63.5 -> -64


def _generated_angular_data(n=512, dim=17):
    """flat_tree.rs:522-533."""
    i = np.arange(n * dim, dtype=np.float64)
    data = (np.sin(((i * 37 + 11).astype(np.float32) * np.float32(0.017)).astype(np.float32))
            + np.float32(0.31) * np.cos(((i * 13 + 5).astype(np.float32) * np.float32(0.029)).astype(np.float32)))
    data = data.astype(np.float32).reshape(n, dim)
    return (data / np.sqrt((data * data).sum(1, dtype=np.float32))[:, None]).astype(np.float32)


def _round_away(v):
    return np.sign(v) * np.floor(np.abs(v) + np.float32(0.5))


def test_generated_angular_tree_routing_fidelity():
    """flat_tree.rs:520-570: with a leaf budget of 3, the int8 tree visits the same leaves as the FP32 tree for at
    least 95 % of the queries and 98 % of the leaves on average."""
    data = _generated_angular_data()
    F = O.build_tree(data, 24, O.rng_state(73), angular=True)
    identical, overlap = 0, 0.0
    for qid in range(128):
        q = data[qid]
        scale = np.float32(np.abs(q).max()) / np.float32(127.0)
        codes = np.clip(_round_away(q / scale), -127, 127).astype(np.int8)
        fp, _ = F.search_leaves(0, q, O.rng_state(991), 3)
        ql = F.search_leaves_quantized(0, codes, float(scale), O.rng_state(991), 3)
        identical += fp == ql
        overlap += sum(1 for leaf in fp if leaf in ql) / len(fp)
    assert identical / 128 >= 0.95 and overlap / 128 >= 0.98


def test_quantized_tree_preserves_leaves_and_finds_itself():
    """flat_tree.rs:498-520 (topology and leaves kept); an index routed by int8 trees still finds each data point
    itself (index.rs:811-873) and gives nearly the FP32-routed answers."""
    rng = np.random.default_rng(4)
    X = rng.standard_normal((3000, 24)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    r = O.Index(X, metric="sqeuclidean", n_neighbors=10, n_trees=4, leaf_size=30, n_iters=5, angular=True)
    fi, fd = r.query(X[:200], k=10, epsilon=0.1)
    r.quantize_trees()
    qi, qd = r.query(X[:200], k=10, epsilon=0.1)
    order = r.vertex_order
    assert np.mean(qi[:, 0] == np.arange(200)) >= 0.99
    same = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(fi, qi)])
    assert same >= 0.95

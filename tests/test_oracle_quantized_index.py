"""The oracle's quantized index (QuantizedI8Index, crates/rinnd/src/lib.rs:133-398; the quantized search of
core/index.rs:55-126, 283-330 and core/search/greedy.rs:101-146, 365-408), pinned by hand-computed codes, the
reference's own quantization tests (lib.rs:1714-1907) and brute force over the codes.

Codes are round-half-away-from-zero of x * multiplier (Rust's f32::round), restated here in float64 on the
float32 product, which is exact for |x * multiplier| <= 127.5.
"""
import numpy as np
import pytest

from oracle import oracle as O


def half_away(v):
    v = np.asarray(v, np.float64)
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def codes_of(X, encoding):
    """Independent restatement of the encodings with numpy (lib.rs:166-224)."""
    X = np.asarray(X, np.float32)
    if encoding == "sq8u":
        g = np.float32(np.abs(X).max())
        mult = np.full(len(X), np.float32(127.0) / g if g > 0 else np.float32(0), np.float32)
    else:
        mx = np.abs(X).max(axis=1).astype(np.float32)
        with np.errstate(divide="ignore"):
            mult = np.where(mx > 0, np.float32(127.0) / mx, np.float32(0)).astype(np.float32)
    prod = (X * mult[:, None]).astype(np.float32)
    return np.clip(half_away(prod), -127, 127).astype(np.int8), mult


def normalized_fixture(n, dim):
    """lib.rs:1718-1724's rows (sin/cos of the flat index), normalized."""
    i = np.arange(n * dim, dtype=np.float64)
    x = (np.sin(((i * 37 + 11) * 0.071).astype(np.float32)) + np.cos(((i * 13) * 0.037).astype(np.float32)) * 0.3)
    x = x.astype(np.float32).reshape(n, dim)
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


def test_codes_known_answer():
    # max 0.5 -> multiplier 254: 0.5 -> 127, -0.25 -> -63.5 -> -64 (half away from zero), 0.1 -> 25.4 -> 25
    X = np.array([[0.5, -0.25, 0.1], [0.0, 0.0, 0.0], [-2.0, 1.0, 0.0]], np.float32)
    c = O.encode_i8_rows(X, "int8")
    codes, inv, gm, gd = c.arrays()
    np.testing.assert_array_equal(codes[0], [127, -64, 25])
    np.testing.assert_array_equal(codes[1], [0, 0, 0])
    np.testing.assert_array_equal(codes[2], [-127, 64, 0])  # -2 * 63.5 = -127, 1 * 63.5 = 63.5 -> 64
    assert inv[0] == np.float32(1.0) / np.sqrt(np.float32(127 * 127 + 64 * 64 + 25 * 25))
    assert inv[1] == 0.0  # zero norm
    # sq8p: the row's dequantisation scale 1 / multiplier; sq8u: one global multiplier 127 / max|X| = 63.5
    _, inv_p, _, _ = O.encode_i8_rows(X, "sq8p").arrays()
    assert inv_p[0] == np.float32(1.0) / np.float32(254.0) and inv_p[1] == 0.0
    codes_u, inv_u, gm_u, gd_u = O.encode_i8_rows(X, "sq8u").arrays()
    assert gm_u == 63.5 and gd_u == np.float32(1.0) / np.float32(63.5)
    # ties round away from zero, not to even: multiplier 1 keeps 2.5 -> 3 and -0.5 -> -1
    np.testing.assert_array_equal(O.encode_i8_rows(np.array([[127.0, 2.5, -0.5]], np.float32), "sq8p").arrays()[0],
                                  [[127, 3, -1]])
    np.testing.assert_array_equal(codes_u[0], [32, -16, 6])  # 31.75 -> 32, -15.875 -> -16, 6.35 -> 6
    np.testing.assert_array_equal(inv_u, np.full(3, gd_u, np.float32))


@pytest.mark.parametrize("encoding", ["int8", "sq8u", "sq8p"])
def test_codes_match_numpy_restatement(encoding):
    rng = np.random.default_rng(4)
    X = rng.standard_normal((300, 37)).astype(np.float32)
    X[5] = 0.0
    X[7, 3] = 40.0  # one large value: sq8u squeezes every other row
    codes, inv, gm, gd = O.encode_i8_rows(X, encoding).arrays()
    want, mult = codes_of(X, encoding)
    np.testing.assert_array_equal(codes, want)
    if encoding == "int8":
        ns = (want.astype(np.int64) ** 2).sum(1).astype(np.float32)
        with np.errstate(divide="ignore"):
            want_inv = np.where(ns > 0, np.float32(1.0) / np.sqrt(ns), np.float32(0)).astype(np.float32)
    else:
        with np.errstate(divide="ignore"):
            want_inv = np.where(mult > 0, np.float32(1.0) / mult, np.float32(0)).astype(np.float32)
    np.testing.assert_array_equal(inv.view(np.uint32), want_inv.view(np.uint32))


def test_sq8p_reconstruction_bound_and_storage():
    """lib.rs:1748-1768: |x - code * scale| <= 0.501 scale + 1e-7 per coordinate; n x dim code bytes."""
    X = normalized_fixture(48, 5)
    codes, inv, _, _ = O.encode_i8_rows(X, "sq8p").arrays()
    assert codes.size == 48 * 5 and inv.size == 48
    err = np.abs(X - codes.astype(np.float32) * inv[:, None])
    assert np.all(err <= inv[:, None] * 0.501 + 1e-7)


def test_query_codes_follow_the_row_encoding():
    rng = np.random.default_rng(9)
    X = rng.standard_normal((50, 24)).astype(np.float32)
    q = rng.standard_normal(24).astype(np.float32)
    for enc in ("int8", "sq8u", "sq8p"):
        c = O.encode_i8_rows(X, enc)
        codes, s = c.quantize_query(q)
        if enc == "sq8u":
            _, _, gm, gd = c.arrays()
            want = np.clip(half_away((q * np.float32(gm)).astype(np.float32)), -127, 127)
            assert s == gd
        else:
            want, mult = codes_of(q[None], enc)
            want = want[0]
            if enc == "sq8p":
                assert s == np.float32(1.0) / mult[0]
            else:
                assert s == np.float32(1.0) / np.sqrt(np.float32((want.astype(np.int64) ** 2).sum()))
        np.testing.assert_array_equal(codes, want)


def test_i8_distance_known_answers():
    a = np.array([127, 0, 0], np.int8)
    b = np.array([0, 127, 0], np.int8)
    s = np.float32(1.0) / np.float32(127.0)
    assert O.i8_cosine_distance(a, s, b, s) == 1.0  # orthogonal
    assert O.i8_cosine_distance(a, s, -b, s) == 1.0
    # opposite: similarity -127^2 / 127^2 -> distance 2 (clamp keeps it there if rounding overshoots)
    assert O.i8_cosine_distance(a, s, (-a).astype(np.int8), s) == 2.0
    # clamp at 1: a similarity above 1 gives distance 0, not a negative one
    assert O.i8_cosine_distance(a, 1.0, a, 1.0) == 0.0
    assert O.i8_cosine_distance(a, 1.0, (-a).astype(np.int8), 1.0) == 2.0
    # 3 * 5 + 4 * (-2) = 7 -> 1 - 7 * 0.125 * 0.25
    assert O.i8_cosine_distance(np.array([3, 4], np.int8), 0.125, np.array([5, -2], np.int8), 0.25) == \
        np.float32(1.0) - np.float32(7.0) * np.float32(0.125) * np.float32(0.25)


def _index(X, **kw):
    p = dict(metric="direct_cosine", n_neighbors=15, n_trees=4, n_iters=5, seed=11, threads=2)
    p.update(kw)
    return O.Index(X, **p)


def _brute_i8(X, Q, encoding):
    codes, mult = codes_of(X, encoding)
    c = O.encode_i8_rows(X, encoding)
    _, inv, _, _ = c.arrays()
    D = np.zeros((len(Q), len(X)), np.float32)
    for i, q in enumerate(Q):
        qc, qs = c.quantize_query(q)
        dots = codes.astype(np.int64) @ qc.astype(np.int64)
        sim = (dots.astype(np.float32) * np.float32(qs)).astype(np.float32) * inv
        D[i] = np.float32(1.0) - np.clip(sim, -1, 1)
    return D


@pytest.mark.parametrize("encoding", ["int8", "sq8u", "sq8p"])
def test_exhaustive_search_is_brute_force_over_the_codes(encoding):
    """With an unbounded epsilon the search visits every reachable vertex: its k results are the k smallest
    int8 distances over all rows."""
    rng = np.random.default_rng(21)
    X = rng.standard_normal((400, 16)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    Q = rng.standard_normal((20, 16)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    ix = _index(X)
    codes = ix.encode_i8(encoding)
    oi, od = ix.query_i8(codes, Q, k=8, epsilon=1e30, candidate_width=8, rerank_width=0)
    D = _brute_i8(X, Q, encoding)
    for i in range(len(Q)):
        np.testing.assert_array_equal(od[i], np.sort(D[i])[:8])
        np.testing.assert_array_equal(D[i, oi[i]], od[i])


def test_rerank_gives_the_exact_distances_of_the_first_candidates():
    rng = np.random.default_rng(22)
    X = rng.standard_normal((400, 16)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    Q = rng.standard_normal((15, 16)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    ix = _index(X)
    codes = ix.encode_i8("int8")
    ai, ad = ix.query_i8(codes, Q, k=20, epsilon=0.1, candidate_width=20, rerank_width=0)
    ri, rd = ix.query_i8(codes, Q, k=6, epsilon=0.1, candidate_width=20, rerank_width=12)
    for i in range(len(Q)):
        # the re-ranked results are the 6 best of the first 12 approximate candidates, by exact distance
        exact = np.float32(1.0) - (X[ai[i, :12]] @ Q[i]).astype(np.float32)
        order = np.lexsort((ai[i, :12], exact))
        assert set(ri[i]) <= set(ai[i, :12])
        np.testing.assert_array_equal(ri[i], ai[i, :12][order][:6])
        np.testing.assert_allclose(rd[i], exact[order][:6], atol=2e-6)
        assert np.all(np.diff(rd[i]) >= 0)


def test_width_resolution_rules():
    """index.rs:55-88: candidate width 0 means max(k, rerank_width), and then only a rerank width above k
    re-ranks; an explicit candidate width re-ranks whenever rerank_width > 0."""
    X = normalized_fixture(96, 9)
    ix = _index(X, n_neighbors=10, n_trees=3, n_search_trees=2, n_iters=3, seed=41)
    codes = ix.encode_i8("int8")
    Q = X[11:14]
    same = lambda a, b: (np.testing.assert_array_equal(a[0], b[0]),
                         np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32)))
    same(ix.query_i8(codes, Q, 6, 0.08, 0, 0), ix.query_i8(codes, Q, 6, 0.08, 6, 0))
    same(ix.query_i8(codes, Q, 6, 0.08, 0, 6), ix.query_i8(codes, Q, 6, 0.08, 6, 0))  # rerank 6 <= k: none
    same(ix.query_i8(codes, Q, 6, 0.08, 0, 12), ix.query_i8(codes, Q, 6, 0.08, 12, 12))
    # batch == one by one (lib.rs:1850-1871)
    bi, bd = ix.query_i8(codes, Q, 6, 0.08, 18, 12)
    for j in range(3):
        oi, od = ix.query_i8(codes, Q[j:j + 1], 6, 0.08, 18, 12)
        np.testing.assert_array_equal(bi[j], oi[0])
    # widths are independent (lib.rs:1833-1848): all return k sorted results
    for cw, rw in ((6, 0), (18, 0), (18, 12)):
        oi, od = ix.query_i8(codes, Q, 6, 0.08, cw, rw)
        assert oi.shape == (3, 6) and np.all(oi >= 0) and np.all(np.diff(od, axis=1) >= 0)


def test_int8_trees_route_the_quantized_search():
    """With int8 routing trees the descent takes the query's codes (greedy.rs:101-146): an exhaustive search still
    returns the brute-force answer, and routing changes some narrow searches."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((1500, 24)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    Q = rng.standard_normal((200, 24)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    ix = _index(X, n_neighbors=8, n_trees=3, n_iters=4)
    codes = ix.encode_i8("sq8p")
    a = ix.query_i8(codes, Q, 4, 0.0, 4, 0)
    ix.quantize_trees()
    b = ix.query_i8(codes, Q, 4, 0.0, 4, 0)
    c = ix.query_i8(codes, Q[:20], 5, 1e30, 5, 0)
    D = _brute_i8(X, Q[:20], "sq8p")
    for i in range(20):
        np.testing.assert_array_equal(c[1][i], np.sort(D[i])[:5])
    assert (a[0] != b[0]).any()

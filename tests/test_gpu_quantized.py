"""The FP8 e4m3 candidate search with exact FP32 re-ranking (the quantized query of core/index.rs:283-330
with its width rules, core/index.rs:55-124, and the Python surface of rinnd/src/lib.rs:825-870, 1148-1172).

The codes are this build's FP8 counterpart of the reference's int8 codes, so the approximate distances
have no reference bits to match; what is pinned: re-ranked distances equal the reference arithmetic's
FP32 distance of the returned id (oracle pair distance), re-ranked rows are sorted, approximate distances
stay within the FP8 rounding bound of the exact ones, recall against brute force, and the width errors.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import RINND
from rinnd_amd import _lib as L

pytestmark = pytest.mark.gpu
INVALID_ARG = 1  # RINND_ERR_INVALID_ARG


def quantized(g, Q, k, eps, cw, rw):
    nq = Q.shape[0]
    oi = np.zeros((nq, k), np.int32)
    od = np.zeros((nq, k), np.float32)
    st = L.lib().rinnd_index_query_quantized(g._h, L.ptr(Q), nq, k, C.c_float(eps), cw, rw, L.ptr(oi), L.ptr(od))
    return st, oi, od


def brute(X, Q, k):
    D = ((Q[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    return np.sort(D, axis=1)[:, :k]


@pytest.fixture(scope="module")
def index():
    rng = np.random.default_rng(7)
    centres = rng.random((40, 48)).astype(np.float32)
    X = (centres[rng.integers(0, 40, 6000)] + 0.08 * rng.standard_normal((6000, 48))).astype(np.float32)
    Q = (centres[rng.integers(0, 40, 200)] + 0.08 * rng.standard_normal((200, 48))).astype(np.float32)
    g = RINND(X, metric="sqeuclidean", n_neighbors=15, n_trees=6, n_iters=6, random_state=3)
    return X, Q, g


@pytest.mark.parametrize("cw,rw", [(32, 32), (40, 20), (0, 24)])
def test_reranked_distances_are_the_fp32_distances(index, cw, rw):
    X, Q, g = index
    k = 10
    st, oi, od = quantized(g, Q, k, 0.1, cw, rw)
    assert st == 0
    assert np.all(oi >= 0)
    for q in range(0, len(Q), 7):
        exact = O.pair_distances("sqeuclidean", np.concatenate([Q[q:q + 1], X]),
                                 np.zeros(k, np.int64), oi[q].astype(np.int64) + 1)
        np.testing.assert_array_equal(od[q].view(np.uint32), exact.astype(np.float32).view(np.uint32))
        assert np.all(np.diff(od[q]) >= 0)
        assert len(set(oi[q].tolist())) == k


def test_recall_close_to_the_fp32_search(index):
    X, Q, g = index
    k = 10
    true_d = brute(X, Q, k)
    qi, qd = g.query(Q, k=k, epsilon=0.1)
    _, fi, fd = quantized(g, Q, k, 0.1, 32, 32)

    def recall(ids):
        d = ((Q[:, None, :].astype(np.float64) - X[ids]) ** 2).sum(-1)
        return np.mean(d <= true_d[:, -1:] * (1 + 1e-9))
    r32, r8 = recall(qi), recall(fi)
    assert r8 >= 0.9 and r8 >= r32 - 0.03, (r8, r32)


def test_approximate_distances_within_fp8_rounding(index):
    """Without re-ranking the distances come from e4m3 codes of both the row and the query (one scale each): per
    coordinate the rounding error is at most 2^-4 of the value (3 mantissa bits, round to nearest) in the normal
    range, plus a subnormal step of 2^-9 of the scale below it.  With e = |x - x~| + |q - q~| (triangle inequality),
    |d~ - d| <= 2 sqrt(d) e + e^2, plus the FP32 rounding of the parts (|q~|^2 + |x~|^2 - 2 q~.x~)."""
    X, Q, g = index
    k = 10
    st, oi, od = quantized(g, Q, k, 0.1, 24, 0)
    assert st == 0

    def err(v):
        scale = np.abs(v).max(-1, keepdims=True) / 240.0
        return np.sqrt(((np.abs(v) / 16.0 + scale * 2.0 ** -9) ** 2).sum(-1))
    for q in range(0, len(Q), 11):
        xs = X[oi[q]].astype(np.float64)
        qv = Q[q].astype(np.float64)
        exact = ((qv[None, :] - xs) ** 2).sum(-1)
        e = err(xs) + err(qv[None, :])
        parts = (qv * qv).sum() + (xs * xs).sum(-1)
        bound = 2 * np.sqrt(exact) * e + e ** 2 + 1e-6 * parts + 1e-6
        assert np.all(np.abs(od[q] - exact) <= bound), (od[q], exact, bound)


def test_width_rules(index):
    X, Q, g = index
    assert quantized(g, Q[:2], 10, 0.1, 8, 0)[0] == INVALID_ARG      # candidate width < k
    assert quantized(g, Q[:2], 10, 0.1, 16, 20)[0] == INVALID_ARG    # rerank > candidate width
    assert quantized(g, Q[:2], 10, 0.1, 16, 5)[0] == INVALID_ARG     # 0 < rerank < k
    assert quantized(g, Q[:2], 10, 0.1, 0, 5)[0] == 0                            # resolved to k, no rerank
    st, oi, od = quantized(g, Q[:2], 10, 0.1, 0, 12)                              # width 12, rerank 12 > k
    assert st == 0 and np.all(oi >= 0)


def test_python_quantized_cosine_mode():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((3000, 32)).astype(np.float32)
    g = RINND(X, metric="cosine", normalize=True, cosine_distance_mode="int8", n_neighbors=10, n_trees=4,
              n_iters=4, quantized_candidate_width=20, quantized_rerank_width=20, random_state=5)
    i, d = g.query(X[:20], k=5)
    assert i.shape == (20, 5) and np.all((d >= 0.0) & (d <= 2.0))
    assert np.mean(i[:, 0] == np.arange(20)) >= 0.9  # each row finds itself
    i2, d2 = g.query_quantized(X[:20], k=5, candidate_width=30, rerank_width=30)
    assert np.all(np.diff(d2, axis=1) >= 0)
    with pytest.raises(ValueError, match="must be at least k"):
        g.query_quantized(X[:2], k=5, candidate_width=3)
    with pytest.raises(ValueError, match="must not exceed candidate_width"):
        g.query_quantized(X[:2], k=5, candidate_width=8, rerank_width=9)
    with pytest.raises(ValueError, match="zero or at least k"):
        g.query_quantized(X[:2], k=5, candidate_width=8, rerank_width=2)
    info = g.storage_info()
    assert info["encoding"] == "int8_per_vector_cosine" and info["code_bytes"] == 3000 * 32
    plain = RINND(X, n_neighbors=10, n_trees=4, n_iters=3)
    with pytest.raises(ValueError, match="quantized cosine_distance_mode"):
        plain.query_quantized(X[:2], k=5)

"""Committed vectors for the 10,000 x 128 config (tests/golden/cfg1_oracle.json, written by
tests/golden/make_golden.py from the oracle): the oracle must keep reproducing them (CPU), and the GPU path
must reproduce them bit for bit (GPU)."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "cfg1_oracle.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def data():
    X = np.random.default_rng(0).random((10_000, 128), dtype=np.float32)
    Q = np.random.default_rng(1).random((100, 128), dtype=np.float32)
    return X, Q


def check_nn(g, idx, dist, flags, changes, rng_state):
    assert [int(c) for c in changes] == g["changes"]
    assert idx[:5].tolist() == g["rows_0_4_indices"]
    assert dist[:5].view(np.uint32).tolist() == g["rows_0_4_distance_bits"]
    assert sha(idx) == g["indices_sha256"]
    assert sha(dist) == g["distances_sha256"]
    assert sha(flags) == g["flags_sha256"]
    assert [int(v) for v in rng_state] == g["rng_state"]


@pytest.mark.parametrize("T", [1, 8])
def test_oracle_reproduces_golden(T):
    from oracle import oracle as O
    X, _ = data()
    r = O.nn_descent(X, 15, 8, 60, 15, 10, 0.001, threads=T)
    check_nn(GOLD[f"nn_descent_T{T}"], r["indices"], r["distances"], r["flags"], r["changes"], r["rng_state"])


@pytest.mark.gpu
@pytest.mark.parametrize("T", [1, 8])
def test_gpu_reproduces_golden(ctx, T):
    from rinnd_amd import _lib as L
    X, _ = data()
    n, k = X.shape[0], 15
    p = L.make_params(k, 8, 60, 15, 10, 0.001, n_threads=T)
    st = L.rng_seed(42)
    idx = np.zeros((n, k), np.int32)
    dist = np.zeros((n, k), np.float32)
    flags = np.zeros((n, k), np.uint8)
    stats = L.BuildStats()
    h = C.c_void_p()
    rc = L.lib().rinnd_nn_descent(ctx.h, L.ptr(X), n, X.shape[1], L.DIST["sqeuclidean"], C.byref(p), L.ptr(st),
                                  L.ptr(idx), L.ptr(dist), L.ptr(flags), C.byref(h), C.byref(stats))
    ctx.check(rc, "rinnd_nn_descent")
    L.lib().rinnd_forest_free(h)
    check_nn(GOLD[f"nn_descent_T{T}"], idx, dist, flags, stats.to_dict()["changes"], st)


@pytest.mark.gpu
def test_gpu_index_reproduces_golden():
    from rinnd_amd import RINND
    X, Q = data()
    g = GOLD["index_T8"]
    ix = RINND(X, metric="euclidean", n_neighbors=15, n_trees=8, leaf_size=60, max_candidates=15, n_iters=10,
               n_threads=8)
    gi, gd = ix.neighbor_graph
    assert sha(gi) == g["graph_indices_sha256"]
    assert sha(gd) == g["graph_distances_sha256"]
    qi, qd = ix.query(Q, k=10, epsilon=0.1)
    assert qi[:5].tolist() == g["query_rows_0_4_indices"]
    assert qd[:5].view(np.uint32).tolist() == g["query_rows_0_4_distance_bits"]
    assert sha(qi) == g["query_indices_sha256"]
    assert sha(qd) == g["query_distances_sha256"]

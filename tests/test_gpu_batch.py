"""The batched build (rinnd_nn_descent_batch): many independent indexes in one call must each equal the build of
that index alone - rinnd_nn_descent with reference draws and one candidate thread, which the oracle pins - bit for
bit: heap indices, distances, flags, iteration count, convergence, and the advanced FastRng state.  The reference
builds one index per call (core/index.rs:601-733); per index it keeps its own n_trees and n_iters (the batching
form derives both from n: n_trees = clamp(round(2 log10 n), 3, 12), n_iters = max(ceil(log2 n), 5), index.rs:628).

Also: the pass replayed from the captured hipGraph equals the eager pass (every replay feeds the next pass, so the
final heaps check every replay), and the join tile-shape tuner (which needs enough rows to time) leaves results
unchanged.
"""
import ctypes as C
import functools
import math

import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import _lib as L
from rinnd_amd import datasets

pytestmark = pytest.mark.gpu

K, LEAF = 10, 60


def per_index(n):
    n_trees = min(12, max(3, int(round(2 * math.log10(max(n, 1))))))
    n_iters = max(int(math.ceil(math.log2(max(n, 2)))), 5)
    return n_trees, n_iters


def make_batch(sizes, seed=5):
    parts = [datasets.cfg4_index(n, 64, seed=seed, i=i) for i, n in enumerate(sizes)]
    X = np.concatenate(parts) if parts else np.zeros((0, 64), np.float32)
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return X, off


def standalone(ctx, X, n_trees, n_iters, seed):
    n, d = X.shape
    p = L.make_params(K, n_trees, LEAF, K, n_iters, 0.001, angular=True, n_threads=1, rng_mode="reference")
    st = L.rng_seed(seed)
    idx = np.zeros((n, K), np.int32)
    dist = np.zeros((n, K), np.float32)
    flags = np.zeros((n, K), np.uint8)
    stats = L.BuildStats()
    h = C.c_void_p()
    rc = L.lib().rinnd_nn_descent(ctx.h, L.ptr(X), n, d, L.DIST["cosine"], C.byref(p), L.ptr(st), L.ptr(idx),
                                  L.ptr(dist), L.ptr(flags), C.byref(h), C.byref(stats))
    ctx.check(rc, "rinnd_nn_descent")
    L.lib().rinnd_forest_free(h)
    s = stats.to_dict()
    return idx, dist, flags, s["iterations"], s["converged"], st


def run_batch(ctx, sizes, seeds, graph=True, n_iters=None, data_seed=5):
    X, off = make_batch(sizes, data_seed)
    nt = [per_index(n)[0] for n in sizes]
    ni = [per_index(n)[1] for n in sizes] if n_iters is None else n_iters
    p = L.make_params(K, 0, LEAF, K, 0, 0.001, angular=True, n_threads=1, rng_mode="reference")
    out = L.nn_descent_batch(ctx, X, off, "cosine", p, seeds, n_trees=nt, n_iters=ni, graph=graph)
    return X, off, nt, ni, out


def check_against_standalone(ctx, X, off, nt, ni, seeds, out, which):
    idx, dist, flags, its, conv, states, _ = out
    for s in which:
        a, b = int(off[s]), int(off[s + 1])
        if b == a:
            assert its[s] == 0 and conv[s] == 0
            np.testing.assert_array_equal(states[s], L.rng_seed(int(seeds[s])))
            continue
        wi, wd, wf, wit, wconv, wst = standalone(ctx, X[a:b], nt[s], ni[s], int(seeds[s]))
        print(f"[batch] index {s}: n={b - a} iterations={its[s]}", flush=True)
        np.testing.assert_array_equal(idx[a:b], wi)
        np.testing.assert_array_equal(dist[a:b].view(np.uint32), wd.view(np.uint32))
        np.testing.assert_array_equal(flags[a:b], wf)
        assert its[s] == wit and conv[s] == wconv
        np.testing.assert_array_equal(states[s], wst)


SIZES = [2000, 37, 5731, 0, 1, 12000, 60, 61, 3333, 950]


def test_batch_equals_standalone_builds(ctx):
    seeds = [100 + i for i in range(len(SIZES))]
    print(f"[batch] sizes {SIZES}", flush=True)
    X, off, nt, ni, out = run_batch(ctx, SIZES, seeds)
    st = out[6]
    assert st["passes"] == max(out[3])
    assert st["graph_replays"] == st["passes"] - 1
    check_against_standalone(ctx, X, off, nt, ni, seeds, out, range(len(SIZES)))


def test_batch_index_matches_oracle(ctx):
    """Two indexes of the batch against the oracle directly (the reference's algorithm restated on the CPU)."""
    sizes, seeds = [1500, 2600], [7, 8]
    X, off, nt, ni, out = run_batch(ctx, sizes, seeds)
    idx, dist, flags, its, conv, states, _ = out
    for s in range(2):
        a, b = int(off[s]), int(off[s + 1])
        r = O.nn_descent(X[a:b], K, nt[s], LEAF, K, ni[s], 0.001, metric="cosine", angular=True, seed=seeds[s])
        np.testing.assert_array_equal(idx[a:b], r["indices"])
        np.testing.assert_array_equal(dist[a:b].view(np.uint32), r["distances"].view(np.uint32))
        np.testing.assert_array_equal(flags[a:b], r["flags"])
        assert its[s] == len(r["changes"])
        np.testing.assert_array_equal(states[s], r["rng_state"])


def test_graph_replays_equal_eager_passes(ctx):
    sizes = [4000, 900, 7000, 2500]
    seeds = [1, 2, 3, 4]
    _, _, _, _, g = run_batch(ctx, sizes, seeds, graph=True)
    _, _, _, _, e = run_batch(ctx, sizes, seeds, graph=False)
    assert g[6]["graph_replays"] >= 4 and e[6]["graph_replays"] == 0
    for a, b in zip(g[:6], e[:6]):
        np.testing.assert_array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    assert g[6]["changes"] == e[6]["changes"]


def test_iteration_limits_and_empty_batch(ctx):
    sizes = [3000, 3000, 3000]
    seeds = [9, 9, 9]
    X, off, nt, ni, out = run_batch(ctx, sizes, seeds, n_iters=[0, 1, 3])
    assert list(out[3]) == [0, 1, 3]
    check_against_standalone(ctx, X, off, nt, [0, 1, 3], seeds, out, range(3))
    p = L.make_params(K, 3, LEAF, K, 5, 0.001, angular=True)
    r = L.nn_descent_batch(ctx, np.zeros((0, 64), np.float32), np.zeros(1, np.int64), "cosine", p, [])
    assert r[0].shape == (0, K)


def test_unsupported_settings_raise(ctx):
    X, off = make_batch([500, 700])
    for kw in ({"n_threads": 8}, {"rng_mode": "counter"}):
        p = L.make_params(K, 3, LEAF, K, 5, 0.001, angular=True, **kw)
        with pytest.raises(L.RinndError, match="UNSUPPORTED|one candidate thread"):
            L.nn_descent_batch(ctx, X, off, "cosine", p, [1, 2])
    p = L.make_params(K, 3, LEAF, K, 5, 0.001, angular=True)
    with pytest.raises(ValueError, match="non-decreasing"):
        L.nn_descent_batch(ctx, X, np.array([0, 800, 700], np.int64), "cosine", p, [1, 2])


@functools.lru_cache(maxsize=1)
def tuned_batch_sizes():
    return [int(v) for v in datasets.cfg4_sizes(48, 2_000, 20_000, seed=11)]


def test_tuned_join_batch(ctx):
    """Enough rows (> 400k) for the join tile-shape tuner to time every shape on the first pass."""
    sizes = tuned_batch_sizes()
    seeds = list(range(500, 500 + len(sizes)))
    print(f"[batch] {len(sizes)} indexes, {sum(sizes)} rows", flush=True)
    X, off, nt, ni, out = run_batch(ctx, sizes, seeds)
    st = out[6]
    assert st["join_threads_per_row"] in (64, 128, 256) and st["join_rows_per_group"] in (1, 2, 4)
    check_against_standalone(ctx, X, off, nt, ni, seeds, out, [0, len(sizes) // 2, len(sizes) - 1])

"""GPU parity of the NN-Descent build path against the oracle (CPU restatement of the reference).

Bit-exact on the FP32 path: distances, RP forest (hyperplanes, offsets, children, leaf order), and
the full nn_descent heap (indices, distances, flags) plus per-iteration change counts, for the
reference's sequential candidate path (T=1) and its Rayon block path (T=8).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import _lib as L

pytestmark = pytest.mark.gpu


def uniform(n, d, seed=0):
    return np.random.default_rng(seed).random((n, d), dtype=np.float32)


def gpu_pair_distances(ctx, metric, X, a, b):
    out = np.zeros(len(a), np.float32)
    st = L.lib().rinnd_pair_distances(ctx.h, L.DIST[metric], L.ptr(X), X.shape[0], X.shape[1],
                                      L.ptr(a), L.ptr(b), len(a), L.ptr(out))
    ctx.check(st, "rinnd_pair_distances")
    return out


@pytest.mark.parametrize("metric", list(L.DIST))
@pytest.mark.parametrize("d", [1, 7, 8, 13, 32, 37, 64, 96, 128, 130])
def test_pair_distances_bit_exact(ctx, metric, d):
    rng = np.random.default_rng(d)
    X = (rng.standard_normal((64, d)) * 3).astype(np.float32)
    if metric in ("alt_dot", "direct_cosine"):
        X /= np.linalg.norm(X, axis=1, keepdims=True) + np.float32(1e-6)
    a = rng.integers(0, 64, 500).astype(np.int32)
    b = rng.integers(0, 64, 500).astype(np.int32)
    got = gpu_pair_distances(ctx, metric, X, a, b)
    want = O.pair_distances(metric, X, a, b)
    if metric in ("alt_dot", "alt_cosine"):
        # log2f: device libm vs glibc may differ by an ulp; the dot/norm inputs are bit-exact
        np.testing.assert_allclose(got, want, rtol=4e-7, atol=0)
    else:
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def gpu_forest(ctx, X, n_trees, leaf, seed, angular=False):
    st = L.rng_seed(seed)
    h = C.c_void_p()
    rc = L.lib().rinnd_build_forest(ctx.h, L.ptr(X), X.shape[0], X.shape[1], n_trees, leaf, int(angular), 200,
                                    0, L.ptr(st), C.byref(h))
    ctx.check(rc, "rinnd_build_forest")
    return L.Forest(h), st


@pytest.mark.parametrize("angular", [False, True])
@pytest.mark.parametrize("n,d,leaf", [(1, 8, 5), (50, 3, 4), (2000, 32, 20), (3000, 37, 60),
                                      # large enough for several workgroups per tree on the large nodes
                                      (12000, 33, 100), (40000, 64, 60)])
def test_forest_matches_reference_streams(ctx, n, d, leaf, angular):
    X = uniform(n, d, seed=n + d)
    if angular:
        X = X - 0.5
    F, st_gpu = gpu_forest(ctx, X, 4, leaf, 42, angular)
    st_ref = O.rng_state(42)
    R = O.build_forest(X, 4, leaf, st_ref, angular=angular, threads=4)
    np.testing.assert_array_equal(st_gpu, st_ref)
    assert F.n_trees == R.n_trees == 4
    for t in range(4):
        g = F.tree(t, n)
        r = R.tree(t)
        assert g["n_nodes"] == r["n_nodes"]
        np.testing.assert_array_equal(g["indices"], r["indices"])
        np.testing.assert_array_equal(g["children"], r["children"])
        np.testing.assert_array_equal(g["offsets"].view(np.uint32), r["offsets"].view(np.uint32))
        np.testing.assert_array_equal(g["hyperplanes"].view(np.uint32), R.hyperplanes(t, d).view(np.uint32))


@pytest.mark.parametrize("n,d,kind", [(30000, 16, "mixture"), (20000, 128, "mixture"), (20000, 8, "grid"),
                                      (6000, 3, "grid")])
def test_forest_integer_data_with_margin_ties(ctx, n, d, kind):
    """Integer rows: points exactly on a bisector give |margin| < 1e-8 ties, which take draws in the two-pointer
    examination order (builder.rs:319-347) - a few per node in the headline generator's data, many in coarse
    integer grids (both ways of deciding them), in large nodes shared by several workgroups too."""
    from rinnd_amd import datasets
    if kind == "mixture":
        X = datasets.cfg2(n=n, d=d)
    else:
        X = np.random.default_rng(d).integers(0, 6, (n, d)).astype(np.float32)
    F, st_gpu = gpu_forest(ctx, X, 4, 60, 42)
    st_ref = O.rng_state(42)
    R = O.build_forest(X, 4, 60, st_ref, threads=4)
    np.testing.assert_array_equal(st_gpu, st_ref)
    for t in range(4):
        g, r = F.tree(t, n), R.tree(t)
        assert g["n_nodes"] == r["n_nodes"]
        np.testing.assert_array_equal(g["indices"], r["indices"])
        np.testing.assert_array_equal(g["children"], r["children"])
        np.testing.assert_array_equal(g["offsets"].view(np.uint32), r["offsets"].view(np.uint32))


@pytest.mark.parametrize("n", [300, 10000])
def test_forest_handles_duplicate_points(ctx, n):
    # all-equal rows: zero hyperplanes, margins tie -> next_bool() in examination order (at n = 10000 the ties
    # arise in nodes large enough to be shared by several workgroups)
    X = np.ones((n, 16), np.float32)
    X[n // 2:] = 2.0
    F, st_gpu = gpu_forest(ctx, X, 3, 10, 7)
    st_ref = O.rng_state(7)
    R = O.build_forest(X, 3, 10, st_ref)
    np.testing.assert_array_equal(st_gpu, st_ref)
    for t in range(3):
        g, r = F.tree(t, n), R.tree(t)
        np.testing.assert_array_equal(g["indices"], r["indices"])
        np.testing.assert_array_equal(g["children"], r["children"])


def gpu_forest_counter(ctx, X, n_trees, leaf, seed, angular=False, max_depth=200):
    st = L.rng_seed(seed)
    h = C.c_void_p()
    rc = L.lib().rinnd_build_forest(ctx.h, L.ptr(X), X.shape[0], X.shape[1], n_trees, leaf, int(angular), max_depth,
                                    L.RNG["counter"], L.ptr(st), C.byref(h))
    ctx.check(rc, "rinnd_build_forest(counter)")
    return L.Forest(h)


def check_tree_structure(X, tr, leaf, max_depth=200):
    """Every point in exactly one leaf, leaves tile tree.indices, each member lies on its side of every
    ancestor hyperplane (left iff margin < 0; near-zero margins may go either way).  Returns the number
    of points in leaves above leaf_size that the depth limit does not explain (degenerate splits)."""
    n = X.shape[0]
    idx, ch, hp, off = tr["indices"], tr["children"], tr["hyperplanes"].astype(np.float64), tr["offsets"]
    assert sorted(idx.tolist()) == list(range(n))
    leaves = []
    stack = [(0, 0, [])]
    seen = set()
    while stack:
        node, depth, path = stack.pop()
        assert node not in seen
        seen.add(node)
        a, b = int(ch[node, 0]), int(ch[node, 1])
        if b <= 0:  # leaf [-a, -b); internal nodes hold child ids >= 1
            leaves.append((-a, -b, depth, path, node))
            continue
        stack.append((b, depth + 1, path + [(node, 1)]))
        stack.append((a, depth + 1, path + [(node, 0)]))
    assert len(seen) == tr["n_nodes"]
    leaves.sort()
    assert leaves[0][0] == 0 and leaves[-1][1] == n
    assert all(p[1] == q[0] for p, q in zip(leaves, leaves[1:]))
    Xd = X.astype(np.float64)
    oversized = 0
    for s, e, depth, path, node in leaves:
        rows = Xd[idx[s:e]]
        assert not np.any(hp[node]) and off[node] == 0
        if e - s > leaf and depth < max_depth:
            oversized += e - s
        for anc, side in path:
            m = rows @ hp[anc] + off[anc]
            tol = 1e-4 * (1 + np.abs(rows).sum(1))
            if side == 0:
                assert np.all(m < tol)
            else:
                assert np.all(m > -tol)
    return oversized


@pytest.mark.parametrize("angular", [False, True])
@pytest.mark.parametrize("n,d,leaf", [(1, 8, 5), (50, 3, 4), (3000, 36, 30), (4001, 24, 25), (20000, 128, 60),
                                      (6000, 200, 40), (3000, 600, 50)])
def test_counter_forest_structure(ctx, n, d, leaf, angular):
    X = uniform(n, d, seed=n + d + 1)
    if angular:
        X = X - 0.5
    F = gpu_forest_counter(ctx, X, 5, leaf, 11, angular)
    G = gpu_forest_counter(ctx, X, 5, leaf, 11, angular)
    for t in range(5):
        tr = F.tree(t, n)
        assert check_tree_structure(X, tr, leaf) == 0
        g = G.tree(t, n)  # deterministic
        np.testing.assert_array_equal(tr["indices"], g["indices"])
        np.testing.assert_array_equal(tr["children"], g["children"])
    if n > 100:  # different trees
        assert not np.array_equal(F.tree(0, n)["indices"], F.tree(1, n)["indices"])


@pytest.mark.parametrize("n_trees,d", [(20, 64), (37, 24)])
def test_counter_forest_many_trees(ctx, n_trees, d):
    """More trees than one 16-lane batch of the side tests (and a partial last batch)."""
    X = uniform(5000, d, seed=n_trees)
    F = gpu_forest_counter(ctx, X, n_trees, 40, 3)
    for t in (0, 15, 16, n_trees - 1):
        assert check_tree_structure(X, F.tree(t, 5000), 40) == 0


def test_counter_forest_duplicates_and_depth_limit(ctx):
    X = np.ones((400, 16), np.float32)
    X[200:] = 3.0
    F = gpu_forest_counter(ctx, X, 3, 10, 5)
    for t in range(3):
        assert check_tree_structure(X, F.tree(t, 400), 10) <= 40  # all-tie nodes split by random bits
    X = uniform(5000, 8, seed=2)
    F = gpu_forest_counter(ctx, X, 2, 10, 5, max_depth=3)
    for t in range(2):
        tr = F.tree(t, 5000)
        assert check_tree_structure(X, tr, 10, max_depth=3) == 0
        assert tr["n_nodes"] <= 15


def gpu_nn_descent(ctx, X, k, n_trees, leaf, C_, iters, delta=0.001, metric="sqeuclidean", threads=1,
                   rng_mode="reference", seed=42, angular=False):
    n = X.shape[0]
    p = L.make_params(k, n_trees, leaf, C_, iters, delta, angular=angular, n_threads=threads, rng_mode=rng_mode)
    st = L.rng_seed(seed)
    idx = np.zeros((n, k), np.int32)
    dist = np.zeros((n, k), np.float32)
    flags = np.zeros((n, k), np.uint8)
    stats = L.BuildStats()
    h = C.c_void_p()
    rc = L.lib().rinnd_nn_descent(ctx.h, L.ptr(X), n, X.shape[1], L.DIST[metric], C.byref(p), L.ptr(st),
                                  L.ptr(idx), L.ptr(dist), L.ptr(flags), C.byref(h), C.byref(stats))
    ctx.check(rc, "rinnd_nn_descent")
    return idx, dist, flags, stats.to_dict(), st, L.Forest(h)


@pytest.mark.parametrize("threads", [1, 8])
@pytest.mark.parametrize("n,d,k,leaf", [(600, 16, 10, 20), (2500, 32, 15, 40), (1500, 24, 20, 40),
                                        (1200, 16, 113, 150), (1200, 16, 128, 150), (900, 8, 256, 300)])
def test_nn_descent_bit_exact(ctx, n, d, k, leaf, threads):
    """k <= 16 keeps the candidate rows' own ids in registers; k = 20 takes the general duplicate scan; k = 113
    is the largest heap whose 64 lane columns fit the ordered apply's LDS, k >= 114 takes the wave-per-destination
    apply (and the grouped leaf path)."""
    X = uniform(n, d, seed=n)
    C_ = min(k, 60)
    gi, gd, gf, gs, gst, _ = gpu_nn_descent(ctx, X, k, 4, leaf, C_, 6, threads=threads)
    r = O.nn_descent(X, k, 4, leaf, C_, 6, threads=threads)
    assert gs["changes"] == list(r["changes"])
    np.testing.assert_array_equal(gi, r["indices"])
    np.testing.assert_array_equal(gd.view(np.uint32), r["distances"].view(np.uint32))
    np.testing.assert_array_equal(gf, r["flags"])
    np.testing.assert_array_equal(gst, r["rng_state"])


@pytest.mark.parametrize("leaf,d", [(200, 128), (140, 64), (40, 300), (150, 300), (256, 520)])
def test_leaf_init_paths(ctx, leaf, d):
    """Leaf initialisation through the two-pass kernels: rows of a leaf staged whole in LDS, or (150 rows of 300
    dims, 256 rows of 520) staged 64 dimensions at a time with the pairs' accumulators kept in registers - a last
    partial chunk, a scalar tail (300 = 37 x 8 + 4) and leaves with more 4 x 4 tiles than threads (several passes
    over the rows)."""
    X = uniform(4000, d, seed=leaf)
    gi, gd, gf, gs, gst, _ = gpu_nn_descent(ctx, X, 15, 2, leaf, 15, 1, threads=4)
    r = O.nn_descent(X, 15, 2, leaf, 15, 1, threads=4)
    np.testing.assert_array_equal(gi, r["indices"])
    np.testing.assert_array_equal(gd.view(np.uint32), r["distances"].view(np.uint32))
    np.testing.assert_array_equal(gf, r["flags"])


@pytest.mark.parametrize("metric,angular", [("cosine", True), ("inner_product", True), ("alt_inner_product", True)])
def test_nn_descent_other_metrics(ctx, metric, angular):
    X = uniform(800, 24, seed=3) - np.float32(0.3)
    gi, gd, gf, gs, _, _ = gpu_nn_descent(ctx, X, 8, 3, 30, 8, 5, metric=metric, angular=angular, threads=4)
    r = O.nn_descent(X, 8, 3, 30, 8, 5, metric=metric, angular=angular, threads=4)
    np.testing.assert_array_equal(gi, r["indices"])
    np.testing.assert_array_equal(gd.view(np.uint32), r["distances"].view(np.uint32))


def test_nn_descent_cfg1_bit_exact(ctx):
    """10,000 x 128 U[0,1), k=15, leaf 60, 8 trees, <=10 iterations, delta 0.001 (T=1 and T=8)."""
    X = uniform(10000, 128, seed=0)
    for threads in (1, 8):
        gi, gd, gf, gs, gst, _ = gpu_nn_descent(ctx, X, 15, 8, 60, 15, 10, threads=threads)
        r = O.nn_descent(X, 15, 8, 60, 15, 10, threads=threads)
        assert gs["changes"] == list(r["changes"])
        np.testing.assert_array_equal(gi, r["indices"])
        np.testing.assert_array_equal(gd.view(np.uint32), r["distances"].view(np.uint32))
        np.testing.assert_array_equal(gf, r["flags"])


def test_counter_forest_large(ctx):
    """A 70,000-point counter forest (dense point-major levels, then sparse ones) is a partition in the
    caller's ids and deterministic, and the graph built on it finds the true neighbours about as well as
    the reference algorithm, with the distances of the returned ids."""
    n, d = 70000, 16
    X = uniform(n, d, seed=9)
    F = gpu_forest_counter(ctx, X, 5, 60, 13)
    G = gpu_forest_counter(ctx, X, 5, 60, 13)
    for t in (0, 4):
        tr = F.tree(t, n)
        assert check_tree_structure(X, tr, 60) == 0
        np.testing.assert_array_equal(tr["indices"], G.tree(t, n)["indices"])
    gi, gd, _, _, _, Fg = gpu_nn_descent(ctx, X, 10, 5, 60, 10, 8, rng_mode="counter")
    for t in range(Fg.n_trees):  # every tree mapped back to the caller's ids
        assert check_tree_structure(X, Fg.tree(t, n), 60) == 0
    rows = np.arange(0, n, 211)
    X64 = X.astype(np.float64)
    D = (X64[rows] ** 2).sum(1)[:, None] + (X64 ** 2).sum(1)[None, :] - 2.0 * X64[rows] @ X64.T
    D[np.arange(len(rows)), rows] = np.inf
    true = np.argsort(D, axis=1)[:, :10]
    def recall(ids):
        return np.mean([len(set(ids[r].tolist()) & set(true[i].tolist())) / 10 for i, r in enumerate(rows)])
    ref = O.nn_descent(X, 10, 5, 60, 10, 8, 0.001, threads=8)["indices"]  # the reference algorithm's quality
    assert recall(gi) >= recall(ref) - 0.03, (recall(gi), recall(ref))
    dd = ((X64[rows, None, :] - X64[gi[rows]]) ** 2).sum(-1)
    np.testing.assert_allclose(gd[rows], dd, rtol=1e-4, atol=1e-6)


def test_counter_mode_quality(ctx):
    X = uniform(3000, 16, seed=5)
    gi, gd, gf, gs, _, _ = gpu_nn_descent(ctx, X, 10, 4, 30, 10, 8, rng_mode="counter")
    D = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    true = np.argsort(D, axis=1)[:, :10]
    rec = np.mean([len(set(gi[i]) & set(true[i])) / 10 for i in range(len(X))])
    assert rec > 0.9
    assert np.all(gi != np.arange(len(X))[:, None])


@pytest.mark.parametrize("k", [1, 15, 23, 64, 100])  # one row per wave (S > 64) and 64 / S rows per wave
def test_sort_rows_matches_sort_all(ctx, k):
    rng = np.random.default_rng(k)
    n = 501
    idx = rng.integers(-1, 1000, (n, k)).astype(np.int32)
    dist = rng.integers(0, 6, (n, k)).astype(np.float32)  # many ties: stability matters
    dist[idx < 0] = np.inf
    flags = rng.integers(0, 2, (n, k)).astype(np.uint8)
    ri, rd, rf = idx.copy(), dist.copy(), flags.copy()
    O.sort_all(ri, rd, rf)
    gi, gd, gf = idx.copy(), dist.copy(), flags.copy()
    ctx.check(L.lib().rinnd_sort_rows(ctx.h, L.ptr(gi), L.ptr(gd), L.ptr(gf), n, k), "rinnd_sort_rows")
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_array_equal(gd, rd)
    np.testing.assert_array_equal(gf, rf)


def _build(ctx, X, k, trees, leaf, C_, iters, rng_mode, precision, seed=7, threads=4, max_depth=200):
    n, d = X.shape
    p = L.make_params(k, trees, leaf, C_, iters, 0.001, n_threads=threads, rng_mode=rng_mode,
                      join_precision=precision, max_depth=max_depth)
    st = L.rng_seed(seed)
    idx = np.zeros((n, k), np.int32)
    dist = np.zeros((n, k), np.float32)
    flags = np.zeros((n, k), np.uint8)
    stats = L.BuildStats()
    h = C.c_void_p()
    rc = L.lib().rinnd_nn_descent(ctx.h, L.ptr(X), n, d, L.DIST["sqeuclidean"], C.byref(p), L.ptr(st), L.ptr(idx),
                                  L.ptr(dist), L.ptr(flags), C.byref(h), C.byref(stats))
    ctx.check(rc, f"rinnd_nn_descent({precision})")
    L.lib().rinnd_forest_free(h)
    return idx, dist, flags, stats.to_dict(), st


def _integer_mixture(rng, n, d, centres, sigma):
    c = rng.random((centres, d)) * 100
    X = np.clip(np.floor(c[rng.integers(0, centres, n)] + sigma * rng.standard_normal((n, d)) + 0.5), 0, 255)
    return X.astype(np.float32)


def _assert_same_build(got, want_idx, want_dist, want_flags, want_changes):
    assert got[3]["changes"] == list(want_changes)
    np.testing.assert_array_equal(got[0], want_idx)
    np.testing.assert_array_equal(got[1].view(np.uint32), want_dist.view(np.uint32))
    np.testing.assert_array_equal(got[2], want_flags)


@pytest.mark.parametrize("leaf,k", [(256, 15), (100, 30), (17, 8)])
def test_bf16_exact_leaf_init_matches_oracle(ctx, leaf, k):
    """Integer data (exact in bf16), reference draws: the MFMA leaf distances (whole leaves for the first trees,
    survivor pairs only after them) and the BF16 join give the oracle's heaps, flags, change counts and RNG
    state bit for bit, for leaves of up to 256 members."""
    X = _integer_mixture(np.random.default_rng(leaf), 6000, 128, 30, 6)
    b = _build(ctx, X, k, 7, leaf, k, 4, "reference", "bf16")
    r = O.nn_descent(X, k, 7, leaf, k, 4, 0.001, seed=7, threads=4)
    _assert_same_build(b, r["indices"], r["distances"], r["flags"], r["changes"])
    np.testing.assert_array_equal(b[4], r["rng_state"])
    assert sum(b[3]["join_rescored"]) == 0


@pytest.mark.parametrize("k,leaf,depth", [(10, 40, 3), (20, 256, 2), (15, 60, 5)])
def test_bf16_exact_oversized_leaves(ctx, k, leaf, depth):
    """A depth limit leaves leaves above leaf_size (and above the survivor kernel's 256 members): in the
    survivor trees the push pass evaluates those leaves itself, in the reference's order, and the build still
    equals the oracle's, duplicate rows included."""
    rng = np.random.default_rng(k)
    X = _integer_mixture(rng, 3000, 32, 12, 4)
    X[rng.choice(3000, 300, replace=False)] = X[5]  # 300 copies of one row
    b = _build(ctx, X, k, 7, leaf, k, 4, "reference", "bf16", max_depth=depth)
    r = O.nn_descent(X, k, 7, leaf, k, 4, 0.001, max_depth=depth, seed=7, threads=4)
    _assert_same_build(b, r["indices"], r["distances"], r["flags"], r["changes"])
    np.testing.assert_array_equal(b[4], r["rng_state"])


def test_bf16_exact_many_survivor_trees(ctx):
    """Eighteen survivor trees: several groups (a full one and a partial last one), each replayed by one push
    pass against the roots at its start."""
    X = _integer_mixture(np.random.default_rng(3), 3000, 64, 20, 5)
    b = _build(ctx, X, 12, 21, 30, 12, 3, "reference", "bf16")
    r = O.nn_descent(X, 12, 21, 30, 12, 3, 0.001, seed=7, threads=4)
    _assert_same_build(b, r["indices"], r["distances"], r["flags"], r["changes"])
    np.testing.assert_array_equal(b[4], r["rng_state"])


@pytest.mark.parametrize("leaf,k", [(256, 15), (100, 30), (17, 8)])
def test_bf16_exact_counter_mode(ctx, leaf, k):
    """Counter draws (no reference counterpart for the draws): the BF16 build equals the FP32 build, and every
    stored distance is the oracle's FP32 distance of its pair, with distinct ids and no self."""
    X = _integer_mixture(np.random.default_rng(leaf), 6000, 128, 30, 6)
    a = _build(ctx, X, k, 7, leaf, k, 4, "counter", "fp32")
    b = _build(ctx, X, k, 7, leaf, k, 4, "counter", "bf16")
    _assert_same_build(b, a[0], a[1], a[2], a[3]["changes"])
    i, j = np.nonzero(b[0] >= 0)
    np.testing.assert_array_equal(b[1][i, j].view(np.uint32),
                                  O.pair_distances("sqeuclidean", X, i.astype(np.int32), b[0][i, j]).view(np.uint32))
    assert not np.any(b[0] == np.arange(len(X))[:, None])
    assert all(len(set(r[r >= 0].tolist())) == int((r >= 0).sum()) for r in b[0])


@pytest.mark.parametrize("kind", ["integer", "float"])
@pytest.mark.parametrize("C_", [10, 15, 30])
def test_bf16_join_matches_oracle(ctx, kind, C_):
    """The BF16-MFMA join with FP32 re-evaluation inside the proven margin reproduces the oracle's heaps bit for
    bit; exact (integer) data need no re-evaluation, float data re-evaluate a fraction of the pairs."""
    rng = np.random.default_rng(C_)
    n, d = 4000, 96 if C_ != 30 else 160
    X = _integer_mixture(rng, n, d, 20, 5) if kind == "integer" else rng.random((n, d), dtype=np.float32)
    k = C_
    b = _build(ctx, X, k, 3, 40, C_, 5, "reference", "bf16", seed=42)
    r = O.nn_descent(X, k, 3, 40, C_, 5, 0.001, seed=42, threads=4)
    _assert_same_build(b, r["indices"], r["distances"], r["flags"], r["changes"])
    s = b[3]
    if kind == "integer":
        assert sum(s["join_rescored"]) == 0
    else:
        assert 0 < sum(s["join_rescored"]) < sum(s["join_pairs"])

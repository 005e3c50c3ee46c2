"""Pins the oracle (the CPU restatement of the reference) before it is used as the GPU checker.

Checks, per part:
  - the reference's own unit-test values (file:line cited at each test);
  - published vectors of the third-party algorithms (SplitMix64, xoshiro256++);
  - independent re-derivations where those leave the arithmetic open: a numpy emulation of the
    AVX2 lane/horizontal-sum order, exact integer closed forms, brute force on small inputs,
    and the order argument (results independent of the Rayon thread count T).
No reference test pins a FastRng stream, a split choice or an end-to-end graph: those parts are
pinned only through the published algorithms they are built from.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

FMAX = np.float32(3.4028235e38)


# ---------------------------------------------------------------- FastRng (core/rng.rs:15-80)
def test_splitmix64_published_vector():
    # SplitMix64 from state 0 (Vigna's reference splitmix64.c): seed_from_u64(0) state words
    st = O.rng_state(0)
    assert [int(x) for x in st] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                                    0xF88BB8A8724C81EC]


def test_xoshiro256pp_published_vector():
    # xoshiro256++ from state {1,2,3,4} (rand_xoshiro test vector, from the reference C code)
    st = np.array([1, 2, 3, 4], np.uint64)
    got = [int(x) for x in O.rng_u64(st, 10)]
    assert got == [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
                   9973669472204895162, 14011001112246962877, 12406186145184390807, 15849039046786891736,
                   10450023813501588000]
    assert got[0] == ((1 + 4) << 23) + 1  # rotl(s0 + s3, 23) + s0 by hand


def test_rand_conversions_structure():
    a, b, c, d = (O.rng_state(7) for _ in range(4))
    u = O.rng_u64(a, 4096)
    f = O.rng_float(b, 4096)
    bl = O.rng_bool(c, 4096)
    # f32: 24 high bits of the u32 half -> exactly representable multiples of 2^-24 in [0, 1)
    assert np.all((f >= 0) & (f < 1))
    assert np.all(f * np.float32(2 ** 24) == np.floor(f * np.float32(2 ** 24)))
    np.testing.assert_array_equal((f * np.float32(2 ** 24)).astype(np.uint64), u >> np.uint64(40))
    # bool: sign of the u32 == top bit of the u64
    np.testing.assert_array_equal(bl, (u >> np.uint64(63)) == 1)
    # gen_range(0..2^k): zone = (2^k << lz(2^k)) - 1 = 2^63 - 1, so draws whose bit 63-k is set are
    # rejected; accepted draws return the top k bits
    for k in (1, 5, 20):
        e = O.rng_state(7)
        idx = O.rng_index(e, 1 << k, 1000)
        acc = [int(v) >> (64 - k) for v in u if not (int(v) >> (63 - k)) & 1][:1000]
        assert list(map(int, idx)) == acc


def test_next_index_range_and_rejection():
    st = O.rng_state(3)
    for n in (1, 2, 3, 7, 1000, (1 << 40) + 3):
        v = O.rng_index(st, n, 2000)
        assert np.all(v < n)
    # range 3: zone = 0xBFFF...F; accepted draws give floor(v*3 / 2^64)
    s1, s2 = O.rng_state(11), O.rng_state(11)
    raw = O.rng_u64(s1, 4000)
    got = O.rng_index(s2, 3, 1000)
    zone = (3 << 62) - 1
    acc = [(int(v) * 3) >> 64 for v in raw if ((int(v) * 3) & ((1 << 64) - 1)) <= zone][:1000]
    assert list(map(int, got)) == acc


# ---------------------------------------------------------------- distances
def test_reference_distance_kats():
    d = O.distance
    assert abs(d("sqeuclidean", [1, 0, 0], [0, 1, 0]) - 2.0) < 1e-6            # euclidean.rs:188-194
    assert abs(d("sqeuclidean", [0, 0, 0], [3, 4, 0]) - 25.0) < 1e-6           # euclidean.rs:213-223
    a = np.arange(1024, dtype=np.float32) * np.float32(0.01)
    assert abs(d("sqeuclidean", a, a + np.float32(0.5)) - 256.0) < 1e-3       # euclidean.rs:226-234
    assert abs(d("cosine", [1, 2, 3], [1, 2, 3])) < 1e-6                      # cosine.rs:148-153
    assert abs(d("cosine", [1, 0, 0], [0, 1, 0]) - 1.0) < 1e-6
    assert abs(d("cosine", [1, 0, 0], [-1, 0, 0]) - 2.0) < 1e-6
    assert abs(d("cosine", [1, 2, 3], [2, 4, 6])) < 1e-6
    assert abs(d("cosine", [1, 0], [1, 1]) - (1 - np.float32(1 / math.sqrt(2)))) < 1e-5
    assert abs(d("inner_product", [1, 2, 3], [4, 5, 6]) + 32.0) < 1e-6        # inner_product.rs:150-158
    assert abs(d("inner_product", [1, 0, 0], [-1, 0, 0]) - 1.0) < 1e-6
    assert abs(d("inner_product", [3, 4], [3, 4]) + 25.0) < 1e-6
    n3 = np.float32(1 / np.sqrt(np.float32(3)))
    assert abs(d("alt_dot", [n3] * 3, [n3] * 3)) < 1e-5                        # alternatives.rs:301-307
    assert abs(d("direct_cosine", [1, 0, 0], [0, 1, 0]) - 1.0) < 1e-6         # alternatives.rs:328-336
    assert abs(d("direct_cosine", [1, 0, 0], [-1, 0, 0]) - 2.0) < 1e-6
    assert abs(d("alt_inner_product", [1, 2, 3], [1, 1, 1]) - 1 / 6) < 1e-6    # alternatives.rs:339-345
    assert d("alt_inner_product", [1, 0], [-1, 0]) == FMAX
    assert abs(O.correct("alt_cosine", 1.0) - 0.5) < 1e-6                      # alternatives.rs:348-353
    assert abs(O.correct("alt_inner_product", 0.5) + 2.0) < 1e-6              # alternatives.rs:356-362
    assert O.correct("alt_inner_product", float(FMAX)) == 0.0


def _fma32(a, b, c):
    # a*b is exact in float64; one float64 rounding of (a*b + c) then float32 rounding
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def _emulate_l2(a, b):
    """Independent numpy restatement of l2_sqr_avx2's lane and horizontal-sum order."""
    n = len(a)
    acc = [np.float32(0)] * 8
    for c in range(n // 8):
        for l in range(8):
            t = np.float32(a[c * 8 + l] - b[c * 8 + l])
            acc[l] = _fma32(t, t, acc[l])
    s = [np.float32(acc[i] + acc[i + 4]) for i in range(4)]
    r = np.float32(np.float32(s[0] + s[1]) + np.float32(s[2] + s[3]))
    for i in range(n // 8 * 8, n):
        t = np.float32(a[i] - b[i])
        r = np.float32(r + np.float32(t * t))
    return r


def _emulate_normdot(a, b):
    n = len(a)
    s = [[np.float32(0)] * 8 for _ in range(4)]
    for c in range(n // 32):
        for q in range(4):
            for l in range(8):
                i = c * 32 + q * 8 + l
                s[q][l] = _fma32(a[i], b[i], s[q][l])
    v = [np.float32(np.float32(s[0][l] + s[1][l]) + np.float32(s[2][l] + s[3][l])) for l in range(8)]
    t = [np.float32(v[i] + v[i + 4]) for i in range(4)]
    r = np.float32(np.float32(t[0] + t[2]) + np.float32(t[1] + t[3]))
    for i in range(n // 32 * 32, n):
        r = np.float32(r + np.float32(a[i] * b[i]))
    return r


@pytest.mark.parametrize("dim", [1, 7, 8, 13, 31, 32, 33, 64, 100, 128])
def test_avx2_order_emulation(dim):
    rng = np.random.default_rng(dim)
    for _ in range(40):
        a = rng.standard_normal(dim).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        assert O.distance("sqeuclidean", a, b) == _emulate_l2(a, b)
        an, bn = a / np.linalg.norm(a), b / np.linalg.norm(b)
        dot = _emulate_normdot(an.astype(np.float32), bn.astype(np.float32))
        want = 1.0 - min(max(float(dot), -1.0), 1.0)
        assert O.distance("direct_cosine", an, bn) == np.float32(want)


def test_integer_data_is_order_free():
    # integer coordinates 0..255 in d=128: every partial sum is an integer < 2^24, so any order is exact
    rng = np.random.default_rng(5)
    X = rng.integers(0, 256, (200, 128)).astype(np.float32)
    a = rng.integers(0, 200, 500).astype(np.int32)
    b = rng.integers(0, 200, 500).astype(np.int32)
    got = O.pair_distances("sqeuclidean", X, a, b)
    exact = ((X[a].astype(np.int64) - X[b].astype(np.int64)) ** 2).sum(1)
    np.testing.assert_array_equal(got.astype(np.int64), exact)


def test_alternative_metrics_against_closed_forms():
    rng = np.random.default_rng(1)
    for dim in (1, 7, 8, 31, 32, 100, 128):  # alternatives.rs:309-325 remainders
        a = np.sin(np.arange(dim) * 13 + 5).astype(np.float32)
        b = np.cos(np.arange(dim) * 7 + 3).astype(np.float32)
        dot = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
        got = O.distance("alt_dot", a, b)
        if dot <= 0:
            assert got == FMAX
        else:
            assert abs(got + math.log2(dot)) < 1e-4
    a = rng.random(40).astype(np.float32)
    b = rng.random(40).astype(np.float32)
    want = math.log2(math.sqrt(float((a.astype(np.float64) ** 2).sum() * (b.astype(np.float64) ** 2).sum()))
                     / float(np.dot(a.astype(np.float64), b.astype(np.float64))))
    assert abs(O.distance("alt_cosine", a, b) - want) < 1e-5
    assert O.distance("alt_cosine", np.zeros(4), np.zeros(4)) == 0.0
    assert O.distance("alt_cosine", np.zeros(4), np.ones(4)) == FMAX


# ---------------------------------------------------------------- heaps
def test_deheap_sort_kat():
    h = O.Heap(5)  # neighbor_heap.rs:405-420
    for i, d in [(1, 0.5), (2, 0.2), (3, 0.8), (4, 0.1), (5, 0.3)]:
        h.simple_push(i, d)
    idx, dist, flags = h.idx[None].copy(), h.dist[None].copy(), h.flags[None].copy()
    O.sort_all(idx, dist, flags)
    assert list(idx[0]) == [4, 2, 5, 1, 3]


def test_heap_push_kats():
    h = O.Heap(3)  # neighbor_heap.rs:358-376
    assert h.simple_push(5, 1.0) and h.simple_push(3, 0.5) and h.simple_push(7, 0.8)
    assert h.dist[0] == np.float32(1.0)
    assert h.simple_push(9, 0.3)
    assert abs(h.dist[0] - 0.8) < 1e-6
    assert not h.simple_push(11, 2.0)
    h = O.Heap(3)  # checked push rejects duplicates (neighbor_heap.rs:379-391)
    assert h.checked_flagged_push(5, 1.0, 1) and h.checked_flagged_push(3, 0.5, 0)
    assert not h.checked_flagged_push(5, 0.1, 1)
    assert not h.checked_flagged_push(3, 0.3, 1)
    h = O.Heap(3)  # flags (neighbor_heap.rs:393-403)
    h.checked_flagged_push(5, 1.0, True)
    h.checked_flagged_push(3, 0.5, False)
    h.checked_flagged_push(7, 0.8, True)
    assert int(h.flags.sum()) == 2


def test_heap_replay_is_k_smallest():
    # brute force: sequential pushes keep the k smallest distinct ids, max at the root
    rng = np.random.default_rng(2)
    for _ in range(200):
        k = int(rng.integers(1, 12))
        h = O.Heap(k)
        seen = {}
        for _ in range(int(rng.integers(0, 60))):
            x = int(rng.integers(0, 30))
            d = seen.setdefault(x, float(np.float32(rng.random())))
            h.checked_flagged_push(x, d, True)
        got = sorted((float(d), int(i)) for d, i in zip(h.dist, h.idx) if i >= 0)
        want = sorted((d, i) for i, d in seen.items())[:k]
        assert [g[0] for g in got] == [w[0] for w in want]
        for i in range(k):
            for c in (2 * i + 1, 2 * i + 2):
                if c < k:
                    assert h.dist[i] >= h.dist[c]


def test_candidate_push_kats():
    pri = np.full(3, np.inf, np.float32)  # candidates.rs:490-505
    idx = np.full(3, -1, np.int32)
    for p, x in [(0.5, 1), (0.3, 2), (0.7, 3)]:
        O.cand_push(pri, idx, p, x)
    assert set(idx) == {1, 2, 3}
    pri = np.full(3, np.inf, np.float32)  # duplicates (507-518)
    idx = np.full(3, -1, np.int32)
    for p in (0.5, 0.3, 0.7):
        O.cand_push(pri, idx, p, 1)
    assert list(idx).count(1) == 1
    pri = np.full(3, np.inf, np.float32)  # overflow (520-534)
    idx = np.full(3, -1, np.int32)
    for p, x in [(0.9, 1), (0.8, 2), (0.7, 3), (0.5, 4)]:
        O.cand_push(pri, idx, p, x)
    assert 4 in idx and 1 not in idx


def test_bounded_heap_kats():
    oi, od, ins = O.bounded_heap(3, [0.5, 0.2, 0.8, 0.1, 0.9], [1, 2, 3, 4, 5])  # candidate_heap.rs:198-215
    assert list(ins) == [True, True, True, True, False]
    assert abs(od[-1] - 0.5) < 1e-6
    oi, od, _ = O.bounded_heap(5, [0.5, 0.2, 0.8, 0.1, 0.3], [5, 2, 8, 1, 3])  # 218-233
    assert list(oi) == [1, 2, 3, 5, 8]
    oi, od, _ = O.bounded_heap(3, [0.5, 0.6, 0.7, 0.1, 0.2, 0.3], [1, 2, 3, 4, 5, 6])  # 236-255
    assert list(oi) == [4, 5, 6]


def test_bounded_heap_tie_layout():
    # equal distances: into_sorted keeps BinaryHeap layout order among ties (stable sort by distance)
    oi, od, _ = O.bounded_heap(4, [1.0, 1.0, 1.0, 1.0], [3, 1, 2, 0])
    # layout after pushes of (1,3),(1,1),(1,2),(1,0) with (d, id) max-heap order: [3, 1, 2, 0]
    assert list(oi) == [3, 1, 2, 0]


# ---------------------------------------------------------------- RP trees
def _test_tree():
    return O.forest_from_arrays([[1.0, 0.0], [0.0, 0.0], [0.0, 0.0]], [0.0, 0.0, 0.0],
                                [[1, 2], [0, -2], [-2, -4]], [0, 1, 2, 3])


def test_flat_tree_kats():
    f = _test_tree()  # flat_tree.rs:408-478
    st = O.rng_state(42)
    assert f.search_leaves(0, [-1.0, 0.0], st, 1)[0] == [(0, 2)]
    assert f.search_leaves(0, [1.0, 0.0], st, 1)[0] == [(2, 4)]
    leaves, ev = f.search_leaves(0, [1.0, 0.0], O.rng_state(42), 2)
    assert leaves == [(2, 4), (0, 2)] and ev == 1
    assert sorted(f.leaves(0)) == [(0, 2), (2, 4)]


def test_tree_builder_invariants():
    # builder.rs:367-430: every point in exactly one leaf; leaves partition tree.indices
    rng = np.random.default_rng(0)
    for n, d, leaf in [(6, 2, 2), (50, 2, 10), (10, 2, 3), (1000, 16, 20)]:
        X = rng.random((n, d), dtype=np.float32)
        f = O.build_forest(X, 3, leaf, O.rng_state(42), threads=2)
        for t in range(3):
            tr = f.tree(t)
            assert sorted(tr["indices"]) == list(range(n))
            lv = f.leaves(t)
            assert lv[0][0] == 0 and lv[-1][1] == n
            assert all(a[1] == b[0] for a, b in zip(lv, lv[1:]))
            assert all(e - s <= leaf for s, e in lv)


def test_hyperplane_closed_form():
    # the split of the root is the normalised bisector of the two chosen points
    rng = np.random.default_rng(4)
    X = rng.random((200, 8), dtype=np.float32)
    st = O.rng_state(9)
    f = O.build_tree(X, 20, st.copy())
    hp = f.hyperplanes(0, 8)[0]
    off = f.tree(0)["offsets"][0]
    assert abs(np.linalg.norm(hp) - 1) < 1e-5
    st2 = st.copy()
    i1 = int(O.rng_index(st2, 200, 1)[0])
    i2 = int(O.rng_index(st2, 200, 1)[0])
    while i2 == i1:
        i2 = int(O.rng_index(st2, 200, 1)[0])
    h = (X[i2] - X[i1]).astype(np.float64)
    np.testing.assert_allclose(hp, h / np.linalg.norm(h), atol=1e-6)
    mid = (X[i1] + X[i2]).astype(np.float64) / 2
    assert abs(off + np.dot(h / np.linalg.norm(h), mid)) < 1e-5


# ---------------------------------------------------------------- NN-Descent
def _uniform(n, d, seed):
    return np.random.default_rng(seed).random((n, d), dtype=np.float32)


def test_update_iteration_independent_of_threads():
    # the order argument: per-vertex pushes in generation order regardless of T and row blocks
    X = _uniform(3000, 16, 1)
    r = O.nn_descent(X, 10, 3, 30, 10, 1, threads=1)
    heap0 = (r["indices"].copy(), r["distances"].copy(), r["flags"].copy())
    new, old = O.build_candidates(heap0, 10, O.rng_state(5), threads=1)
    outs = []
    for T in (1, 3, 8):
        h = tuple(a.copy() for a in heap0)
        ch = O.update_iteration(h, new, old, X, threads=T)
        outs.append((ch, h))
    for ch, h in outs[1:]:
        assert ch == outs[0][0]
        for a, b in zip(h, outs[0][1]):
            np.testing.assert_array_equal(a, b)


def test_init_from_leaves_independent_of_threads():
    X = _uniform(2500, 8, 2)
    f = O.build_forest(X, 4, 25, O.rng_state(1))
    res = []
    for T in (1, 4, 8):
        h = O.empty_heap(2500, 12)
        O.init_from_leaves(h, f, X, threads=T)
        res.append(h)
    for h in res[1:]:
        for a, b in zip(h, res[0]):
            np.testing.assert_array_equal(a, b)


def test_candidates_sequential_path_layout():
    # T=1: two draws per valid edge from the main stream, forward then reverse (candidates.rs:84-131)
    X = _uniform(300, 4, 3)
    r = O.nn_descent(X, 5, 2, 10, 5, 1, threads=1)
    heap = (r["indices"].copy(), r["distances"].copy(), r["flags"].copy())
    st = O.rng_state(77)
    new, old = O.build_candidates(tuple(a.copy() for a in heap), 5, st.copy(), threads=1)
    # python replay
    n, k = heap[0].shape
    pri = O.rng_float(st.copy(), 2 * n * k)
    npri = np.full((n, 5), np.inf, np.float32)
    opri = np.full((n, 5), np.inf, np.float32)
    nidx = np.full((n, 5), -1, np.int32)
    oidx = np.full((n, 5), -1, np.int32)
    r_ = 0
    for i in range(n):
        for j in range(k):
            nb = heap[0][i, j]
            if nb < 0:
                continue
            isnew = heap[2][i, j] != 0
            P, I = (npri, nidx) if isnew else (opri, oidx)
            O.cand_push(P[i], I[i], pri[2 * r_], nb)
            O.cand_push(P[nb], I[nb], pri[2 * r_ + 1], i)
            r_ += 1
    np.testing.assert_array_equal(new, nidx)
    np.testing.assert_array_equal(old, oidx)


def test_nn_descent_invariants_and_recall():
    X = _uniform(1500, 12, 4)  # nndescent/mod.rs:508-568
    r = O.nn_descent(X, 10, 4, 30, 10, 10, threads=4)
    idx, dist = r["indices"], r["distances"]
    assert np.all(idx >= 0)
    assert np.all(np.isfinite(dist))
    assert not np.any(idx == np.arange(len(X))[:, None])
    D = ((X[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    true = np.argsort(D, axis=1)[:, :10]
    rec = np.mean([len(set(idx[i]) & set(true[i])) / 10 for i in range(len(X))])
    assert rec > 0.97
    # heap distances are the reference distances of their ids
    i, j = np.nonzero(idx >= 0)
    np.testing.assert_array_equal(dist[i, j], O.pair_distances("sqeuclidean", X, i.astype(np.int32), idx[i, j]))


# ---------------------------------------------------------------- search graph + index
def test_weighted_union_kats():
    # search_graph.rs:628-646
    fwd = [[(4, 0.1), (1, 0.4), (2, 0.3)], [], [], [], []]
    rev = [[(3, 0.2), (4, 0.9)], [], [], [], []]
    indptr, indices = O.weighted_union(fwd, rev, 2)
    assert list(indices[indptr[0]:indptr[1]]) == [2, 3]
    # search_graph.rs:648-666
    fwd = [[(0, 0.0), (-1, 0.1), (3, 0.2), (2, 0.2)], [], [], []]
    rev = [[(1, float("inf"))], [], [], []]
    indptr, indices = O.weighted_union(fwd, rev, 1)
    assert list(indices) == [2]


def test_search_graph_invariants():
    X = _uniform(800, 8, 6)
    r = O.nn_descent(X, 10, 3, 30, 10, 6, threads=2)
    i, d, f = r["indices"].copy(), r["distances"].copy(), r["flags"].copy()
    O.sort_all(i, d, f)
    indptr, indices, md = O.search_graph(i, d, X, mult=1.5)
    maxdeg = 15
    for v in range(len(X)):
        row = indices[indptr[v]:indptr[v + 1]]
        assert len(row) <= maxdeg
        assert np.all(np.diff(row) > 0)
        assert v not in row
    assert md == d[d > 0].min()


def test_index_queries_find_themselves_and_single_equals_batch():
    X = np.sin(np.arange(100 * 10, dtype=np.float32) * np.float32(0.1)).reshape(100, 10).astype(np.float32)
    ix = O.Index(X, n_neighbors=10, n_trees=2, n_iters=3, correction="sqrt", n_search_trees=2)
    vo = ix.vertex_order
    assert sorted(vo) == list(range(100))
    f = ix.forest()
    assert f.tree(0)["indices"].tolist() == list(range(100))  # index.rs:835-836
    qi, qd = ix.query(X[:5], k=5)
    assert qd[0, 0] < 1.0  # index.rs:780-788
    for q in (37,):
        i1, d1 = ix.query_one(X[q], k=10)
        assert q in i1
        bi, bd = ix.query(X[q][None], k=10)
        np.testing.assert_array_equal(i1, bi[0])

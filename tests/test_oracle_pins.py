"""Pins of the oracle's search-graph diversification, greedy search and parallel candidate sampling.

Each case is worked out by hand (small 1-D / 2-D point sets whose squared distances are exact in float32) so
that a plausible mistake in the restatement fails it: a wrong occlusion rule (d(c,j) < d(i,c), or <=), a
dropped MIN_POSITIVE guard, break vs skip on -1 entries, a missing max_degree cut
(core/graph/search_graph.rs:432-523); a wrong epsilon bound (max + eps*max), <= instead of <, stopping
before the result heap is full, drawing fill_random ids until full instead of k - |seeds| times
(core/search/greedy.rs:148-201); and the push order of the Rayon candidate path
(core/nndescent/candidates.rs:153-351: per-thread seeds, own-block reverse pushes immediate, other blocks'
buckets applied after, in source-thread order).
"""
import numpy as np
import pytest

from oracle import oracle as O

F32_MIN_POSITIVE = np.float32(1.17549435e-38)


def pts(*xy):
    return np.array(xy, np.float32).reshape(len(xy), -1)


# ---------------------------------------------------------------- diversify_row / diversify_sorted_csr
def test_diversify_occlusion_rule():
    # i = (0,0) is point 0; its sorted row: 1 at (1,0) d=1, 2 at (-1.5,0) d=2.25, 3 at (2,0) d=4.
    # 2 is kept (d(1,2) = 6.25 >= 2.25); 3 is occluded by 1 (d(1,3) = 1 < 4).  The transposed rule
    # d(c,j) < d(i,c) would keep 3 (1 < 1 is false).
    X = pts((0, 0), (1, 0), (-1.5, 0), (2, 0))
    ids, ds = O.diversify([1, 2, 3], [1.0, 2.25, 4.0], X, max_degree=10)
    assert ids.tolist() == [1, 2]
    assert ds.tolist() == [1.0, 2.25]


def test_diversify_is_strict():
    # c = (1,0) kept with d=1; j = (0.5,1): d(i,j) = 1.25 = d(c,j) -> not occluded (strict <)
    X = pts((0, 0), (1, 0), (0.5, 1))
    ids, _ = O.diversify([1, 2], [1.0, 1.25], X, max_degree=10)
    assert ids.tolist() == [1, 2]
    # moved a little closer to c, it is occluded
    X = pts((0, 0), (1, 0), (0.75, 1))  # d(i,j) = 1.5625, d(c,j) = 1.0625
    ids, _ = O.diversify([1, 2], [1.0, 1.5625], X, max_degree=10)
    assert ids.tolist() == [1]


@pytest.mark.parametrize("d_ic,kept", [(0.0, [1, 2]), (1e-39, [1, 2]), (float(F32_MIN_POSITIVE), [1, 2]),
                                       (2e-38, [1])])
def test_diversify_min_positive_guard(d_ic, kept):
    # the row's own distances decide the guard: c's stored d_ic <= MIN_POSITIVE never occludes, even though
    # the data put c at distance 0.25 from j (d_cj = 0.25 < d_ij = 1)
    X = pts((0, 0), (5, 0), (4.5, 0))
    ids, _ = O.diversify([1, 2], [np.float32(d_ic), 1.0], X, max_degree=10)
    assert ids.tolist() == kept


def test_diversify_negative_entries_break_forward_and_skip_reverse():
    X = pts((0, 0), (1, 0), (0, 3))  # 2 is not occluded by 1: d(1,2) = 10 >= 9
    fwd, _ = O.diversify([1, -1, 2], [1.0, np.inf, 9.0], X, max_degree=10, stop_at_neg=True)
    assert fwd.tolist() == [1]
    rev, _ = O.diversify([1, -1, 2], [1.0, np.inf, 9.0], X, max_degree=10, order=[0, 1, 2], stop_at_neg=False)
    assert rev.tolist() == [1, 2]


def test_diversify_max_degree_and_order():
    # five points on different axes, none occludes another; the cut keeps the first max_degree in the given order
    X = pts((0, 0), (1, 0), (0, 1.5), (-2, 0), (0, -2.5), (3, 3))
    row = [1, 2, 3, 4, 5]
    dist = [1.0, 2.25, 4.0, 6.25, 18.0]
    ids, _ = O.diversify(row, dist, X, max_degree=3)
    assert ids.tolist() == [1, 2, 3]
    ids, ds = O.diversify(row, dist, X, max_degree=3, order=[3, 0, 2, 1, 4], stop_at_neg=False)
    assert ids.tolist() == [4, 1, 3]
    assert ds.tolist() == [6.25, 1.0, 4.0]


def test_from_dense_diversified_composition():
    """search_graph.rs:263-360 from the pinned pieces: forward rows diversified in row order, transpose_csr in
    source-row order (526-567), reverse rows ordered by distance then diversified without the -1 break,
    weighted union (pinned by search_graph.rs:573-667), min_distance over the forward edges (0 -> EPSILON)."""
    rng = np.random.default_rng(4)
    n, k, dim = 300, 8, 3
    X = rng.random((n, dim), dtype=np.float32)
    X[7] = X[3]  # a duplicate: a zero forward distance
    r = O.nn_descent(X, k, 3, 20, k, 6, threads=2)
    i, d, f = r["indices"].copy(), r["distances"].copy(), r["flags"].copy()
    O.sort_all(i, d, f)
    maxdeg = int(np.round(np.float32(1.5) * np.float32(k)))
    fwd = [list(zip(*O.diversify(i[v], d[v], X, maxdeg))) for v in range(n)]
    rev_rows = [[] for _ in range(n)]
    for v in range(n):
        for (j, dj) in fwd[v]:
            rev_rows[j].append((v, dj))
    rev = []
    for v in range(n):
        row = rev_rows[v]
        if not row:
            rev.append([])
            continue
        ids = np.array([e[0] for e in row], np.int32)
        ds = np.array([e[1] for e in row], np.float32)
        order = np.argsort(ds, kind="stable")  # no exact ties among these random distances
        rev.append(list(zip(*O.diversify(ids, ds, X, maxdeg, order=order, stop_at_neg=False))))
    want_ptr, want_idx = O.weighted_union([[(int(a), float(b)) for a, b in row] for row in fwd],
                                          [[(int(a), float(b)) for a, b in row] for row in rev], maxdeg)
    got_ptr, got_idx, md = O.search_graph(i, d, X, mult=1.5, threads=3)
    np.testing.assert_array_equal(got_ptr, want_ptr)
    np.testing.assert_array_equal(got_idx, want_idx)
    fwd_d = np.array([b for row in fwd for _, b in row], np.float32)
    assert np.any(fwd_d == 0)
    assert md == np.float32(np.finfo(np.float32).eps)


# ---------------------------------------------------------------- greedy search (greedy.rs:148-201)
def csr(n, edges):
    rows = [[] for _ in range(n)]
    for a, b in edges:
        rows[a].append(b)
    indptr = np.zeros(n + 1, np.int32)
    for v in range(n):
        indptr[v + 1] = indptr[v] + len(rows[v])
    return indptr, np.array([b for r in rows for b in r] or [0], np.int32)


# query (0,0); A=(2,0) d=4 is the only seed (k=1: no random fill); A -> B -> C with C=(0.5,0) d=0.25.
# B is pushed (and later expanded) only if d_B < bound = 4 + eps*(4 - min_distance).
@pytest.mark.parametrize("B,eps,min_d,found_c", [
    ((2.1, 0), 0.2, 1.0, True),    # bound 4.6 > 4.41
    ((2.1, 0), 0.2, 2.5, False),   # bound 4.3 < 4.41 (max + eps*max = 4.8 would find C)
    ((1.5, 1.5), 0.25, 2.0, False),  # bound 4.5 == d_B: not pushed (strict <)
    ((1.5, 1.4), 0.25, 2.0, True),   # d_B = 4.21 < 4.5
])
def test_greedy_epsilon_bound_decides_expansion(B, eps, min_d, found_c):
    X = pts((2, 0), B, (0.5, 0))
    indptr, indices = csr(3, [(0, 1), (1, 2)])
    st = O.rng_state(1)
    ids, ds = O.greedy(np.zeros(2, np.float32), X, indptr, indices, [0], 1, eps, min_d, st)
    assert ids.tolist() == ([2] if found_c else [0])


def test_greedy_stops_only_when_the_heap_is_full():
    # 1-D: A=1 (d=1), B=3 (d=9), C=0.5 (d=0.25); B -> C.  min_distance 20 makes the bound 9 + 0.1*(9-20) = 7.9,
    # below B's distance.  k=2: the heap {A,B} is full and B (9 >= 7.9) ends the search.  k=3: one random fill
    # draw (an already visited id for this seed) leaves the heap at two entries, so B is still expanded.
    X = pts((1,), (3,), (0.5,))
    indptr, indices = csr(3, [(1, 2)])
    q = np.zeros(1, np.float32)
    ids, _ = O.greedy(q, X, indptr, indices, [0, 1], 2, 0.1, 20.0, O.rng_state(1))
    assert ids.tolist() == [0, 1]
    seed = next(s for s in range(1000) if int(O.rng_index(O.rng_state(s), 3, 1)[0]) in (0, 1))
    ids, ds = O.greedy(q, X, indptr, indices, [0, 1], 3, 0.1, 20.0, O.rng_state(seed))
    assert ids.tolist() == [2, 0, 1]
    assert ds.tolist() == [0.25, 1.0, 9.0]


def test_greedy_push_is_strictly_below_the_bound():
    # A=(2,0) d=4 is the only seed, k=3 and both random fill draws hit A (chosen seed): the heap holds {A} only,
    # so the bound is 4 + 0.25*(4-2) = 4.5.  B=(1.5,1.5) has d = 4.5 exactly: not pushed.  (With a full heap the
    # difference between < and <= cannot show: an equal-distance vertex only enters the frontier, and popping it
    # ends the search.)
    X = pts((2, 0), (1.5, 1.5))
    indptr, indices = csr(2, [(0, 1)])
    seed = next(s for s in range(1000) if O.rng_index(O.rng_state(s), 2, 2).tolist() == [0, 0])
    ids, ds = O.greedy(np.zeros(2, np.float32), X, indptr, indices, [0], 3, 0.25, 2.0, O.rng_state(seed))
    assert ids.tolist() == [0, -1, -1]
    X = pts((2, 0), (1.5, 1.4))  # d_B = 4.21
    ids, ds = O.greedy(np.zeros(2, np.float32), X, indptr, indices, [0], 3, 0.25, 2.0, O.rng_state(seed))
    assert ids.tolist() == [0, 1, -1]


def test_fill_random_draws_k_minus_seeds_ids():
    # ten 1-D points x = 1..10, no edges: the result is the seed plus the fill draws (visited ones skipped,
    # not redrawn), sorted by distance; the query's stream advances by exactly k - |seeds| next_index draws
    X = np.arange(1, 11, dtype=np.float32)[:, None]
    indptr, indices = csr(10, [])
    k = 4

    def draws(s):
        return [int(v) for v in O.rng_index(O.rng_state(s), 10, k - 1)]
    seed = next(s for s in range(5000) if len(set(draws(s)) | {5}) < k)  # a repeated or already-seen draw
    st = O.rng_state(seed)
    ids, ds = O.greedy(np.zeros(1, np.float32), X, indptr, indices, [5], k, 0.1, 1.0, st)
    want = sorted(set(draws(seed)) | {5})
    assert ids.tolist() == want + [-1] * (k - len(want))
    assert np.all(np.isinf(ds[len(want):]))
    after = O.rng_state(seed)
    O.rng_index(after, 10, k - 1)
    np.testing.assert_array_equal(st, after)


def test_greedy_large_epsilon_is_exact_on_a_connected_graph():
    rng = np.random.default_rng(2)
    n, k = 400, 10
    X = rng.random((n, 3), dtype=np.float32)
    edges = [(v, (v + 1) % n) for v in range(n)] + [(v, int(u)) for v in range(n) for u in rng.integers(0, n, 3)]
    indptr, indices = csr(n, edges)
    for qi in range(5):
        q = rng.random(3, dtype=np.float32)
        ids, ds = O.greedy(q, X, indptr, indices, [int(rng.integers(0, n))], k, 1e6, 0.0, O.rng_state(qi))
        D = ((X.astype(np.float64) - q.astype(np.float64)) ** 2).sum(1)
        assert ids.tolist() == np.argsort(D)[:k].tolist()
        assert np.all(np.diff(ds) >= 0)


# ---------------------------------------------------------------- Rayon candidate path (T > 1)
@pytest.mark.parametrize("T,n", [(3, 301), (8, 700), (5, 256)])
def test_candidates_parallel_path_replay(T, n):
    X = np.random.default_rng(n).random((n, 4), dtype=np.float32)
    r = O.nn_descent(X, 6, 2, 12, 6, 1, threads=1)
    idx, dist, flags = r["indices"].copy(), r["distances"].copy(), r["flags"].copy()
    C_ = 5
    st = O.rng_state(99 + T)
    heap = (idx.copy(), dist.copy(), flags.copy())
    new, old = O.build_candidates(heap, C_, st.copy(), threads=T)
    # replay: T seeds from the main stream, thread t owns rows [t*bs, (t+1)*bs); one draw per valid edge shared
    # by the forward push and the reverse push; reverse pushes into the thread's own block at once, into other
    # blocks afterwards in source-thread order
    k = idx.shape[1]
    bs = (n + T - 1) // T
    seeds = O.rng_u64(st.copy(), T)
    npri = np.full((n, C_), np.inf, np.float32)
    opri = np.full((n, C_), np.inf, np.float32)
    nidx = np.full((n, C_), -1, np.int32)
    oidx = np.full((n, C_), -1, np.int32)
    buckets = [[[] for _ in range(T)] for _ in range(T)]
    for t in range(T):
        b0, b1 = t * bs, min((t + 1) * bs, n)
        if b0 >= n:
            continue
        valid = int((idx[b0:b1] >= 0).sum())
        pri = O.rng_float(O.rng_state(int(seeds[t])), valid)
        p = 0
        for i in range(b0, b1):
            for j in range(k):
                nb = int(idx[i, j])
                if nb < 0:
                    continue
                isnew = flags[i, j] != 0
                P, I = (npri, nidx) if isnew else (opri, oidx)
                O.cand_push(P[i], I[i], pri[p], nb)
                if nb // bs == t:
                    O.cand_push(P[nb], I[nb], pri[p], i)
                else:
                    buckets[t][nb // bs].append((nb, i, pri[p], isnew))
                p += 1
    for t in range(T):
        for src in range(T):
            if src == t:
                continue
            for nb, i, pr, isnew in buckets[src][t]:
                P, I = (npri, nidx) if isnew else (opri, oidx)
                O.cand_push(P[nb], I[nb], pr, i)
    np.testing.assert_array_equal(new, nidx)
    np.testing.assert_array_equal(old, oidx)
    # mark_old_flags (candidates.rs:354-417): a new entry that became a new candidate is flagged old
    want_flags = flags.copy()
    for i in range(n):
        for j in range(k):
            if idx[i, j] >= 0 and flags[i, j] and idx[i, j] in nidx[i]:
                want_flags[i, j] = 0
    np.testing.assert_array_equal(heap[2], want_flags)

"""The query service (rinnd_server_*): requests of mixed k and epsilon, submitted asynchronously (from several host
threads) and answered in shared launches, must return exactly what rinnd_index_query (rinnd_index_query_quantized
for a quantized server) returns for the same request - the request's query ids 0..nq-1 seed its random fill
(core/index.rs:157-202, 406-439).  Also: admission limits (beam memory, batch size) and the ticket protocol.
"""
import ctypes as C
import threading

import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import RINND
from rinnd_amd import _lib as L

pytestmark = pytest.mark.gpu
INVALID_ARG, UNSUPPORTED = 1, 4


def direct(g, Q, k, eps, quantized=None):
    nq = Q.shape[0]
    oi = np.zeros((nq, k), np.int32)
    od = np.zeros((nq, k), np.float32)
    if quantized is None:
        st = L.lib().rinnd_index_query(g._h, L.ptr(Q), nq, k, C.c_float(eps), L.ptr(oi), L.ptr(od))
    else:
        st = L.lib().rinnd_index_query_quantized(g._h, L.ptr(Q), nq, k, C.c_float(eps), quantized[0], quantized[1],
                                                 L.ptr(oi), L.ptr(od))
    assert st == 0
    return oi, od


@pytest.fixture(scope="module")
def setup():
    rng = np.random.default_rng(11)
    centres = rng.random((30, 40)).astype(np.float32)
    X = (centres[rng.integers(0, 30, 8000)] + 0.1 * rng.standard_normal((8000, 40))).astype(np.float32)
    Q = (centres[rng.integers(0, 30, 3000)] + 0.1 * rng.standard_normal((3000, 40))).astype(np.float32)
    g = RINND(X, metric="euclidean", n_neighbors=20, n_trees=5, n_iters=6, random_state=2)
    return X, Q, g


def test_mixed_requests_equal_direct_queries(setup):
    import torch
    X, Q, g = setup
    srv = L.Server(g._h, max_batch=1024, max_inflight=3, max_k=128)
    rng = np.random.default_rng(0)
    reqs, s0 = [], 0
    for i in range(40):
        nq = int(rng.choice([1, 3, 17, 64, 250, 700]))
        k = int(rng.choice([10, 100]))
        eps = float(rng.choice([0.05, 0.2]))
        q = Q[s0 % 2000: s0 % 2000 + nq]
        s0 += nq
        on_device = i % 3 == 0
        if on_device:
            qd = torch.from_numpy(q).cuda()
            oi = torch.full((len(q), k), -7, dtype=torch.int32, device="cuda")
            od = torch.full((len(q), k), -7.0, dtype=torch.float32, device="cuda")
        else:
            qd, oi, od = q, np.full((len(q), k), -7, np.int32), np.full((len(q), k), -7.0, np.float32)
        t = srv.submit(qd, len(q), k, eps, oi, od)
        reqs.append((t, q, k, eps, qd, oi, od))
    print(f"[serve] {len(reqs)} requests submitted", flush=True)
    for t, q, k, eps, qd, oi, od in reqs:
        srv.wait(t)
        gi = oi.cpu().numpy() if hasattr(oi, "cpu") else oi
        gd = od.cpu().numpy() if hasattr(od, "cpu") else od
        wi, wd = direct(g, q, k, eps)
        np.testing.assert_array_equal(gi, wi)
        np.testing.assert_array_equal(gd.view(np.uint32), wd.view(np.uint32))
    st = srv.stats()
    assert st["requests"] == 40 and st["queries"] == sum(len(r[1]) for r in reqs)
    assert st["launches"] < 40  # requests shared launches
    srv.close()


def test_concurrent_submitters(setup):
    X, Q, g = setup
    srv = L.Server(g._h, max_batch=512, max_inflight=2, max_k=32)
    results, errors = {}, []

    def client(cid):
        try:
            rng = np.random.default_rng(cid)
            for j in range(12):
                nq = int(rng.integers(1, 200))
                s = int(rng.integers(0, 2800))
                q = np.ascontiguousarray(Q[s:s + nq])
                oi = np.zeros((len(q), 10), np.int32)
                od = np.zeros((len(q), 10), np.float32)
                t = srv.submit(q, len(q), 10, 0.1, oi, od)
                if j % 2:
                    while not srv.poll(t):
                        pass
                else:
                    srv.wait(t)
                results[(cid, j)] = (q, oi, od)
        except Exception as e:  # reported below
            errors.append(e)
    threads = [threading.Thread(target=client, args=(c,)) for c in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for (cid, j), (q, oi, od) in results.items():
        wi, wd = direct(g, q, 10, 0.1)
        np.testing.assert_array_equal(oi, wi)
        np.testing.assert_array_equal(od.view(np.uint32), wd.view(np.uint32))
    srv.close()


def test_quantized_server_equals_quantized_query():
    """FP8 e4m3 candidates on MFMA, FP32 re-ranking of the first 32: equal to rinnd_index_query_quantized, and the
    re-ranked distances are the reference arithmetic's FP32 distances of the returned ids."""
    rng = np.random.default_rng(5)
    c = rng.standard_normal((40, 32)).astype(np.float32)
    X = c[rng.integers(0, 40, 6000)] + 0.3 * rng.standard_normal((6000, 32)).astype(np.float32)
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    Q = c[rng.integers(0, 40, 400)] + 0.3 * rng.standard_normal((400, 32)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    g = RINND(X, metric="cosine", normalize=True, cosine_distance_mode="int8", n_neighbors=20, n_trees=4, n_iters=5,
              quantized_candidate_width=32, quantized_rerank_width=32, random_state=4)
    srv = L.Server(g._h, max_batch=256, quantized=True, candidate_width=32, rerank_width=32, max_k=64)
    reqs = []
    for i, (nq, k, eps, widths) in enumerate([(50, 10, 0.05, None), (200, 10, 0.2, None), (7, 20, 0.2, (48, 40)),
                                              (143, 10, 0.05, (0, 24))]):
        q = np.ascontiguousarray(Q[i * 50:i * 50 + nq])
        oi, od = np.zeros((nq, k), np.int32), np.zeros((nq, k), np.float32)
        if widths is None:  # the server's widths
            t = srv.submit(q, nq, k, eps, oi, od)
        else:  # per-request widths, as query_quantized takes them per call
            t = srv.submit(q, nq, k, eps, oi, od, candidate_width=widths[0], rerank_width=widths[1])
        reqs.append((t, q, k, eps, oi, od, widths or (32, 32)))
    for t, q, k, eps, oi, od, widths in reqs:
        srv.wait(t)
        wi, wd = direct(g, q, k, eps, quantized=widths)
        np.testing.assert_array_equal(oi, wi)
        np.testing.assert_array_equal(od.view(np.uint32), wd.view(np.uint32))
        from rinnd_amd.api import normalize_rows
        Xn = normalize_rows(X)  # the rows the index holds (RINND normalises them); the C ABI takes q as given
        for r in range(0, len(q), 9):  # DirectNormalizedCosine in the reference's arithmetic
            exact = O.pair_distances("direct_cosine", np.concatenate([q[r:r + 1], Xn]),
                                     np.zeros(k, np.int64), oi[r].astype(np.int64) + 1)
            np.testing.assert_array_equal(od[r].view(np.uint32), exact.view(np.uint32))
    srv.close()


def test_int8_server_equals_int8_query():
    """A server over the index's int8 codes (the reference's quantized cosine index): each request, with the
    server's or its own widths, equals rinnd_index_query_i8 on the same rows; a released index refuses re-ranking
    requests and FP32 / FP8 servers."""
    rng = np.random.default_rng(6)
    c = rng.standard_normal((40, 32)).astype(np.float32)
    X = c[rng.integers(0, 40, 6000)] + 0.3 * rng.standard_normal((6000, 32)).astype(np.float32)
    Q = c[rng.integers(0, 40, 400)] + 0.3 * rng.standard_normal((400, 32)).astype(np.float32)
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    g = RINND(X, metric="cosine", normalize=True, cosine_distance_mode="sq8p", n_neighbors=20, n_trees=4, n_iters=5,
              random_state=4)
    srv = L.Server(g._h, max_batch=256, quantized="int8", max_k=64)
    reqs = []
    for i, (nq, k, eps, widths) in enumerate([(50, 10, 0.05, None), (200, 10, 0.2, None), (7, 20, 0.2, (48, 40)),
                                              (143, 10, 0.05, (0, 24))]):
        q = np.ascontiguousarray(Q[i * 50:i * 50 + nq])
        oi, od = np.zeros((nq, k), np.int32), np.zeros((nq, k), np.float32)
        if widths is None:
            t = srv.submit(q, nq, k, eps, oi, od)
        else:
            t = srv.submit(q, nq, k, eps, oi, od, candidate_width=widths[0], rerank_width=widths[1])
        reqs.append((t, q, k, eps, oi, od, widths or (0, 0)))
    for t, q, k, eps, oi, od, widths in reqs:
        srv.wait(t)
        wi, wd = np.zeros((len(q), k), np.int32), np.zeros((len(q), k), np.float32)
        assert L.lib().rinnd_index_query_i8(g._h, L.ptr(q), len(q), k, C.c_float(eps), widths[0], widths[1],
                                            L.ptr(wi), L.ptr(wd)) == 0
        np.testing.assert_array_equal(oi, wi)
        np.testing.assert_array_equal(od.view(np.uint32), wd.view(np.uint32))
    srv.close()
    assert L.lib().rinnd_index_release_data(g._h) == 0
    srv = L.Server(g._h, max_batch=64, quantized="int8", max_k=64)
    q = np.ascontiguousarray(Q[:5])
    oi, od = np.zeros((5, 10), np.int32), np.zeros((5, 10), np.float32)
    srv.wait(srv.submit(q, 5, 10, 0.1, oi, od))
    with pytest.raises(ValueError, match="FP32 vector data was released"):
        srv.submit(q, 5, 10, 0.1, oi, od, candidate_width=20, rerank_width=10)
    srv.close()
    for mode in (False, "fp8"):
        with pytest.raises(L.RinndError):
            L.Server(g._h, max_batch=64, quantized=mode)


def test_admission_limits_and_tickets(setup):
    X, Q, g = setup
    srv = L.Server(g._h, max_batch=100, max_k=50)
    q = np.ascontiguousarray(Q[:10])
    oi, od = np.zeros((10, 60), np.int32), np.zeros((10, 60), np.float32)
    with pytest.raises(L.RinndError, match="beam memory"):
        srv.submit(q, 10, 60, 0.1, oi, od)  # k above max_k: RINND_ERR_UNSUPPORTED
    big = np.ascontiguousarray(Q[:101])
    with pytest.raises(ValueError, match="max_batch"):
        srv.submit(big, 101, 10, 0.1, np.zeros((101, 10), np.int32), np.zeros((101, 10), np.float32))
    oi10, od10 = oi[:, :10].copy(), od[:, :10].copy()  # the buffers must outlive the request
    t = srv.submit(q, 10, 10, 0.1, oi10, od10)
    srv.wait(t)
    with pytest.raises(ValueError, match="unknown ticket"):
        srv.wait(t)  # released by the first wait
    t0 = srv.submit(q, 0, 10, 0.1, oi, od)  # an empty request completes at once
    assert srv.poll(t0)
    srv.close()


def test_search_width_floor(setup):
    """min_width: a request of k results searches with width max(k, min_width) and returns the first k of those
    results, i.e. the first k columns of rinnd_index_query at that width; requests wider than the floor search as
    usual.  Limits: min_width <= max_k, and FP32 servers only."""
    X, Q, g = setup
    srv = L.Server(g._h, max_batch=512, max_inflight=2, max_k=64, min_width=40)
    reqs = []
    for i, (nq, k, eps) in enumerate([(300, 10, 0.05), (200, 40, 0.2), (100, 50, 0.1), (7, 1, 0.2)]):
        q = Q[i * 500: i * 500 + nq]
        oi, od = np.full((nq, k), -7, np.int32), np.full((nq, k), -7.0, np.float32)
        reqs.append((srv.submit(q, nq, k, eps, oi, od), q, k, eps, oi, od))
    for t, q, k, eps, oi, od in reqs:
        srv.wait(t)
        wi, wd = direct(g, q, max(k, 40), eps)
        np.testing.assert_array_equal(oi, wi[:, :k])
        np.testing.assert_array_equal(od.view(np.uint32), wd[:, :k].view(np.uint32))
    srv.close()
    with pytest.raises(Exception):
        L.Server(g._h, max_batch=64, max_k=32, min_width=40)
    with pytest.raises(Exception):
        L.Server(g._h, max_batch=64, max_k=64, quantized=True, min_width=40)

"""rinnd_index_query_noalloc (the query that allocates nothing) against rinnd_index_query, through the C ABI.

The two entry points share the search kernel; the no-allocation one keeps the per-query state in a fixed
device array instead of the index's workspace, so results must be identical, bit for bit, call after call
(the array keeps the previous call's contents), with the second call smaller than the first.
"""
import numpy as np
import pytest

from rinnd_amd import RINND
from rinnd_amd import _lib as L

pytestmark = pytest.mark.gpu
INVALID_ARG, UNSUPPORTED = 1, 4


def run(entry, g, Q, k, eps):
    import torch
    nq = Q.shape[0]
    oi = torch.full((nq, k), -7, dtype=torch.int32, device="cuda")
    od = torch.full((nq, k), -7.0, dtype=torch.float32, device="cuda")
    st = getattr(L.lib(), entry)(g._h, L.ptr(Q), nq, k, L.C.c_float(eps), L.ptr(oi), L.ptr(od))
    torch.cuda.synchronize()
    return st, oi.cpu().numpy(), od.cpu().numpy()


def check_same(g, Q, k, eps):
    st_a, ia, da = run("rinnd_index_query", g, Q, k, eps)
    st_b, ib, db = run("rinnd_index_query_noalloc", g, Q, k, eps)
    assert st_a == 0 and st_b == 0, (st_a, st_b, L.lib().rinnd_last_error(g._ctx.h))
    np.testing.assert_array_equal(ib, ia)
    np.testing.assert_array_equal(db.view(np.uint32), da.view(np.uint32))


@pytest.mark.parametrize("dim", [32, 30])  # 30: query rows padded inside the device array
def test_noalloc_query_equals_query_twice(dim):
    import torch
    rng = np.random.default_rng(dim)
    X = rng.random((3000, dim), dtype=np.float32)
    g = RINND(X, n_neighbors=10, n_trees=4, n_iters=5, random_state=3)
    Q1 = torch.from_numpy(rng.random((500, dim), dtype=np.float32)).cuda()
    Q2 = torch.from_numpy(rng.random((37, dim), dtype=np.float32)).cuda()
    print(f"[no_alloc] 3000x{dim}: 500 queries, then 37", flush=True)
    check_same(g, Q1, 10, 0.1)
    check_same(g, Q2, 7, 0.2)


def test_noalloc_query_more_queries_than_slots():
    """200,000 points: 724 KB of state per query slot, so the 128 MiB array holds 185 slots and each slot runs
    about eleven of the 2,000 queries in turn (visited epochs advance within a slot)."""
    import torch
    rng = np.random.default_rng(5)
    X = rng.random((200_000, 16), dtype=np.float32)
    g = RINND(X, n_neighbors=10, n_trees=4, n_iters=5, random_state=1, rng_mode="counter")
    print("[no_alloc] 200000x16: 2000 queries, then 300", flush=True)
    check_same(g, torch.from_numpy(rng.random((2000, 16), dtype=np.float32)).cuda(), 10, 0.1)
    check_same(g, torch.from_numpy(rng.random((300, 16), dtype=np.float32)).cuda(), 10, 0.1)


def test_noalloc_query_rejections():
    import torch
    rng = np.random.default_rng(0)
    X = rng.random((1000, 30), dtype=np.float32)
    g = RINND(X, n_neighbors=8, n_trees=3, n_iters=4)
    Qh = rng.random((4, 30), dtype=np.float32)
    oi = np.zeros((4, 8), np.int32)
    od = np.zeros((4, 8), np.float32)
    assert L.lib().rinnd_index_query_noalloc(g._h, L.ptr(Qh), 4, 8, L.C.c_float(0.1), L.ptr(oi),
                                             L.ptr(od)) == INVALID_ARG  # host pointers
    # 1.2 M padded rows of 32 floats do not fit in the 128 MiB array: rejected before any work
    Qbig = torch.zeros((1_200_000, 30), dtype=torch.float32, device="cuda")
    st, _, _ = run("rinnd_index_query_noalloc", g, Qbig, 1, 0.1)
    assert st == UNSUPPORTED
    assert b"128 MiB" in L.lib().rinnd_last_error(g._ctx.h) or b"bytes" in L.lib().rinnd_last_error(g._ctx.h)
    # the entry point still works afterwards
    check_same(g, torch.from_numpy(rng.random((64, 30), dtype=np.float32)).cuda(), 8, 0.1)

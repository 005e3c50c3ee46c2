"""The headline workload (1,000,000 x 128 integer Gaussian mixture, squared Euclidean, k=15) on the GPU.

- A 50,000-row draw of the same generator, in the reference's sampling mode, at the benchmark's parameters
  (16 trees, leaf 256, <= 20 iterations), at the reference defaults for this size (12 trees,
  leaf max(60, min(256, 5k)) = 75, n_iters max(5, ceil(log2 n)) = 16) and with delta = 0 so that five
  iterations run: heap indices, distance bits, flags, per-iteration change counts and the caller's RNG state
  equal the oracle's, for T = 1 and T = 16 and for both join precisions (FP32 and BF16-MFMA).
- The full 1,000,000 rows in the GPU-native counter mode (no reference counterpart for its draws): the
  BF16 build's graph equals the FP32 build's bit for bit, every stored distance of 10,000 sampled rows equals
  the oracle's pair distance bits, ids are distinct and never the row itself, and tie-aware recall@15 against
  an FP64 brute force is >= 0.98.
"""
import ctypes as C
import functools

import numpy as np
import pytest

from oracle import oracle as O
from rinnd_amd import _lib as L
from rinnd_amd import datasets

pytestmark = pytest.mark.gpu

K = 15
PARAMS = {  # name -> (n_trees, leaf, n_iters, delta)
    "bench": (16, 256, 20, 0.001),
    "reference_defaults": (12, 75, 16, 0.001),
    "five_iterations": (12, 75, 5, 0.0),
}


@functools.lru_cache(maxsize=1)
def sample():
    return datasets.cfg2(n=50_000)


@functools.lru_cache(maxsize=None)
def oracle_run(name, T):
    trees, leaf, iters, delta = PARAMS[name]
    return O.nn_descent(sample(), K, trees, leaf, K, iters, delta, threads=T)


def gpu_build(ctx, X, trees, leaf, iters, delta, T, rng_mode, precision, seed=42, out=None):
    n, d = X.shape
    p = L.make_params(K, trees, leaf, K, iters, delta, n_threads=T, rng_mode=rng_mode, join_precision=precision)
    st = L.rng_seed(seed)
    if out is None:
        out = (np.zeros((n, K), np.int32), np.zeros((n, K), np.float32), np.zeros((n, K), np.uint8))
    stats = L.BuildStats()
    h = C.c_void_p()
    rc = L.lib().rinnd_nn_descent(ctx.h, L.ptr(X), n, d, L.DIST["sqeuclidean"], C.byref(p), L.ptr(st),
                                  L.ptr(out[0]), L.ptr(out[1]), L.ptr(out[2]), C.byref(h), C.byref(stats))
    ctx.check(rc, "rinnd_nn_descent")
    L.lib().rinnd_forest_free(h)
    return out, stats.to_dict(), st


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("T", [1, 16])
@pytest.mark.parametrize("name", list(PARAMS))
def test_headline_sample_reference_mode_bit_exact(ctx, name, T, precision):
    X = sample()
    trees, leaf, iters, delta = PARAMS[name]
    print(f"[headline] 50000x128 {name} T={T} {precision}", flush=True)
    (gi, gd, gf), s, st = gpu_build(ctx, X, trees, leaf, iters, delta, T, "reference", precision)
    r = oracle_run(name, T)
    assert s["changes"] == list(r["changes"])
    np.testing.assert_array_equal(gi, r["indices"])
    np.testing.assert_array_equal(gd.view(np.uint32), r["distances"].view(np.uint32))
    np.testing.assert_array_equal(gf, r["flags"])
    np.testing.assert_array_equal(st, r["rng_state"])
    if precision == "bf16":
        assert sum(s["join_rescored"]) == 0  # integer data: the screened BF16 distance is the FP32 distance


def test_headline_full_scale_counter_mode(ctx):
    torch = pytest.importorskip("torch")
    n = 1_000_000
    print("[headline] 1000000x128 counter mode, FP32 and BF16 builds", flush=True)
    X = datasets.cfg2(n=n)
    Xd = torch.from_numpy(X).cuda()
    outs = {}
    for prec in ("fp32", "bf16"):
        o = (torch.empty((n, K), dtype=torch.int32, device="cuda"),
             torch.empty((n, K), dtype=torch.float32, device="cuda"),
             torch.empty((n, K), dtype=torch.uint8, device="cuda"))
        _, s, _ = gpu_build(ctx, Xd, 16, 256, 20, 0.001, 8, "counter", prec, out=o)
        ctx.check(L.lib().rinnd_sort_rows(ctx.h, L.ptr(o[0]), L.ptr(o[1]), L.ptr(o[2]), n, K), "rinnd_sort_rows")
        torch.cuda.synchronize()
        outs[prec] = [t.cpu().numpy() for t in o]
        print(f"[headline] {prec}: {s['iterations']} iterations, changes {s['changes']}", flush=True)
    a, b = outs["fp32"], outs["bf16"]
    np.testing.assert_array_equal(b[0], a[0])
    np.testing.assert_array_equal(b[1].view(np.uint32), a[1].view(np.uint32))
    np.testing.assert_array_equal(b[2], a[2])
    gi, gd = b[0], b[1]
    rows = np.random.default_rng(3).choice(n, size=10_000, replace=False)
    ri = gi[rows]
    assert np.all(ri >= 0)
    assert not np.any(ri == rows[:, None])
    assert all(len(set(r.tolist())) == K for r in ri)
    src = np.repeat(rows, K).astype(np.int32)
    want = O.pair_distances("sqeuclidean", X, src, ri.reshape(-1))
    np.testing.assert_array_equal(gd[rows].reshape(-1).view(np.uint32), want.view(np.uint32))
    assert np.all(np.diff(gd[rows], axis=1) >= 0)  # sort_all
    # tie-aware recall@15 against FP64 brute force on the GPU
    X64 = Xd.double()
    sq = (X64 * X64).sum(1)
    hits = 0.0
    for s0 in range(0, len(rows), 1000):
        r = torch.as_tensor(rows[s0:s0 + 1000], device="cuda")
        D = sq[r][:, None] + sq[None, :] - 2.0 * (X64[r] @ X64.T)
        D[torch.arange(len(r), device="cuda"), r] = float("inf")
        kth = torch.topk(D, K, dim=1, largest=False).values[:, -1:]
        f = torch.as_tensor(ri[s0:s0 + 1000].astype(np.int64), device="cuda")
        df = ((X64[f] - X64[r][:, None, :]) ** 2).sum(-1)
        hits += float((df <= kth * (1 + 1e-12)).double().sum())
    recall = hits / (len(rows) * K)
    print(f"[headline] recall@15 {recall:.4f}", flush=True)
    assert recall >= 0.98

"""Synthetic workloads of the benchmark configurations (seeded numpy PCG64 streams).

cfg1  10,000 x 128 uniform[0,1) float32                                    (seed 0; queries seed 1)
cfg2  1,000,000 x 128: 1,000-centre Gaussian mixture, centres U[0,100)^128, sigma 5, rounded half
      away from zero and clipped to 0..255, stored as float32              (seed 2)
cfg3  1,000,000 x 960: 100-centre mixture, centres U[0,1]^960, sigma 0.1, clipped to [0,1]  (seed 4)
cfg4  512 indexes, n log-uniform in [2,000, 100,000], dim 64, N(0,1) rows L2-normalised  (seed 5)
cfg5  2,000,000 x 96 unit-norm rows: 500 centres N(0, I), x = centre + 0.2 N(0, I) (sigma relative to the
      centres' unit per-coordinate spread), then normalised; queries from the same mixture  (seed 6)
"""
from __future__ import annotations

import numpy as np


def cfg1(n=10_000, d=128, seed=0):
    return np.random.default_rng(seed).random((n, d), dtype=np.float32)


def cfg1_queries(nq=100, d=128, seed=1):
    return np.random.default_rng(seed).random((nq, d), dtype=np.float32)


def _round_half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def cfg2(n=1_000_000, d=128, centers=1000, sigma=5.0, seed=2, chunk=250_000):
    rng = np.random.default_rng(seed)
    c = rng.random((centers, d)) * 100.0
    labels = rng.integers(0, centers, n)
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = c[labels[s:e]] + sigma * rng.standard_normal((e - s, d))
        out[s:e] = np.clip(_round_half_away(x), 0, 255)
    return out


def cfg3(n=1_000_000, d=960, centers=100, sigma=0.1, seed=4, chunk=100_000):
    rng = np.random.default_rng(seed)
    c = rng.random((centers, d), dtype=np.float32)
    labels = rng.integers(0, centers, n)
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = c[labels[s:e]] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        out[s:e] = np.clip(x, 0.0, 1.0)
    return out


def cfg4_sizes(count=512, lo=2_000, hi=100_000, seed=5):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)


def cfg4_index(n, d=64, seed=5, i=0):
    x = np.random.default_rng([seed, i]).standard_normal((int(n), d), dtype=np.float32)
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def _mixture_unit(n, d, centers, sigma, rng, c=None, chunk=250_000):
    if c is None:
        c = rng.standard_normal((centers, d)).astype(np.float32)
    labels = rng.integers(0, centers, n)
    out = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        x = c[labels[s:e]] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        out[s:e] = x / np.linalg.norm(x, axis=1, keepdims=True)
    return out, c


def cfg5(n=2_000_000, d=96, centers=500, sigma=0.2, seed=6):
    rng = np.random.default_rng(seed)
    data, c = _mixture_unit(n, d, centers, sigma, rng)
    return data, c


def cfg5_queries(centres, nq, sigma=0.2, seed=7):
    rng = np.random.default_rng(seed)
    q, _ = _mixture_unit(nq, centres.shape[1], centres.shape[0], sigma, rng, c=centres)
    return q

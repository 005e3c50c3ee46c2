"""CPU checks of the boundary: the C-ABI library loads and exports every symbol include/rinnd_hip.h
declares (no compute calls), and the host-side Python mirror validates arguments like the reference."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rinnd_hip.h")
LIB = os.path.join(ROOT, "rinnd_amd", "librinnd_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rinnd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("rinnd_nn_descent", "rinnd_index_build", "rinnd_index_query", "rinnd_sort_rows",
              "rinnd_search_graph_build", "rinnd_build_forest", "rinnd_pair_distances"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        from rinnd_amd import build
        build.build()
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    from rinnd_amd import _lib
    assert sorted(_lib.EXPORTS) == [s for s in declared_symbols() if s in _lib.EXPORTS]
    assert set(_lib.EXPORTS) >= set(declared_symbols())


def test_host_rng_helpers_match_oracle():
    # rinnd_rng_seed / rinnd_rng_jump are host-only entry points
    from oracle import oracle as O
    from rinnd_amd import _lib
    st = _lib.rng_seed(2024)
    np.testing.assert_array_equal(st, O.rng_state(2024))
    for steps in (0, 1, 63, 64, 1000, 1 << 20, 123456789):
        ref = O.rng_state(2024)
        if steps:
            O.rng_u64(ref, steps)
        np.testing.assert_array_equal(_lib.rng_jump(st, steps), ref)


def test_python_mirror_argument_errors():
    from rinnd_amd import RINND
    X = np.arange(32 * 4, dtype=np.float32).reshape(32, 4)
    with pytest.raises(ValueError, match="Unknown metric"):
        RINND(X, metric="not-a-metric")
    with pytest.raises(ValueError, match="cosine_distance_mode"):
        RINND(X, metric="cosine", cosine_distance_mode="bogus")
    with pytest.raises(ValueError, match="require metric='cosine' and normalize=True"):
        RINND(X, metric="cosine", cosine_distance_mode="direct")
    with pytest.raises(ValueError, match="quantized_rerank_width"):
        RINND(X, quantized_candidate_width=4, quantized_rerank_width=8)


def test_normalize_rows_matches_reference_arithmetic():
    from rinnd_amd.api import normalize_rows
    rng = np.random.default_rng(0)
    X = rng.standard_normal((50, 37)).astype(np.float32)
    got = normalize_rows(X)
    for r in range(50):  # py:1131-1146 with a sequential f32 sum
        s = np.float32(0)
        for v in X[r]:
            s = np.float32(s + np.float32(v * v))
        inv = np.float32(1) / np.sqrt(s)
        np.testing.assert_array_equal(got[r], (X[r] * inv).astype(np.float32))
    Z = np.zeros((2, 3), np.float32)
    np.testing.assert_array_equal(normalize_rows(Z), Z)


def test_default_tree_count_formula():
    from rinnd_amd.api import _round_half_away
    import math
    # py:1201-1204: clamp(round(2*log10(n)), 3, 12) with f64 round-half-away
    for n, want in [(32, 3), (100, 4), (10_000, 8), (1_000_000, 12), (31_623, 9)]:
        assert min(12, max(3, _round_half_away(2 * math.log10(n)))) == want

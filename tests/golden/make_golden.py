"""Writes tests/golden/cfg1_oracle.json: restatement-generated reference vectors for the 10,000 x 128 config.

The reference (Rust) cannot be built or imported in this image, so these vectors come from the oracle
(oracle/rinnd_oracle.cpp, the C++ restatement) and are unverified against a Rust build.  They pin the
oracle and the GPU path against silent drift: SHA-256 of the full outputs plus a few rows in clear.

Config: 10,000 x 128 float32 uniform [0,1) (numpy PCG64 seed 0), squared Euclidean, k=15, 8 trees, leaf 60,
max_candidates 15, <= 10 iterations, delta 0.001, seed 42, T = 1 and T = 8; index (euclidean: sqrt
correction) queried with 100 uniform queries (PCG64 seed 1), k=10, epsilon=0.1.

Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "cfg1_oracle.json")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def data():
    X = np.random.default_rng(0).random((10_000, 128), dtype=np.float32)
    Q = np.random.default_rng(1).random((100, 128), dtype=np.float32)
    return X, Q


def vectors():
    X, Q = data()
    out = {"config": "10000x128 U[0,1) seed 0; k=15, 8 trees, leaf 60, C=15, <=10 iters, delta 0.001, seed 42"}
    for T in (1, 8):
        r = O.nn_descent(X, 15, 8, 60, 15, 10, 0.001, threads=T)
        out[f"nn_descent_T{T}"] = {
            "changes": [int(c) for c in r["changes"]],
            "indices_sha256": sha(r["indices"]), "distances_sha256": sha(r["distances"]),
            "flags_sha256": sha(r["flags"]),
            "rng_state": [int(v) for v in r["rng_state"]],
            "rows_0_4_indices": r["indices"][:5].tolist(),
            "rows_0_4_distance_bits": r["distances"][:5].view(np.uint32).tolist(),
        }
    ix = O.Index(X, metric="sqeuclidean", correction="sqrt", n_neighbors=15, n_trees=8, leaf_size=60,
                 max_candidates=15, n_iters=10, threads=8)
    gi, gd = ix.neighbor_graph
    qi, qd = ix.query(Q, k=10, epsilon=0.1)
    out["index_T8"] = {
        "graph_indices_sha256": sha(gi), "graph_distances_sha256": sha(gd),
        "query_indices_sha256": sha(qi), "query_distances_sha256": sha(qd),
        "query_rows_0_4_indices": qi[:5].tolist(), "query_rows_0_4_distance_bits": qd[:5].view(np.uint32).tolist(),
    }
    return out


if __name__ == "__main__":
    v = vectors()
    with open(OUT, "w") as f:
        json.dump(v, f, indent=1)
    print(f"wrote {OUT}")

"""MI355X (gfx950) NN-Descent k-NN graph construction and greedy graph search.

A drop-in for lmcinnes/rinnd's build/query path: `RINND` mirrors the reference's Python class
(crates/rinnd/src/lib.rs:615-1128) and runs on hand-written HIP kernels through the C ABI in
include/rinnd_hip.h.
"""
from .api import RINND, simd_info, version  # noqa: F401

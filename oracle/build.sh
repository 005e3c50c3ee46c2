#!/usr/bin/env bash
# Builds the oracle (CPU restatement of the reference; test infrastructure only).
# Flags keep the reference's arithmetic: AVX2+FMA intrinsics, no FP contraction, no fast-math.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
g++ -O3 -std=c++17 -mavx2 -mfma -ffp-contract=off -fno-fast-math -fPIC -shared -pthread \
    -o "$here/liboracle.so" "$here/rinnd_oracle.cpp"

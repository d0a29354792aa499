#!/bin/bash
# Builds the FP64 CPU oracle (test infrastructure) into oracle/liboracle_sbslpm.so.
# -ffp-contract=off keeps every multiply-add rounded as written, like the
# reference's host build on x86-64 without FMA.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
g++ -std=c++17 -O2 -fopenmp -ffp-contract=off -fPIC -shared \
    -o "$here/liboracle_sbslpm.so.tmp" "$here/sbslpm_oracle.cpp"
mv "$here/liboracle_sbslpm.so.tmp" "$here/liboracle_sbslpm.so"

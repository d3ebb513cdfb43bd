#!/usr/bin/env bash
# Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.
# -ffp-contract=off keeps every a*b+c as a rounded multiply then a rounded add,
# matching the operation order the reference's arithmetic is restated in.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
g++ -O2 -std=c++17 -fPIC -shared -fopenmp -ffp-contract=off -fno-fast-math \
    -Wall -Wextra -Wno-unused-parameter \
    -o "$here/liboracle.so" "$here/brt_oracle.cpp"

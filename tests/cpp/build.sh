#!/bin/bash
# Builds the C++ TpAccumulator-mirror test driver against dca_amd/libg4acc.so (host code, g++).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
mkdir -p "$HERE/_build"
g++ -std=c++17 -O2 -Wall -I "$ROOT/include" "$HERE/tp_accumulator_wrapper_test.cpp" \
    -L "$ROOT/dca_amd" -lg4acc -Wl,-rpath,'$ORIGIN/../../../dca_amd' -Wl,-rpath-link,/opt/rocm/lib \
    -o "$HERE/_build/tp_accumulator_wrapper_test.tmp"
mv "$HERE/_build/tp_accumulator_wrapper_test.tmp" "$HERE/_build/tp_accumulator_wrapper_test"

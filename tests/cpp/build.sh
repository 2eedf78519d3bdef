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
# The mirror on the general model (complex M and sign, left/right labels, not spin symmetric).
g++ -std=c++17 -O2 -Wall -I "$ROOT/include" "$HERE/tp_accumulator_general_test.cpp" \
    -L "$ROOT/dca_amd" -lg4acc -Wl,-rpath,'$ORIGIN/../../../dca_amd' -Wl,-rpath-link,/opt/rocm/lib \
    -o "$HERE/_build/tp_accumulator_general_test.tmp"
mv "$HERE/_build/tp_accumulator_general_test.tmp" "$HERE/_build/tp_accumulator_general_test"
# Host-only checks of the mirror (no library calls, no GPU).
g++ -std=c++17 -O1 -Wall -I "$ROOT/include" "$HERE/host_logic_test.cpp" -o "$HERE/_build/host_logic_test.tmp"
mv "$HERE/_build/host_logic_test.tmp" "$HERE/_build/host_logic_test"
# The register site transforms of the M -> G kernels, run on the CPU.
g++ -std=c++17 -O2 -Wall -Wno-unknown-pragmas -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include \
    "$HERE/site_fft_test.cpp" -o "$HERE/_build/site_fft_test.tmp"
mv "$HERE/_build/site_fft_test.tmp" "$HERE/_build/site_fft_test"

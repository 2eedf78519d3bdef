#!/bin/bash
# Builds dca_amd/libg4acc.so for gfx950 with hipcc (cross-compiles without a GPU).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/libg4acc.so"
SRC=("$HERE"/csrc/g4acc_api.hip "$HERE"/csrc/mtog_kernels.hip "$HERE"/csrc/g4_kernels.hip)
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -Wall
       -Xclang -target-feature -Xclang -packed-fp32-ops)
/opt/rocm/bin/hipcc "${FLAGS[@]}" "${SRC[@]}" -o "$OUT.tmp" 2> >(grep -v "packed-fp32-ops' is not a recognized feature" >&2)
mv "$OUT.tmp" "$OUT"
echo "built $OUT"

#!/bin/bash
# Builds dca_amd/libg4acc.so for gfx950 with hipcc (cross-compiles without a GPU).
# Each source compiles to its own object in parallel; objects go to dca_amd/_build/.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/libg4acc.so"
OBJ="$HERE/_build"
mkdir -p "$OBJ"
SRC=(g4acc_api mtog_kernels mtog_fused mtog_large g4_kernels g4_tiled g4_mfma g4_band g4_rank1 sp_acc)
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall
       -Xclang -target-feature -Xclang -packed-fp32-ops)
filter() { grep -v "packed-fp32-ops' is not a recognized feature" >&2 || true; }
pids=()
# The matrix-core M -> G and rank-one kernels keep their accumulators in VGPRs (no AGPR copies).
extra() { case "$1" in mtog_fused|g4_rank1) echo "-mllvm -amdgpu-mfma-vgpr-form=1" ;; esac; }
for s in "${SRC[@]}"; do
  /opt/rocm/bin/hipcc "${FLAGS[@]}" $(extra "$s") -c "$HERE/csrc/$s.hip" -o "$OBJ/$s.o" 2> >(filter) &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
/opt/rocm/bin/hipcc --offload-arch=gfx950 -shared -fPIC $(printf "%s.o " "${SRC[@]/#/$OBJ/}") -o "$OUT.tmp" 2> >(filter)
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
# G4 files in the reference's HDF5 format (host C over the image's HDF5 1.10 C library).
HDF5=/opt/conda
if [ -f "$HDF5/include/hdf5.h" ] && [ -f "$HDF5/lib/libhdf5.so" ]; then
  gcc -O2 -std=c11 -fPIC -shared -Wall -I "$HDF5/include" "$HERE/csrc/g4io.c" -L "$HDF5/lib" -lhdf5 \
      -Wl,-rpath,"$HDF5/lib" -o "$HERE/libg4acc_io.so.tmp"
  mv "$HERE/libg4acc_io.so.tmp" "$HERE/libg4acc_io.so"
  echo "built $HERE/libg4acc_io.so"
else
  echo "HDF5 C library not found under $HDF5: libg4acc_io.so not built" >&2
fi

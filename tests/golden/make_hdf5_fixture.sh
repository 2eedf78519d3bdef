#!/bin/bash
# Extracts the G4 dataset of the reference's CT-AUX two-particle integration baseline
# (test/integration/cluster_solver/stdthread_qmci/stdthread_ctaux_tp_test_shared_baseline.hdf5,
# written by the reference's HDF5Writer) with its attributes into a small fixture, unchanged:
# tests/golden/ctaux_tp_baseline_G4.hdf5. Run in a container that holds the reference.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC=/root/reference/test/integration/cluster_solver/stdthread_qmci/stdthread_ctaux_tp_test_shared_baseline.hdf5
OUT="$HERE/ctaux_tp_baseline_G4.hdf5"
rm -f "$OUT"
/opt/conda/bin/h5copy -p -i "$SRC" -o "$OUT" -s /functions/G4 -d /functions/G4
ls -l "$OUT"

#!/bin/sh
# Rebuilds the HDF5 fixtures under tests/golden/h5 with the HDF5 library of
# the build container (/opt/conda, HDF5 1.10).  The fixtures are committed;
# nothing else reads this library.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=${TMPDIR:-/tmp}/nnx_make_h5
gcc -O1 -o "$OUT" "$HERE/make_h5.c" -I/opt/conda/include -L/opt/conda/lib -lhdf5 -Wl,-rpath,/opt/conda/lib -lm
mkdir -p "$HERE/h5"
"$OUT" "$HERE/h5"
rm -f "$OUT"

#!/bin/sh
# Compiles the reference's pure-C++ element layer in place from the reference
# tree (never copied into this repository) together with ref_shim.cpp into
# oracle/_ref/libseref.so.  The solve side of the reference (Matrix, Vector,
# Utils, src/linsys, src/qg) needs PETSc, which is not installed, so it is not
# built.  HDF5 (used only by the reference's Mesh::Save / Field::Save/Read) is
# linked from the image's own /opt/conda installation.
# Used only to write golden vectors (tests/golden/make_golden.py) and to
# cross-check the C oracle; nothing in the product links it.
set -e
REF=${SE_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
HDF5=${SE_HDF5_PREFIX:-/opt/conda}
if [ ! -d "$REF/src/base" ]; then
    echo "build_ref: reference tree not present, skipping"
    exit 0
fi
if [ ! -f "$HDF5/include/hdf5.h" ]; then
    echo "build_ref: hdf5 headers not present, skipping"
    exit 0
fi
mkdir -p "$OUT"
SRCS=""
for f in QuadPoint LinAlg Jacobi Element Legendre Chebyshev Mesh BCs Field Operator RHSOp Advector; do
    SRCS="$SRCS $REF/src/base/$f.cpp"
done
g++ -O2 -fPIC -shared -std=c++11 -w -ffp-contract=off \
    -I"$REF/src/base" -I"$HDF5/include" \
    $SRCS "$HERE/ref_shim.cpp" \
    -o "$OUT/libseref.so" -L"$HDF5/lib" -lhdf5 -Wl,-rpath,"$HDF5/lib"
echo "build_ref: wrote $OUT/libseref.so"

#!/bin/sh
# Builds the drop-in check program: the reference's own Mesh / BCs / Field /
# element classes compiled in place from the reference tree (never copied
# here) into oracle/_ref/libsebase.so, the drop-in classes
# (se_amd/dropin/se_dropin.cpp) compiled against those headers, the C oracle
# and libse_amd.so.  Output: oracle/_ref/dropin_check (the binary links the
# compiled reference, so it lives with the other reference builds).
set -e
REF=${SE_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="$ROOT/oracle/_ref"
HDF5=${SE_HDF5_PREFIX:-/opt/conda}
if [ ! -d "$REF/src/base" ] || [ ! -f "$HDF5/include/hdf5.h" ]; then
    echo "dropin build: reference tree or hdf5 not present, skipping"
    exit 0
fi
mkdir -p "$OUT"
SRCS=""
for f in QuadPoint LinAlg Jacobi Element Legendre Chebyshev Mesh BCs Field; do
    SRCS="$SRCS $REF/src/base/$f.cpp"
done
g++ -O2 -fPIC -shared -std=c++11 -w -I"$REF/src/base" -I"$HDF5/include" $SRCS \
    -o "$OUT/libsebase.so" "$HDF5/lib/libhdf5.so" -Wl,-rpath,"$HDF5/lib"
# the drop-in headers come first: they replace the reference's Advector.h
g++ -O2 -std=c++11 -w -I"$ROOT/se_amd/dropin" -I"$REF/src/base" -I"$ROOT/include" -I"$ROOT/oracle" \
    "$HERE/dropin_check.cpp" "$ROOT/se_amd/dropin/se_dropin.cpp" \
    -o "$OUT/dropin_check" \
    -L"$OUT" -lsebase -L"$ROOT/oracle" -loracle -L"$ROOT/se_amd/lib" -lse_amd \
    -Wl,-rpath,'$ORIGIN' -Wl,-rpath,'$ORIGIN/..' -Wl,-rpath,'$ORIGIN/../../se_amd/lib' \
    -Wl,-rpath-link,"$HDF5/lib"
echo "dropin build: wrote $OUT/dropin_check"
# field-file interop with the reference's Field::Save / Read (needs libse_amd_io.so)
if [ -f "$ROOT/se_amd/lib/libse_amd_io.so" ]; then
    g++ -O2 -std=c++11 -w -I"$REF/src/base" -I"$ROOT/include" -I"$HDF5/include" "$HERE/io_check.cpp" \
        -o "$OUT/io_check" -L"$OUT" -lsebase -L"$ROOT/se_amd/lib" -lse_amd_io "$HDF5/lib/libhdf5.so" \
        -Wl,-rpath,"$HDF5/lib" \
        -Wl,-rpath,'$ORIGIN' -Wl,-rpath,'$ORIGIN/../../se_amd/lib' -Wl,-rpath-link,"$HDF5/lib"
    echo "dropin build: wrote $OUT/io_check"
fi

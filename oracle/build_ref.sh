#!/usr/bin/env bash
# Builds oracle/_ref/libref.so from the reference's own sources where they lie
# (/root/reference), for the parts of the CPU path that compile without CUDA
# headers: include/Differ.hpp, src/Function.cpp, src/FileReader.cpp,
# src/Mesh.cpp, src/Global.cpp.  Mesh.cpp needs mpi.h; the MPICH shipped in
# /opt/conda provides it.  Flowfield.cpp, Reaction.cpp, TimeAdv.cpp and
# Euler.cpp include <cuda_runtime.h>, which this image lacks: they are not built.
# Same flags as the reference build (its Makefile:6).  Output goes only to
# oracle/_ref/.  Exits 0 without building when the reference is absent.
set -euo pipefail
here=$(cd -- "$(dirname -- "${BASH_SOURCE[0]}")" && pwd)
ref=${EU3D_REFERENCE:-/root/reference}
if [ ! -d "$ref/src" ]; then
    echo "reference sources not present; oracle/_ref not built" >&2
    exit 0
fi
mpi_inc=/opt/conda/include
mpi_lib=/opt/conda/lib
if [ ! -f "$mpi_inc/mpi.h" ]; then
    echo "mpi.h not found; oracle/_ref not built" >&2
    exit 0
fi
mkdir -p "$here/_ref"
g++ -std=c++17 -O3 -ffp-contract=off -fPIC -shared -w \
    -I"$ref/include" -I"$mpi_inc" \
    "$here/ref_harness.cpp" "$ref/src/Function.cpp" "$ref/src/FileReader.cpp" \
    "$ref/src/Mesh.cpp" "$ref/src/Global.cpp" \
    -L"$mpi_lib" -Wl,-rpath,"$mpi_lib" -lmpi \
    -o "$here/_ref/libref.so.tmp"
mv -f "$here/_ref/libref.so.tmp" "$here/_ref/libref.so"
echo "built $here/_ref/libref.so"

#!/bin/sh
# Builds oracle/_ref/libmkref.so from the reference's own, unmodified sources
# (core/TransferFunction.cpp, core/Methods.cpp, core/RenderInfo.cpp,
# core/PlaneInfo.cpp, core/fpng/fpng.cpp) plus the harness in this directory.
# Test infrastructure only: the output writes and cross-checks golden vectors
# and times the reference's PNG encoder as a CPU baseline; the product never
# loads it.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${MONKEYGL_REFERENCE:-/root/reference}/core
OUT="$HERE/_ref"
mkdir -p "$OUT"
g++ -std=c++11 -O2 -msse4.1 -mpclmul -fPIC -shared -w -I"$REF" \
    "$HERE/ref_harness.cpp" "$REF/TransferFunction.cpp" "$REF/Methods.cpp" "$REF/RenderInfo.cpp" \
    "$REF/PlaneInfo.cpp" "$REF/fpng/fpng.cpp" \
    -o "$OUT/libmkref.so.tmp"
mv "$OUT/libmkref.so.tmp" "$OUT/libmkref.so"
echo "built $OUT/libmkref.so"

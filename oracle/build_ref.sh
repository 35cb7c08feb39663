#!/bin/sh
# Builds oracle/_ref/libmkref.so from the reference's own, unmodified sources
# (core/TransferFunction.cpp, core/Methods.cpp, core/RenderInfo.cpp,
# core/PlaneInfo.cpp) plus the
# harness in this directory. Test infrastructure only: the output is used to
# write and cross-check golden vectors, never by the product.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${MONKEYGL_REFERENCE:-/root/reference}/core
OUT="$HERE/_ref"
mkdir -p "$OUT"
g++ -std=c++11 -O2 -fPIC -shared -w -I"$REF" \
    "$HERE/ref_harness.cpp" "$REF/TransferFunction.cpp" "$REF/Methods.cpp" "$REF/RenderInfo.cpp" \
    "$REF/PlaneInfo.cpp" \
    -o "$OUT/libmkref.so.tmp"
mv "$OUT/libmkref.so.tmp" "$OUT/libmkref.so"
echo "built $OUT/libmkref.so"

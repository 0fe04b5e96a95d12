#!/bin/sh
# Builds oracle/_ref/libbvhref.so from the reference's vendored lib/bvh
# headers, where they lie in the reference tree. Outputs only into
# oracle/_ref/. The reference tree is needed only here, at build time.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${BTRC_REFERENCE:-/root/reference}
if [ ! -d "$REF/lib/bvh/include/bvh" ]; then
    echo "reference lib/bvh not found under $REF; oracle/_ref not built" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
g++ -O3 -std=c++17 -fPIC -shared -fopenmp -ffp-contract=off \
    -I"$REF/lib/bvh/include" "$HERE/ref_bvh.cpp" -o "$HERE/_ref/libbvhref.so.tmp"
mv "$HERE/_ref/libbvhref.so.tmp" "$HERE/_ref/libbvhref.so"

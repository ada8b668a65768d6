#!/bin/sh
# Builds oracle/_ref/libnbsref_f64.so and libnbsref_f32.so from the
# reference's own CPU sources (octree.cpp, particle.cpp, vector.cpp,
# thread_functions.cpp), compiled where they lie, plus ref_harness.cpp.
# Flags: the reference release build's -O2 and -std=c++11, without
# -march=native (so no FMA contraction: the arithmetic the oracle restates).
# Output goes only to oracle/_ref/.  Needs the reference tree; skipped
# (exit 0, message) when it is absent.
set -e
REF=${NBS_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/octree.cpp" ]; then
  echo "build_ref: reference sources not found under $REF; skipping"
  exit 0
fi
mkdir -p "$OUT"
for P in DOUBLE:f64 FLOAT:f32; do
  DEF=${P%%:*}
  TAG=${P##*:}
  g++ -std=c++11 -O2 -ffp-contract=off -fPIC -shared -D$DEF -w \
    -I"$REF/src" \
    "$REF/src/octree.cpp" "$REF/src/particle.cpp" "$REF/src/vector.cpp" \
    "$REF/src/thread_functions.cpp" "$HERE/ref_harness.cpp" \
    -o "$OUT/libnbsref_$TAG.so.tmp" -lpthread
  mv "$OUT/libnbsref_$TAG.so.tmp" "$OUT/libnbsref_$TAG.so"
done
echo "build_ref: built $OUT/libnbsref_f64.so $OUT/libnbsref_f32.so"

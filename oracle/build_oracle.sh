#!/bin/sh
# Builds the oracle (the CPU restatement in nbs_oracle.cpp) into
# oracle/_build/libnbs_oracle.so.  -O2 without -march=native and with
# -ffp-contract=off: the reference's arithmetic without FMA contraction.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$HERE/_build"
g++ -std=c++17 -O2 -ffp-contract=off -fPIC -shared -pthread \
  "$HERE/nbs_oracle.cpp" -o "$HERE/_build/libnbs_oracle.so.tmp"
mv "$HERE/_build/libnbs_oracle.so.tmp" "$HERE/_build/libnbs_oracle.so"
echo "build_oracle: built $HERE/_build/libnbs_oracle.so"

#!/usr/bin/env bash
# Builds the CPU oracle (test infrastructure) into oracle/_build/.
#   libkd_oracle.so       kd_math.h k() (bit-identical to the device)
#   libkd_oracle_glibc.so same code, k() through glibc log/asin (sensitivity check)
# The reference itself (Mojo 1.1.0 / MAX, osx-arm64 lock) cannot be built in
# this image, so there is no oracle/_ref build.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="$here/_build"
mkdir -p "$out"
CFLAGS="-O3 -std=gnu11 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function"
gcc $CFLAGS -o "$out/libkd_oracle.so.tmp" "$here/kd_oracle.c" -lm -lpthread
mv "$out/libkd_oracle.so.tmp" "$out/libkd_oracle.so"
gcc $CFLAGS -DKD_ORACLE_GLIBC_K=1 -o "$out/libkd_oracle_glibc.so.tmp" "$here/kd_oracle.c" -lm -lpthread
mv "$out/libkd_oracle_glibc.so.tmp" "$out/libkd_oracle_glibc.so"

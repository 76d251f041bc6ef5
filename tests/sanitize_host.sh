#!/bin/bash
# AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU tests.
#
# Rebuilds the CPU oracle (oracle/stc_oracle.c) and the C++ host library
# (staccato_amd/csrc/host/*.cpp -> libstaccato_host.so) with
# -fsanitize=address,undefined into a temporary folder, points the tests at
# those builds (STC_ORACLE_LIB, STC_HOST_LIB) and runs the CPU suite with the
# sanitizer runtimes preloaded; any report aborts the test process.  Host code
# only: gcc/g++, no GPU code is built with a sanitizer.
#
#   bash tests/sanitize_host.sh [extra pytest arguments]
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=$(mktemp -d)
trap 'rm -rf "$OUT"' EXIT
SAN="-fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer -g -O1"
gcc -std=gnu99 -fPIC -fopenmp $SAN -shared -o "$OUT/libstc_oracle.so" \
    "$ROOT/oracle/stc_oracle.c" -lm -ldl
g++ -std=c++17 -fPIC $SAN -shared -o "$OUT/libstaccato_host.so" \
    "$ROOT/staccato_amd/csrc/host/KrylovROMSweep.cpp" "$ROOT/staccato_amd/csrc/host/IJCSACoupling.cpp" \
    -L"$ROOT/staccato_amd/lib" -lstaccato_sweep -Wl,-rpath,"$ROOT/staccato_amd/lib"
PRE="$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so)"
cd "$ROOT"
env STC_ORACLE_LIB="$OUT/libstc_oracle.so" STC_HOST_LIB="$OUT/libstaccato_host.so" \
    ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:halt_on_error=1 \
    UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
    LD_PRELOAD="$PRE" python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"

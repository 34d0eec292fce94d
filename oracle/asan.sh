#!/bin/bash
# Builds the C oracle together with oracle/asan_main.c (a driver that calls
# every oracle function on a small synthetic scene) with AddressSanitizer and
# UndefinedBehaviorSanitizer, and runs it.  CPU only; output in oracle/_build/.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$HERE/_build"
gcc -std=c11 -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined \
    -fno-sanitize-recover=undefined -ffp-contract=off -fno-fast-math \
    -o "$HERE/_build/oracle_asan" "$HERE/sailor_oracle.c" "$HERE/asan_main.c" -lm
ASAN_OPTIONS=detect_leaks=0:abort_on_error=0 "$HERE/_build/oracle_asan"

#!/bin/bash
# CPU sanitizer lane: the CPU test suite with the oracle and the library's
# host-only stitch code built with AddressSanitizer and
# UndefinedBehaviorSanitizer (host code only; no GPU is used).
#   bash tests/sanitize_cpu.sh [pytest args]
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
make -s -C "$ROOT/oracle" SAN=1
make -s -C "$ROOT/cuphu_amd/csrc" san
ASAN_RT=$(gcc -print-file-name=libasan.so)
UBSAN_RT=$(gcc -print-file-name=libubsan.so)
cd "$ROOT"
CUPHU_ORACLE_LIB="$ROOT/oracle/_build_san/libcuphu_oracle.so" \
CUPHU_LIB="$ROOT/cuphu_amd/csrc/_build_san/libcuphu_hip.so" \
ASAN_OPTIONS=detect_leaks=0:verify_asan_link_order=0:abort_on_error=1 \
UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
LD_PRELOAD="$ASAN_RT:$UBSAN_RT" \
python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"

#!/bin/bash
# Runs the CPU test suite with the oracle built under AddressSanitizer and
# UndefinedBehaviorSanitizer (oracle/_san/libsseoracle.so).  Python itself is
# not instrumented, so the ASan runtime is preloaded; leak checking is off
# (the interpreter keeps its allocations until exit).  Any sanitizer report
# aborts the run with a non-zero status.
#   bash oracle/run_sanitized.sh [pytest arguments]
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
root="$(dirname "$here")"
make -s -C "$here" sanitize
asan="$(gcc -print-file-name=libasan.so)"
cd "$root"
SSE_ORACLE_LIB="$here/_san/libsseoracle.so" \
ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1" \
UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1" \
LD_PRELOAD="$asan${LD_PRELOAD:+:$LD_PRELOAD}" \
python -m pytest tests -x -q -m "not gpu" -p no:cacheprovider "$@"

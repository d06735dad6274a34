#!/bin/bash
# The CPU test suite on AddressSanitizer + UndefinedBehaviorSanitizer builds of the oracle and
# of libgpsat.so's host code (CPU only; device code is not instrumented).  Run from the
# repository root:  bash tests/run_sanitized.sh
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
make -s -C "$ROOT/oracle" sanitize
LIB="$(python3 "$ROOT/gpupsat_amd/_build.py" --sanitize | tail -1)"
ASAN_RT="$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)"
export GPSAT_LIB="$LIB" GPSAT_ORACLE_LIB="$ROOT/oracle/liboracle_san.so"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1" UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"
cd "$ROOT"
LD_PRELOAD="$ASAN_RT" python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"

#!/bin/bash
# The CPU test suite under AddressSanitizer and UndefinedBehaviorSanitizer,
# in two passes (no GPU needed; each pass loads one sanitized component):
#   1. the oracle (oracle/_san/liboracle.so, gcc's runtime preloaded);
#   2. the library's host code (spnngp_amd/_san/libnngp_hip.so, its device
#      code built as usual, clang's runtime preloaded).
# Any sanitizer report aborts the test process, which fails the pass.
#
#   bash tests/run_sanitized.sh            # build both, run both passes
#   bash tests/run_sanitized.sh --no-build # reuse the sanitized builds
set -o pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
cd "$ROOT" || exit 1
if [ "$1" != "--no-build" ]; then
  make -s -C oracle sanitize || exit 1
  python -m spnngp_amd.build --sanitize || exit 1
fi
GCC_ASAN="$(gcc -print-file-name=libasan.so)"
CLANG_ASAN="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -n 1)"
# leaks of the Python interpreter itself are not ours to report
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
PYTEST=(python -m pytest tests/ -x -q -m "not gpu" -p no:cacheprovider)

echo "[sanitized] pass 1: oracle"
NNGP_ORACLE_LIB="$ROOT/oracle/_san/liboracle.so" LD_PRELOAD="$GCC_ASAN" \
  "${PYTEST[@]}" || exit $?

echo "[sanitized] pass 2: library host code"
NNGP_LIB="$ROOT/spnngp_amd/_san/libnngp_hip.so" LD_PRELOAD="$CLANG_ASAN" \
  "${PYTEST[@]}" || exit $?
echo "[sanitized] both passes clean"

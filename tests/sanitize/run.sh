#!/bin/bash
# Builds libfip.so, the CLI and the FP64 oracle with AddressSanitizer and
# UndefinedBehaviorSanitizer on their host code (Makefile here) and runs the
# CPU tests that drive the FITS codec, the CLI's argument and file handling,
# the ingest layout checks and the oracle against those builds.  Any
# sanitizer report aborts the test process, so a clean run ends with pytest's
# own summary.  CPU only: no test here launches a kernel.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
make -s -j8 -C "$HERE"
RT="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
cd "$ROOT"
LD_PRELOAD="$RT" \
ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1" \
UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1" \
FIP_LIBRARY="$HERE/_build/libfip.so" \
FIP_ORACLE_LIB="$HERE/_build/libfip_oracle.so" \
FIP_CLI="$HERE/_build/sharedlibrary_gpu" \
python -m pytest -q -p no:cacheprovider -m "not gpu" \
    tests/test_fits_cli_cpu.py tests/test_abi_cpu.py tests/test_oracle_math.py tests/test_oracle_imaging.py \
    "$@"

#!/bin/bash
# Host-code sanitizer run (CPU only, no GPU needed): builds the hand-written host
# sources of the library (config/CSV/binary I/O, synthetic generator, plug-in
# adapter) and the C oracle with gcc's AddressSanitizer + UndefinedBehavior-
# Sanitizer, then runs the CPU tests that exercise them against those builds.
# Any sanitizer report aborts the run (halt_on_error); leak checking is off
# because the Python interpreter itself keeps allocations alive at exit.
set -euo pipefail
ROOT="$(cd "$(dirname "${BASH_SOURCE[0]}")/.." && pwd)"
make -s -j8 -C "$ROOT/gpu-montecarlo-risk_amd/csrc" all sanitize
make -s -C "$ROOT/oracle" all sanitize
export MCR_LIB_PATH="$ROOT/gpu-montecarlo-risk_amd/_lib/sanitize/libmcr_hip.so"
export ORACLE_LIB_PATH="$ROOT/oracle/_build/sanitize/liboracle.so"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1"
export UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"
# the sanitizer runtimes first (ASan must be), any preload already set kept after them
export LD_PRELOAD="$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so)${LD_PRELOAD:+:$LD_PRELOAD}"
cd "$ROOT"
python -m pytest -q -p no:cacheprovider -m "not gpu" tests/test_io.py tests/test_io_properties.py \
  tests/test_abi.py tests/test_oracle_rng.py tests/test_oracle_paths.py "$@"

#!/bin/bash
# The CPU test suite against AddressSanitizer + UndefinedBehaviorSanitizer
# builds of the host code: the oracle (oracle/_san/liboracle.so) and the
# library's model, model C ABI, server and image export
# (cpsfdtd_amd/lib/sanitize/libcpsfdtd.so).  CPU only, no GPU.
#
#   bash tests/run_sanitized.sh [pytest arguments]
#
# gcc's libasan and libubsan are preloaded into the Python process (Python
# itself is not instrumented); leak checking is off because the interpreter
# keeps its own allocations to the end.  Any report ends the run with a
# non-zero status.
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
cd "$ROOT"
make -s -C oracle sanitize
python -m cpsfdtd_amd.build --sanitize -q > /dev/null
ASAN_LIB="$(gcc -print-file-name=libasan.so)"
UBSAN_LIB="$(gcc -print-file-name=libubsan.so)"
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1:verify_asan_link_order=0"
export UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"
export ORC_LIB="$ROOT/oracle/_san/liboracle.so"
export CFD_LIB="$ROOT/cpsfdtd_amd/lib/sanitize/libcpsfdtd.so"
# the instrumented builds are the ones a test process maps
LD_PRELOAD="$ASAN_LIB:$UBSAN_LIB" python - <<'PY'
import sys
sys.path[:0] = ["oracle", "."]
import pyoracle, cpsfdtd_amd
pyoracle.lib(); cpsfdtd_amd.library()
maps = open("/proc/self/maps").read()
for lib in ("_san/liboracle.so", "sanitize/libcpsfdtd.so", "libasan.so", "libubsan.so"):
    assert lib in maps, "not loaded: " + lib
print("sanitized builds loaded:", pyoracle.LIB_PATH, cpsfdtd_amd.LIB_PATH)
PY
LD_PRELOAD="$ASAN_LIB:$UBSAN_LIB" python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider "$@"

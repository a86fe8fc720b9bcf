#!/usr/bin/env bash
# Host sanitizer run: builds libvbgemm with AddressSanitizer and
# UndefinedBehaviorSanitizer on its host code (C ABI, planner, launch glue;
# device code unchanged) and the CPU oracle with the same sanitizers, then runs
# the CPU test files that exercise them (host-only planning, argument checks,
# the reference planner, the oracle) with both sanitized libraries loaded.
# Any ASan report or UBSan finding aborts the run with a non-zero status.
# No GPU is used.  The sanitized libraries are removed afterwards.
#
#   bash tests/sanitize_host.sh
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
JOBS="${MAX_JOBS:-8}"
make -C "$ROOT/batchgemm_amd" -j"$JOBS" asan
make -C "$ROOT/oracle" -B asan
RT="$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)"
cd "$ROOT"
status=0
# detect_leaks=0: the Python interpreter itself does not free everything at exit
san_env=(LD_PRELOAD="$RT"
         ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1"
         UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"
         VBG_LIBRARY="$ROOT/batchgemm_amd/libvbgemm_asan.so"
         ORC_LIBRARY="$ROOT/oracle/liboracle_asan.so")
# the run must load the sanitized builds and the ASan runtime, not the defaults
env "${san_env[@]}" python - <<'EOF' || status=$?
import sys
sys.path.insert(0, ".")
import batchgemm_amd
from oracle import oracle
batchgemm_amd.lib(); oracle.lib()
maps = open("/proc/self/maps").read()
for lib in ("libvbgemm_asan.so", "liboracle_asan.so", "libclang_rt.asan"):
    assert lib in maps, lib + " not loaded"
print("sanitized libraries loaded")
EOF
[ $status -eq 0 ] && { env "${san_env[@]}" python -m pytest tests/test_host.py tests/test_oracle.py -q \
    -m "not gpu" -p no:cacheprovider || status=$?; }
rm -f "$ROOT/batchgemm_amd/libvbgemm_asan.so" "$ROOT/oracle/liboracle_asan.so"
exit $status

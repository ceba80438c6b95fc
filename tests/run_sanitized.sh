#!/bin/bash
# The CPU-side tests (pytest -m "not gpu": the oracle, the scene generator,
# the C-ABI library's argument and error paths, the host logic) against
# AddressSanitizer + UndefinedBehaviorSanitizer builds: the host code of
# libgprm (gpuprm_amd/libgprm_san.so) and the oracle (oracle/lib*_san.so),
# both built by clang and run with clang's ASan runtime (which carries the
# UBSan handlers) preloaded into Python.  Leak checking is off: the
# interpreter itself keeps its allocations to the end.
set -euo pipefail
cd "$(dirname "$0")/.."
python -c "from gpuprm_amd import build; build.build_sanitized()"
RT=$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)
export GPRM_LIBRARY=$PWD/gpuprm_amd/libgprm_san.so
export GPRM_ORACLE_SUFFIX=_san
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
LD_PRELOAD=$RT python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"

"""Build the native libraries in-tree.

``build_all()`` compiles the HIP library ``gpuprm_amd/libgprm.so`` for gfx950
with hipcc, and the CPU oracle (``oracle/``, test infrastructure) with gcc.
When the reference sources are present, ``oracle/_ref/libref.so`` (the
reference's own CPU path, used as checker and CPU baseline) is built too.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpuprm_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libgprm.so")
REFERENCE_LIB = "/root/reference/prm/src/lib"

# Floating point must follow the reference's host path operation by
# operation: no FMA contraction, IEEE division/sqrt, denormals kept, NaN/inf
# semantics intact.  Packed FP32 instructions are disabled on purpose.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-finite-math-only",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
SOURCES = ["kernels.hip", "gprm_api.cpp", "batch.cpp", "scene.cpp", "prm.cpp", "surrogate.hip", "noalloc.hip", "planner.hip"]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_hip() -> str:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIP_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


LIB_SAN = os.path.join(PKG, "libgprm_san.so")
# AddressSanitizer and UndefinedBehaviorSanitizer on the host code only (the
# device code is built as usual); load with clang's ASan runtime preloaded.
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer", "-g"]


def build_sanitized() -> str:
    """libgprm_san.so (host code sanitized) and the sanitized oracle, for the
    CPU-side tests (tests/run_sanitized.sh)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    flags = [f for f in HIP_FLAGS if f != "-O3"] + ["-O1"]
    _run([hipcc] + flags + SAN_FLAGS + ["-o", LIB_SAN] + [os.path.join(CSRC, s) for s in SOURCES])
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "san"])
    return LIB_SAN


def build_oracle() -> None:
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "all"])


def build_reference() -> bool:
    if not os.path.isdir(REFERENCE_LIB):
        return False
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "ref", "REF=" + REFERENCE_LIB])
    return True


def build_all() -> None:
    build_hip()
    build_oracle()
    build_reference()


if __name__ == "__main__":
    build_all()
    print("built", LIB, file=sys.stderr)

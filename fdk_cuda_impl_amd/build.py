"""Build recipe of ``libyk_hip.so`` (hipcc, gfx950).

Every translation unit under ``csrc/`` is compiled by hipcc for gfx950 with
packed-FP32 instruction selection disabled, then linked into one shared
library next to this file.  Usage: ``python -m fdk_cuda_impl_amd.build`` or
``build()`` from Python.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
REPO = os.path.dirname(HERE)
OBJ_DIR = os.path.join(REPO, "_build", "obj")
LIB_PATH = os.path.join(HERE, "libyk_hip.so")
SESSION_CHECK = os.path.join(HERE, "yk_session_check")
VALU_COSTS = os.path.join(HERE, "yk_valu_costs")
VMEM_COSTS = os.path.join(HERE, "yk_vmem_costs")
ARCH = os.environ.get("YK_OFFLOAD_ARCH", "gfx950")

SOURCES = ["capi.cpp", "geometry.cpp", "fdk.cpp", "projector.cpp", "algebraic.cpp", "cgls.cpp", "diagnostics.cpp", "session.cpp",
           "kernels_filter.hip", "kernels_bp.hip", "kernels_proj.hip", "kernels_misc.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
          "-I", os.path.join(REPO, "include")]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{res.stdout}\n{res.stderr}")
    noise = "is not a recognized feature for this target"
    warn = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if warn.strip():
        print(warn, file=sys.stderr)
    return obj


def build(jobs: int | None = None) -> str:
    """Compile every source and link libyk_hip.so; returns the library path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    build_session_check()
    build_valu_costs()
    build_vmem_costs()
    return LIB_PATH


def build_session_check() -> str:
    """Compile tests/session_check.cpp (the Session-API scenarios) against
    include/yk_session.hpp and libyk_hip.so, next to the library."""
    src = os.path.join(REPO, "tests", "session_check.cpp")
    out = SESSION_CHECK
    cmd = [hipcc(), "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-I", os.path.join(REPO, "include"),
           "-x", "hip", src, "-o", out + ".tmp", "-L", HERE, "-lyk_hip", f"-Wl,-rpath,{HERE}",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"session check build failed\n{res.stdout}\n{res.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build_valu_costs() -> str:
    """Compile tests/valu_costs.hip, the VALU instruction-cost probe behind
    profiles/valu_instruction_costs_gfx950.txt, next to the library."""
    src = os.path.join(REPO, "tests", "valu_costs.hip")
    cmd = [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-x", "hip", src, "-o", VALU_COSTS + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"VALU cost probe build failed\n{res.stdout}\n{res.stderr}")
    os.replace(VALU_COSTS + ".tmp", VALU_COSTS)
    return VALU_COSTS


def build_vmem_costs() -> str:
    """Compile tests/vmem_costs.hip, the vector-memory / LDS cost probe behind
    profiles/vector_memory_costs_gfx950.txt, next to the library."""
    src = os.path.join(REPO, "tests", "vmem_costs.hip")
    cmd = [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-x", "hip", src, "-o", VMEM_COSTS + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"vector-memory cost probe build failed\n{res.stdout}\n{res.stderr}")
    os.replace(VMEM_COSTS + ".tmp", VMEM_COSTS)
    return VMEM_COSTS


if __name__ == "__main__":
    print(build())

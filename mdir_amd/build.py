"""Builds libmdrt_hip.so for gfx950 in-tree (hipcc, no JIT cache).

    python -m mdir_amd.build

Every HIP source is compiled with -Xclang -target-feature -Xclang
-packed-fp32-ops (no packed FP32 instructions are generated).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmdrt_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result", "-Wno-unused-value"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    obj_engine = os.path.join(LIBDIR, "engine.o")
    obj_rt = os.path.join(LIBDIR, "mdrt_hip.o")
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    run([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, "engine.hip"), "-o", obj_engine])
    run(["gcc", "-O2", "-fPIC", "-std=c11", "-Wall", "-D__HIP_PLATFORM_AMD__",
         f"-I{ROCM}/include", "-c", os.path.join(CSRC, "runtime", "mdrt_hip.c"),
         "-o", obj_rt])
    tmp = LIB + ".tmp"
    run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", obj_engine, obj_rt,
         "-o", tmp, f"-L{ROCM}/lib", "-lamdhip64"])
    os.replace(tmp, LIB)
    build_abi_module()
    return LIB


ABI_MODULE = os.path.join(LIBDIR, "abi_module.co")


def build_abi_module() -> str:
    """The small code object the ABI tests load through mgpuModuleLoad
    (tests/abi_module.hip), as a compiled program hands its kernels over."""
    src = os.path.join(ROOT, "tests", "abi_module.hip")
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    run([hipcc, "--genco", f"--offload-arch={ARCH}", "-O3", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", src, "-o", ABI_MODULE])
    return ABI_MODULE


if __name__ == "__main__":
    build()
    sys.exit(0)

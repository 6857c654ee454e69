"""The C-ABI library loads and exports every symbol include/mdrt_hip.h
declares (no GPU needed: nothing is called)."""
import ctypes as C
import os
import re

from mdir_amd import runtime as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mdrt_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "while", "for", "return", "sizeof"}
    return sorted({n for n in names if n not in skip and
                   (n.startswith("mgpu") or n.startswith("mdrt") or n.startswith("_mlir"))})


def test_header_lists_the_runtime_abi():
    names = declared_functions()
    for must in ["mgpuModuleLoad", "mgpuLaunchKernel", "mgpuMemAlloc", "mgpuMemcpy",
                 "mgpuStreamCreate", "_mlir_ciface_mdrtCudaFFTForward3DF32",
                 "mdrtFlagStart", "mdrtHipSystemCreate", "mdrtHipRun"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(rt.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    bound = {name for name, _, _ in rt.SIGNATURES}
    assert set(declared_functions()) <= bound
    rt.lib()  # binds every signature without calling anything

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


def run_isolated(code):
    """Runs `code` in a fresh interpreter with the library loaded; returns
    (return code, stderr). The misuse checks below run before any device
    call, so they need no GPU."""
    import subprocess
    import sys
    prog = ("import sys, ctypes as C; sys.path.insert(0, %r)\n"
            "from mdir_amd import runtime as rt\nL = rt.lib()\n" % ROOT) + code
    p = subprocess.run([sys.executable, "-c", prog], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stderr


def test_pool_refuses_a_foreign_stream():
    """Allocation and free on a stream other than the runtime's first one
    abort (mdrt_cuda.c:2129-2145): the pool hands a freed block to the next
    request without a wait, which is sound only in one stream's order."""
    rc, err = run_isolated("L.mgpuMemAlloc(256, C.c_void_p(0x1000), False)")
    assert rc == -6, (rc, err)
    assert "allocation of device memory on a stream other than" in err
    rc, err = run_isolated("L.mgpuMemFree(C.c_void_p(0x2000), C.c_void_p(0x1000))")
    assert rc == -6, (rc, err)
    assert "free of device memory on a stream other than" in err


def test_restore_without_use_is_refused():
    rc, err = run_isolated("L.mdrtRestoreStream(None)")
    assert rc == -6 and "without a matching mdrtUseStream" in err, (rc, err)
    rc, err = run_isolated("L.mdrtSideEnd()")
    assert rc == -6 and "no side stream is open" in err, (rc, err)

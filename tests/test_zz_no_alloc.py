"""The no-alloc forward transform (mdrtHipFFTForward3DF32NoAlloc): through
the C ABI, its half spectrum equals that of the plan-based entry point it
varies (_mlir_ciface_mdrtCudaFFTForward3DF32) on two calls in a row, the
second on a smaller grid, and both meet the oracle's FFT; sizes beyond its
fixed workspace are refused with an error, not run."""
import ctypes as C

import numpy as np
import pytest

from mdir_amd import runtime as rt
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_no_alloc_forward_matches_the_plan_entry_point():
    L = rt.lib()
    s = L.mgpuStreamCreate()
    rng = np.random.default_rng(12)
    for K in ((48, 40, 60), (16, 20, 12)):
        say(f"case: no-alloc forward transform {K}")
        real = rng.standard_normal(K).astype(np.float32)
        h = K[2] // 2 + 1
        nc = K[0] * K[1] * h * 2
        d_real = L.mgpuMemAlloc(real.nbytes, s, False)
        d_plan = L.mgpuMemAlloc(nc * 4, s, False)
        d_free = L.mgpuMemAlloc(nc * 4, s, False)
        L.mgpuMemcpy(d_real, real.ctypes.data, real.nbytes, s)
        rb = rt.Buffer1(d_real, d_real, 0, real.size, 1)
        cb = rt.Buffer1(d_plan, d_plan, 0, nc, 1)
        L._mlir_ciface_mdrtCudaFFTForward3DF32(C.byref(rb), C.byref(cb), *K)
        rt.check(L.mdrtHipFFTForward3DF32NoAlloc(d_real, d_free, K[0], K[1], K[2], s))
        L.mgpuStreamSynchronize(s)
        a = np.zeros(nc, np.float32)
        b = np.zeros(nc, np.float32)
        L.mgpuMemcpy(a.ctypes.data, d_plan, a.nbytes, s)
        L.mgpuMemcpy(b.ctypes.data, d_free, b.nbytes, s)
        diff = np.abs(a - b).max() / np.abs(a).max()
        say(f"  largest difference to the plan entry point {diff:.3e} of the largest value")
        assert diff <= 3e-7
        ref = O.fft3d_r2c(real.astype(np.float64))
        got = b.reshape(K[0], K[1], h, 2)
        got = got[..., 0].astype(np.float64) + 1j * got[..., 1]
        assert np.abs(got - ref).max() <= 3e-6 * np.abs(ref).max()
        for p in (d_real, d_plan, d_free):
            L.mgpuMemFree(p, s)


def test_no_alloc_refuses_sizes_beyond_its_workspace():
    L = rt.lib()
    s = L.mgpuStreamCreate()
    d = L.mgpuMemAlloc(4096, s, False)
    assert L.mdrtHipFFTForward3DF32NoAlloc(d, d, 2048, 4, 4, s) != 0
    assert b"1024" in L.mdrtHipLastError()
    assert L.mdrtHipFFTForward3DF32NoAlloc(d, d, 22, 4, 4, s) != 0   # factor 11
    assert L.mdrtHipFFTForward3DF32NoAlloc(d, d, 8, 4, 5, s) != 0    # odd last edge
    L.mgpuMemFree(d, s)

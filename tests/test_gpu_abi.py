"""The runtime ABI as a compiled program drives it (runtime/mdrt_cuda.c's
entry points): a gfx950 code object loaded, its kernels found and launched
through mgpuModuleLoad / mgpuModuleGetFunction / mgpuLaunchKernel; the two
side streams forked from and joined to the first stream; a caller's stream
lent with mdrtUseStream (nested, with the legacy default stream); scoped
copies; flags set by a kernel in mapped host memory; host results; and the
Gaussian backward transform of the PME chain against the oracle's influence
function."""
import ctypes as C
import os

import numpy as np
import pytest

from mdir_amd import build as B
from mdir_amd import runtime as rt
from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 1 << 20
BLOCK = 256


def say(msg):
    print(msg, flush=True)


class Module:
    """The ABI test code object, loaded through the runtime."""

    def __init__(self):
        L = rt.lib()
        assert os.path.exists(B.ABI_MODULE), "build the library first (mdir_amd/build.py)"
        self.image = open(B.ABI_MODULE, "rb").read()
        self.buf = C.create_string_buffer(self.image, len(self.image))
        self.module = L.mgpuModuleLoad(C.cast(self.buf, C.c_void_p), len(self.image))
        assert self.module
        self.fn = {}
        for name in ("abi_axpy", "abi_fill_slow", "abi_scale", "abi_add", "abi_set_flag"):
            self.fn[name] = L.mgpuModuleGetFunction(self.module, name.encode())
            assert self.fn[name]

    def launch(self, name, n, stream, *args):
        """Launches `name` over n elements; args are ctypes values."""
        L = rt.lib()
        holders = [C.c_void_p(a) if isinstance(a, int) and not isinstance(a, bool) else a
                   for a in args]
        params = (C.c_void_p * len(holders))(*[C.cast(C.pointer(h), C.c_void_p)
                                               for h in holders])
        grid = max(1, (n + BLOCK - 1) // BLOCK)
        L.mgpuLaunchKernel(self.fn[name], grid, 1, 1, BLOCK, 1, 1, 0, stream,
                           params, None, len(holders))

    def close(self):
        rt.lib().mgpuModuleUnload(self.module)


def ptr(p):
    return C.c_void_p(p)


def to_host(p, n, dtype, stream):
    out = np.empty(n, dtype)
    rt.lib().mgpuMemcpy(out.ctypes.data, p, out.nbytes, stream)
    return out


def test_module_load_and_launch():
    say("case: mgpuModuleLoad + mgpuLaunchKernel (abi_axpy, 1M floats)")
    L = rt.lib()
    m = Module()
    s = L.mgpuStreamCreate()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(N).astype(np.float32)
    y = rng.standard_normal(N).astype(np.float32)
    dx = L.mgpuMemAlloc(x.nbytes, s, False)
    dy = L.mgpuMemAlloc(y.nbytes, s, False)
    L.mgpuMemcpy(dx, x.ctypes.data, x.nbytes, s)
    L.mgpuMemcpy(dy, y.ctypes.data, y.nbytes, s)
    m.launch("abi_axpy", N, s, ptr(dy), ptr(dx), C.c_float(2.5), C.c_int(N))
    got = to_host(dy, N, np.float32, s)
    # the kernel may fuse a x + y into one rounding: compare with the exact
    # value within one f32 ulp
    exact = 2.5 * x.astype(np.float64) + y.astype(np.float64)
    assert np.all(np.abs(got - exact) <= np.spacing(np.abs(exact).astype(np.float32)))
    L.mgpuMemFree(dx, s)
    L.mgpuMemFree(dy, s)
    m.close()


def test_side_streams_fork_and_join():
    """A slow kernel on the first stream writes A; side stream 1 computes
    B = 2A and side stream 2 C = 3A; after the join the first stream adds
    them: every element is 5 value only if both side streams waited for A
    and the first stream waited for both."""
    say("case: side streams 1 and 2 fork from and join the first stream")
    L = rt.lib()
    m = Module()
    trunk = L.mgpuStreamCreate()
    nb = N * 4
    A, Bv, Cv, D = (L.mgpuMemAlloc(nb, trunk, False) for _ in range(4))
    for p in (A, Bv, Cv, D):
        L.mgpuMemset32(p, 0, N, trunk)
    L.mgpuStreamSynchronize(trunk)
    for value in (7, 11):
        m.launch("abi_fill_slow", N, trunk, ptr(A), C.c_int(value), C.c_int(N),
                 C.c_int(4000), C.c_uint(0))
        L.mdrtSideBegin()
        s1 = L.mgpuStreamCreate()
        assert s1 and s1 != trunk
        m.launch("abi_scale", N, s1, ptr(Bv), ptr(A), C.c_int(2), C.c_int(N))
        L.mdrtSideEnd()
        assert L.mgpuStreamCreate() == trunk
        L.mdrtSide2Begin()
        s2 = L.mgpuStreamCreate()
        assert s2 and s2 not in (trunk, s1)
        m.launch("abi_scale", N, s2, ptr(Cv), ptr(A), C.c_int(3), C.c_int(N))
        L.mdrtSide2End()
        L.mdrtSideJoin()
        m.launch("abi_add", N, trunk, ptr(D), ptr(Bv), ptr(Cv), C.c_int(N))
        got = to_host(D, N, np.int32, trunk)
        say(f"  value {value}: D in [{got.min()}, {got.max()}]")
        assert (got == 5 * value).all()
    for p in (A, Bv, Cv, D):
        L.mgpuMemFree(p, trunk)
    m.close()


def test_use_stream_nested_and_scoped_copy():
    """A caller's stream lent to the runtime: work, side streams and scoped
    copies follow it; a nested scope lends the legacy default stream; each
    restore brings back the stream its scope replaced."""
    say("case: mdrtUseStream with a caller stream and the default stream")
    L = rt.lib()
    hip = C.CDLL("libamdhip64.so.7")
    m = Module()
    own = L.mgpuStreamCreate()
    nb = N * 4
    A = L.mgpuMemAlloc(nb, own, False)
    Bv = L.mgpuMemAlloc(nb, own, False)
    L.mgpuMemset32(A, 0, N, own)
    L.mgpuMemset32(Bv, 0, N, own)
    caller = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(caller)) == 0
    previous = L.mdrtUseStream(caller)
    assert previous == own
    assert L.mgpuStreamCreate() == caller.value
    m.launch("abi_fill_slow", N, caller, ptr(A), C.c_int(123), C.c_int(N),
             C.c_int(4000), C.c_uint(0))
    L.mdrtSideBegin()
    side = L.mgpuStreamCreate()
    assert side not in (caller.value, own)
    m.launch("abi_scale", N, side, ptr(Bv), ptr(A), C.c_int(4), C.c_int(N))
    L.mdrtSideEnd()
    nested = L.mdrtUseStream(None)
    assert nested == caller.value
    assert not L.mgpuStreamCreate()
    got = np.empty(N, np.int32)
    L.mdrtCudaMemcpyScoped(got.ctypes.data, Bv, got.nbytes)
    assert (got == 492).all()
    L.mdrtRestoreStream(nested)
    assert L.mgpuStreamCreate() == caller.value
    L.mdrtRestoreStream(previous)
    assert L.mgpuStreamCreate() == own
    L.mgpuStreamSynchronize(own)
    assert hip.hipStreamSynchronize(caller) == 0
    assert hip.hipStreamDestroy(caller) == 0
    L.mgpuMemFree(A, own)
    L.mgpuMemFree(Bv, own)
    m.close()


def test_flags_set_by_a_kernel_and_host_results():
    say("case: a kernel sets mapped and device flags; host results")
    L = rt.lib()
    m = Module()
    s = L.mgpuStreamCreate()
    flag = rt.Buffer1()
    L._mlir_ciface_mdrtFlagMemory(C.byref(flag), 17)
    m.launch("abi_set_flag", 1, s, ptr(flag.aligned), C.c_int(5))
    slot = L.mdrtFlagStart(flag.aligned)
    assert L.mdrtFlagFinish(slot) == 5
    assert L.mdrtFlagTake(flag.aligned) == 0          # cleared by the read
    m.launch("abi_set_flag", 1, s, ptr(flag.aligned), C.c_int(9))
    L.mgpuStreamSynchronize(s)
    assert L.mdrtFlagTake(flag.aligned) == 9
    # a flag in device memory: copied to a staging word
    dflag = L.mgpuMemAlloc(4, s, False)
    m.launch("abi_set_flag", 1, s, ptr(dflag), C.c_int(-3))
    slots = [L.mdrtFlagStart(dflag) for _ in range(3)]
    assert [L.mdrtFlagFinish(k) for k in slots] == [-3, -3, -3]   # not cleared
    L.mgpuMemFree(dflag, s)
    # host results: zeroed, kept per place, replaced when a larger one is asked
    r = rt.Buffer1()
    L._mlir_ciface_mdrtHostResults(C.byref(r), 4, 10)
    assert r.size == 10 and r.aligned
    vals = np.ctypeslib.as_array(C.cast(r.aligned, C.POINTER(C.c_double)), (10,))
    assert (vals == 0).all()
    vals[:] = np.arange(10)
    again = rt.Buffer1()
    L._mlir_ciface_mdrtHostResults(C.byref(again), 4, 8)
    assert again.aligned == r.aligned
    big = rt.Buffer1()
    L._mlir_ciface_mdrtHostResults(C.byref(big), 4, 5000)
    assert big.size == 5000
    m.close()


def test_gaussian_backward_against_oracle_influence():
    """_mlir_ciface_mdrtCudaFFTGaussianBackwardF32: the backward transform
    of a spectrum times the influence function, with the influence given as
    separable tables (t_a = exp(-pi^2 m_a^2 / beta^2) |b_a(k)|^2, t_1 times
    f/(pi V); m_a the signed wave numbers k/L_a), against the oracle's
    influence function on the same grid (pme_callback.cu:5-21)."""
    say("case: Gaussian backward F32, 24 x 20 x 30 grid, order 4")
    L = rt.lib()
    w = S.tip3p_box(64, 1.3, seed=5, grid=(24, 20, 30))
    K = w.pme_grid
    bc = O.pme_influence(w)
    kmax = max(K)
    tab = np.zeros((6, kmax), np.float64)
    for a in range(3):
        Ka = K[a]
        mods = bspline_moduli(Ka, w.pme_order)
        ks = np.arange(Ka)
        ks = np.where(ks <= Ka // 2, ks, ks - Ka)
        m_a = ks / w.box[a]
        t = np.exp(-np.pi ** 2 * m_a ** 2 / w.beta ** 2) * mods
        if a == 0:
            t *= S.COULOMB / (np.pi * np.prod(w.box))
        tab[a, :Ka] = t
        tab[3 + a, :Ka] = m_a
    rng = np.random.default_rng(8)
    # the half spectrum of a real grid (Hermitian where the transform needs it)
    spec = O.fft3d_r2c(rng.standard_normal(K))
    s = L.mgpuStreamCreate()
    dspec = L.mgpuMemAlloc(spec.size * 8, s, False)
    dreal = L.mgpuMemAlloc(int(np.prod(K)) * 4, s, False)
    dtab = L.mgpuMemAlloc(tab.size * 4, s, False)
    packed = np.stack([spec.real, spec.imag], -1).astype(np.float32)
    tab32 = tab.astype(np.float32)
    L.mgpuMemcpy(dspec, packed.ctypes.data, packed.nbytes, s)
    L.mgpuMemcpy(dtab, tab32.ctypes.data, tab32.nbytes, s)
    bs = rt.Buffer1(dspec, dspec, 0, spec.size * 2, 1)
    br = rt.Buffer1(dreal, dreal, 0, int(np.prod(K)), 1)
    bt = rt.Buffer1(dtab, dtab, 0, tab.size, 1)
    L._mlir_ciface_mdrtCudaFFTGaussianBackwardF32(C.byref(bs), C.byref(br), C.byref(bt), *K)
    got = to_host(dreal, int(np.prod(K)), np.float32, s).reshape(K).astype(np.float64)
    want = O.fft3d_c2r(spec * bc, K[2])
    err = np.abs(got - want).max() / np.abs(want).max()
    say(f"  max error {err:.3e} of the largest value")
    assert err < 3e-5
    for p in (dspec, dreal, dtab):
        L.mgpuMemFree(p, s)


def cardinal_bspline(n, x):
    """M_n(x) by the recursion of Essmann et al. (1995), eq. 4.1:
    M_2(x) = 1 - |x - 1| on [0, 2], M_n(x) = (x M_{n-1}(x) + (n - x)
    M_{n-1}(x - 1)) / (n - 1)."""
    if n == 2:
        return max(0.0, 1.0 - abs(x - 1.0))
    return (x * cardinal_bspline(n - 1, x) + (n - x) * cardinal_bspline(n - 1, x - 1)) / (n - 1)


def bspline_moduli(K, order):
    """|b(k)|^2 = 1 / |sum_{j=0}^{n-2} M_n(j+1) e^{2 pi i k j / K}|^2."""
    vals = np.array([cardinal_bspline(order, j + 1.0) for j in range(order - 1)])
    k = np.arange(K)[:, None]
    j = np.arange(order - 1)[None, :]
    z = np.sum(vals[None, :] * np.exp(2j * np.pi * k * j / K), axis=1)
    return 1.0 / np.abs(z) ** 2

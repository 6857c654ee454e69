"""The runtime ABI on the GPU: FFT entry points against the oracle's FFT,
memory, streams and flags; NVE dynamics with SETTLE."""
import ctypes as C

import numpy as np
import pytest

from mdir_amd import runtime as rt
from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def dev_array(host):
    L = rt.lib()
    nbytes = host.nbytes
    p = L.mgpuMemAlloc(nbytes, L.mgpuStreamCreate(), False)
    L.mgpuMemcpy(p, host.ctypes.data, nbytes, L.mgpuStreamCreate())
    return p


def buf(p, n):
    return rt.Buffer1(p, p, 0, n, 1)


@pytest.mark.parametrize("K,single", [((16, 16, 16), 1), ((64, 64, 64), 1),
                                      ((24, 20, 30), 1), ((35, 35, 35), 1),
                                      ((32, 24, 18), 0), ((12, 14, 10), 0)])
def test_fft_abi_against_oracle(K, single):
    say(f"case: FFT ABI {K} single={single}")
    L = rt.lib()
    rng = np.random.default_rng(sum(K))
    real = rng.standard_normal(K)
    dt = np.float32 if single else np.float64
    h = K[2] // 2 + 1
    r_dev = dev_array(real.astype(dt))
    c_dev = L.mgpuMemAlloc(K[0] * K[1] * h * 2 * np.dtype(dt).itemsize, L.mgpuStreamCreate(), False)
    rb, cb = buf(r_dev, real.size), buf(c_dev, K[0] * K[1] * h * 2)
    if single:
        L._mlir_ciface_mdrtCudaFFTForward3DF32(C.byref(rb), C.byref(cb), *K)
    else:
        L._mlir_ciface_mdrtCudaFFTForward3D(C.byref(rb), C.byref(cb), *K)
    out = np.zeros((K[0], K[1], h, 2), dt)
    L.mgpuMemcpy(out.ctypes.data, c_dev, out.nbytes, L.mgpuStreamCreate())
    got = out[..., 0].astype(np.float64) + 1j * out[..., 1]
    ref = O.fft3d_r2c(real)
    tol = 3e-6 if single else 1e-12
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    # backward of the reference spectrum
    spec = np.stack([ref.real, ref.imag], -1).astype(dt)
    L.mgpuMemcpy(c_dev, spec.ctypes.data, spec.nbytes, L.mgpuStreamCreate())
    if single:
        L._mlir_ciface_mdrtCudaFFTBackward3DF32(C.byref(cb), C.byref(rb), *K)
    else:
        L._mlir_ciface_mdrtCudaFFTBackward3D(C.byref(cb), C.byref(rb), *K)
    back = np.zeros(K, dt)
    L.mgpuMemcpy(back.ctypes.data, r_dev, back.nbytes, L.mgpuStreamCreate())
    n = np.prod(K)
    assert np.abs(back / n - real).max() <= tol * 10 * np.abs(real).max()
    L.mgpuMemFree(r_dev, None)
    L.mgpuMemFree(c_dev, None)


def test_memory_pool_and_flags():
    say("case: pool, memset, flags")
    L = rt.lib()
    s = L.mgpuStreamCreate()
    p = L.mgpuMemAlloc(4096, s, False)
    L.mgpuMemset32(p, 0x01020304, 1024, s)
    h = np.zeros(1024, np.uint32)
    L.mgpuMemcpy(h.ctypes.data, p, 4096, s)
    assert (h == 0x01020304).all()
    L.mgpuMemFree(p, s)
    q = L.mgpuMemAlloc(4096, s, False)
    assert q == p  # the freed block serves the same request
    L.mgpuMemFree(q, s)
    counts = np.zeros(8, np.int64)
    L.mdrtDeviceMemoryCounts(counts.ctypes.data_as(C.POINTER(C.c_int64)))
    assert counts[2] >= 1
    flag = rt.Buffer1()
    L._mlir_ciface_mdrtFlagMemory(C.byref(flag), 3)
    one = np.array([7], np.int32)
    L.mgpuMemcpy(flag.aligned, one.ctypes.data, 4, s)
    slot = L.mdrtFlagStart(flag.aligned)
    assert L.mdrtFlagFinish(slot) == 7
    assert L.mdrtFlagTake(flag.aligned) == 0


@pytest.mark.parametrize("skin,prune", [(0.1, 0.0), (0.2, 0.05)])
def test_nve_settle_water(skin, prune):
    """NVE with SETTLE: constraints hold to round-off, and the conserved
    energy shows no drift beyond its fluctuation (slope of a least-squares
    line through 21 samples over 2,000 steps after 200 of equilibration),
    with the single list and with the pruned dual list."""
    say(f"case: NVE TIP3P 512 waters, 2,200 steps of 2 fs, skin {skin}, pruned {prune}")
    from mdir_amd.simulation import Simulation
    w = S.tip3p_box(512, 2.5, seed=21, grid=(24, 24, 24))
    sim = Simulation(w, skin=skin, prune_skin=prune)
    sim.minimize(50, 0.01)
    relaxed = w.copy()
    relaxed.positions = sim.positions
    sim.velocities = S.initial_velocities(relaxed, 42, 300.0)
    sim.run(200, 0.002, graph=50)
    times, energies = [], []
    for k in range(21):
        if k:
            sim.run(100, 0.002, graph=50)
        _, e = sim.forces()
        times.append(0.2 * k)
        energies.append(e.sum())
    x = sim.positions
    first = w.waters
    doh = np.linalg.norm(x[first + 1] - x[first], axis=1)
    dhh = np.linalg.norm(x[first + 2] - x[first + 1], axis=1)
    assert np.abs(doh - 0.09572).max() < 1e-10
    assert np.abs(dhh - 0.15139).max() < 1e-10
    slope = np.polyfit(times, energies, 1)[0] / w.n
    spread = np.std(energies) / w.n
    say(f"  drift {slope:.3e} kJ/mol/ps per atom, fluctuation {spread:.3e} kJ/mol per atom")
    assert abs(slope) < 2e-3

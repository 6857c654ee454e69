"""The MFMA transforms (DFT matrix products in split BF16x3 on the matrix
cores: whole planes in LDS for edges of 32 and 64, tiles of lines for the
other grids) against the FP32 Stockham FFT of the same library and the FP64
oracle.

Bars: the reciprocal forces of the two transforms within 3e-5 of the
largest (the reference's mixed-precision bar for its axis transforms,
test/Integration/pme-axis-fft.py:45-75); total forces within the force
path's 5e-6 relative RMS of FP64; reciprocal energy within 2e-6."""
import numpy as np
import pytest

from mdir_amd import runtime as rt
from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("case", [("tip3p 512, 32^3", 512, 2.5, 32),
                                  ("tip3p 2000, 64^3", 2000, 3.9, 64)])
def test_mfma_transform_forces(case):
    from mdir_amd.simulation import Simulation
    name, nw, edge, K = case
    say(f"case: MFMA plane transform, {name}")
    w = S.tip3p_box(nw, edge, seed=31, grid=(K, K, K))
    sim = Simulation(w)
    Ff, ef = sim.force_terms(rt.TERM_RECIP)
    sim.set_pme_transform(1)
    Fm, em = sim.force_terms(rt.TERM_RECIP)
    d = np.abs(Fm - Ff).max() / np.abs(Ff).max()
    say(f"  reciprocal forces, MFMA vs FFT: max difference / largest {d:.2e}; "
        f"energy {em[O.E_COUL_RECIP]:.6f} vs {ef[O.E_COUL_RECIP]:.6f}")
    assert d < 3e-5
    F, e = sim.forces()
    Fo, eo = O.Oracle(w).total()
    err = O.rel_rms(F, Fo)
    Fr, er = O.Oracle(w).pme()
    errr = O.rel_rms(Fm, Fr)
    say(f"  total rel rms force error {err:.2e}; reciprocal term {errr:.2e}")
    assert err < 5e-6
    assert abs(e[O.E_COUL_RECIP] - er) <= 2e-6 * abs(er)


@pytest.mark.parametrize("case", [("tip3p 2000, 72^3", (72, 72, 72), True),
                                  ("tip3p 2000, 80 x 96 x 72", (80, 96, 72), True),
                                  ("tip3p 2000, 216^3", (216, 216, 216), False),
                                  ("tip3p 2000, 48^3", (48, 48, 48), True),
                                  ("tip3p 2000, 36 x 56 x 40", (36, 56, 40), True)])
def test_mfma_axis_transform_forces(case):
    """Edges other than 32 and 64: the y and x DFTs as MFMA products over
    tiles of 32 kz lines, the lines padded to a multiple of 32 (72 -> 96,
    80 -> 96, 216 -> 224: 216 = 2^3 3^3, the 999,000-atom box's edge; 48,
    36, 56, 40 on grids whose planes would fit in LDS, which the variant
    switches to the tiles' layout), against the Stockham FFT; with the FP64
    oracle where it runs in seconds."""
    from mdir_amd.simulation import Simulation
    name, grid, oracle = case
    say(f"case: MFMA axis transform, {name}")
    w = S.tip3p_box(2000, 3.9, seed=31, grid=grid)
    sim = Simulation(w)
    Ff, ef = sim.force_terms(rt.TERM_RECIP)
    sim.set_pme_transform(1)
    Fm, em = sim.force_terms(rt.TERM_RECIP)
    d = np.abs(Fm - Ff).max() / np.abs(Ff).max()
    say(f"  reciprocal forces, MFMA vs FFT: max difference / largest {d:.2e}; "
        f"energy {em[O.E_COUL_RECIP]:.6f} vs {ef[O.E_COUL_RECIP]:.6f}")
    assert d < 3e-5
    assert abs(em[O.E_COUL_RECIP] - ef[O.E_COUL_RECIP]) <= 2e-6 * abs(ef[O.E_COUL_RECIP])
    if oracle:
        F, e = sim.forces()
        Fo, eo = O.Oracle(w).total()
        err = O.rel_rms(F, Fo)
        Fr, er = O.Oracle(w).pme()
        say(f"  total rel rms force error {err:.2e}; reciprocal term {O.rel_rms(Fm, Fr):.2e}")
        assert err < 5e-6
        assert abs(e[O.E_COUL_RECIP] - er) <= 2e-6 * abs(er)


def test_mfma_transform_refuses_wide_axes():
    from mdir_amd.simulation import Simulation
    w = S.tip3p_box(512, 2.5, seed=3, grid=(48, 240, 48))
    sim = Simulation(w)
    with pytest.raises(RuntimeError, match="MFMA DFT"):
        sim.set_pme_transform(1)


def test_mfma_layout_switch_and_back():
    """A 48^3 grid: the plane kernel (FFT), the MFMA tiles (another
    half-spectrum layout for the spread and gather), and the FFT again give
    the same reciprocal forces as before the switches, to the bit."""
    from mdir_amd.simulation import Simulation
    w = S.tip3p_box(512, 2.5, seed=4, grid=(48, 48, 48))
    sim = Simulation(w)
    F0, e0 = sim.force_terms(rt.TERM_RECIP)
    sim.set_pme_transform(1)
    F1, _ = sim.force_terms(rt.TERM_RECIP)
    sim.set_pme_transform(0)
    F2, e2 = sim.force_terms(rt.TERM_RECIP)
    assert np.abs(F1 - F0).max() < 3e-5 * np.abs(F0).max()
    assert np.array_equal(F2, F0) and e2[O.E_COUL_RECIP] == e0[O.E_COUL_RECIP]


def test_mfma_transform_frames():
    """The frame evaluator's forward and adjoint with the MFMA transform:
    the same forces and charge products as with the FFT, to the
    transforms' difference."""
    from mdir_amd.frames import FrameEvaluator
    say("case: frames with the MFMA transform, 960 atoms, 32^3")
    s = S.solvated_solutes(nwaters=300, nsolutes=12, seed=6, grid=(32, 32, 32))
    X = S.perturbed_frames(s, 3, seed=2)
    G = np.random.default_rng(5).normal(size=X.shape)
    ev = FrameEvaluator(s, max_frames=3)
    a = ev.evaluate(X)
    pa = a.vjp(positions=G)
    rt.check(rt.lib().mdrtHipFramesSetPmeTransform(ev._h, 1))
    b = ev.evaluate(X)
    pb = b.vjp(positions=G)
    for n in range(3):
        assert O.rel_rms(b.forces[n], O.Oracle(s).total(X[n])[0]) < 5e-6
    d = np.abs(pb["charges"] - pa["charges"]).max() / np.abs(pa["charges"]).max()
    say(f"  charge products, MFMA vs FFT: {d:.2e}")
    assert d < 3e-5

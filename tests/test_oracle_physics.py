"""The oracle's physics pinned by what the mathematics fixes: finite
differences, closed forms, invariants, an independent Ewald sum."""
import math

import numpy as np
import pytest

from fixtures import load_opc3
from mdir_amd import systems as S
from oracle import oracle as O


def fd_check(oracle, x, term_energy, atoms, h=1e-5, rtol=2e-6):
    """Central differences of an energy against the oracle's forces."""
    F, _ = oracle(x)
    for i in atoms:
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            d = -(term_energy(xp) - term_energy(xm)) / (2 * h)
            assert d == pytest.approx(F[i, k], rel=rtol, abs=1e-6 * max(1, abs(F[i]).max()))


@pytest.mark.parametrize("modifier", [S.LJ_NONE, S.LJ_POTENTIAL_SHIFT, S.LJ_FORCE_SWITCH,
                                      S.LJ_POTENTIAL_SWITCH])
def test_lj_forces_are_gradients(modifier):
    a = S.argon_fcc(cells=3, lattice=0.58, cutoff=0.8, switch_from=0.6,
                    lj_modifier=modifier)
    rng = np.random.default_rng(3)
    x = a.positions + rng.normal(0, 0.02, a.positions.shape)
    o = O.Oracle(a)
    fd_check(lambda y: o.total(y), x, lambda y: o.total(y)[1].sum(), atoms=[0, 5, 17])


def test_lj_modifier_boundary_values():
    """U(rc) = 0 for POTENTIAL_SHIFT, FORCE_SWITCH and POTENTIAL_SWITCH;
    F(rc) = 0 for the two switches (TermLennardJones.cpp:26-168)."""
    for modifier, zero_force in ((S.LJ_POTENTIAL_SHIFT, False), (S.LJ_FORCE_SWITCH, True),
                                 (S.LJ_POTENTIAL_SWITCH, True)):
        for r in (0.8 - 1e-10,):
            box = np.full(3, 5.0)
            x = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
            s = S.System(positions=x, masses=np.ones(2), charges=np.zeros(2),
                         types=np.zeros(2, np.int32), sigma=np.array([[0.3]]),
                         epsilon=np.array([[1.0]]), box=box, cutoff=0.8,
                         switch_from=0.6, lj_modifier=modifier)
            F, e = O.Oracle(s).total()
            assert abs(e[O.E_LJ]) < 1e-9
            if zero_force:
                assert abs(F[0, 0]) < 1e-6


def test_dispersion_correction_closed_form():
    """One type, no exclusions: E = -(2 pi/3) N^2 C6 / (V rc^3)."""
    a = S.argon_config1()
    _, e = O.Oracle(a).total()
    c6 = 4 * 0.996 * 0.3405 ** 6
    V = 3.47 ** 3
    n = a.n
    assert e[O.E_DISP_CORR] == pytest.approx(-2 * math.pi / 3 * n * n * c6 / (V * 1.0 ** 3),
                                             rel=1e-12)


def test_coulomb_forces_are_gradients():
    w = S.tip3p_box(27, 1.9, seed=1, grid=(20, 20, 20), cutoff=0.8)
    o = O.Oracle(w)
    x = w.positions.copy()
    fd_check(lambda y: o.total(y), x, lambda y: o.total(y)[1].sum(), atoms=[0, 1, 40],
             h=1e-5, rtol=5e-6)


def test_pme_matches_ewald_forces():
    """PME at order 8 on a fine grid agrees with the Ewald sum, energy and
    forces (method of test/Driver/Inputs/ewald_reference.py)."""
    s = load_opc3(grid=(64, 64, 64), order=8)
    o = O.Oracle(s)
    F, e = o.pme()
    Fe, ee = o.ewald_reciprocal()
    assert e == pytest.approx(ee, rel=2e-6)
    assert O.rel_rms(F, Fe) < 1e-5


def test_bspline_identities():
    for order in (4, 6, 8):
        for w in (0.0, 0.13, 0.5, 0.999):
            M, dM = O.bspline(w, order)
            assert M.sum() == pytest.approx(1.0, abs=1e-14)
            assert dM.sum() == pytest.approx(0.0, abs=1e-13)
    w = 0.37
    M, dM = O.bspline(w, 4)
    closed = [(1 - w) ** 3 / 6, (3 * w ** 3 - 6 * w ** 2 + 4) / 6,
              (-3 * w ** 3 + 3 * w ** 2 + 3 * w + 1) / 6, w ** 3 / 6]
    assert np.allclose(M, closed, atol=1e-15)
    h = 1e-6
    Mp, _ = O.bspline(w + h, 4)
    Mm, _ = O.bspline(w - h, 4)
    assert np.allclose((Mp - Mm) / (2 * h), dM, atol=1e-8)


def test_spline_moduli_definition():
    K, n = 24, 4
    mod = np.zeros(K)
    O.lib().orc_spline_moduli(K, n, mod.ctypes.data_as(O.PD))
    M = [1 / 6, 4 / 6, 1 / 6]  # M_4 at 1, 2, 3
    for k in range(K):
        s = sum(M[j] * np.exp(2j * np.pi * k * j / K) for j in range(3))
        assert mod[k] == pytest.approx(1 / abs(s) ** 2, rel=1e-13)


def test_pme_virial_trace_against_volume_derivative():
    """tr W = -dE/dln(V)-style check: sum_m E_m (1 - 2 pi^2 m^2/beta^2) equals
    -3 V dE/dV at fixed fractional coordinates."""
    s = load_opc3(grid=(32, 32, 32), order=6)
    o = O.Oracle(s)
    _, e, w = o.pme(virial=True)
    eps = 1e-6
    es = []
    for f in (1 + eps, 1 - eps):
        t = s.copy()
        t.box = s.box * f
        t.positions = s.positions * f
        es.append(O.Oracle(t).pme()[1])
    dE_dlnL = (es[0] - es[1]) / (2 * eps)
    assert (w[0] + w[1] + w[2]) == pytest.approx(-dE_dlnL, rel=1e-6)


def test_pairs14_special_values():
    """A lone scaled 1-4 pair: U_LJ(sigma) = 0 and U_LJ(2^(1/6) sigma) =
    -eps s_LJ with zero force (the minimum); Coulomb qq/r with its -qq/r^2
    force (TermPairs14.cpp:11-17, the pair excluded from the nonbonded sum)."""
    sig, eps = 0.31, 0.7
    for r, e_lj in ((sig, 0.0), (2 ** (1 / 6) * sig, -0.5 * eps)):
        x = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
        s = S.System(positions=x, masses=np.ones(2), charges=np.zeros(2),
                     types=np.zeros(2, np.int32), sigma=np.array([[0.3]]),
                     epsilon=np.array([[0.0]]), box=np.full(3, 5.0), cutoff=1.0,
                     exclusions=np.array([[0, 1]], np.int32),
                     pairs14=np.array([[0, 1]], np.int32), pairs14_sigma=np.array([sig]),
                     pairs14_eps=np.array([0.5 * eps]), pairs14_qq=np.array([0.0]))
        F, e = O.Oracle(s).total()
        assert e[O.E_LJ14] == pytest.approx(e_lj, abs=1e-12)
        if e_lj != 0.0:
            assert abs(F[0, 0]) < 1e-9
    x = np.array([[1.0, 1.0, 1.0], [1.0, 1.0 + 0.4, 1.0]])
    s = S.System(positions=x, masses=np.ones(2), charges=np.zeros(2),
                 types=np.zeros(2, np.int32), sigma=np.array([[0.3]]),
                 epsilon=np.array([[0.0]]), box=np.full(3, 5.0), cutoff=1.0,
                 exclusions=np.array([[0, 1]], np.int32),
                 pairs14=np.array([[0, 1]], np.int32), pairs14_sigma=np.array([0.3]),
                 pairs14_eps=np.array([0.0]), pairs14_qq=np.array([-7.0]))
    F, e = O.Oracle(s).total()
    assert e[O.E_COUL14] == pytest.approx(-7.0 / 0.4, rel=1e-14)
    assert F[1, 1] == pytest.approx(-7.0 / 0.16, rel=1e-12)  # attraction pulls j back


@pytest.mark.parametrize("modifier", [S.LJ_NONE, S.LJ_FORCE_SWITCH])
def test_chain_forces_are_gradients(modifier):
    """Exclusions, excluded-pair Ewald correction, 1-4 pairs and PME together:
    forces are the negative gradient of the total energy."""
    from fixtures import chain_box
    c = chain_box(nchains=20, edge=2.0, grid=(20, 20, 20), lj_modifier=modifier,
                  cutoff=0.8, switch_from=0.6)
    o = O.Oracle(c)
    fd_check(lambda y: o.total(y), c.positions.copy(), lambda y: o.total(y)[1].sum(),
             atoms=[0, 3, 41], h=1e-5, rtol=5e-6)


def test_cpu_baseline_integrator_follows_the_oracle():
    """The benchmark's CPU baseline (matrix reused until an atom moved half
    the buffer, f32 pair rows on wrapped f32 positions with its own erfc
    and exp, fixed-point PME spread) follows the oracle's velocity Verlet
    with the mixed-mode pair rows (matrix rebuilt at every step) to f32
    rounding over a few steps of an unrelaxed box, and rebuilds its matrix
    only when an atom moved half the buffer."""
    from mdir_amd import systems as S
    w = S.tip3p_box(216, 1.9, seed=4, grid=(20, 20, 20), cutoff=0.8)
    o = O.Oracle(w)
    v0 = o.initial_velocities(3, 300.0)
    xa, va, _ = o.nve(w.positions, v0, 0.001, 4, mixed=True)
    md = o.md(skin=0.15)
    xb, vb = md.run(w.positions, v0, 0.001, 4)
    assert np.abs(xa - xb).max() < 1e-7
    assert np.abs(va - vb).max() < 1e-4
    assert md.builds == 1
    md.close()


def test_erfc_fit_of_the_pair_kernels():
    """The pair kernels' erfc, t exp(-x^2) P(t) with t = 1/(1 +
    0.475 x), meets erfc to 1.2e-7 relative over [0, 3.6], and the fit of
    tests/golden/erfc_fit.py reproduces coefficients at least as good."""
    import importlib.util
    import os
    from scipy.special import erfc
    path = os.path.join(os.path.dirname(__file__), "golden", "erfc_fit.py")
    spec = importlib.util.spec_from_file_location("erfc_fit", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    x = np.linspace(0.0, 3.6, 50001)
    assert np.abs(mod.relative_error(mod.KERNEL, x)).max() < 1.2e-7
    assert np.abs(mod.relative_error(mod.fit(), x)).max() < 1.0e-7
    # the fit is of erfc, not of something close to it
    assert abs(mod.approx(mod.KERNEL, np.array([1.0]))[0] - erfc(1.0)) < 2e-8

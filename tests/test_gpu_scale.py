"""The other workloads of the force path on the GPU: the small argon NVE run,
a PME grid too wide for the plane transform (the tiled y / x passes), the
999,000-atom water box at full size through size-independent checks, and
an 8.2 nm box, the size from which the pair loop works on group origins."""
import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
KINETIC = 8  # MDRT_E_KINETIC


def say(msg):
    print(msg, flush=True)


def test_argon_nve_config():
    """864 argon atoms, LJ cutoff 1.0 nm with dispersion correction,
    Maxwell-Boltzmann velocities at 120 K, NVE 2 fs for 2,000 steps:
    forces match the oracle and the total energy stays within a small
    fraction of the kinetic energy."""
    from mdir_amd.simulation import Simulation
    say("case: argon 864, NVE 2,000 steps of 2 fs")
    a = S.argon_config1()
    # off the lattice, where forces do not cancel by symmetry
    a.positions = a.positions + np.random.default_rng(4).normal(0, 0.01, a.positions.shape)
    sim = Simulation(a)
    F, e = sim.forces()
    Fo, eo = O.Oracle(a).total()
    assert O.rel_rms(F, Fo) < 5e-6
    assert abs(e[O.E_LJ] - eo[O.E_LJ]) <= 2e-6 * abs(eo[O.E_LJ]) + 5e-7 * a.n
    sim.velocities = S.initial_velocities(a, 7, 120.0)
    _, e0 = sim.forces()
    total0 = e0.sum()
    ke0 = e0[KINETIC]
    sim.run(2000, 0.002, graph=50)
    _, e1 = sim.forces()
    say(f"  total energy {total0:.4f} -> {e1.sum():.4f} kJ/mol, kinetic {ke0:.2f}")
    assert abs(e1.sum() - total0) < 2e-2 * ke0


def test_wide_grid_tiled_passes():
    """A 72 x 72 x 36 grid (the xy plane exceeds what the plane transform
    holds): forces and the reciprocal energy against the oracle."""
    from mdir_amd.simulation import Simulation
    say("case: TIP3P 2,000 waters, 3.9 nm, grid 72 x 72 x 36")
    w = S.tip3p_box(2000, 3.9, seed=17, grid=(72, 72, 36))
    F, e = Simulation(w).forces()
    Fo, eo = O.Oracle(w).total()
    err = O.rel_rms(F, Fo)
    say(f"  rel rms force error {err:.3e}; recip {e[O.E_COUL_RECIP]:.6f} vs {eo[O.E_COUL_RECIP]:.6f}")
    assert err < 5e-6
    assert abs(e[O.E_COUL_RECIP] - eo[O.E_COUL_RECIP]) <= 2e-6 * abs(eo[O.E_COUL_RECIP]) + 5e-7 * w.n


def test_million_atom_box():
    """333,000 waters (999,000 atoms), 21.5 nm, PME 216^3: the list, the
    bricks and the graphs at full size.  The force error of each term
    against the FP64 oracle (nonbonded pairs, excluded pairs, PME
    reciprocal) and of the total; a net force near zero; SETTLE keeps the
    geometry through 20 steps.

    Bar: 5e-6 relative RMS per term, as for the 6.2 nm box: boxes with a
    half edge from 4 nm give the pair loop f32 positions relative to group
    origins, whose spacing does not grow with the box (about the box
    centre it would be 9.5e-7 nm at 10.75 nm, and the nonbonded error
    8.2e-6).  The nonbonded energies to 2e-6 relative."""
    from mdir_amd import runtime as rt
    from mdir_amd.simulation import Simulation
    say("case: 999,000 atoms, 21.5 nm, PME 216^3")
    w = S.tip3p_box(333000, 21.5, seed=5, grid=(216, 216, 216), jitter=0.02)
    sim = Simulation(w)
    F, e = sim.forces()
    assert np.isfinite(F).all() and np.isfinite(e).all()
    o = O.Oracle(w)
    Fo, eo = o.total()
    Fn, _ = o.nonbonded()
    Fr, er = o.pme()
    terms = {"nonbonded pairs": (rt.TERM_PAIRS, Fn), "PME reciprocal": (rt.TERM_RECIP, Fr),
             "excluded pairs": (rt.TERM_SPECIAL, Fo - Fn - Fr)}
    for name, (bit, Fw) in terms.items():
        Ft, _ = sim.force_terms(bit)
        err = O.rel_rms(Ft, Fw)
        say(f"  {name}: rel rms force error {err:.2e}")
        assert err < 5e-6
    err = O.rel_rms(F, Fo)
    say(f"  total: rel rms force error {err:.2e}; recip energy {e[O.E_COUL_RECIP]:.4f} vs {er:.4f}")
    assert err < 5e-6
    for k in (O.E_LJ, O.E_COUL_DIRECT):
        say(f"  energy term {k}: {e[k]:.4f} vs {eo[k]:.4f}")
        assert abs(e[k] - eo[k]) <= 2e-6 * abs(eo[k]) + 5e-7 * w.n
    # the split-BF16x3 MFMA DFT variant of the y and x transforms (216 padded
    # to 224) against the FP32 FFT and the FP64 oracle
    Ff, ef = sim.force_terms(rt.TERM_RECIP)
    sim.set_pme_transform(1)
    Fm, em = sim.force_terms(rt.TERM_RECIP)
    sim.set_pme_transform(0)
    d = np.abs(Fm - Ff).max() / np.abs(Ff).max()
    err = O.rel_rms(Fm, Fr)
    say(f"  MFMA DFT: reciprocal forces vs FFT, max difference / largest {d:.2e}; "
        f"rel rms error vs FP64 {err:.2e}; energy {em[O.E_COUL_RECIP]:.4f} vs {er:.4f}")
    assert d < 3e-5 and err < 5e-6
    assert abs(em[O.E_COUL_RECIP] - er) <= 2e-6 * abs(er)
    net = np.abs(F.sum(axis=0)).max() / np.sqrt((F * F).sum(axis=1).mean())
    say(f"  net force / rms atom force {net:.3e}")
    assert net < 1e-3 * np.sqrt(w.n)
    st = sim.stats()
    say(f"  stats {st}")
    sim.velocities = S.initial_velocities(w, 3, 300.0) * 0.1
    sim.run(20, 0.002, graph=10)
    x = sim.positions
    f = w.waters
    doh = np.linalg.norm(x[f + 1] - x[f], axis=1)
    assert np.abs(doh - 0.09572).max() < 1e-9


def test_group_origins_nve():
    """An 8.2 nm box (half edge 4.1 nm: the pair loop runs on positions
    relative to group origins): forces and nonbonded energies against the
    oracle, then NVE with SETTLE, whose steps move the positions in place
    relative to the origins of the last list build."""
    from mdir_amd.simulation import Simulation
    say("case: TIP3P 18,000 waters, 8.2 nm, grid 80^3 (group origins)")
    w = S.tip3p_box(18000, 8.2, seed=23, grid=(80, 80, 80), jitter=0.02)
    base = Simulation(w)
    base.minimize(50, 0.01)
    w = w.copy()
    w.positions = base.positions
    base.close()
    sim = Simulation(w)
    F, e = sim.forces()
    Fo, eo = O.Oracle(w).total()
    err = O.rel_rms(F, Fo)
    say(f"  rel rms force error {err:.2e}")
    assert err < 5e-6
    for k in (O.E_LJ, O.E_COUL_DIRECT, O.E_COUL_RECIP):
        say(f"  energy term {k}: {e[k]:.4f} vs {eo[k]:.4f}")
        assert abs(e[k] - eo[k]) <= 2e-6 * abs(eo[k]) + 5e-7 * w.n
    sim.velocities = S.initial_velocities(w, 9, 300.0)
    # the first steps constrain the random velocities (about 250 kJ/mol of
    # kinetic energy along the bonds goes); conservation from there
    sim.run(100, 0.002, graph=50)
    e0 = sum(sim.energies().values())
    sim.run(500, 0.002, graph=50)
    e1 = sum(sim.energies().values())
    st = sim.stats()
    say(f"  total energy {e0:.3f} -> {e1:.3f}; stats {st}")
    assert st["builds"] > 2
    assert abs(e1 - e0) < 1e-4 * abs(e0) + 5.0
    # the moved positions still give the oracle's forces
    F, _ = sim.forces()
    w2 = w.copy()
    w2.positions = sim.positions
    Fo, _ = O.Oracle(w2).total()
    assert O.rel_rms(F, Fo) < 5e-6
    sim.close()

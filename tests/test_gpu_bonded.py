"""Bonded tuples and constraint clusters on the GPU, through the C ABI,
against the oracle: the reference's alanine dipeptide in TIP3P water
(test/Driver/Inputs/dipeptide, kept under tests/golden/dipeptide/) with
PME, its bonds, angles, dihedrals and 1-4 pairs, SETTLE waters and the
bonds to hydrogen as M-SHAKE clusters."""
import os

import numpy as np
import pytest

from mdir_amd import runtime as rt
from mdir_amd import systems as S
from mdir_amd.amber import load_amber
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dipeptide")
DEV = {O.E_BOND: 9, O.E_UREY_BRADLEY: 10, O.E_ANGLE: 11, O.E_DIHEDRAL: 12,
       O.E_IMPROPER: 13, O.E_LJ14: 6, O.E_COUL14: 7}


def say(msg):
    print(msg, flush=True)


def dipeptide(**kw):
    return load_amber(os.path.join(GOLDEN, "dipeptide.prmtop"),
                      os.path.join(GOLDEN, "dipeptide.inpcrd"), **kw)


def test_bonded_forces_and_energies():
    """The tuples run in f64 and land in 64-bit words at 2^32 per kJ/mol/nm:
    the bonded forces alone (MDRT_TERM_BONDED) equal the oracle's to the
    word, and each tuple energy to f64 rounding; the whole force (f32 pairs
    and PME) meets the 5e-6 relative RMS bar."""
    from mdir_amd.simulation import Simulation
    say("case: dipeptide in 382 waters, PME, bonded tuples")
    s = dipeptide(rigid_water=True, hydrogen_bonds=False, pme=True)
    sim = Simulation(s)
    o = O.Oracle(s)
    Fb, eb = sim.force_terms(rt.TERM_BONDED)
    Fo, eo = o.bonded()
    assert np.abs(Fb - Fo).max() < 2.0 ** -31 * 4, np.abs(Fb - Fo).max()
    for k in (O.E_BOND, O.E_ANGLE, O.E_DIHEDRAL):
        assert abs(eb[DEV[k]] - eo[k]) < 1e-10 * max(1.0, abs(eo[k])), (k, eb[DEV[k]], eo[k])
    F, e = sim.forces()
    Fo, eo = o.total()
    rel = np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum())
    say(f"  total force error {rel:.2e} (relative RMS)")
    assert rel < 5e-6
    for k in (O.E_BOND, O.E_ANGLE, O.E_DIHEDRAL, O.E_LJ14, O.E_COUL14):
        assert abs(e[DEV[k]] - eo[k]) < 1e-7 * max(1.0, abs(eo[k])), (k, e[DEV[k]], eo[k])


@pytest.mark.parametrize("langevin", [False, True])
def test_constrained_steps_against_the_oracle(langevin):
    """Ten steps (velocity Verlet, or g-BAOAB with the same Philox noise)
    with SETTLE waters and M-SHAKE clusters from the same state as the
    oracle's; the bar is the one the water integrator test derives from the
    force error, and every constrained bond holds to 1e-10 nm."""
    from mdir_amd.simulation import Simulation
    say(f"case: dipeptide with hydrogen-bond clusters, 10 steps, langevin={langevin}")
    s = dipeptide(rigid_water=True, hydrogen_bonds=True, pme=True)
    o = O.Oracle(s)
    x0 = o.shake_positions(s.positions, s.positions)
    v0 = S.initial_velocities(s, 3, 300.0)
    v0 = o.rattle_velocities(x0, o.settle_velocities(x0, v0))
    s = s.copy()
    s.positions = x0
    sim = Simulation(s)
    sim.velocities = v0
    F, _ = sim.forces()
    Fo, _ = o.total(x0)
    a_err = (np.linalg.norm(F - Fo, axis=1) / s.masses).max()
    dt = 0.001
    if langevin:
        sim.set_langevin(300.0, 2.0, seed=99)
        sim.run(10, dt, graph=0)
        xo, vo, _ = o.langevin(x0, v0, dt, 10, 99, 2.0, 300.0)
    else:
        sim.run(10, dt, graph=0)
        xo, vo, _ = o.nve(x0, v0, dt, 10)
    x, v = sim.positions, sim.velocities
    k = 10
    bx = 8.0 * a_err * dt * dt * k * (k + 1) / 2 + 1e-12
    bv = 8.0 * a_err * dt * k + 1e-10
    dx, dv = np.abs(x - xo).max(), np.abs(v - vo).max()
    say(f"  |dx| {dx:.3e} (bar {bx:.3e}), |dv| {dv:.3e} (bar {bv:.3e})")
    assert dx < bx and dv < bv
    cl, length = s.clusters()
    worst = 0.0
    for first, size, off in cl:
        c = first + off
        for h in range(first, first + size):
            if h != c:
                worst = max(worst, abs(np.linalg.norm(x[h] - x[c]) - length[h]))
                assert abs((x[h] - x[c]) @ (v[h] - v[c])) < 1e-9
    assert worst < 1e-10, worst


def test_constrained_nve_conserves_energy():
    """20,000 steps of 1 fs (graphs of 50) of the dipeptide with SETTLE
    waters and hydrogen-bond clusters, the Lennard-Jones and direct Coulomb
    energies shifted to zero at the cutoff (a plain cutoff adds a random
    walk of about 1 kJ/mol per 2 ps here from pairs crossing it, the
    integrator's error hidden under it).  The total energy, sampled every
    1,000 steps, fluctuates by about 0.3 kJ/mol, so a slope fitted over
    20 ps is known to about 0.015 kJ/mol/ps; the bar is 5e-5 kJ/mol/ps per
    atom (0.058 kJ/mol/ps, five times the reference's NVE drift on JAC,
    10-performance.md:549), far below what a wrong kick, constraint or
    velocity correction gives (several kJ/mol/ps); the bonds stay at their
    lengths."""
    from mdir_amd.simulation import Simulation
    say("case: dipeptide NVE, 20000 steps")
    s = dipeptide(rigid_water=True, hydrogen_bonds=True, pme=True)
    s.lj_modifier = S.LJ_POTENTIAL_SHIFT
    s.coulomb_shift = True
    o = O.Oracle(s)
    x0 = o.shake_positions(s.positions, s.positions)
    s = s.copy()
    s.positions = x0
    sim = Simulation(s, skin=0.1)
    sim.minimize(50, 0.005)
    v0 = S.initial_velocities(s, 8, 300.0)
    x1 = sim.positions
    sim.velocities = o.rattle_velocities(x1, o.settle_velocities(x1, v0))
    _, e0 = sim.forces()
    totals = [e0.sum()]
    for _ in range(20):
        sim.run(1000, 0.001, graph=50)
        totals.append(sim.forces()[1].sum())
    dev = np.abs(np.array(totals) - totals[0]).max()
    slope = np.polyfit(np.arange(21) * 1.0, totals, 1)[0]
    say(f"  total {totals[0]:.4f} -> {totals[-1]:.4f} kJ/mol (largest change {dev:.4f}, "
        f"slope {slope:.4f} kJ/mol/ps), kinetic {e0[rt.KINETIC]:.2f}")
    assert abs(slope) < 5e-5 * s.n and dev < 1.5
    x = sim.positions
    cl, length = s.clusters()
    for first, size, off in cl:
        c = first + off
        for h in range(first, first + size):
            if h != c:
                assert abs(np.linalg.norm(x[h] - x[c]) - length[h]) < 1e-10


def test_frame_batches_refuse_tuples():
    from mdir_amd.frames import FrameEvaluator
    s = dipeptide(rigid_water=True, hydrogen_bonds=False, pme=True)
    with pytest.raises(RuntimeError, match="bonded tuples"):
        FrameEvaluator(s, max_frames=2)

"""Pressure coupling on the GPU: the Monte Carlo volume move's scaling of
the units' centres of mass, the energies at the new volume against the
oracle, the Metropolis decision recomputed from the oracle's energies and
the same Philox numbers, and NPT runs of TIP3P water (Langevin 300 K,
1 bar and 4,000 bar)."""
import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BAR = 1e5 * 1e-27 * 6.02214076e23 * 1e-3   # kJ mol^-1 nm^-3
WATER_G = 18.01528                          # g/mol (TIP3P masses: O 15.9994, H 1.008)


def say(msg):
    print(msg, flush=True)


def water_box():
    from mdir_amd.simulation import Simulation
    w = S.tip3p_box(512, 2.5, seed=21, grid=(24, 24, 24))
    base = Simulation(w)
    base.minimize(50, 0.01)
    w = w.copy()
    w.positions = base.positions
    base.close()
    return w


def scaled_units(w, x, s):
    """Each water's (or free atom's) centre of mass times s; shapes kept."""
    y = x.copy()
    units = [list(range(f, f + 3)) for f in w.waters]
    inw = np.zeros(w.n, bool)
    inw[np.concatenate([np.arange(f, f + 3) for f in w.waters])] = True
    units += [[i] for i in range(w.n) if not inw[i]]
    for u in units:
        m = w.masses[u]
        com = (m[:, None] * x[u]).sum(0) / m.sum()
        y[u] += (s - 1.0) * com
    return y


def test_volume_scaling_and_energies_against_oracle():
    from mdir_amd.simulation import Simulation
    say("case: box scaled by 1.012 and back, 512 waters")
    w = water_box()
    sim = Simulation(w, skin=0.1)
    x0 = sim.positions
    s = np.array([1.012, 1.012, 1.012])
    sim.scale_box(s)
    x1 = sim.positions
    want = scaled_units(w, x0, s)
    assert np.abs(x1 - want).max() < 1e-12
    assert np.allclose(sim.box, w.box * s, rtol=0, atol=1e-15)
    w1 = w.copy()
    w1.positions = x1
    w1.box = w.box * s
    F, e = sim.forces()
    Fo, eo = O.Oracle(w1).total()
    err = O.rel_rms(F, Fo)
    say(f"  scaled box {w1.box[0]:.5f} nm: rel rms force error {err:.3e}; "
        f"potential {e[:8].sum():.5f} vs {eo[:8].sum():.5f}")
    assert err < 5e-6
    assert abs(e[:8].sum() - eo[:8].sum()) < 2e-6 * abs(eo[:8]).sum() + 5e-7 * w.n
    sim.scale_box(1.0 / s)
    assert np.abs(sim.positions - x0).max() < 1e-12
    sim.close()


def test_barostat_move_against_oracle():
    """Moves at 1 bar and 300 K: U0 and U1 against the oracle's energies of
    the old and the scaled state, W = U1 - U0 + P dV - N kT ln(V1/V0), the
    volume change from the Philox words of (seed, step, 0, stream 1), and
    the decision (kept iff W <= 0 or u < exp(-W/kT))."""
    from mdir_amd.simulation import Simulation
    say("case: Monte Carlo volume moves, 512 waters")
    w = water_box()
    sim = Simulation(w, skin=0.1)
    kT = S.BOLTZMANN * 300.0
    seed = 77
    nunits = len(w.waters)
    for trial in range(4):
        sim.step_count = 1000 + trial
        x0, box0 = sim.positions, sim.box
        v0 = box0.prod()
        dv_max = sim.barostat_stats()["max_dv"] or 0.01 * v0
        kept, info = sim.barostat_move(1.0, 300.0, seed)
        step = 1000 + trial
        wd = O.philox4x32_10([step & 0xFFFFFFFF, step >> 32, 0, 1 << 24],
                             [seed & 0xFFFFFFFF, seed >> 32])
        u = lambda hi, lo: ((((hi << 32) | lo) >> 11) + 0.5) * 2.0 ** -53
        dv = dv_max * (2 * u(wd[0], wd[1]) - 1)
        assert info["dV"] == pytest.approx(dv, rel=1e-12)
        s = ((v0 + dv) / v0) ** (1.0 / 3.0)
        wa = w.copy()
        wa.positions, wa.box = x0, box0
        wb = w.copy()
        wb.positions, wb.box = scaled_units(w, x0, s), box0 * s
        u0 = O.Oracle(wa).total()[1][:8].sum()
        u1 = O.Oracle(wb).total()[1][:8].sum()
        tol = 2e-6 * 2.5e4
        assert abs(info["U0"] - u0) < tol and abs(info["U1"] - u1) < tol
        work = (u1 - u0) + BAR * dv - nunits * kT * np.log((v0 + dv) / v0)
        assert info["W"] == pytest.approx(work, abs=2 * tol)
        p = min(1.0, np.exp(-work / kT))
        ub = u(wd[2], wd[3])
        say(f"  move {trial}: dV {dv:+.4f} nm^3, W {work:+.3f} kJ/mol, p {p:.3f}, u {ub:.3f}, "
            f"kept {kept}")
        if abs(ub - p) > 1e-3:
            assert kept == (ub < p)
        if kept:
            assert np.allclose(sim.box, box0 * s, rtol=1e-14)
        else:
            assert np.array_equal(sim.box, box0)
            assert np.abs(sim.positions - x0).max() == 0.0
    sim.close()


def test_npt_water_density():
    """Langevin 300 K (1/ps) with a volume move every 20 steps: 512 TIP3P
    waters settle at a liquid density at 1 bar (TIP3P with a 0.9 nm cutoff
    and dispersion correction: near 0.98 g/cm^3), and 4,000 bar compresses
    the box by several per cent.  Moves are accepted at 15-85 %."""
    from mdir_amd.simulation import Simulation
    w = water_box()
    mass = len(w.waters) * WATER_G / 6.02214076e23   # g
    dens = {}
    for pressure in (1.0, 4000.0):
        say(f"case: NPT 512 waters at {pressure:g} bar, 300 K, 16 ps")
        sim = Simulation(w, skin=0.1)
        sim.velocities = S.initial_velocities(w, 5, 300.0)
        sim.set_langevin(300.0, 1.0, 31)
        sim.run_npt(2000, 0.002, graph=20, every=20, pressure=pressure, temperature=300.0,
                    seed=9)
        vols = []
        for k in range(30):
            sim.run_npt(200, 0.002, graph=20, every=20, pressure=pressure, temperature=300.0,
                        seed=9)
            vols.append(sim.box.prod())
        st = sim.barostat_stats()
        rho = mass / (np.mean(vols) * 1e-21)
        dens[pressure] = rho
        say(f"  density {rho:.4f} g/cm^3 (volume {np.mean(vols):.3f} +- {np.std(vols):.3f} "
            f"nm^3), moves kept {st['kept']}/{st['tried']}")
        assert 0.15 < st["kept"] / st["tried"] < 0.85
        x = sim.positions
        f = w.waters
        assert np.abs(np.linalg.norm(x[f + 1] - x[f], axis=1) - w.water_geometry[0]).max() < 1e-10
        sim.close()
    assert 0.93 < dens[1.0] < 1.04
    assert dens[4000.0] > 1.04 * dens[1.0]

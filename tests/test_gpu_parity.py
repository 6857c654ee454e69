"""Parity of the workloads at their own sizes, and of the step path itself,
through the C ABI against the FP64 oracle.

- The 23,700-atom box of the benchmark (7,900 TIP3P waters, 6.2 nm, PME
  64^3, rc 0.9 nm) after 200 minimisation steps: forces, energy terms, the
  pair set at rc bit for bit, steps with SETTLE, and the forces of the list
  the steps hold.
- The integrator (velocity Verlet, SETTLE, the velocity constraint) over
  1 and 10 steps from one state against the oracle's nve, bounded by what
  the measured force error can move the state.
- The list between rebuilds (the counterpart of the reference's
  MDIR_CHECK_PRUNING, runtime/mdrt.c:48-79): the forces the steps compute
  through the held list, and the pairs within rc in the list and in the
  pruned rows, in a 2.5 nm box and in the 8.2 nm box whose pair loop and
  prune work on group origins.
- The argon example (test/Driver/argon.test, examples/argon/argon.toml):
  the total energy changes by less than 1e-4 of its value over 2,000 steps
  of 5 fs.
- The differentiation workload at its size: 64 frames of 3,150 atoms."""
import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
KCAL = 4.184


def say(msg):
    print(msg, flush=True)


def oracle_pair_keys(system, x, reach, excluded=True):
    """The reference matrix's unordered pairs at `reach` (exclusions
    removed by its filter) as sorted keys i << 32 | j, i < j."""
    counts, index, _ = O.neighbors_matrix(x, system.box, reach)
    if excluded:
        es, ea = system.exclusion_csr()
        counts, index = O.exclusion_filter(counts, index, es, ea)
    rows = np.repeat(np.arange(len(counts)), counts)
    cols = index[np.arange(index.shape[1])[None, :] < counts[:, None]]
    lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
    return np.unique((lo.astype(np.int64) << 32) | hi)


def keys_of(pairs):
    return (pairs[:, 0].astype(np.int64) << 32) | pairs[:, 1]


def energy_bar(eo, k, n):
    """2e-6 relative plus 5e-7 kJ/mol per atom (the pair terms run on f32
    positions: a strongly cancelling sum gets a per-atom floor)."""
    return 2e-6 * abs(eo[k]) + 5e-7 * n


def water_bonds_ok(w, x, tol=1e-10):
    f = w.waters
    doh1 = np.linalg.norm(x[f + 1] - x[f], axis=1)
    doh2 = np.linalg.norm(x[f + 2] - x[f], axis=1)
    dhh = np.linalg.norm(x[f + 2] - x[f + 1], axis=1)
    d_oh, d_hh = w.water_geometry
    return max(np.abs(doh1 - d_oh).max(), np.abs(doh2 - d_oh).max(),
               np.abs(dhh - d_hh).max()) < tol


def bond_velocity(w, x, v):
    """Largest relative velocity along a rigid bond (nm/ps)."""
    f = w.waters
    worst = 0.0
    for a, b in ((0, 1), (0, 2), (1, 2)):
        e = x[f + b] - x[f + a]
        e /= np.linalg.norm(e, axis=1)[:, None]
        worst = max(worst, np.abs(np.sum(e * (v[f + b] - v[f + a]), axis=1)).max())
    return worst


def test_config2_box_at_full_size():
    """The benchmark's box: 7,900 TIP3P waters in 6.2 nm, PME 64^3 order 4,
    rc 0.9 nm, list buffer 0.15 nm, relaxed by 200 steepest-descent steps.
    Forces within 5e-6 relative RMS of FP64, energies per term, the pair set
    within rc bit for bit (exclusions removed as the reference filter
    does), 50 graph-replayed steps with SETTLE, and the forces of the list
    the steps hold after them."""
    from mdir_amd.simulation import Simulation
    say("case: 7,900 TIP3P waters, 6.2 nm, 64^3, minimised 200 steps")
    w = S.tip3p_config2()
    sim = Simulation(w, skin=0.15)
    sim.minimize(200, 0.01)
    w = w.copy()
    w.positions = sim.positions
    F, e = sim.forces()
    o = O.Oracle(w)
    Fo, eo = o.total()
    err = O.rel_rms(F, Fo)
    say(f"  rel rms force error {err:.3e}")
    assert err < 5e-6
    for k in (O.E_LJ, O.E_COUL_DIRECT, O.E_COUL_EXCL, O.E_COUL_RECIP, O.E_COUL_SELF):
        say(f"  term {O.TERM_NAMES[k]}: {e[k]:.6f} vs {eo[k]:.6f}")
        assert abs(e[k] - eo[k]) <= energy_bar(eo, k, w.n), (k, e[k], eo[k])
    pairs = sim.neighbor_pairs(w.cutoff, excluded=True)
    want = oracle_pair_keys(w, w.positions, w.cutoff)
    say(f"  pairs within rc: {len(pairs)} (oracle {len(want)})")
    assert np.array_equal(np.sort(keys_of(pairs)), want)
    sim.velocities = S.initial_velocities(w, 42, 300.0)
    sim.run(50, 0.002, graph=50)
    x = sim.positions
    assert water_bonds_ok(w, x)
    Fl, el, counts = sim.check_list()
    Fo2, eo2 = o.total(x)
    err2 = O.rel_rms(Fl, Fo2)
    say(f"  after 50 steps: list forces rel rms {err2:.3e}, counts {counts}")
    assert err2 < 5e-6
    assert counts["list"] == counts["pair_loop"] == len(oracle_pair_keys(w, x, w.cutoff))


@pytest.mark.parametrize("case", ["water", "argon"])
def test_integrator_against_oracle_steps(case):
    """One and ten velocity-Verlet steps (with SETTLE and the velocity
    constraint for water) from the same positions and velocities as the
    oracle's nve (FP64 forces).  The difference can only come from the force
    error: with a_err the largest |F - F_oracle| / m over the atoms at the
    start, k steps move positions by at most about a_err dt^2 k(k+1)/2 and
    velocities by a_err dt k; the bars are eight times that (SETTLE's
    mass-weighted projection can hand an oxygen's error to a hydrogen
    enlarged by up to sqrt(m_water / m_H), about 4.2).  A dropped or
    doubled half kick would move velocities by dt |F| / (2m), which the
    test checks is at least 100 times the velocity bar."""
    from mdir_amd.simulation import Simulation
    dt = 0.002
    if case == "water":
        say("case: integrator against the oracle, TIP3P 512 waters, 2.5 nm")
        w = S.tip3p_box(512, 2.5, seed=21, grid=(24, 24, 24))
        base = Simulation(w)
        base.minimize(50, 0.01)
        w = w.copy()
        w.positions = base.positions
        base.close()
        v0 = O.Oracle(w).initial_velocities(42, 300.0)
    else:
        say("case: integrator against the oracle, argon 864, rc 1.0 + dispersion")
        w = S.argon_config1()
        w.positions = w.positions + np.random.default_rng(4).normal(0, 0.01, w.positions.shape)
        v0 = O.Oracle(w).initial_velocities(7, 120.0)
    x0 = w.positions.copy()
    sim = Simulation(w, skin=0.1)
    sim.velocities = v0
    F, _ = sim.forces()
    o = O.Oracle(w)
    Fo, _ = o.total(x0)
    a_err = (np.linalg.norm(F - Fo, axis=1) / w.masses).max()
    kick = (0.5 * dt * np.linalg.norm(Fo, axis=1) / w.masses).max()
    say(f"  largest acceleration error {a_err:.3e} nm/ps^2; half kick {kick:.3e} nm/ps")
    done = 0
    for k in (1, 10):
        sim.run(k - done, dt, graph=0)
        done = k
        x, v = sim.positions, sim.velocities
        xo, vo, _ = o.nve(x0, v0, dt, k)
        dx, dv = np.abs(x - xo).max(), np.abs(v - vo).max()
        bx = 8.0 * a_err * dt * dt * k * (k + 1) / 2 + 1e-12
        bv = 8.0 * a_err * dt * k + 1e-10
        say(f"  {k} steps: |dx| {dx:.3e} (bar {bx:.3e}) nm, |dv| {dv:.3e} (bar {bv:.3e}) nm/ps")
        assert kick > 100 * bv
        assert dx < bx
        assert dv < bv
        if case == "water":
            assert water_bonds_ok(w, x)
            bvel = bond_velocity(w, x, v)
            say(f"  largest velocity along a bond {bvel:.3e} nm/ps")
            assert bvel < 1e-9
    sim.close()


@pytest.mark.parametrize("box,skin,prune", [("small", 0.1, 0.0), ("small", 0.2, 0.05),
                                            ("origins", 0.2, 0.05)])
def test_list_between_rebuilds(box, skin, prune):
    """After runs that reuse the list (fewer builds than steps), the forces
    the steps compute through the held list meet the oracle's at the
    current positions, and the list and the rows the pair loop runs over
    (the pruned rows with a dual list) hold every pair within rc: as many
    as the reference matrix finds there.  The 8.2 nm box runs the pair
    loop and the prune on positions relative to group origins."""
    from mdir_amd.simulation import Simulation
    if box == "small":
        say(f"case: list check, TIP3P 512 waters 2.5 nm, skin {skin}, pruned {prune}")
        w = S.tip3p_box(512, 2.5, seed=11, grid=(24, 24, 24))
        rounds, per = 8, 10
    else:
        say(f"case: list check, TIP3P 18,000 waters 8.2 nm (group origins), skin {skin}, "
            f"pruned {prune}")
        w = S.tip3p_box(18000, 8.2, seed=23, grid=(80, 80, 80), jitter=0.02)
        rounds, per = 4, 25
    base = Simulation(w)
    base.minimize(50, 0.01)
    w = w.copy()
    w.positions = base.positions
    base.close()
    sim = Simulation(w, skin=skin, prune_skin=prune)
    sim.velocities = S.initial_velocities(w, 5, 300.0)
    o = O.Oracle(w)
    b0 = sim.stats()["builds"]
    for r in range(rounds):
        sim.run(per, 0.002, graph=per)
        x = sim.positions
        F, e, counts = sim.check_list()
        Fo, _ = o.total(x)
        err = O.rel_rms(F, Fo)
        want = len(oracle_pair_keys(w, x, w.cutoff))
        say(f"  after {(r + 1) * per} steps: rel rms {err:.3e}, pairs {counts} (oracle {want})")
        assert err < 5e-6
        assert counts["builds"] == 0
        assert counts["list"] == counts["pair_loop"]
        if counts["list"] != want:
            # The pair set is defined by an f32 test with the margin
            # 3e-6 (Lx+Ly+Lz); pairs the two counts disagree on must lie in
            # that band, where f32 positions of a large box (ulp 5e-7 nm at
            # 8 nm) decide; every pair inside rc itself must be in the list.
            have = keys_of(sim.list_pairs(w.cutoff))
            ref = oracle_pair_keys(w, x, w.cutoff)
            margin = 3e-6 * w.box.sum()
            held = set(have.tolist())
            for k in np.setxor1d(have, ref).tolist():
                i, j = k >> 32, k & 0xFFFFFFFF
                d = x[i] - x[j]
                d -= w.box * np.round(d / w.box)
                r = np.linalg.norm(d)
                say(f"  pair ({i}, {j}) at {r:.9f} nm: in list {k in held}")
                assert w.cutoff - 2e-6 < r < w.cutoff + margin + 2e-6
            assert abs(counts["list"] - want) <= 3
    st = sim.stats()
    say(f"  stats {st}")
    assert st["builds"] - b0 < rounds * per
    if prune > 0:
        assert st["prunes"] > 1
    sim.close()


def test_argon_example_energy_conservation():
    """test/Driver/argon.test with examples/argon/argon.toml: 864 argon
    atoms on the lattice of argon.mlir, LJ sigma 3.4 A, epsilon 0.238465
    kcal/mol, potential switch 7.5-8.5 A, pair list at 9.5 A; velocities at
    183 K from seed 314159 (470.7558 kcal/mol); NVE velocity Verlet, 2,000
    steps of 5 fs.  The reference requires a total-energy change below
    1e-4 of its value (argon.test:14)."""
    from mdir_amd.simulation import Simulation
    say("case: argon example, 2,000 steps of 5 fs from 183 K")
    a = S.argon_fcc()
    sim = Simulation(a, skin=0.1)
    sim.velocities = S.initial_velocities(a, 314159, 183.0)
    _, e0 = sim.forces()
    say(f"  start: potential {e0[O.E_LJ] / KCAL:.4f}, kinetic {e0[8] / KCAL:.4f} kcal/mol")
    assert abs(e0[O.E_LJ] / KCAL - (-1304.88)) < 0.005
    assert abs(e0[8] / KCAL - 470.7558) < 1e-4
    total0 = e0.sum()
    sim.run(2000, 0.005, graph=50)
    _, e1 = sim.forces()
    change = abs(e1.sum() - total0) / abs(total0)
    say(f"  total {total0 / KCAL:.5f} -> {e1.sum() / KCAL:.5f} kcal/mol: {change:.3e} of its value")
    assert change < 1e-4


def test_force_matching_workload_at_full_size():
    """64 frames of the 3,150-atom workload (1,000 TIP3P waters, 30
    five-site solutes of 4 LJ types, PME 32^3, rc 0.9 nm): batched forces
    equal one-frame evaluations bit for bit; the forces of three frames
    meet the oracle; the force-matching gradient of a 4-frame subset at
    six tunables against FP64 central differences of the oracle (h = 1e-4)
    within 5e-4 of the largest entry of each tunable (the frames' mixed
    precision bar, as for the small systems)."""
    from mdir_amd.frames import FrameEvaluator
    say("case: 64 frames x 3,150 atoms")
    s = S.solvated_solutes()
    X = S.perturbed_frames(s, 64, seed=9)
    q1, s1, e1 = S.perturbed_parameters(s)
    ev = FrameEvaluator(s, max_frames=64)
    out = ev.evaluate(X)
    one = FrameEvaluator(s, max_frames=1)
    for k in (0, 31, 63):
        single = one.evaluate(X[k:k + 1])
        assert np.array_equal(single.forces[0], out.forces[k])
        Fo, _ = O.Oracle(s).total(X[k])
        err = O.rel_rms(out.forces[k], Fo)
        say(f"  frame {k}: rel rms force error {err:.3e}")
        assert err < 5e-6
    one.close()
    ref = S.with_parameters(s, q1, s1, e1)
    sub = X[[0, 21, 42, 63]]
    Fref = np.stack([O.Oracle(ref).total(sub[n])[0] for n in range(4)])
    ev4 = FrameEvaluator(s, max_frames=4)
    loss, grad = ev4.force_matching(sub, Fref)
    gnorm = np.sqrt(sum(np.sum(g * g) for g in grad.values()))
    say(f"  loss {loss}; gradient norm {gnorm:.4e}")

    def L(sys_):
        return sum(np.sum((O.Oracle(sys_).total(sub[n])[0] - Fref[n]) ** 2) for n in range(4))
    h = 1e-4
    q = s.charges.copy()
    sig, eps = S.type_parameters(s)
    picks = [("charges", 0), ("charges", 3000), ("charges", s.n - 1),
             ("sigma", 0), ("sigma", 3), ("epsilon", 4)]
    for name, i in picks:
        def at(d):
            v = {"charges": q.copy(), "sigma": sig.copy(), "epsilon": eps.copy()}
            v[name][i] += d
            return L(S.with_parameters(s, v["charges"], v["sigma"], v["epsilon"]))
        fd = (at(h) - at(-h)) / (2 * h)
        scale = np.abs(grad[name]).max()
        rel = abs(grad[name][i] - fd) / scale
        say(f"  d loss / d {name}[{i}]: {grad[name][i]:.6e} vs {fd:.6e} ({rel:.2e} of the largest)")
        assert rel < 5e-4
    ev.close()
    ev4.close()


@pytest.mark.parametrize("case", ["water", "argon"])
def test_langevin_against_oracle_steps(case):
    """g-BAOAB Langevin steps on the GPU against the oracle's from the same
    state, seed and step numbers: the noise is the same Philox draw on both
    sides, so the difference comes from the force error alone, bounded as
    for velocity Verlet (eight times a_err dt^2 k(k+1)/2 for positions,
    a_err dt k for velocities).  A noise of another step, particle or
    stream would move velocities by about spread/sqrt(m), which the test
    checks is far above the bar."""
    from mdir_amd.simulation import Simulation
    dt, gamma, T, seed = 0.002, 1.0, 300.0, 2718
    if case == "water":
        say("case: Langevin against the oracle, TIP3P 512 waters, 2.5 nm")
        w = S.tip3p_box(512, 2.5, seed=21, grid=(24, 24, 24))
        base = Simulation(w)
        base.minimize(50, 0.01)
        w = w.copy()
        w.positions = base.positions
        base.close()
    else:
        say("case: Langevin against the oracle, argon 864")
        w = S.argon_config1()
        w.positions = w.positions + np.random.default_rng(4).normal(0, 0.01, w.positions.shape)
        T = 120.0
    o = O.Oracle(w)
    v0 = o.initial_velocities(7, T)
    x0 = w.positions.copy()
    sim = Simulation(w, skin=0.1)
    sim.velocities = v0
    sim.step_count = 40
    sim.set_langevin(T, gamma, seed)
    F, _ = sim.forces()
    Fo, _ = o.total(x0)
    a_err = (np.linalg.norm(F - Fo, axis=1) / w.masses).max()
    c = np.exp(-gamma * dt)
    noise = np.sqrt((1 - c * c) * S.BOLTZMANN * T / w.masses.max())
    done = 0
    for k, graph in ((1, 0), (10, 0), (30, 10)):
        sim.run(k - done, dt, graph=graph)
        done = k
        assert sim.step_count == 40 + k
        x, v = sim.positions, sim.velocities
        xo, vo, _ = o.langevin(x0, v0, dt, k, seed=seed, friction=gamma, temperature=T,
                               first_step=40)
        dx, dv = np.abs(x - xo).max(), np.abs(v - vo).max()
        bx = 8.0 * a_err * dt * dt * k * (k + 1) / 2 + 1e-12
        bv = 8.0 * a_err * dt * k + 1e-10
        say(f"  {k} steps: |dx| {dx:.3e} (bar {bx:.3e}), |dv| {dv:.3e} (bar {bv:.3e}); "
            f"noise per step {noise:.3e} nm/ps")
        assert noise > 20 * bv
        assert dx < bx and dv < bv
        if case == "water":
            assert water_bonds_ok(w, x)
            assert bond_velocity(w, x, v) < 1e-9
    sim.close()


def test_langevin_thermostat_holds_the_temperature():
    """512 TIP3P waters started at 200 K under a 300 K bath (friction 5/ps)
    reach and hold the bath temperature: the mean over the last 6 ps of the
    instantaneous temperature is within 3 % of 300 K."""
    from mdir_amd.simulation import Simulation
    say("case: Langevin thermostat, 512 waters, 200 K -> 300 K bath")
    w = S.tip3p_box(512, 2.5, seed=21, grid=(24, 24, 24))
    sim = Simulation(w, skin=0.1)
    sim.minimize(50, 0.01)
    relaxed = w.copy()
    relaxed.positions = sim.positions
    sim.velocities = S.initial_velocities(relaxed, 3, 200.0)
    sim.set_langevin(300.0, 5.0, 99)
    dof = w.degrees_of_freedom()
    temps = []
    for k in range(40):
        ke = sim.run(100, 0.002, graph=50)
        temps.append(2.0 * ke / (dof * S.BOLTZMANN))
    late = np.mean(temps[10:])
    say(f"  temperature after 2 ps {temps[9]:.1f} K, mean of the last 6 ps {late:.1f} K")
    assert abs(late / 300.0 - 1.0) < 0.03
    assert water_bonds_ok(w, sim.positions)

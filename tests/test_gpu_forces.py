"""Parity of the HIP force path with the FP64 oracle (through the C ABI).

Bar: relative RMS force error <= 5e-6 against FP64 for the
mixed-precision path, energies per term to the stated tolerances below,
pair sets bit-exact."""
import numpy as np
import pytest

from fixtures import KCAL, lcg_positions, load_opc3
from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def sim(system, skin=0.1):
    from mdir_amd.simulation import Simulation
    return Simulation(system, skin=skin)


def say(msg):
    print(msg, flush=True)


def test_argon_lattice_energy_and_forces():
    """examples/argon lattice: -5459.64 kJ/mol (test/Driver/argon.test:8-12)."""
    say("case: argon lattice 864, potential switch 0.75-0.85")
    a = S.argon_fcc()
    F, e = sim(a).forces()
    Fo, eo = O.Oracle(a).total()
    assert abs(e[0] - (-5459.64)) < 0.01
    assert abs(e[0] - eo[O.E_LJ]) < 1e-5 * abs(eo[O.E_LJ])
    # perfect lattice: forces vanish by symmetry
    assert np.abs(F).max() < 1e-3


def test_argon_config1_perturbed_forces():
    say("case: argon config 1 (rc 1.0, dispersion), perturbed lattice")
    a = S.argon_config1()
    rng = np.random.default_rng(7)
    a.positions = a.positions + rng.normal(0, 0.01, a.positions.shape)
    F, e = sim(a).forces()
    Fo, eo = O.Oracle(a).total()
    assert O.rel_rms(F, Fo) < 5e-6
    assert abs(e[0] - eo[O.E_LJ]) < 2e-6 * abs(eo[O.E_LJ])
    assert abs(e[5] - eo[O.E_DISP_CORR]) < 1e-9 * abs(eo[O.E_DISP_CORR])


@pytest.mark.parametrize("modifier", [S.LJ_NONE, S.LJ_POTENTIAL_SHIFT,
                                      S.LJ_FORCE_SWITCH, S.LJ_POTENTIAL_SWITCH])
def test_lj_modifiers(modifier):
    say(f"case: argon modifiers {modifier}")
    a = S.argon_fcc(cells=5, lattice=0.58, cutoff=1.0, switch_from=0.8,
                    lj_modifier=modifier)
    rng = np.random.default_rng(modifier)
    a.positions = a.positions + rng.normal(0, 0.02, a.positions.shape)
    F, e = sim(a).forces()
    Fo, eo = O.Oracle(a).total()
    assert O.rel_rms(F, Fo) < 5e-6
    assert abs(e[0] - eo[O.E_LJ]) < 5e-6 * abs(eo[O.E_LJ])


def test_opc3_terms_golden():
    """test/Driver/pme.test:20-23 (kcal/mol): direct -14.3572978, excluded
    535.1260682, reciprocal 4.6972413 (64^3, order 8), self -539.7928271."""
    say("case: OPC three waters, 64^3 order 8")
    s = load_opc3()
    F, e = sim(s).forces()
    assert abs(e[1] / KCAL - (-14.3572978)) < 2e-6 * 14.36
    assert abs(e[2] / KCAL - 535.1260682) < 1e-9 * 535.13
    assert abs(e[3] / KCAL - 4.6972413) < 5e-6 * 4.697
    assert abs(e[4] / KCAL - (-539.7928271)) < 1e-9 * 539.8
    Fo, eo = O.Oracle(s).total()
    assert O.rel_rms(F, Fo) < 5e-6


def test_opc3_default_grid_reciprocal():
    """pme.test:17-18: 4.693822 kcal/mol at the 35^3 order-4 grid."""
    say("case: OPC three waters, 35^3 order 4")
    s = load_opc3(grid=(35, 35, 35), order=4)
    _, e = sim(s).forces()
    assert abs(e[3] / KCAL - 4.693822) < 5e-6 * 4.694


def test_tip3p_small_box_forces():
    say("case: TIP3P 512 waters, 2.5 nm, 24^3")
    w = S.tip3p_box(512, 2.5, seed=11, grid=(24, 24, 24))
    F, e = sim(w).forces()
    o = O.Oracle(w)
    Fo, eo = o.total()
    err = O.rel_rms(F, Fo)
    say(f"  rel rms force error {err:.3e}")
    assert err < 5e-6
    # Energy bar: 2e-6 relative plus 5e-7 kJ/mol per atom.  The pair terms
    # run on f32 positions (about 1.2e-7 nm at 1.25 nm from the box centre),
    # which bounds a strongly cancelling sum like the water LJ energy by a
    # per-atom floor rather than a fraction of the net.
    for k in (O.E_LJ, O.E_COUL_DIRECT, O.E_COUL_EXCL, O.E_COUL_RECIP, O.E_COUL_SELF):
        bar = 2e-6 * abs(eo[k]) + 5e-7 * w.n
        say(f"  term {O.TERM_NAMES[k]}: {e[k]:.8f} vs {eo[k]:.8f}")
        assert abs(e[k] - eo[k]) <= bar, (k, e[k], eo[k])


@pytest.mark.parametrize("modifier", [S.LJ_NONE, S.LJ_FORCE_SWITCH, S.LJ_POTENTIAL_SHIFT])
def test_chain_box_scaled_pairs(modifier):
    """Exclusions beyond water, the excluded-pair Ewald correction and scaled
    1-4 pairs under each LJ modifier, with PME order 6."""
    from fixtures import chain_box
    say(f"case: 216 four-atom chains, 3.6 nm, 36^3 order 6, modifier {modifier}")
    c = chain_box(nchains=216, edge=3.6, grid=(36, 36, 36), order=6, lj_modifier=modifier,
                  cutoff=0.9, switch_from=0.7)
    F, e = sim(c).forces()
    Fo, eo = O.Oracle(c).total()
    err = O.rel_rms(F, Fo)
    say(f"  rel rms force error {err:.3e}")
    assert err < 5e-6
    for k in (O.E_LJ, O.E_COUL_DIRECT, O.E_COUL_EXCL, O.E_COUL_RECIP, O.E_COUL_SELF,
              O.E_LJ14, O.E_COUL14):
        bar = 2e-6 * abs(eo[k]) + 5e-7 * c.n
        say(f"  term {O.TERM_NAMES[k]}: {e[k]:.8f} vs {eo[k]:.8f}")
        assert abs(e[k] - eo[k]) <= bar, (k, e[k], eo[k])


def test_profile_step_follows_the_dynamics():
    """Kernel-by-kernel profiling (mdrtHipProfileStep) advances the state as
    the step sequence does, with the pruned list too: its build prunes the
    rows the pair kernel then reads."""
    from mdir_amd.simulation import Simulation
    say("case: profiled steps against run steps, TIP3P 512 waters, skin 0.2, pruned 0.05")
    w = S.tip3p_box(512, 2.5, seed=11, grid=(24, 24, 24))
    v = O.Oracle(w).initial_velocities(42, 300.0)
    a = Simulation(w, skin=0.2, prune_skin=0.05)
    b = Simulation(w, skin=0.2, prune_skin=0.05)
    a.velocities = v
    b.velocities = v
    a.profile_step(30)
    b.run(30, 0.002, graph=0)
    d = np.abs(a.positions - b.positions).max()
    say(f"  largest position difference {d:.3e} nm")
    assert d < 1e-9
    a.close()
    b.close()


def test_pruned_list():
    """The dual list: the pair loop over the pruned rows gives the oracle's
    forces, keeps fewer tiles than the list, re-prunes as atoms move, and
    runs bit-identically twice."""
    from mdir_amd.simulation import Simulation
    say("case: pruned list, TIP3P 512 waters, skin 0.2, pruned buffer 0.05")
    w = S.tip3p_box(512, 2.5, seed=11, grid=(24, 24, 24))
    a = Simulation(w, skin=0.2, prune_skin=0.05)
    F, e = a.forces()
    Fo, eo = O.Oracle(w).total()
    err = O.rel_rms(F, Fo)
    st = a.stats()
    say(f"  rel rms force error {err:.3e}; tiles {st['tiles']} -> {st['pair_loop_tiles']}")
    assert err < 5e-6
    assert 0 < st["pair_loop_tiles"] < st["tiles"]
    b = Simulation(w, skin=0.2, prune_skin=0.05)
    v = O.Oracle(w).initial_velocities(42, 300.0)
    a.velocities = v
    b.velocities = v
    a.run(100, 0.002, graph=25)
    b.run(100, 0.002, graph=25)
    assert np.array_equal(a.positions, b.positions)
    assert a.stats()["prunes"] > 1
    F, _ = a.forces()
    o = O.Oracle(w)
    Fo, _ = o.total(a.positions)
    assert O.rel_rms(F, Fo) < 5e-6


def test_pme_against_ewald_sum():
    say("case: PME reciprocal vs Ewald sum, 64 waters")
    w = S.tip3p_box(64, 2.2, seed=3, grid=(32, 32, 32), order=6, cutoff=0.9)
    F, e = sim(w).forces()
    o = O.Oracle(w)
    Fe, ee = o.ewald_reciprocal()
    assert abs(e[3] - ee) < 1e-4 * abs(ee)


@pytest.mark.parametrize("count,edge,reach,entries", [
    (200, 6.0, 1.5, 2666), (2000, 12.0, 1.5, 32704), (2000, 7.0, 1.0, 48778),
    (2000, 5.8, 0.8, 43958)])
def test_neighbor_pair_set_bit_exact(count, edge, reach, entries):
    """The pair set equals the reference matrix's, pair by pair; entry counts
    are the golden values of test/Runtime/neighbors-matrix*.mlir."""
    say(f"case: pair set {count} {edge} {reach}")
    x = lcg_positions(count, edge)
    s = S.System(positions=x, masses=np.ones(count), charges=np.zeros(count),
                 types=np.zeros(count, np.int32), sigma=np.zeros((1, 1)),
                 epsilon=np.zeros((1, 1)), box=np.full(3, edge), cutoff=0.5)
    pairs = sim(s, skin=0.05).neighbor_pairs(reach)
    assert 2 * len(pairs) == entries
    counts, index, _ = O.neighbors_matrix(x, [edge] * 3, reach)
    ref = set()
    for i in range(count):
        for k in range(counts[i]):
            j = int(index[i, k])
            ref.add((min(i, j), max(i, j)))
    assert set(map(tuple, pairs.tolist())) == ref


def test_neighbor_exclusion_masks_bit_exact():
    """Excluded pairs are dropped exactly as the reference's exclusion filter
    drops them (MDExecKernels.cpp:1748-1823)."""
    say("case: exclusion masks, TIP3P 512")
    w = S.tip3p_box(512, 2.5, seed=5, grid=(24, 24, 24))
    reach = 1.0
    pairs = sim(w).neighbor_pairs(reach, excluded=True)
    counts, index, _ = O.neighbors_matrix(w.positions, w.box, reach)
    es, ea = w.exclusion_csr()
    counts, index = O.exclusion_filter(counts, index, es, ea)
    ref = set()
    for i in range(w.n):
        for k in range(counts[i]):
            j = int(index[i, k])
            ref.add((min(i, j), max(i, j)))
    assert set(map(tuple, pairs.tolist())) == ref


def test_determinism_bitwise():
    say("case: determinism, two evaluations and two runs")
    w = S.tip3p_box(512, 2.5, seed=9, grid=(24, 24, 24))
    a, b = sim(w), sim(w)
    Fa, _ = a.forces()
    Fb, _ = b.forces()
    assert np.array_equal(Fa, Fb)
    v = O.Oracle(w).initial_velocities(42, 300.0)
    a.velocities = v
    b.velocities = v
    a.run(20, 0.002)
    b.run(20, 0.002, graph=10)
    assert np.array_equal(a.positions, b.positions)
    assert np.array_equal(a.velocities, b.velocities)


def _special_forces(system, x):
    """Excluded-pair correction plus 1-4 pairs from the oracle: the total
    less the nonbonded pairs and the reciprocal sum."""
    o = O.Oracle(system)
    F, _ = o.total(x)
    Fn, _ = o.nonbonded(x)
    Fr = o.pme(x)[0] if system.beta > 0 else 0.0
    return F - Fn - Fr


@pytest.mark.parametrize("case", ["tip3p", "chains"])
def test_force_error_per_term(case):
    """Each force term alone against its FP64 oracle term: the nonbonded
    pairs (LJ + direct Coulomb), the excluded and 1-4 pairs, the PME
    reciprocal sum; and the terms add up to the total bit for bit in
    fixed point (to 2^-32 kJ/mol/nm per term)."""
    from mdir_amd import runtime as rt
    from fixtures import chain_box
    say(f"case: per-term forces, {case}")
    s = S.tip3p_box(2000, 3.9, seed=21) if case == "tip3p" else chain_box(nchains=120)
    sm = sim(s)
    o = O.Oracle(s)
    want = {rt.TERM_PAIRS: o.nonbonded()[0], rt.TERM_RECIP: o.pme()[0],
            rt.TERM_SPECIAL: _special_forces(s, s.positions)}
    total, _ = sm.forces()
    parts = 0.0
    for term, Fo in want.items():
        F, _ = sm.force_terms(term)
        err = O.rel_rms(F, Fo)
        say(f"  term {term}: rel rms force error {err:.2e} (rms force {np.sqrt((Fo * Fo).sum(1).mean()):.3e})")
        assert err < 5e-6
        parts = parts + F
    assert np.abs(parts - total).max() <= 3 * 2.0 ** -32

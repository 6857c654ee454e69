"""The launch-shape trial (mdrtHipAutotune): every PME brick shape and plane
transform it times gives the forces of the FP64 oracle, the state is
restored, and the shape it keeps runs the dynamics."""
import numpy as np
import pytest

from mdir_amd import systems as S
from mdir_amd.simulation import Simulation
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def water(grid):
    w = S.tip3p_box(512, 2.5, seed=21, grid=grid)
    base = Simulation(w, skin=0.1)
    base.minimize(50, 0.01)
    w = w.copy()
    w.positions = base.positions
    base.close()
    return w


@pytest.mark.parametrize("grid,what", [((24, 24, 24), Simulation.TUNE_BRICKS),
                                       ((32, 32, 32), Simulation.TUNE_BRICKS
                                        | Simulation.TUNE_TRANSFORM
                                        | Simulation.TUNE_SEGMENTS)])
def test_autotune_keeps_state_and_forces(grid, what):
    say(f"case: autotune TIP3P 512 waters, grid {grid}, what {what}")
    w = water(grid)
    sim = Simulation(w, skin=0.1)
    sim.velocities = S.initial_velocities(w, 5, 300.0)
    x0, v0 = sim.positions, sim.velocities
    F0, _ = sim.forces()
    times = sim.autotune(what, steps=20, graphs=2)
    say(f"  ms per step: {times}")
    # every brick shape that fits a z edge of 24 or 32 (8 z lines at most)
    assert {tl for tl, _, _ in times} == {8, 4, 2, 1}
    if what & Simulation.TUNE_TRANSFORM:
        assert {tr for _, tr, _ in times} == {0, 1}
    if what & Simulation.TUNE_SEGMENTS:
        assert {sg for _, _, sg in times} == {16, 24, 48}
        assert len(times) == 24
    else:
        assert {sg for _, _, sg in times} == {16}
    assert all(t > 0 for t in times.values())
    # positions and velocities restored bit for bit
    assert np.array_equal(sim.positions, x0)
    assert np.array_equal(sim.velocities, v0)
    # the kept shape: the oracle's forces, and the untuned forces up to the
    # f32 rounding of the gather's per-brick partial sums (9e-9 measured at
    # 24^3, 1.04e-7 at 32^3), or, with the MFMA transform kept, its
    # split-BF16x3 error
    F1, _ = sim.forces()
    Fo, _ = O.Oracle(w).total()
    assert O.rel_rms(F1, Fo) < 5e-6
    assert O.rel_rms(F1, F0) < (2e-6 if what & Simulation.TUNE_TRANSFORM else 5e-7)
    # and it runs: 200 NVE steps conserve energy to the usual bar
    e0 = sum(sim.energies().values())
    sim.run(200, 0.002, graph=50)
    e1 = sum(sim.energies().values())
    say(f"  total energy {e0:.3f} -> {e1:.3f}")
    assert abs(e1 - e0) < 5e-4 * abs(e0) + 5.0
    sim.close()


def test_frames_autotune():
    """The frame evaluator's trial: every shape it times, then the kept one,
    gives the force-matching loss and gradient of the untuned evaluator."""
    from mdir_amd.frames import FrameEvaluator
    say("case: frames autotune, 300 waters + 12 solutes, 4 frames, grid 32")
    system = S.solvated_solutes(nwaters=300, nsolutes=12, grid=(32, 32, 32))
    X = S.perturbed_frames(system, 4, seed=3)
    q1, s1, e1 = S.perturbed_parameters(system)
    ev = FrameEvaluator(system, max_frames=4)
    theta = ev.tunables
    ev.update(q1, s1, e1)
    Fref = ev.evaluate(X).forces
    ev.update(theta["charges"], theta["sigma"], theta["epsilon"])
    loss0, g0 = ev.force_matching(X, Fref)
    times = ev.autotune(X, Fref, what=7, repeats=2)
    say(f"  ms per call: {times}")
    assert {tl for tl, _, _ in times} == {8, 4, 2, 1}
    assert {tr for _, tr, _ in times} == {0, 1}
    assert {sg for _, _, sg in times} == {16, 24, 48}
    loss1, g1 = ev.force_matching(X, Fref)
    # brick shapes differ by the f32 rounding of per-brick gather sums, the
    # MFMA transform (if kept) by its split-BF16x3 error, segment lengths by
    # the order of the per-wave partial sums
    assert np.allclose(loss1, loss0, rtol=1e-4)
    for name in ("charges", "sigma", "epsilon"):
        a, b = g1[name], g0[name]
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b), name
    ev.close()


def test_every_launch_shape_matches_the_oracle():
    """Each of the 24 launch shapes, set directly, gives the oracle's forces;
    the Stockham shapes agree with each other to the f32 rounding of the
    gather's per-brick sums, their energies to the order of the f64
    partial sums."""
    say("case: every launch shape, TIP3P 512 waters, grid 32")
    w = water((32, 32, 32))
    Fo, _ = O.Oracle(w).total()
    sim = Simulation(w, skin=0.1)
    ref = None
    for seg in Simulation.SEGMENTS:
        for bricks in (8, 4, 2, 1):
            for transform in (0, 1):
                sim.launch_shape = (seg, bricks, transform)
                assert sim.launch_shape == (seg, bricks, transform)
                F, e = sim.forces()
                err = O.rel_rms(F, Fo)
                assert err < 5e-6, (seg, bricks, transform, err)
                if transform == 0:
                    if ref is None:
                        ref = (F, e)
                    else:
                        # measured up to 1.04e-7 between brick widths at 32^3
                        assert O.rel_rms(F, ref[0]) < 5e-7
                        assert np.allclose(e, ref[1], rtol=1e-9, atol=1e-6)
    with pytest.raises(RuntimeError):
        sim.launch_shape = (32, 8, 0)   # not a segment length
    with pytest.raises(RuntimeError):
        sim.launch_shape = (16, 3, 0)   # not a brick width
    # a non-default shape runs the dynamics (captured graphs)
    sim.launch_shape = (24, 4, 0)
    sim.velocities = S.initial_velocities(w, 5, 300.0)
    e0 = sum(sim.energies().values())
    sim.run(200, 0.002, graph=50)
    e1 = sum(sim.energies().values())
    assert abs(e1 - e0) < 5e-4 * abs(e0) + 5.0
    sim.close()

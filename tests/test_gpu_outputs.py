"""Outputs on the GPU, through the C ABI: the staging entry points and kept
memory, DCD trajectories written from staged copies (byte for byte the
oracle's file of the same positions), and checkpoints whose continuation
takes the saved run's steps bit for bit, trajectory included (the
reference's test/Driver/Inputs/python_checkpoints.py asks the same: "equal
to the bit; energy file and DCD byte for byte")."""
import ctypes as C
import os
import time

import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import dcd as odcd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def water(n=600, seed=31):
    w = S.tip3p_box(n, S.water_edge(n), seed=seed)
    return w


def started(w, skin=0.1, prune_skin=0.0, langevin=False):
    from mdir_amd.simulation import Simulation
    sim = Simulation(w, skin=skin, prune_skin=prune_skin)
    sim.minimize(30, 0.01)
    sim.velocities = S.initial_velocities(w, 11, 300.0)
    if langevin:
        sim.set_langevin(300.0, 1.0, seed=77)
    return sim


def test_staging_and_kept_memory():
    from mdir_amd import runtime as rt
    L = rt.lib()
    say("case: staging slot of 3 parts, kept memory, a word held back 200 ms")
    n = 1 << 16
    a = np.arange(n, dtype=np.float64)
    d = L.mdrtDeviceAllocateKept(a.nbytes)
    d2 = L.mdrtDeviceAllocateKept(a.nbytes)
    assert d and d2
    stream = L.mgpuStreamCreate()
    L.mgpuMemcpy(C.c_void_p(d), a.ctypes.data_as(C.c_void_p), a.nbytes, stream)
    L.mdrtDeviceCopyWithin(C.c_void_p(d2), C.c_void_p(d), a.nbytes)
    back = np.zeros_like(a)
    L.mdrtDeviceCopyToHost(back.ctypes.data_as(C.c_void_p), C.c_void_p(d2), a.nbytes)
    assert np.array_equal(back, a)

    slot = L.mdrtStagingCreate()
    assert L.mdrtStagingCopy(slot, 4, C.c_void_p(d), 8) is None  # parts 0-3 only
    call = C.c_char_p()
    for rnd, delay in enumerate(("", "200", "")):
        os.environ["MDRT_STAGING_MARK_DELAY"] = delay
        b = a * (rnd + 2)
        L.mgpuMemcpy(C.c_void_p(d), b.ctypes.data_as(C.c_void_p), b.nbytes, stream)
        hosts = [L.mdrtStagingCopy(slot, p, C.c_void_p(d + 8 * 1024 * p), 8 * 1024 * (p + 1))
                 for p in range(3)]
        L.mdrtStagingEnd(slot)
        t0 = time.time()
        assert L.mdrtStagingWait(slot, C.byref(call)) == 0
        waited = time.time() - t0
        for p, h in enumerate(hosts):
            got = np.ctypeslib.as_array(C.cast(C.c_void_p(h), C.POINTER(C.c_double)),
                                        shape=(1024 * (p + 1),))
            assert np.array_equal(got, b[1024 * p: 1024 * p + 1024 * (p + 1)]), (rnd, p)
        if delay:  # the word lands after the copies' event: the wait took the event
            assert waited < 0.2, waited
    os.environ.pop("MDRT_STAGING_MARK_DELAY", None)
    time.sleep(0.3)  # the held-back word of round 1 lands long after round 2's
    assert L.mdrtStagingWait(slot, C.byref(call)) == 0
    L.mdrtStagingDestroy(slot)
    L.mdrtDeviceFreeKept(C.c_void_p(d))
    L.mdrtDeviceFreeKept(C.c_void_p(d2))


def test_trajectory_matches_the_oracle_file(tmp_path):
    """A run with a TrajectoryReporter every 10 steps writes the oracle's
    file of the positions a twin run reaches at those steps."""
    from mdir_amd.output import TrajectoryReporter, read_dcd
    say("case: 600 waters, 40 steps, a frame every 10")
    w = water()
    a, b = started(w), started(w)
    path = str(tmp_path / "run.dcd")
    a.reporters = [TrajectoryReporter(path, 10)]
    a.run(40, 0.002, graph=50)
    a.close_reporters()
    xs, boxes = [], []
    for _ in range(4):
        b.run(10, 0.002, graph=50)
        xs.append(b.positions)
        boxes.append(b.box)
    assert np.array_equal(a.positions, b.positions)
    want = odcd.write(np.array(xs), np.array(boxes), first=10, period=10, timestep_ps=0.002)
    got = open(path, "rb").read()
    assert got == want
    numbers, _, pos, _ = read_dcd(path)
    assert numbers[0] == 4 and pos.shape == (4, w.n, 3)


def dipeptide_started(langevin=False, **kw):
    """The reference's alanine dipeptide in water (tests/golden/dipeptide)
    with bonded tuples and hydrogen-bond clusters."""
    from mdir_amd.amber import load_amber
    from mdir_amd.simulation import Simulation
    from oracle import oracle as O
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dipeptide")
    s = load_amber(os.path.join(g, "dipeptide.prmtop"), os.path.join(g, "dipeptide.inpcrd"),
                   hydrogen_bonds=True)
    o = O.Oracle(s)
    s.positions = o.shake_positions(s.positions, s.positions)
    sim = Simulation(s, **kw)
    v = S.initial_velocities(s, 4, 300.0)
    sim.velocities = o.rattle_velocities(s.positions, o.settle_velocities(s.positions, v))
    if langevin:
        sim.set_langevin(300.0, 1.0, seed=5)
    return sim, s


@pytest.mark.parametrize("case", ["nve", "langevin", "pruned", "dipeptide"])
def test_checkpoint_continues_bit_for_bit(tmp_path, case):
    """Whole: 60 steps, frames every 10, checkpoints every 20.  Part: 30
    steps of the same, stopped; a new simulation from the checkpoint of
    step 20 continues the trajectory after its 2 frames and runs 40 steps.
    The state at step 60 and the trajectory equal the whole run's, bit for
    bit and byte for byte."""
    from mdir_amd.output import CheckpointReporter, TrajectoryReporter, read_checkpoint
    from mdir_amd.simulation import Simulation
    say(f"case: checkpoint continuation, {case}")
    w = water(500, seed=37)
    kw = dict(skin=0.15, prune_skin=0.05) if case == "pruned" else dict(skin=0.1)
    lang = case == "langevin"
    whole_dir, part_dir = tmp_path / "whole", tmp_path / "part"
    whole_dir.mkdir()
    part_dir.mkdir()

    if case == "dipeptide":
        whole, w = dipeptide_started(**kw)
    else:
        whole = started(w, langevin=lang, **kw)
    whole.reporters = [TrajectoryReporter(str(whole_dir / "run.dcd"), 10),
                       CheckpointReporter(str(whole_dir / "run.ckpt"), 20)]
    whole.run(60, 0.002, graph=50)
    whole.close_reporters()

    part = dipeptide_started(**kw)[0] if case == "dipeptide" else started(w, langevin=lang, **kw)
    part.reporters = [TrajectoryReporter(str(part_dir / "run.dcd"), 10),
                      CheckpointReporter(str(part_dir / "run.ckpt"), 20)]
    part.run(30, 0.002, graph=50)
    part.close_reporters()
    ck = read_checkpoint(str(part_dir / "run.ckpt"))
    assert ck.step == 20 and ck.frames == 2 and ck.trajectory == "run.dcd"
    assert ck.list_kept
    assert os.path.exists(str(whole_dir / "run.ckpt.prev"))  # step 40's, kept beside 60's

    cont = Simulation.from_checkpoint(w, str(part_dir / "run.ckpt"), **kw)
    assert cont.step_count == 20
    cont.reporters = [TrajectoryReporter(str(part_dir / "run.dcd"), 10)]
    cont.run(40, 0.002, graph=50)
    cont.close_reporters()
    assert cont.step_count == 60
    assert np.array_equal(cont.positions, whole.positions)
    assert np.array_equal(cont.velocities, whole.velocities)
    assert (part_dir / "run.dcd").read_bytes() == (whole_dir / "run.dcd").read_bytes()
    # the whole run's last checkpoint is the continued run's state
    last = read_checkpoint(str(whole_dir / "run.ckpt"))
    assert last.step == 60 and np.array_equal(last.positions, cont.positions)

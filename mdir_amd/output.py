"""Outputs of a run: DCD trajectories and checkpoints.

Mirrors the reference's Python reporters and checkpoints
(python/Bindings.cpp:1318-1390 TrajectoryReporter and CheckpointReporter,
:1703 Simulation.save_checkpoint, :1110-1179 Checkpoint and
read_checkpoint; the state kept, include/mdir/Driver/Checkpoint.h:36-136):

    sim.reporters = [TrajectoryReporter("run.dcd", 10), CheckpointReporter("run.ckpt", 10)]
    sim.run(20)
    sim.close_reporters()
    state = read_checkpoint("run.ckpt")
    cont = Simulation.from_checkpoint(system, "run.ckpt")   # continues run.dcd too

The trajectory is the reference's DCD (lib/Driver/Trajectory.cpp), written
by libmdrt_hip.so from staged copies on a thread of its own.  Checkpoints
hold the reference's quantities in a NumPy .npz container: the reference
writes H5MD through the HDF5 library, which this image does not have.  A
checkpoint also holds the positions the run's neighbor list was built (and
pruned) at, so that a continuation takes the steps the saved run would
have taken, bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import runtime as rt

CHECKPOINT_FORMAT = 1
AKMA_PS = 0.04888821  # ps per AKMA time unit (the DCD header's time step)


def has_checkpoint_support() -> bool:
    """Checkpoints need no library beyond NumPy here (the reference's
    hasCheckpointSupport, include/mdir/Driver/Checkpoint.h:140, asks for
    HDF5)."""
    return True


class TrajectoryReporter:
    """A DCD frame every `period` steps (python/Bindings.cpp:1340-1357)."""

    def __init__(self, file: str, period: int):
        if not str(file).lower().endswith(".dcd"):
            raise ValueError(f"TrajectoryReporter writes DCD; cannot write '{file}'")
        if int(period) <= 0:
            raise ValueError("the period must be a positive number of steps")
        self.file, self.period = str(file), int(period)
        self._h = None

    def _open(self, sim, keep: int):
        step = sim.step_count
        first = (step // self.period + 1) * self.period
        h = rt.lib().mdrtHipTrajectoryOpen(sim._h, self.file.encode(), first, self.period,
                                            float(sim.timestep), int(keep))
        if not h:
            raise RuntimeError("mdrt_hip: " + rt.lib().mdrtHipLastError().decode())
        self._h = h

    def _report(self, sim, step: int):
        rt.check(rt.lib().mdrtHipTrajectoryFrame(self._h, int(step)))

    def flush(self):
        if self._h:
            rt.check(rt.lib().mdrtHipTrajectoryFlush(self._h))

    @property
    def frames(self) -> int:
        return int(rt.lib().mdrtHipTrajectoryFrames(self._h)) if self._h else 0

    def close(self):
        if self._h:
            h, self._h = self._h, None
            rt.check(rt.lib().mdrtHipTrajectoryClose(h))


class CheckpointReporter:
    """A checkpoint every `period` steps (python/Bindings.cpp:1383-1390)."""

    def __init__(self, file: str, period: int):
        if not str(file):
            raise ValueError("CheckpointReporter takes the name of a file")
        if int(period) <= 0:
            raise ValueError("the period must be a positive number of steps")
        self.file, self.period = str(file), int(period)

    def _report(self, sim, step: int):
        sim.save_checkpoint(self.file)

    def flush(self):
        pass

    def close(self):
        pass


@dataclass
class Checkpoint:
    """The state of a run, in nm, ps, amu, kJ/mol (the fields of the
    reference's Checkpoint, include/mdir/Driver/Checkpoint.h:36-136, that
    this path has), plus the positions of the run's list."""
    step: int
    time: float
    positions: np.ndarray
    velocities: np.ndarray
    forces: np.ndarray | None
    masses: np.ndarray
    species: np.ndarray
    box: np.ndarray
    periodic: bool = True
    integrator: str = "velocity-verlet"
    velocity_offset: float = 0.0
    precision: str = "mixed"
    timestep: float = 0.0
    seed: int = 0
    first_step: int = 0
    part: int = 1
    trajectory: str = ""
    frames: int = 0
    bath: float = 0.0
    creator: str = "mdir_amd"
    front_end: str = "python"
    temperature: float = 0.0
    friction: float = 0.0
    list_origin: np.ndarray | None = None
    prune_origin: np.ndarray | None = None
    list_kept: bool = False
    barostat_dv: float = 0.0
    barostat_counts: np.ndarray = field(default_factory=lambda: np.zeros(4, np.int64))

    @property
    def cell(self) -> np.ndarray:
        return self.box

    def __repr__(self):
        return (f"Checkpoint(step={self.step}, particles={len(self.masses)}, "
                f"integrator={self.integrator})")


_ARRAYS = ("positions", "velocities", "forces", "masses", "species", "box", "list_origin",
           "prune_origin", "barostat_counts")
_SCALARS = ("step", "time", "periodic", "integrator", "velocity_offset", "precision",
            "timestep", "seed", "first_step", "part", "trajectory", "frames", "bath",
            "creator", "front_end", "temperature", "friction", "list_kept", "barostat_dv")


def write_checkpoint(path: str, ck: Checkpoint):
    """Writes `ck` to `path` (a NumPy .npz container, no pickles); an
    existing file at `path` becomes `path`.prev first, as the reference's
    writer keeps the one before (lib/Driver/Checkpoint.cpp)."""
    data = {"format": np.array(CHECKPOINT_FORMAT, np.int64)}
    for k in _ARRAYS:
        v = getattr(ck, k)
        if v is not None:
            data[k] = np.ascontiguousarray(v)
    for k in _SCALARS:
        data[k] = np.array(getattr(ck, k))
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        np.savez(f, **data)
        f.flush()
        os.fsync(f.fileno())
    if os.path.exists(path):
        os.replace(path, path + ".prev")
    os.replace(tmp, path)


def read_checkpoint(path: str) -> Checkpoint:
    """The checkpoint at `path` (python/Bindings.cpp:1177 read_checkpoint)."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["format"]) != CHECKPOINT_FORMAT:
            raise ValueError(f"'{path}': checkpoint format {int(z['format'])}, "
                             f"this reader knows {CHECKPOINT_FORMAT}")
        kw = {}
        for k in _ARRAYS:
            kw[k] = z[k].copy() if k in z.files else None
        for k in _SCALARS:
            v = z[k][()]
            kw[k] = v.item() if hasattr(v, "item") else v
    kw["integrator"] = str(kw["integrator"])
    kw["precision"] = str(kw["precision"])
    kw["trajectory"] = str(kw["trajectory"])
    kw["creator"] = str(kw["creator"])
    kw["front_end"] = str(kw["front_end"])
    kw["periodic"] = bool(kw["periodic"])
    kw["list_kept"] = bool(kw["list_kept"])
    if kw["barostat_counts"] is None:
        kw["barostat_counts"] = np.zeros(4, np.int64)
    return Checkpoint(**kw)


def state_of(sim) -> Checkpoint:
    """The checkpoint of `sim` now."""
    n = sim.n
    x, v, F = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    origin, prune_origin = np.zeros((n, 3)), np.zeros((n, 3))
    st = rt.RunState()
    rt.check(rt.lib().mdrtHipSaveState(sim._h, rt.ptr(x, C.c_double), rt.ptr(v, C.c_double),
                                       rt.ptr(F, C.c_double), rt.ptr(origin, C.c_double),
                                       rt.ptr(prune_origin, C.c_double), C.byref(st)))
    traj = next((r for r in sim.reporters if isinstance(r, TrajectoryReporter)), None)
    if traj is not None:
        traj.flush()  # the checkpoint counts the frames written before it
    step = int(st.step_count)
    return Checkpoint(
        step=step, time=step * float(sim.timestep), positions=x, velocities=v, forces=F,
        masses=np.asarray(sim.system.masses, np.float64).copy(),
        species=np.asarray(sim.system.types, np.int32).copy(), box=np.array(st.box[:]),
        integrator="langevin-baoab" if sim.friction > 0 else "velocity-verlet",
        timestep=float(sim.timestep), seed=int(sim.seed), first_step=sim.first_step,
        part=sim.part, trajectory=os.path.basename(traj.file) if traj else "",
        frames=traj.frames if traj else 0, temperature=float(sim.temperature),
        friction=float(sim.friction), list_origin=origin, prune_origin=prune_origin,
        list_kept=bool(st.list_kept), barostat_dv=float(st.barostat_dv),
        barostat_counts=np.array(st.barostat_counts[:], np.int64))


def restore(sim, ck: Checkpoint):
    """Puts the state of `ck` into `sim` (the same system)."""
    n = sim.n
    if ck.positions.shape != (n, 3):
        raise ValueError(f"the checkpoint holds {len(ck.positions)} particles, "
                         f"the system {n}")
    if not np.array_equal(ck.masses, np.asarray(sim.system.masses, np.float64)):
        raise ValueError("the checkpoint's masses are not those of the system")
    if ck.friction > 0:
        sim.set_langevin(ck.temperature, ck.friction, ck.seed)
    st = rt.RunState()
    st.step_count = int(ck.step)
    for k in range(3):
        st.box[k] = float(ck.box[k])
    st.barostat_dv = float(ck.barostat_dv)
    for k in range(4):
        st.barostat_counts[k] = int(ck.barostat_counts[k])
    st.list_kept = int(ck.list_kept and ck.list_origin is not None)
    x = np.ascontiguousarray(ck.positions, np.float64)
    v = np.ascontiguousarray(ck.velocities, np.float64)
    origin = None if ck.list_origin is None else np.ascontiguousarray(ck.list_origin, np.float64)
    prune = None if ck.prune_origin is None else np.ascontiguousarray(ck.prune_origin, np.float64)
    rt.check(rt.lib().mdrtHipRestoreState(
        sim._h, rt.ptr(x, C.c_double), rt.ptr(v, C.c_double),
        rt.ptr(origin, C.c_double) if origin is not None else None,
        rt.ptr(prune, C.c_double) if prune is not None else None, C.byref(st)))
    if ck.timestep > 0:
        sim.timestep = float(ck.timestep)
    sim.first_step = int(ck.first_step)
    sim.part = int(ck.part) + 1
    sim._continued_trajectory = (ck.trajectory, int(ck.frames))


def read_dcd(path: str):
    """The frames of a DCD file: (header numbers (20 int32), title lines,
    positions (frames, n, 3) f32 in Angstrom, cells (frames, 6) f64 or
    None).  A reader of the record layout alone (CHARMM / NAMD DCD: each
    record framed by its byte length), so that it reads any writer's file
    of this kind, not only this library's."""
    with open(path, "rb") as f:
        data = f.read()
    at = 0

    def record():
        nonlocal at
        size, = struct.unpack_from("<i", data, at)
        body = data[at + 4: at + 4 + size]
        end, = struct.unpack_from("<i", data, at + 4 + size)
        if end != size or len(body) != size:
            raise ValueError(f"{path}: a broken record at byte {at}")
        at += 8 + size
        return body

    head = record()
    if head[:4] != b"CORD":
        raise ValueError(f"{path}: not a DCD file")
    numbers = struct.unpack_from("<20i", head, 4)
    title = record()
    lines, = struct.unpack_from("<i", title, 0)
    titles = [title[4 + 80 * k: 4 + 80 * (k + 1)].decode("ascii", "replace").rstrip()
              for k in range(lines)]
    n, = struct.unpack("<i", record())
    has_cell = numbers[10] != 0
    xyz, cells = [], []
    while at < len(data):
        if has_cell:
            cells.append(np.frombuffer(record(), "<f8"))
        comps = [np.frombuffer(record(), "<f4") for _ in range(3)]
        if any(len(c) != n for c in comps):
            raise ValueError(f"{path}: a frame of the wrong size")
        xyz.append(np.stack(comps, axis=1))
    pos = np.array(xyz, np.float32).reshape(len(xyz), n, 3)
    return numbers, titles, pos, (np.array(cells) if has_cell else None)

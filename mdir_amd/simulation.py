"""Simulation: a system on the GPU, its forces and energies, and NVE steps.

Mirrors the part of the reference's Python API that drives the per-step
force path (python/Bindings.cpp:1625 Simulation, :2230 run; Simulation.cpp
:1679 run, :1537 runPart): construct from a system, run a number of steps,
read the state and the energies.  The arithmetic runs in libmdrt_hip.so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import runtime as rt
from .systems import System


def shape_table(ms) -> dict:
    """{(brick z lines, transform, segment entries): time} from the trial's
    [segment][transform][brick] array (NaN: not run)."""
    return {(8 >> (k % 4), (k // 4) % 2, Simulation.SEGMENTS[k // 8]): float(v)
            for k, v in enumerate(ms) if np.isfinite(v)}


class Simulation:
    def __init__(self, system: System, skin: float = 0.1, prune_skin: float = 0.0):
        self.system = system
        s = system
        L = rt.lib()
        d, self._arrays = rt.system_desc(system, skin, prune_skin)
        self._desc = d
        h = L.mdrtHipSystemCreate(C.byref(d))
        if not h:
            raise RuntimeError("mdrt_hip: " + L.mdrtHipLastError().decode())
        self._h = h
        self._arrays = []
        self.n = s.n
        # run settings a checkpoint records
        self.timestep = 0.002
        self.temperature, self.friction, self.seed = 0.0, 0.0, 0
        self.first_step, self.part = 0, 1
        self._reporters: list = []
        self._continued_trajectory = ("", 0)

    @classmethod
    def from_checkpoint(cls, system: System, checkpoint, skin: float = 0.1,
                        prune_skin: float = 0.0) -> "Simulation":
        """A simulation of `system` that continues the run of `checkpoint`
        (a path or a Checkpoint): the reference's
        mdir.Simulation(program, checkpoint=path).  A TrajectoryReporter on
        the file the checkpoint names continues it after the frames the
        checkpoint counts."""
        from .output import read_checkpoint, restore
        ck = read_checkpoint(checkpoint) if isinstance(checkpoint, (str, os.PathLike)) \
            else checkpoint
        sim = cls(system, skin, prune_skin)
        restore(sim, ck)
        return sim

    def close(self):
        if getattr(self, "_h", None):
            try:
                self.close_reporters()
            finally:
                rt.lib().mdrtHipSystemDestroy(self._h)
                self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----
    @property
    def positions(self) -> np.ndarray:
        x = np.zeros((self.n, 3))
        rt.check(rt.lib().mdrtHipGetPositions(self._h, rt.ptr(x, C.c_double)))
        return x

    @positions.setter
    def positions(self, x):
        x = np.ascontiguousarray(x, np.float64)
        rt.check(rt.lib().mdrtHipSetPositions(self._h, rt.ptr(x, C.c_double)))

    @property
    def velocities(self) -> np.ndarray:
        v = np.zeros((self.n, 3))
        rt.check(rt.lib().mdrtHipGetVelocities(self._h, rt.ptr(v, C.c_double)))
        return v

    @velocities.setter
    def velocities(self, v):
        v = np.ascontiguousarray(v, np.float64)
        rt.check(rt.lib().mdrtHipSetVelocities(self._h, rt.ptr(v, C.c_double)))

    # ---- evaluation ----
    def forces(self):
        """Forces (n, 3) and the energy terms (rt.NUM_ENERGIES,)."""
        F = np.zeros((self.n, 3))
        e = np.zeros(rt.NUM_ENERGIES)
        rt.check(rt.lib().mdrtHipComputeForces(self._h, rt.ptr(F, C.c_double),
                                               rt.ptr(e, C.c_double)))
        return F, e

    def force_terms(self, terms: int):
        """Forces and energies of the terms `terms` names (bits
        runtime.TERM_PAIRS, TERM_SPECIAL, TERM_RECIP, TERM_BONDED) only."""
        F = np.zeros((self.n, 3))
        e = np.zeros(rt.NUM_ENERGIES)
        rt.check(rt.lib().mdrtHipComputeForceTerms(self._h, int(terms), rt.ptr(F, C.c_double),
                                                   rt.ptr(e, C.c_double)))
        return F, e

    def set_pme_transform(self, mode: int):
        """0: the LDS Stockham FFT; 1: DFT matrix products on MFMA in split
        BF16x3 (planes of edges 32 or 64 in LDS; other x and y edges up to
        224 over tiles of lines)."""
        rt.check(rt.lib().mdrtHipSetPmeTransform(self._h, int(mode)))

    TUNE_BRICKS = 1
    TUNE_TRANSFORM = 2
    TUNE_SEGMENTS = 4
    SEGMENTS = (16, 24, 48)

    @property
    def launch_shape(self) -> tuple:
        """(segment entries, PME brick z lines, plane transform)."""
        out = np.zeros(3, np.int32)
        rt.check(rt.lib().mdrtHipGetLaunchShape(self._h, rt.ptr(out, C.c_int32)))
        return tuple(int(v) for v in out)

    @launch_shape.setter
    def launch_shape(self, shape):
        seg, bricks, transform = shape
        rt.check(rt.lib().mdrtHipSetLaunchShape(self._h, int(seg), int(bricks), int(transform)))

    def autotune(self, what: int = TUNE_BRICKS | TUNE_TRANSFORM | TUNE_SEGMENTS,
                 steps: int = 50, graphs: int = 4, dt: float = 0.002) -> dict:
        """Times the launch shapes on this system and keeps the fastest (the
        launch-shape trial of tools/mdir/TowersTrial.h); positions and
        velocities are restored. Returns {(brick z lines, transform, segment
        entries): ms per step} for the shapes run."""
        ms = np.full(24, np.nan)
        rt.check(rt.lib().mdrtHipAutotune(self._h, int(what), int(steps), int(graphs),
                                          float(dt), rt.ptr(ms, C.c_double)))
        return shape_table(ms)

    def energies(self) -> dict:
        _, e = self.forces()
        return dict(zip(rt.ENERGY_NAMES, e))

    def set_langevin(self, temperature: float, friction: float, seed: int = 0):
        """Langevin dynamics (g-BAOAB) for the following run() calls:
        friction in 1/ps (0: back to velocity Verlet), temperature in K; the
        noise of each atom and step comes from Philox under `seed`."""
        rt.check(rt.lib().mdrtHipSetLangevin(self._h, float(temperature), float(friction),
                                             int(seed) & 0xFFFFFFFFFFFFFFFF))
        self.temperature, self.friction = float(temperature), float(friction)
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF

    @property
    def box(self) -> np.ndarray:
        b = np.zeros(3)
        rt.check(rt.lib().mdrtHipGetBox(self._h, rt.ptr(b, C.c_double)))
        return b

    def scale_box(self, scale):
        """Scales the units' centres of mass and the box by scale (3,)."""
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(scale, np.float64), (3,)))
        rt.check(rt.lib().mdrtHipScaleBox(self._h, rt.ptr(s, C.c_double)))

    def barostat_move(self, pressure: float, temperature: float, seed: int = 0):
        """One Monte Carlo volume move; returns (kept, {U0, U1, dV, W})."""
        acc = C.c_int32(0)
        info = np.zeros(4)
        rt.check(rt.lib().mdrtHipBarostatMove(self._h, float(pressure), float(temperature),
                                              int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(acc),
                                              rt.ptr(info, C.c_double)))
        return bool(acc.value), dict(zip(("U0", "U1", "dV", "W"), info))

    def run_npt(self, steps: int, dt: float = 0.002, graph: int = 0, every: int = 100,
                pressure: float = 1.0, temperature: float = 300.0, seed: int = 0) -> float:
        """run() with a Monte Carlo volume move after every `every` steps."""
        ke = C.c_double(0.0)
        rt.check(rt.lib().mdrtHipRunNPT(self._h, int(steps), float(dt), int(graph), int(every),
                                        float(pressure), float(temperature),
                                        int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(ke)))
        return ke.value

    def barostat_stats(self) -> dict:
        c = np.zeros(2, np.int64)
        dv = C.c_double(0.0)
        rt.check(rt.lib().mdrtHipBarostatStats(self._h, rt.ptr(c, C.c_int64), C.byref(dv)))
        return {"tried": int(c[0]), "kept": int(c[1]), "max_dv": dv.value}

    @property
    def step_count(self) -> int:
        """The step number of the next step (the noise's step counter)."""
        v = C.c_int64(0)
        rt.check(rt.lib().mdrtHipGetStepCount(self._h, C.byref(v)))
        return v.value

    @step_count.setter
    def step_count(self, step: int):
        rt.check(rt.lib().mdrtHipSetStepCount(self._h, int(step)))

    def run(self, steps: int, dt: float = 0.002, graph: int = 0) -> float:
        """`steps` steps (velocity Verlet, or Langevin after set_langevin);
        returns the kinetic energy at the end.  The reporters report at the
        steps that are multiples of their periods."""
        self.timestep = float(dt)
        if not self._reporters:
            return self._run(steps, dt, graph)
        from .output import TrajectoryReporter
        self._open_reporters()
        step = self.step_count
        end = step + int(steps)
        # frames before checkpoints, which count them
        order = sorted(self._reporters, key=lambda r: not isinstance(r, TrajectoryReporter))
        ke = 0.0
        while step < end:
            nxt = min([end] + [(step // r.period + 1) * r.period for r in order])
            ke = self._run(nxt - step, dt, graph)
            step = nxt
            for r in order:
                if step % r.period == 0:
                    r._report(self, step)
        return ke

    def _run(self, steps: int, dt: float, graph: int) -> float:
        ke = C.c_double(0.0)
        rt.check(rt.lib().mdrtHipRun(self._h, int(steps), float(dt), int(graph),
                                     C.byref(ke)))
        return ke.value

    # ---- reporters and checkpoints ----
    @property
    def reporters(self) -> list:
        return list(self._reporters)

    @reporters.setter
    def reporters(self, items):
        from .output import CheckpointReporter, TrajectoryReporter
        items = list(items)
        kinds = [type(r) for r in items]
        for r in items:
            if not isinstance(r, (TrajectoryReporter, CheckpointReporter)):
                raise TypeError(f"not a reporter: {r!r}")
        if kinds.count(TrajectoryReporter) > 1 or kinds.count(CheckpointReporter) > 1:
            raise ValueError("a simulation takes one reporter of each kind")
        self.close_reporters()
        self._reporters = items

    def _open_reporters(self):
        from .output import TrajectoryReporter
        for r in self._reporters:
            if isinstance(r, TrajectoryReporter) and r._h is None:
                name, frames = self._continued_trajectory
                keep = frames if name and os.path.basename(r.file) == name else -1
                r._open(self, keep)

    def close_reporters(self):
        """Writes what the reporters hold and closes their files."""
        for r in getattr(self, "_reporters", []):
            r.close()

    def state(self):
        """The Checkpoint of the run now."""
        from .output import state_of
        return state_of(self)

    def save_checkpoint(self, path: str):
        """Writes the checkpoint of the run now (python/Bindings.cpp:1703)."""
        from .output import state_of, write_checkpoint
        write_checkpoint(str(path), state_of(self))

    def check_list(self):
        """The check of the list between rebuilds (the reference's
        MDIR_CHECK_PRUNING): after run(), the forces and energies the next
        step computes through the list the steps hold, and a dict of the
        pairs within rc found in the list ("list") and in the rows the pair
        loop runs over ("pair_loop"), and the builds the check did."""
        F = np.zeros((self.n, 3))
        e = np.zeros(rt.NUM_ENERGIES)
        c = np.zeros(3, np.int64)
        rt.check(rt.lib().mdrtHipCheckList(self._h, rt.ptr(F, C.c_double), rt.ptr(e, C.c_double),
                                           rt.ptr(c, C.c_int64)))
        return F, e, {"list": int(c[0]), "pair_loop": int(c[1]), "builds": int(c[2])}

    def list_pairs(self, reach: float, pruned: bool = False,
                   capacity: int | None = None) -> np.ndarray:
        """The pairs (i < j) of the list the steps hold (or of its pruned
        rows) within `reach` at the current positions, without a rebuild."""
        if capacity is None:
            s = self.system
            rho = s.n / float(np.prod(s.box))
            capacity = int(s.n * rho * 4.19 * (reach + 0.05) ** 3 * 0.75) + 4096
        out = np.zeros((capacity, 2), np.int32)
        count = C.c_int64(0)
        rt.check(rt.lib().mdrtHipListPairs(self._h, float(reach), int(pruned), capacity,
                                           rt.ptr(out, C.c_int32), C.byref(count)))
        if count.value > capacity:
            return self.list_pairs(reach, pruned, count.value + 16)
        return out[:count.value].copy()

    def prepare(self, steps: int, dt: float = 0.002, graph: int = 50):
        """Captures the step graphs a later run(steps, dt, graph) replays,
        without running them, so that the capture stays out of a timed run."""
        rt.check(rt.lib().mdrtHipPrepareRun(self._h, int(steps), float(dt), int(graph)))

    def run_measured(self, steps: int, dt: float, graph: int = 50) -> float:
        """Runs like run() and returns the mean duration (ms) of the pair
        kernel, timed live on the device clock in every step."""
        ms = C.c_double(0.0)
        rt.check(rt.lib().mdrtHipRunMeasured(self._h, int(steps), float(dt), int(graph),
                                             C.byref(ms)))
        return ms.value

    def minimize(self, steps: int = 200, max_step: float = 0.01) -> float:
        e = C.c_double(0.0)
        rt.check(rt.lib().mdrtHipMinimize(self._h, int(steps), float(max_step),
                                          C.byref(e)))
        return e.value

    def neighbor_pairs(self, reach: float, excluded: bool = False,
                       capacity: int | None = None) -> np.ndarray:
        """Unordered pairs (i < j) of the reference's matrix test at `reach`."""
        if capacity is None:
            import math
            s = self.system
            rho = s.n / float(np.prod(s.box))
            capacity = int(s.n * rho * 4.19 * (reach + 0.05) ** 3 * 0.75) + 4096
        out = np.zeros((capacity, 2), np.int32)
        count = C.c_int64(0)
        rt.check(rt.lib().mdrtHipNeighborPairs(self._h, float(reach), int(excluded),
                                               capacity, rt.ptr(out, C.c_int32),
                                               C.byref(count)))
        if count.value > capacity:
            return self.neighbor_pairs(reach, excluded, count.value + 16)
        return out[:count.value].copy()

    def stats(self) -> dict:
        s = np.zeros(10, np.int64)
        rt.check(rt.lib().mdrtHipStats(self._h, rt.ptr(s, C.c_int64)))
        keys = ["places", "clusters", "tiles", "largest_row", "row_cap",
                "columns", "column_cap", "builds", "pair_loop_tiles", "prunes"]
        return dict(zip(keys, (int(v) for v in s)))

    PROFILE_KERNELS = ["integrate_first", "nb_bin", "nb_sort", "nb_search",
                       "nb_pairs", "pme_bin", "pme_spread_rows", "pme_yx",
                       "pme_gather_rows", "integrate_second"]

    def profile_step(self, repeats: int = 20) -> dict:
        """ms per launch of each kernel of a step (HIP events on the stream
        the kernels run on); the profiled steps advance the state."""
        ms = np.zeros(len(self.PROFILE_KERNELS))
        rt.check(rt.lib().mdrtHipProfileStep(self._h, int(repeats),
                                             rt.ptr(ms, C.c_double)))
        return dict(zip(self.PROFILE_KERNELS, ms))

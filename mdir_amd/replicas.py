"""Replicas: independent systems of uneven size advanced together, every
chunk of steps of all of them one replayed HIP graph (libmdrt_hip.so,
mdrtHipReplicas*).

The batching workload: replicas of 500 to 8,000 TIP3P waters under Langevin
dynamics, advanced in one captured step graph; the step latency of each
replica inside the batch is read from device-clock stamps at the start and
end of its branch.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import runtime as rt
from .simulation import Simulation
from .systems import System


class ReplicaView(Simulation):
    """A replica of a set: every Simulation call on the set's own handle
    (the set owns it; closing the view does not destroy it)."""

    def __init__(self, owner, handle, system: System):
        self.system = system
        self._h = handle
        self._owner = owner
        self.n = system.n

    def close(self):
        self._h = None

    def __del__(self):
        pass


class ReplicaSet:
    def __init__(self, systems, skin: float = 0.15):
        L = rt.lib()
        descs = (rt.SystemDesc * len(systems))()
        keep = []
        for k, s in enumerate(systems):
            d, arrays = rt.system_desc(s, skin, 0.0)
            descs[k] = d
            keep.append(arrays)
        h = L.mdrtHipReplicasCreate(descs, len(systems))
        if not h:
            raise RuntimeError("mdrt_hip: " + L.mdrtHipLastError().decode())
        self._h = h
        self.systems = list(systems)
        self.replicas = [ReplicaView(self, L.mdrtHipReplicasSystem(h, k), s)
                         for k, s in enumerate(systems)]

    def __len__(self):
        return len(self.replicas)

    def __getitem__(self, k) -> ReplicaView:
        return self.replicas[k]

    def run(self, steps: int, dt: float = 0.002, graph: int = 50):
        """`steps` steps of every replica, each chunk of `graph` steps one
        graph launch for the whole set."""
        rt.check(rt.lib().mdrtHipReplicasRun(self._h, int(steps), float(dt), int(graph)))

    def set_launch(self, mode: int):
        """0: one graph of all replicas per chunk; 1 (the default): a graph
        per replica, each on its own stream."""
        rt.check(rt.lib().mdrtHipReplicasSetLaunch(self._h, int(mode)))

    def latency(self, k: int) -> np.ndarray:
        """ms per step of replica k in each chunk since the last clear."""
        count = C.c_int64(0)
        out = np.zeros(4096)
        rt.check(rt.lib().mdrtHipReplicasLatency(self._h, int(k), out.size,
                                                 rt.ptr(out, C.c_double), C.byref(count)))
        if count.value > out.size:
            out = np.zeros(count.value)
            rt.check(rt.lib().mdrtHipReplicasLatency(self._h, int(k), out.size,
                                                     rt.ptr(out, C.c_double), C.byref(count)))
        return out[:count.value].copy()

    def clear_latency(self):
        rt.check(rt.lib().mdrtHipReplicasClearLatency(self._h))

    def close(self):
        if getattr(self, "_h", None):
            for r in self.replicas:
                r.close()
            rt.lib().mdrtHipReplicasDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BUCKETS = ((500, 2375), (2376, 4250), (4251, 6125), (6126, 8000))


def replica_benchmark(count: int = 256, seed: int = 5, steps: int = 500, chunk: int = 10,
                      warmup: int = 50, minimize: int = 30, autotune: bool = True,
                      dt: float = 0.002, launch: int = 1) -> dict:
    """The batching workload: `count` TIP3P replicas of 500 to 8,000 waters
    (uniform, seeded) at 1.0 g/cm^3, PME rc 0.9 nm with a grid no wider
    than 0.12 nm, order 4, Langevin 300 K (1/ps) at dt, all advanced in one
    graph per chunk of `chunk` steps.  Returns aggregate ns/day and the
    p50/p99 step latency (ms) of the replicas of each size bucket (waters),
    from the per-chunk branch clocks.  With `autotune`, the launch-shape
    trial (PME bricks and pair-loop segments) runs on the median replica of
    each bucket, alone, and its shape is given to the whole bucket."""
    import time
    from . import systems as S
    sizes = S.replica_sizes(count, seed)
    systems = [S.tip3p_replica(int(nw), seed=1000 + k) for k, nw in enumerate(sizes)]
    t_setup = time.perf_counter()
    rs = ReplicaSet(systems)
    for k, r in enumerate(rs):
        if minimize:
            r.minimize(minimize, 0.01)
        relaxed = systems[k].copy()
        relaxed.positions = r.positions
        r.velocities = S.initial_velocities(relaxed, 7 + k, 300.0)
        r.set_langevin(300.0, 1.0, 5000 + k)
    shapes = {}
    if autotune:
        for lo, hi in BUCKETS:
            members = [k for k, nw in enumerate(sizes) if lo <= nw <= hi]
            if not members:
                continue
            mid = sorted(members, key=lambda k: sizes[k])[len(members) // 2]
            rs[mid].autotune(Simulation.TUNE_BRICKS | Simulation.TUNE_SEGMENTS, steps=chunk,
                             graphs=4, dt=dt)
            shape = rs[mid].launch_shape
            shapes[f"{lo}-{hi}"] = shape
            for k in members:
                rs[k].launch_shape = shape
    rs.set_launch(launch)
    rs.run(warmup, dt, chunk)
    setup_s = time.perf_counter() - t_setup
    rs.clear_latency()
    t0 = time.perf_counter()
    rs.run(steps, dt, chunk)
    elapsed = time.perf_counter() - t0
    ns_day = count * steps * dt * 1e-3 / (elapsed / 86400.0)
    buckets = {}
    for lo, hi in BUCKETS:
        lat = np.concatenate([rs.latency(k) for k, nw in enumerate(sizes) if lo <= nw <= hi]
                             or [np.zeros(0)])
        if lat.size:
            buckets[f"{lo}-{hi}"] = {"replicas": int(sum(lo <= nw <= hi for nw in sizes)),
                                     "p50_ms": round(float(np.percentile(lat, 50)), 4),
                                     "p99_ms": round(float(np.percentile(lat, 99)), 4)}
    out = {"metric": "aggregate ns/day of uneven TIP3P replicas (1 GPU)",
           "value": round(ns_day, 1), "unit": "ns/day", "replicas": count,
           "atoms": int(3 * sizes.sum()), "waters_range": [int(sizes.min()), int(sizes.max())],
           "steps": steps, "chunk_steps": chunk, "ms_per_step": round(elapsed / steps * 1e3, 4),
           "ensemble": "NVT Langevin 300 K, 1/ps", "latency_by_waters": buckets,
           "launch": ["one graph of all replicas", "a graph per replica, own stream"][launch],
           "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default"),
           "launch_shapes": {k: list(v) for k, v in shapes.items()},
           "setup_s": round(setup_s, 1)}
    rs.close()
    return out

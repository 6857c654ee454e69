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

"""Frame evaluator: energies, forces and reverse products of many frames of
one system on one GPU (libmdrt_hip.so, include/mdrt_hip.h group 3).

Mirrors the reference's Python frame API (python/mdir/_frames.py:458
FrameEvaluator, :715 evaluate, :272 FrameEnergies, :327 vjp, :220
FrameGradient): evaluate a batch of frames, then take the product of a
cotangent on the energies and on their position gradients with the
derivatives in the tunables,

    p = sum_n [ g_n dU_n/dtheta + sum_i G_ni . d2U_n/(dx_i dtheta) ],

and, with gradients=("positions",), the position part
g_n dU_n/dx + H_n G_n.  The tunables are "charges" (per atom), "sigma" and
"epsilon" (per Lennard-Jones type, Lorentz-Berthelot pairs).  Unlike the
reference (which refuses the charges under PME in the reverse product, and
the reciprocal sum in the position Hessian), both include the PME
reciprocal sum here.

All frames of one call run in one launch per kernel (the frame index is the
grid's z dimension); the force-matching call runs the forward pass, the
residual and the adjoint as one captured HIP graph.
"""
from __future__ import annotations

import ctypes as C
from collections.abc import Mapping

import numpy as np

from . import runtime as rt
from .systems import System

TUNABLES = ("charges", "sigma", "epsilon")


def _frame_array(a, shape, name):
    """(pointer, keep-alive, is_device) of a float64 frame array: a NumPy
    array (host) or a CUDA torch tensor (device)."""
    if hasattr(a, "data_ptr") and hasattr(a, "is_cuda"):
        import torch
        if a.dtype != torch.float64:
            raise TypeError(f"{name}: expected float64, found {a.dtype}")
        if tuple(a.shape) != tuple(shape):
            raise ValueError(f"{name}: expected shape {tuple(shape)}, found {tuple(a.shape)}")
        a = a.contiguous()
        if a.is_cuda:  # the library's stream is not ordered after torch's
            torch.cuda.synchronize(a.device)
        return C.c_void_p(a.data_ptr()), a, a.is_cuda
    a = np.ascontiguousarray(a, np.float64)
    if a.shape != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, found {a.shape}")
    if not np.all(np.isfinite(a)):
        raise ValueError(f"{name}: not finite")
    return C.c_void_p(a.ctypes.data), a, False


class FrameGradient(Mapping):
    """The product in the tunables (read-only arrays by name) and, when
    asked, `positions` (K, N, 3)."""

    def __init__(self, values, positions=None):
        self._values = values
        self.positions = positions

    def __getitem__(self, name):
        return self._values[name]

    def __iter__(self):
        return iter(self._values)

    def __len__(self):
        return len(self._values)

    def __repr__(self):
        return f"FrameGradient({', '.join(self._values)})"


class FrameEnergies:
    """Energies of the evaluated frames: `energy` (K,), `terms`
    (K, NUM_ENERGIES) in the order of runtime.ENERGY_NAMES, `forces`
    (K, N, 3), and `gradients`: "positions" = dU/dx = -forces (K, N, 3) and,
    when asked, "strain" = dU/d(eps) (K, 3, 3), row alpha, column beta, for
    x -> (1 + eps) x, h -> h (1 + eps)^T (the reference's D252)."""

    def __init__(self, evaluator, count, terms, forces, strain=None):
        self._evaluator = evaluator
        self.count = count
        self.version = evaluator._version
        self.terms = terms
        self.energy = terms.sum(axis=1)
        self.forces = forces
        self.gradients = {"positions": -forces} if forces is not None else {}
        if strain is not None:
            self.gradients["strain"] = strain

    def vjp(self, cotangent=None, *, positions=None, gradients=("tunables",)):
        """sum_n cotangent[n] dU_n/dtheta + sum_n sum_i positions[n, i] .
        d2U_n/dx_i dtheta (the reference's FrameEnergies.vjp, _frames.py:327);
        `positions` is the cotangent on dU/dx (minus that on the forces).
        With "positions" in `gradients`, also cotangent[n] dU_n/dx + H_n
        positions[n] as the attribute `positions`."""
        ev = self._evaluator
        if isinstance(gradients, str):
            raise ValueError("gradients is a tuple of names among 'tunables' and 'positions'")
        names = tuple(dict.fromkeys(gradients))
        for name in names:
            if name not in ("tunables", "positions"):
                raise ValueError(f"vjp: unknown gradient {name!r}")
        if cotangent is None and positions is None:
            raise ValueError("vjp: give a cotangent on the energies, on their position "
                             "gradients (positions=), or both")
        if ev._version != self.version:
            raise RuntimeError("vjp: the evaluator has evaluated other frames or changed "
                               "its tunables since")
        K, n = self.count, ev.n
        g = None
        if cotangent is not None:
            g = np.ascontiguousarray(cotangent, np.float64)
            if g.shape != (K,) or not np.all(np.isfinite(g)):
                raise ValueError(f"vjp: expected a finite cotangent of shape ({K},)")
        values = {}
        if "tunables" in names:
            dq = np.zeros(n)
            ds = np.zeros(ev.ntypes)
            de = np.zeros(ev.ntypes)
            if positions is not None:
                pp, keep, _ = _frame_array(positions, (K, n, 3), "positions")
            else:
                pp, keep = None, None
            rt.check(rt.lib().mdrtHipFramesVjp(
                ev._h, K, None if g is None else rt.ptr(g, C.c_double), pp,
                rt.ptr(dq, C.c_double), rt.ptr(ds, C.c_double), rt.ptr(de, C.c_double)))
            del keep
            values = {"charges": dq, "sigma": ds, "epsilon": de}
        dx = None
        if "positions" in names:
            dx = np.zeros((K, n, 3))
            if positions is not None:
                dx += ev.hvp(positions)
            if g is not None:
                dx -= g[:, None, None] * self.forces
        return FrameGradient(values, dx)

    def vjp_pairs(self, pairs, cotangent=None, *, positions=None):
        """The product of vjp in type-pair table entries (the reference's
        sigma and epsilon "by pairs of types", NBFIX entries) instead of the
        per-type parameters: `pairs` names 1 to 4 pairs of types (a, b);
        returns {"sigma": (len(pairs),), "epsilon": (len(pairs),)}."""
        ev = self._evaluator
        if ev._version != self.version:
            raise RuntimeError("vjp_pairs: the evaluator has evaluated other frames since")
        K, n = self.count, ev.n
        tp = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        g = None if cotangent is None else np.ascontiguousarray(cotangent, np.float64)
        if g is not None and g.shape != (K,):
            raise ValueError(f"vjp_pairs: expected a cotangent of shape ({K},)")
        pp, keep = (None, None) if positions is None else \
            _frame_array(positions, (K, n, 3), "positions")[:2]
        ds = np.zeros(len(tp))
        de = np.zeros(len(tp))
        rt.check(rt.lib().mdrtHipFramesVjpPairs(
            ev._h, K, None if g is None else rt.ptr(g, C.c_double), pp, len(tp),
            rt.ptr(tp, C.c_int32), rt.ptr(ds, C.c_double), rt.ptr(de, C.c_double)))
        del keep
        return {"sigma": ds, "epsilon": de}


class FrameEvaluator:
    """Frames of one system (shared topology and box) on the GPU, up to
    `max_frames` per call."""

    def __init__(self, system: System, max_frames: int = 64):
        self.system = system
        self.n = system.n
        self.ntypes = system.ntypes
        self.max_frames = int(max_frames)
        d, keep = rt.system_desc(system)
        h = rt.lib().mdrtHipFramesCreate(C.byref(d), self.max_frames)
        del keep
        if not h:
            raise RuntimeError("mdrt_hip: " + rt.lib().mdrtHipLastError().decode())
        self._h = h
        self._version = 0
        self._loaded = 0

    def close(self):
        if getattr(self, "_h", None):
            rt.lib().mdrtHipFramesDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- tunables ----
    @property
    def tunables(self) -> dict:
        q = np.zeros(self.n)
        s = np.zeros(self.ntypes)
        e = np.zeros(self.ntypes)
        rt.check(rt.lib().mdrtHipFramesGetParameters(
            self._h, rt.ptr(q, C.c_double), rt.ptr(s, C.c_double), rt.ptr(e, C.c_double)))
        return {"charges": q, "sigma": s, "epsilon": e}

    def update(self, charges=None, sigma=None, epsilon=None):
        """Sets any of the tunables; the frames must be evaluated again."""
        def arr(a, size, name):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64)
            if a.shape != (size,):
                raise ValueError(f"{name}: expected shape ({size},)")
            return a
        q = arr(charges, self.n, "charges")
        s = arr(sigma, self.ntypes, "sigma")
        e = arr(epsilon, self.ntypes, "epsilon")
        rt.check(rt.lib().mdrtHipFramesSetParameters(
            self._h, None if q is None else rt.ptr(q, C.c_double),
            None if s is None else rt.ptr(s, C.c_double),
            None if e is None else rt.ptr(e, C.c_double)))
        self._version += 1

    # ---- evaluation ----
    def _count(self, positions):
        K = int(positions.shape[0])
        if not 1 <= K <= self.max_frames:
            raise ValueError(f"frames: 1 to {self.max_frames} per call, found {K}")
        return K

    def evaluate(self, positions, forces_out=None, gradients=("positions",)) -> FrameEnergies:
        """Energies and forces of the frames `positions` (K, N, 3); loads
        them for vjp and hvp.  `forces_out`, a CUDA float64 tensor
        (K, N, 3), receives the forces on the device instead.  "strain" in
        `gradients` adds gradients["strain"] (K, 3, 3)."""
        K = self._count(positions)
        strain = np.zeros((K, 3, 3)) if "strain" in gradients else None
        pp, keep, _ = _frame_array(positions, (K, self.n, 3), "positions")
        terms = np.zeros((K, rt.NUM_ENERGIES))
        if forces_out is not None:
            fp, fkeep, _ = _frame_array(forces_out, (K, self.n, 3), "forces_out")
            F = None
        else:
            F = np.zeros((K, self.n, 3))
            fp, fkeep = C.c_void_p(F.ctypes.data), F
        rt.check(rt.lib().mdrtHipFramesEvaluate(
            self._h, K, pp, rt.ptr(terms, C.c_double), fp,
            None if strain is None else rt.ptr(strain, C.c_double)))
        del keep, fkeep
        self._version += 1
        self._loaded = K
        return FrameEnergies(self, K, terms, F, strain)

    def hvp(self, v, out=None):
        """H_n v_n (K, N, 3) for the loaded frames (the reference's
        frame_hvp, torch.py:341).  `out`, a CUDA float64 tensor, receives
        it on the device instead of a returned array."""
        K = self._loaded
        if K == 0:
            raise RuntimeError("hvp: evaluate frames first")
        vp, keep, _ = _frame_array(v, (K, self.n, 3), "v")
        if out is not None:
            op, okeep, _ = _frame_array(out, (K, self.n, 3), "out")
            res = out
        else:
            res = np.zeros((K, self.n, 3))
            op, okeep = C.c_void_p(res.ctypes.data), res
        rt.check(rt.lib().mdrtHipFramesHvp(self._h, K, vp, op))
        del keep, okeep
        return res

    def force_matching(self, positions, reference_forces, graph: bool = True):
        """Loss per frame sum_i |F_ni - F_ref,ni|^2 (K,) and the gradient of
        the summed loss in the tunables, in one call (forward, residual and
        adjoint captured in one HIP graph per frame count)."""
        K = self._count(positions)
        pp, keep, _ = _frame_array(positions, (K, self.n, 3), "positions")
        rp, rkeep, _ = _frame_array(reference_forces, (K, self.n, 3), "reference_forces")
        loss = np.zeros(K)
        dq = np.zeros(self.n)
        ds = np.zeros(self.ntypes)
        de = np.zeros(self.ntypes)
        rt.check(rt.lib().mdrtHipFramesForceMatching(
            self._h, K, pp, rp, int(bool(graph)), rt.ptr(loss, C.c_double),
            rt.ptr(dq, C.c_double), rt.ptr(ds, C.c_double), rt.ptr(de, C.c_double)))
        del keep, rkeep
        self._version += 1
        self._loaded = K
        return loss, FrameGradient({"charges": dq, "sigma": ds, "epsilon": de})

    def autotune(self, positions, reference_forces, what: int = 7, repeats: int = 5) -> dict:
        """Times force_matching on these frames for each launch shape (bit 0
        of what: PME brick widths, bit 1: plane transforms, bit 2: pair-loop
        segment lengths) and keeps the fastest; returns {(brick z lines,
        transform, segment entries): ms per call}. The evaluator holds no
        evaluated frames afterwards."""
        K = self._count(positions)
        pp, keep, _ = _frame_array(positions, (K, self.n, 3), "positions")
        rp, rkeep, _ = _frame_array(reference_forces, (K, self.n, 3), "reference_forces")
        ms = np.full(24, np.nan)
        rt.check(rt.lib().mdrtHipFramesAutotune(self._h, int(what), K, pp, rp, int(repeats),
                                                rt.ptr(ms, C.c_double)))
        del keep, rkeep
        self._version += 1
        self._loaded = 0
        from .simulation import shape_table
        return shape_table(ms)


def force_matching_benchmark(nframes: int = 64, repeats: int = 20, loop_passes: int = 3,
                             autotune: int = 7):
    """Force-matching frame-gradients per second on the differentiation
    workload (64 frames of 3,150 atoms: 1,000 TIP3P waters and 30 five-site
    Lennard-Jones solutes of 4 types, PME 32^3, rc 0.9 nm): one call
    evaluates the frames, the loss sum |F(theta) - F_ref|^2 against forces
    at moved parameters, and its gradient in every charge and the per-type
    sigma and epsilon.  Positions and reference forces are resident on the
    device before the timed region; a call ends with the gradient on the
    host.  Batched (all frames per call, one captured graph) against the
    plain per-frame loop (one frame per call, same library and graphs).
    With autotune, both evaluators first keep their fastest PME launch shape
    (FrameEvaluator.autotune on the same frames)."""
    import sys
    import time
    from . import systems as S
    L = rt.lib()
    system = S.solvated_solutes()
    n = system.n
    X = S.perturbed_frames(system, nframes, seed=9)
    q1, s1, e1 = S.perturbed_parameters(system)
    ev = FrameEvaluator(system, max_frames=nframes)
    theta = ev.tunables
    ev.update(q1, s1, e1)
    Fref = ev.evaluate(X).forces
    ev.update(theta["charges"], theta["sigma"], theta["epsilon"])
    tuned = ev.autotune(X, Fref, what=autotune, repeats=5) if autotune else {}
    stream = L.mgpuStreamCreate()
    nbytes = X.nbytes
    dx = L.mgpuMemAlloc(nbytes, stream, False)
    df = L.mgpuMemAlloc(nbytes, stream, False)
    Xc = np.ascontiguousarray(X)
    Fc = np.ascontiguousarray(Fref)
    L.mgpuMemcpy(dx, Xc.ctypes.data, nbytes, stream)
    L.mgpuMemcpy(df, Fc.ctypes.data, nbytes, stream)
    L.mgpuStreamSynchronize(stream)
    loss = np.zeros(nframes)
    dq = np.zeros(n)
    ds = np.zeros(system.ntypes)
    de = np.zeros(system.ntypes)
    P = lambda a: rt.ptr(a, C.c_double)

    def call(h, count, px, pf):
        rt.check(L.mdrtHipFramesForceMatching(h, count, C.c_void_p(px), C.c_void_p(pf), 1,
                                              P(loss), P(dq), P(ds), P(de)))
    for _ in range(3):
        call(ev._h, nframes, dx, df)
    t0 = time.perf_counter()
    for _ in range(repeats):
        call(ev._h, nframes, dx, df)
    batched_s = (time.perf_counter() - t0) / repeats
    grad_batched = (dq.copy(), ds.copy(), de.copy())
    one = FrameEvaluator(system, max_frames=1) if loop_passes > 0 else None
    if one is not None and autotune:
        one.autotune(X[:1], Fref[:1], what=autotune, repeats=5)
    frame_bytes = n * 3 * 8
    loop_s = float("nan")
    if one is not None:
        for k in range(3):
            call(one._h, 1, dx + k * frame_bytes, df + k * frame_bytes)
        t0 = time.perf_counter()
        for _ in range(loop_passes):
            for k in range(nframes):
                call(one._h, 1, dx + k * frame_bytes, df + k * frame_bytes)
        loop_s = (time.perf_counter() - t0) / loop_passes
        one.close()
    L.mgpuMemFree(dx, stream)
    L.mgpuMemFree(df, stream)
    L.mgpuStreamSynchronize(stream)
    L.mgpuStreamDestroy(stream)
    ev.close()
    rate = nframes / batched_s
    rate_loop = nframes / loop_s
    print(f"frames: {nframes} x {n} atoms, batched {batched_s * 1e3:.3f} ms "
          f"({rate:.0f} frame-gradients/s), per-frame loop {loop_s * 1e3:.3f} ms "
          f"({rate_loop:.0f}/s), x{rate / rate_loop:.1f}", file=sys.stderr, flush=True)
    return {"metric": "force-matching frame-gradients/sec", "value": round(rate, 1),
            "unit": "frame-gradients/s", "higher_is_better": True,
            "ms_per_batch": round(batched_s * 1e3, 4),
            "per_frame_loop": {"value": round(rate_loop, 1), "ms_per_batch": round(loop_s * 1e3, 4)},
            "speedup_vs_per_frame_loop": round(rate / rate_loop, 2),
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "tip3p_1000_waters_30_solutes_pme32", "frames": nframes,
                       "atoms": n, "box_nm": round(float(system.box[0]), 4),
                       "cutoff_nm": system.cutoff, "pme_grid": list(system.pme_grid),
                       "tunables": {"charges": n, "sigma": system.ntypes,
                                    "epsilon": system.ntypes},
                       "repeats": repeats, "loop_passes": loop_passes},
            "gradient_norm": float(np.sqrt(sum(float((g * g).sum()) for g in grad_batched))),
            "autotune_ms_per_call": {f"seg{sg}_bricks{tl}_transform{tr}": round(v, 4)
                                     for (tl, tr, sg), v in sorted(tuned.items())}}


if __name__ == "__main__":
    import argparse
    import json
    p = argparse.ArgumentParser(description="force-matching frame-gradients per second")
    p.add_argument("--frames", type=int, default=64)
    p.add_argument("--repeats", type=int, default=20)
    p.add_argument("--loop-passes", type=int, default=3)
    p.add_argument("--autotune", type=int, default=7,
                   help="launch-shape trial: bit 0 PME brick shapes, bit 1 plane transforms, "
                        "bit 2 pair-loop segment lengths (0: none)")
    a = p.parse_args()
    print(json.dumps(force_matching_benchmark(a.frames, a.repeats, a.loop_passes, a.autotune)),
          flush=True)

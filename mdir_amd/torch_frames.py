"""PyTorch autograd over the frame evaluator (the reference's torch custom
operators mdir::frame_energies, frame_vjp and frame_hvp,
python/mdir/torch.py:139-364).

    energy, forces = frame_forces(evaluator, positions, charges, sigma, epsilon)

returns the energies (K,) and forces (K, N, 3) of the frames `positions`
(a float64 tensor, CUDA or CPU; import this module before anything loads
libmdrt_hip.so when CUDA tensors are used) at the tunables `charges` (N,), `sigma`
and `epsilon` (ntypes,).  The backward pass takes cotangents on both
outputs and gives gradients in the tunables (the hand-written adjoint
kernels: pair loop, excluded and 1-4 pairs, PME charge adjoint) and in the
positions (the Hessian product of the same kernels), recomputing the pair
geometry instead of keeping it.  A second backward is not provided.
"""
from __future__ import annotations

import numpy as np
import torch

from . import runtime as rt
from .frames import FrameEvaluator

# torch ships its own HIP runtime under the same soname as the system's.
# Whichever loads first serves the process: initialise torch's before this
# package loads libmdrt_hip.so, so that CUDA tensors and the library share
# one runtime (the library then runs on torch's).
if rt._lib is None and torch.cuda.is_available():
    torch.cuda.init()


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


class FrameForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, evaluator: FrameEvaluator, positions, charges, sigma, epsilon):
        evaluator.update(_np(charges), _np(sigma), _np(epsilon))
        x = positions.detach().to(torch.float64).contiguous()
        if x.is_cuda:
            forces = torch.empty_like(x)
            out = evaluator.evaluate(x, forces_out=forces)
        else:
            out = evaluator.evaluate(x.numpy())
            forces = torch.from_numpy(out.forces)
        energy = torch.from_numpy(out.energy.copy()).to(positions.device)
        ctx.evaluator = evaluator
        ctx.version = evaluator._version
        ctx.save_for_backward(x, forces)
        ctx.tunables = (_np(charges), _np(sigma), _np(epsilon))
        ctx.devices = (charges.device, sigma.device, epsilon.device, positions.device)
        return energy, forces

    @staticmethod
    def backward(ctx, g_energy, g_forces):
        ev = ctx.evaluator
        x, forces = ctx.saved_tensors
        if ev._version != ctx.version:  # other frames since: evaluate again
            ev.update(*ctx.tunables)
            ev.evaluate(x if x.is_cuda else x.numpy())
        K = x.shape[0]
        g = np.zeros(K) if g_energy is None else _np(g_energy)
        G = np.zeros(tuple(x.shape)) if g_forces is None else -_np(g_forces)  # on dU/dx
        from .frames import FrameEnergies
        import mdir_amd.runtime as rt
        fe = FrameEnergies(ev, K, np.zeros((K, rt.NUM_ENERGIES)), _np(forces))
        want_x = ctx.needs_input_grad[1]
        p = fe.vjp(g, positions=G, gradients=("tunables", "positions") if want_x
                   else ("tunables",))
        dq = torch.from_numpy(p["charges"]).to(ctx.devices[0])
        ds = torch.from_numpy(p["sigma"]).to(ctx.devices[1])
        de = torch.from_numpy(p["epsilon"]).to(ctx.devices[2])
        dx = torch.from_numpy(p.positions).to(ctx.devices[3]) if want_x else None
        return None, dx, dq, ds, de


def frame_forces(evaluator: FrameEvaluator, positions, charges, sigma, epsilon):
    """Energies (K,) and forces (K, N, 3), differentiable in the positions
    and the tunables."""
    return FrameForces.apply(evaluator, positions, charges, sigma, epsilon)

"""The autograd check of the frame evaluator with CUDA tensors, run as its
own process (torch's HIP runtime initialised before the library loads):
python tests/torch_frames_check.py; exit status 0 when every check holds."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def check(device):
    import torch
    from mdir_amd import systems as S
    from mdir_amd.frames import FrameEvaluator
    from mdir_amd.torch_frames import frame_forces
    from oracle import oracle as O
    s = S.solvated_solutes(nwaters=300, nsolutes=12, seed=12, grid=(24, 24, 24))
    X = S.perturbed_frames(s, 2, seed=7)
    q1, s1, e1 = S.perturbed_parameters(s)
    ref = S.with_parameters(s, q1, s1, e1)
    Fref = np.stack([O.Oracle(ref).total(X[n])[0] for n in range(2)])
    ev = FrameEvaluator(s, max_frames=2)
    dev = torch.device(device)
    sig0, eps0 = S.type_parameters(s)
    q = torch.tensor(s.charges, dtype=torch.float64, device=dev, requires_grad=True)
    sg = torch.tensor(sig0, dtype=torch.float64, device=dev, requires_grad=True)
    ep = torch.tensor(eps0, dtype=torch.float64, device=dev, requires_grad=True)
    x = torch.tensor(X, dtype=torch.float64, device=dev, requires_grad=True)
    fr = torch.tensor(Fref, dtype=torch.float64, device=dev)
    energy, forces = frame_forces(ev, x, q, sg, ep)
    loss = ((forces - fr) ** 2).sum()
    loss.backward()
    lo, grad = ev.force_matching(X, Fref)
    np.testing.assert_allclose(loss.item(), lo.sum(), rtol=1e-12)
    for name, t in (("charges", q), ("sigma", sg), ("epsilon", ep)):
        np.testing.assert_allclose(t.grad.cpu().numpy(), grad[name], rtol=1e-9,
                                   atol=1e-9 * np.abs(grad[name]).max())
    # the position gradient along a direction against the FP64 oracle
    rng = np.random.default_rng(3)
    V = rng.normal(size=X.shape)
    h = 1e-6

    def L(Y):
        return sum(np.sum((O.Oracle(s).total(Y[n])[0] - Fref[n]) ** 2) for n in range(2))
    fd = (L(X + h * V) - L(X - h * V)) / (2 * h)
    got = float((x.grad.cpu().numpy() * V).sum())
    print(f"  {device}: d loss along v {got:.8e} vs central difference {fd:.8e}", flush=True)
    assert abs(got - fd) <= 5e-4 * abs(fd)
    # the energy cotangent: d(sum U)/dq against the force-matching-free vjp
    energy2, _ = frame_forces(ev, x.detach(), q, sg, ep)
    q.grad = None
    energy2.sum().backward()
    p = ev.evaluate(X).vjp(np.ones(2))
    np.testing.assert_allclose(q.grad.cpu().numpy(), p["charges"], rtol=1e-9,
                               atol=1e-9 * np.abs(p["charges"]).max())


if __name__ == "__main__":
    import torch
    torch.cuda.init()  # torch's runtime first
    check(sys.argv[1] if len(sys.argv) > 1 else "cuda:0")
    print("ok", flush=True)

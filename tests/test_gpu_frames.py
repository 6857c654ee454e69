"""The frame evaluator on the GPU: forces of a batch of frames against the
FP64 oracle, and the reverse products against FP64 central differences of
the oracle (the method of the reference's
test/Driver/Inputs/frame_batch_oracle.py:28-31, h = 1e-4).

Tolerances: forces as the force path (relative RMS 5e-6); the products in
the tunables for a given cotangent within 5e-5 of the largest entry of each
tunable; the force-matching gradient, whose cotangent comes from the f32
forces (an error of the forces is amplified by |F| / |F - F_ref|), within
the reference's mixed-precision bar for vjp, 5e-4 of the largest entry
(docs/python-frames.md, the table of the product checks: "1e-7, 5e-4");
Hessian products within 2e-5 relative RMS of the central difference of the
FP64 forces, over the atoms in no pair within the difference's reach of the
cutoff (a pair crossing the cutoff makes a step in the forces)."""
import numpy as np
import pytest

from fixtures import chain_box
from mdir_amd import systems as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
H_THETA = 1e-4
TOL_THETA = 5e-5
TOL_FORCE_MATCHING = 5e-4
TOL_HVP = 2e-5


def say(msg):
    print(msg, flush=True)


def small_solutes(seed=4):
    return S.solvated_solutes(nwaters=300, nsolutes=12, seed=seed, grid=(24, 24, 24))


def oracle_eval(system, x):
    F, e = O.Oracle(system).total(x)
    return F, e.sum()


def product_objective(system, frames, g, G):
    """sum_n [g_n U_n + sum_i G_ni . dU_n/dx_i] from the FP64 oracle."""
    f = 0.0
    for n in range(len(frames)):
        F, U = oracle_eval(system, frames[n])
        f += g[n] * U - np.sum(G[n] * F)
    return f


def central(fun, system, which, index, h=H_THETA):
    q = system.charges.copy()
    s, e = S.type_parameters(system)
    vals = {"charges": q, "sigma": s, "epsilon": e}

    def at(delta):
        v = {k: a.copy() for k, a in vals.items()}
        v[which][index] += delta
        return fun(S.with_parameters(system, v["charges"], v["sigma"], v["epsilon"]))
    return (at(h) - at(-h)) / (2 * h)


def tunable_indices(system, rng, ncharges=6):
    _, e = S.type_parameters(system)
    lj = [a for a in range(system.ntypes) if e[a] > 0]
    q = list(rng.choice(system.n, ncharges, replace=False))
    return {"charges": q, "sigma": lj, "epsilon": lj}


def smooth_atoms(system, x, dx, width):
    """Atoms in no pair whose distance is within `width` of the cutoff at x
    or x + dx: their forces are smooth across a central difference of that
    size (a pair crossing the cutoff changes a force by a finite step)."""
    from scipy.spatial import cKDTree
    L = system.box[0]
    bad = np.zeros(system.n, bool)
    for y in (x, x + dx, x - dx):
        t = cKDTree(np.mod(y, L), boxsize=L)
        near = t.query_pairs(system.cutoff + width, output_type="ndarray")
        if len(near) == 0:
            continue
        d = y[near[:, 0]] - y[near[:, 1]]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        hit = near[np.abs(r - system.cutoff) < width]
        bad[hit.ravel()] = True
    return ~bad


def check_hvp(system, x, v, hv, label, h=1e-6, tol=TOL_HVP):
    Fp, _ = oracle_eval(system, x + h * v)
    Fm, _ = oracle_eval(system, x - h * v)
    want = -(Fp - Fm) / (2 * h)
    ok = smooth_atoms(system, x, h * v, 10 * h * np.abs(v).max())
    err = O.rel_rms(hv[ok], want[ok])
    worst = np.abs(hv[ok] - want[ok]).max() / np.abs(want[ok]).max()  # of the largest
    say(f"  {label}: rel rms error of H v {err:.2e}, max {worst:.2e} "
        f"({(~ok).sum()} atoms near the cutoff left out)")
    assert err < tol


def check_products(got, want, label, tol=TOL_THETA):
    for name in got:
        scale = max(np.abs(got[name]).max(), np.abs(want[name]).max(), 1e-30)
        err = np.abs(got[name] - want[name]).max() / scale
        say(f"  {label} {name}: max error / largest entry {err:.2e} (largest {scale:.4e})")
        assert err < tol, (name, got[name], want[name])


def test_frame_forces_match_oracle():
    from mdir_amd.frames import FrameEvaluator
    say("case: frames 4 x 960 atoms (300 waters + 12 five-site solutes), PME 24^3")
    s = small_solutes()
    X = S.perturbed_frames(s, 4, seed=1)
    ev = FrameEvaluator(s, max_frames=4)
    out = ev.evaluate(X)
    for n in range(4):
        Fo, eo = O.Oracle(s).total(X[n])
        err = O.rel_rms(out.forces[n], Fo)
        say(f"  frame {n}: rel rms force error {err:.2e}")
        assert err < 5e-6
        for k in (O.E_LJ, O.E_COUL_DIRECT, O.E_COUL_EXCL, O.E_COUL_RECIP, O.E_COUL_SELF):
            assert abs(out.terms[n, k] - eo[k]) <= 2e-6 * abs(eo[k]) + 1e-3, (n, k)


def test_batch_matches_single_frames_and_repeats_bit_exact():
    from mdir_amd.frames import FrameEvaluator
    say("case: 5 frames batched against one at a time; repeated calls")
    s = small_solutes(seed=5)
    X = S.perturbed_frames(s, 5, seed=2)
    rng = np.random.default_rng(3)
    G = rng.normal(size=X.shape)
    ev = FrameEvaluator(s, max_frames=5)
    a = ev.evaluate(X)
    pa = a.vjp(positions=G)
    b = ev.evaluate(X)
    pb = b.vjp(positions=G)
    assert np.array_equal(a.forces, b.forces)
    for k in pa:
        assert np.array_equal(pa[k], pb[k])
    one = FrameEvaluator(s, max_frames=1)
    tot = {k: np.zeros_like(v) for k, v in pa.items()}
    for n in range(5):
        r = one.evaluate(X[n:n + 1])
        assert np.array_equal(r.forces[0], a.forces[n])
        p = r.vjp(positions=G[n:n + 1])
        for k in tot:
            tot[k] += p[k]
    for k in tot:
        np.testing.assert_allclose(tot[k], pa[k], rtol=1e-12, atol=1e-9 * np.abs(pa[k]).max())


def test_vjp_against_central_differences():
    from mdir_amd.frames import FrameEvaluator
    say("case: vjp with energy and position cotangents, 3 frames x 960 atoms")
    s = small_solutes(seed=6)
    X = S.perturbed_frames(s, 3, seed=3)
    rng = np.random.default_rng(4)
    G = rng.normal(size=X.shape) * 0.5
    g = rng.normal(size=3)
    ev = FrameEvaluator(s, max_frames=3)
    p = ev.evaluate(X).vjp(g, positions=G)
    idx = tunable_indices(s, rng)
    fun = lambda sys: product_objective(sys, X, g, G)
    got = {k: p[k][idx[k]] for k in idx}
    want = {k: np.array([central(fun, s, k, i) for i in idx[k]]) for k in idx}
    check_products(got, want, "vjp")


def test_force_matching_loss_and_gradient():
    from mdir_amd.frames import FrameEvaluator
    say("case: force matching, 3 frames x 960 atoms, reference forces at moved parameters")
    s = small_solutes(seed=7)
    X = S.perturbed_frames(s, 3, seed=5)
    q1, s1, e1 = S.perturbed_parameters(s)
    ref = S.with_parameters(s, q1, s1, e1)
    Fref = np.stack([O.Oracle(ref).total(X[n])[0] for n in range(3)])
    ev = FrameEvaluator(s, max_frames=3)
    loss, grad = ev.force_matching(X, Fref, graph=True)
    loss2, grad2 = ev.force_matching(X, Fref, graph=True)
    loss3, grad3 = ev.force_matching(X, Fref, graph=False)
    for k in grad:
        assert np.array_equal(grad[k], grad2[k]) and np.array_equal(grad[k], grad3[k])
    assert np.array_equal(loss, loss2) and np.array_equal(loss, loss3)

    def L(sys):
        return sum(np.sum((O.Oracle(sys).total(X[n])[0] - Fref[n]) ** 2) for n in range(3))
    lo = [np.sum((O.Oracle(s).total(X[n])[0] - Fref[n]) ** 2) for n in range(3)]
    say(f"  loss {loss} vs oracle {lo}")
    np.testing.assert_allclose(loss, lo, rtol=1e-4)
    rng = np.random.default_rng(8)
    idx = tunable_indices(s, rng)
    got = {k: grad[k][idx[k]] for k in idx}
    want = {k: np.array([central(L, s, k, i) for i in idx[k]]) for k in idx}
    check_products(got, want, "force matching", TOL_FORCE_MATCHING)


def test_hvp_against_central_differences():
    from mdir_amd.frames import FrameEvaluator
    say("case: Hessian products, 2 frames x 960 atoms (pairs, exclusions, PME)")
    s = small_solutes(seed=8)
    X = S.perturbed_frames(s, 2, seed=6)
    rng = np.random.default_rng(9)
    V = rng.normal(size=X.shape)
    ev = FrameEvaluator(s, max_frames=2)
    ev.evaluate(X)
    Hv = ev.hvp(V)
    for n in range(2):
        check_hvp(s, X[n], V[n], Hv[n], f"frame {n}")


@pytest.mark.parametrize("modifier", [S.LJ_NONE, S.LJ_POTENTIAL_SHIFT, S.LJ_FORCE_SWITCH,
                                      S.LJ_POTENTIAL_SWITCH])
def test_chains_with_14_pairs(modifier):
    """Four-atom chains: excluded pairs and scaled 1-4 pairs in the products,
    every Lennard-Jones modifier, the dispersion correction in the energy
    cotangent."""
    from mdir_amd.frames import FrameEvaluator
    say(f"case: chains with 1-4 pairs, modifier {modifier}")
    c = chain_box(nchains=60, lj_modifier=modifier)
    sig, eps = S.lorentz_berthelot(np.array([0.34, 0.32]), np.array([0.36, 0.45]))
    c = S.with_parameters(c, sigma=np.diag(sig), epsilon=np.diag(eps))
    c.dispersion = modifier in (S.LJ_NONE, S.LJ_POTENTIAL_SHIFT)
    X = S.perturbed_frames(c, 2, seed=10, scale=0.01)
    rng = np.random.default_rng(11)
    G = rng.normal(size=X.shape)
    g = rng.normal(size=2)
    ev = FrameEvaluator(c, max_frames=2)
    out = ev.evaluate(X)
    for n in range(2):
        assert O.rel_rms(out.forces[n], O.Oracle(c).total(X[n])[0]) < 5e-6
    p = out.vjp(g, positions=G)
    idx = {"charges": list(rng.choice(c.n, 6, replace=False)), "sigma": [0, 1],
           "epsilon": [0, 1]}
    fun = lambda sys: product_objective(sys, X, g, G)
    got = {k: p[k][idx[k]] for k in idx}
    want = {k: np.array([central(fun, c, k, i) for i in idx[k]]) for k in idx}
    check_products(got, want, "chains vjp")
    V = rng.normal(size=X.shape)
    Hv = ev.hvp(V)
    for n in range(2):
        check_hvp(c, X[n], V[n], Hv[n], f"frame {n}")


def test_vjp_type_pair_entries():
    """The product in individual type-pair table entries (sigma_ab = sigma_ba,
    eps_ab = eps_ba, the reference's tunables "by pairs of types"), with the
    dispersion correction's share of the energy cotangent, against central
    differences of the oracle with the two table entries moved together."""
    from mdir_amd.frames import FrameEvaluator
    say("case: vjp in type-pair entries (O-O, O-solute, solute pairs), dispersion correction")
    s = small_solutes(seed=9)
    s.dispersion = True
    X = S.perturbed_frames(s, 2, seed=11)
    rng = np.random.default_rng(12)
    G = rng.normal(size=X.shape) * 0.5
    g = rng.normal(size=2)
    ev = FrameEvaluator(s, max_frames=2)
    pairs = [(0, 0), (0, 2), (3, 2), (4, 4)]
    p = ev.evaluate(X).vjp_pairs(pairs, g, positions=G)

    def central_entry(which, a, b, h=H_THETA):
        def at(delta):
            t = s.copy()
            t.sigma, t.epsilon = s.sigma.copy(), s.epsilon.copy()
            tab = t.sigma if which == "sigma" else t.epsilon
            tab[a, b] += delta
            if a != b:
                tab[b, a] += delta
            return product_objective(t, X, g, G)
        return (at(h) - at(-h)) / (2 * h)
    want = {k: np.array([central_entry(k, a, b) for a, b in pairs]) for k in ("sigma", "epsilon")}
    check_products(p, want, "type-pair entries")

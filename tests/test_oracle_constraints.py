"""SETTLE, its velocity constraint and the NVE step of the oracle, pinned by
an independent Newton-converged SHAKE and by invariants."""
import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import oracle as O


def shake(x_old, x_new, masses, d_oh, d_hh, waters, iters=50):
    """Newton iteration for multipliers along the old bonds (the equations
    SETTLE solves in closed form)."""
    x = x_new.copy()
    for o in waters:
        idx = [o, o + 1, o + 2]
        bonds = [(0, 1, d_oh), (0, 2, d_oh), (1, 2, d_hh)]
        im = 1.0 / masses[idx]
        r0 = [x_old[idx[b]] - x_old[idx[a]] for a, b, _ in bonds]
        lam = np.zeros(3)
        base = x_new[idx].copy()
        for _ in range(iters):
            pos = base.copy()
            for k, (a, b, _) in enumerate(bonds):
                pos[a] -= lam[k] * im[a] * r0[k]
                pos[b] += lam[k] * im[b] * r0[k]
            g = np.array([np.dot(pos[b] - pos[a], pos[b] - pos[a]) - d * d for a, b, d in bonds])
            J = np.zeros((3, 3))
            for k, (a, b, _) in enumerate(bonds):
                rk = pos[b] - pos[a]
                for m, (c, d_, _) in enumerate(bonds):
                    dr = np.zeros(3)
                    if d_ == b: dr += im[b] * r0[m]
                    if c == b: dr -= im[b] * r0[m]
                    if d_ == a: dr -= im[a] * r0[m]
                    if c == a: dr += im[a] * r0[m]
                    J[k, m] = 2 * np.dot(rk, dr)
            lam -= np.linalg.solve(J, g)
        pos = base.copy()
        for k, (a, b, _) in enumerate(bonds):
            pos[a] -= lam[k] * im[a] * r0[k]
            pos[b] += lam[k] * im[b] * r0[k]
        x[idx] = pos
    return x


def test_settle_matches_shake():
    w = S.tip3p_box(27, 1.9, seed=4)
    o = O.Oracle(w)
    rng = np.random.default_rng(2)
    x_old = w.positions
    x_new = x_old + rng.normal(0, 0.004, x_old.shape)
    got = o.settle_positions(x_old, x_new)
    ref = shake(x_old, x_new, w.masses, *w.water_geometry, w.waters)
    assert np.abs(got - ref).max() < 1e-12
    f = w.waters
    assert np.allclose(np.linalg.norm(got[f + 1] - got[f], axis=1), 0.09572, atol=1e-13)
    assert np.allclose(np.linalg.norm(got[f + 2] - got[f + 1], axis=1), 0.15139, atol=1e-13)
    # displacement conserves momentum
    dp = (w.masses[:, None] * (got - x_new)).reshape(-1, 3, 3).sum(1)
    assert np.abs(dp).max() < 1e-14


def test_settle_velocities_constraint():
    w = S.tip3p_box(27, 1.9, seed=5)
    o = O.Oracle(w)
    rng = np.random.default_rng(1)
    v = rng.normal(0, 1, w.positions.shape)
    got = o.settle_velocities(w.positions, v)
    x = w.positions
    for o_ in w.waters:
        for a, b in ((0, 1), (0, 2), (1, 2)):
            d = x[o_ + b] - x[o_ + a]
            assert abs(np.dot(got[o_ + b] - got[o_ + a], d)) < 1e-12
    p0 = (w.masses[:, None] * v).sum(0)
    p1 = (w.masses[:, None] * got).sum(0)
    assert np.allclose(p0, p1, atol=1e-12)
    # the product's numpy version agrees
    prod = S.constrain_water_velocities(x, v, w.waters, w.masses[0], w.masses[1])
    assert np.abs(prod - got).max() < 1e-12


def test_initial_velocities_product_equals_oracle():
    w = S.tip3p_box(64, 2.2, seed=3)
    a = S.initial_velocities(w, 42, 300.0)
    b = O.Oracle(w).initial_velocities(42, 300.0)
    assert np.abs(a - b).max() < 1e-13
    ke = 0.5 * np.sum(w.masses[:, None] * a * a)
    assert ke == pytest.approx(0.5 * w.degrees_of_freedom() * S.BOLTZMANN * 300.0, rel=1e-12)


def test_splitmix64_published_value():
    """splitmix64 from seed 0 gives 0xe220a8397b1dcdaf first (Steele et al.)."""
    import ctypes as C
    st = C.c_uint64(0)
    assert O.lib().orc_splitmix64(C.byref(st)) == 0xE220A8397B1DCDAF


def test_nve_water_energy_conservation():
    w = S.tip3p_box(27, 1.9, seed=8, grid=(20, 20, 20), cutoff=0.8)
    o = O.Oracle(w)
    v = o.initial_velocities(3, 300.0)
    x, v, trace = o.nve(w.positions, v, 0.001, 40)
    total = trace.sum(1)
    assert np.ptp(total) < 2e-3 * abs(total).mean() + 0.5
    f = w.waters
    assert np.allclose(np.linalg.norm(x[f + 1] - x[f], axis=1), 0.09572, atol=1e-12)


def test_xoshiro256ss_published_values():
    """xoshiro256** from the state {1, 2, 3, 4}: 11520, 0, 1509978240,
    1215971899390074240 (the published first outputs of the generator)."""
    import ctypes as C
    st = (C.c_uint64 * 4)(1, 2, 3, 4)
    got = [O.lib().orc_xoshiro256ss(st) for _ in range(4)]
    assert got == [11520, 0, 1509978240, 1215971899390074240]

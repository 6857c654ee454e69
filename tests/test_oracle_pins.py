"""Pins of oracle functions that the golden vectors leave open: the
initial-velocity draws against the reference's own generator compiled into
oracle/_ref, the argon example's initial kinetic energy, and the dispersion
correction with exclusions against a brute-force mean over pairs."""
import ctypes as C
import itertools

import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import oracle as O

KCAL = 4.184

needs_ref = pytest.mark.skipif(not O.ref_draws_available(),
                               reason="oracle/_ref/libref_draws.so not built (no reference tree)")


@needs_ref
@pytest.mark.parametrize("seed", [42, 314159])
def test_initial_normals_match_reference_word_for_word(seed):
    """orc_initial_normals draws sequentially; the reference's generator
    (lib/Driver/InitialVelocityDraws.h) seeks to counter 6 i per particle.
    The normals agree bit for bit, in order and in a permuted visit order."""
    n = 864
    mine = O.initial_normals(seed, n)
    ref = O.ref_initial_normals(seed, np.arange(n))
    assert np.array_equal(mine, ref)
    perm = np.random.default_rng(seed).permutation(n)
    assert np.array_equal(mine, O.ref_initial_normals(seed, perm))


@needs_ref
def test_initial_velocities_directions_match_reference():
    """Before the constraint, centring and scaling, the oracle's velocities
    are the reference's normals times sqrt(kT/m): argon (no constraints),
    where only the centring and the common scale follow."""
    a = S.argon_fcc()
    v = O.Oracle(a).initial_velocities(314159, 183.0)
    ref = O.ref_initial_normals(314159, np.arange(a.n))
    ref = ref - ref.mean(axis=0)                    # equal masses: COM at rest
    scale = np.sum(v * ref) / np.sum(ref * ref)
    assert np.abs(v - scale * ref).max() < 1e-12 * np.abs(v).max()


def test_argon_example_initial_kinetic_energy():
    """test/Driver/argon.test:11: 864 argon atoms at 183 K start with
    470.7558 kcal/mol of kinetic energy (3N - 3 degrees of freedom, k_B
    from SI): the scaling of the draws and the degrees of freedom."""
    a = S.argon_fcc()
    v = S.initial_velocities(a, 314159, 183.0)
    ke = 0.5 * np.sum(a.masses[:, None] * v * v) / KCAL
    assert abs(ke - 470.7558) < 6e-5
    assert np.abs((a.masses[:, None] * v).sum(0)).max() < 1e-10


def test_dispersion_correction_with_exclusions_brute_force():
    """E = -(2 pi / 3) N^2 <C6> / (V rc^3) with <C6> the mean of
    4 eps sigma^6 over the ordered pairs of distinct, non-excluded
    particles (TermNonbonded.cpp:601-636), enumerated pair by pair on a
    small system of three types with exclusions between and within types.
    A slip in the exclusion term (its factor 2, or the pair count) fails."""
    rng = np.random.default_rng(5)
    n, nt = 14, 3
    types = rng.integers(0, nt, n).astype(np.int32)
    sig_t = rng.uniform(0.25, 0.4, nt)
    eps_t = rng.uniform(0.1, 1.0, nt)
    sigma = 0.5 * (sig_t[:, None] + sig_t[None, :])
    eps = np.sqrt(eps_t[:, None] * eps_t[None, :])
    all_pairs = list(itertools.combinations(range(n), 2))
    pick = rng.choice(len(all_pairs), 9, replace=False)
    excl = np.array([all_pairs[k] for k in pick], np.int32)
    box = np.array([2.1, 2.3, 2.6])
    rc = 0.95
    excluded = {tuple(p) for p in excl.tolist()} | {(j, i) for i, j in excl.tolist()}
    c6 = lambda i, j: 4.0 * eps[types[i], types[j]] * sigma[types[i], types[j]] ** 6
    terms = [c6(i, j) for i in range(n) for j in range(n) if i != j and (i, j) not in excluded]
    mean = np.mean(terms)
    want = -2.0 * np.pi / (3.0 * box.prod()) * n * n * mean / rc ** 3
    f64 = lambda a: np.ascontiguousarray(a, np.float64)
    got = O.lib().orc_dispersion_energy(
        n, types.ctypes.data_as(C.POINTER(C.c_int32)), nt,
        f64(sigma).ctypes.data_as(C.POINTER(C.c_double)),
        f64(eps).ctypes.data_as(C.POINTER(C.c_double)), len(excl),
        excl.ctypes.data_as(C.POINTER(C.c_int32)),
        f64(box).ctypes.data_as(C.POINTER(C.c_double)), rc)
    assert got == pytest.approx(want, rel=1e-13)
    # the same mean without exclusions differs: the test sees the exclusion term
    plain = [c6(i, j) for i in range(n) for j in range(n) if i != j]
    assert abs(np.mean(plain) - mean) > 1e-6 * abs(mean)

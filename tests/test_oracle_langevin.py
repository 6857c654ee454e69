"""Pins of the oracle's Langevin dynamics: Philox 4x32-10 against the
published known-answer vectors (Random123's kat_vectors) and against the
reference's own mdrtPhilox4x32 (runtime/mdrt.c:289-312, compiled into
oracle/_ref); the normals' statistics; and the g-BAOAB step's limits
(friction 0 is velocity Verlet with SETTLE; the stationary temperature of
free particles in a harmonic-free box is the bath's)."""
import numpy as np
import pytest

from mdir_amd import systems as S
from oracle import oracle as O

# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key, result
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]

needs_ref = pytest.mark.skipif(not O.ref_host_available(),
                               reason="oracle/_ref/libmdrt_host_ref.so not built (no reference tree)")


@pytest.mark.parametrize("counter,key,want", KAT)
def test_philox_known_answers(counter, key, want):
    assert O.philox4x32_10(counter, key) == list(want)


@needs_ref
def test_philox_matches_reference_runtime():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        c = [int(w) for w in rng.integers(0, 2 ** 32, 4, dtype=np.uint64)]
        k = [int(w) for w in rng.integers(0, 2 ** 32, 2, dtype=np.uint64)]
        assert O.philox4x32_10(c, k) == O.ref_philox4x32(c, k)


def test_langevin_normals_follow_the_noise_formula():
    """The three normals of a particle from its Philox block, recomputed
    here from the words by the formula of dyn.noise (uniforms (w + 1/2)
    2^-32, Box-Muller cos and sin of the first pair, cos of the second)."""
    seed, step, pid = 314159, 123456789012, 77
    w = O.philox4x32_10([step & 0xFFFFFFFF, step >> 32, pid, 2 << 24],
                        [seed & 0xFFFFFFFF, seed >> 32])
    u = [(x + 0.5) * 2.0 ** -32 for x in w]
    a, b = np.sqrt(-2 * np.log(u[0])), np.sqrt(-2 * np.log(u[2]))
    want = [a * np.cos(2 * np.pi * u[1]), a * np.sin(2 * np.pi * u[1]),
            b * np.cos(2 * np.pi * u[3])]
    assert np.allclose(O.langevin_normals(seed, step, pid), want, rtol=1e-14, atol=1e-15)


def test_langevin_normals_statistics():
    r = np.array([O.langevin_normals(42, s, p) for s in range(40) for p in range(500)])
    n = r.size
    assert abs(r.mean()) < 5 / np.sqrt(n)
    assert abs(r.var() - 1.0) < 5 * np.sqrt(2.0 / n)
    assert abs(np.mean(r ** 4) - 3.0) < 0.1
    # the three numbers of a particle are uncorrelated
    c = np.corrcoef(r.T)
    assert np.abs(c - np.eye(3)).max() < 5 / np.sqrt(len(r))


def test_langevin_without_friction_is_velocity_verlet():
    """Without constraints and with friction 0 the O step is the identity
    and the two half drifts add up to velocity Verlet's drift."""
    a = S.argon_config1()
    a.positions = a.positions + np.random.default_rng(4).normal(0, 0.01, a.positions.shape)
    o = O.Oracle(a)
    v0 = o.initial_velocities(7, 120.0)
    x1, v1, _ = o.nve(a.positions, v0, 0.002, 3)
    x2, v2, _ = o.langevin(a.positions, v0, 0.002, 3, seed=5, friction=0.0, temperature=120.0)
    assert np.abs(x1 - x2).max() < 1e-13
    assert np.abs(v1 - v2).max() < 1e-11


def test_geodesic_step_without_friction_conserves_energy_and_bonds():
    """With rigid waters and friction 0, g-BAOAB is a symplectic constrained
    integrator (two SETTLE half drifts): bond lengths hold and the total
    energy stays within the fluctuation of velocity Verlet's."""
    w = S.tip3p_box(64, 1.3, seed=3, grid=(16, 16, 16), cutoff=0.6)
    o = O.Oracle(w)
    v0 = o.initial_velocities(9, 300.0)
    x, v, tr = o.langevin(w.positions, v0, 0.002, 200, seed=5, friction=0.0, temperature=300.0)
    _, _, tn = o.nve(w.positions, v0, 0.002, 200)
    f = w.waters
    assert np.abs(np.linalg.norm(x[f + 1] - x[f], axis=1) - w.water_geometry[0]).max() < 1e-10
    total, total_nve = tr.sum(1), tn.sum(1)
    assert np.ptp(total) < 2.0 * np.ptp(total_nve) + 1e-3


def test_langevin_ideal_gas_temperature():
    """Free particles without interactions (charges and epsilon 0): the
    O step alone sets the velocity distribution; after many steps the mean
    kinetic energy per degree of freedom is kT/2."""
    rng = np.random.default_rng(2)
    n = 400
    a = S.System(positions=rng.uniform(0, 3.0, (n, 3)), masses=np.full(n, 20.0),
                 charges=np.zeros(n), types=np.zeros(n, np.int32),
                 sigma=np.zeros((1, 1)), epsilon=np.zeros((1, 1)), box=np.full(3, 3.0),
                 cutoff=0.9, beta=0.0)
    o = O.Oracle(a)
    x, v = a.positions, np.zeros((n, 3))
    x, v, trace = o.langevin(x, v, 0.004, 400, seed=11, friction=5.0, temperature=250.0)
    ke = trace[200:, 1].mean()
    want = 1.5 * n * S.BOLTZMANN * 250.0
    assert abs(ke / want - 1.0) < 0.04

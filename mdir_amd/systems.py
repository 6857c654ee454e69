"""Systems of the force path: topology, parameters and synthetic generators.

A :class:`System` holds what the per-step force path reads: positions,
masses, charges, Lennard-Jones types with per type-pair sigma/epsilon
tables, excluded pairs, scaled 1-4 pairs, rigid (SETTLE) waters, the
orthorhombic box and the nonbonded/PME settings.  Field meanings follow the
reference's Topology and Control (lib/Driver/TermNonbonded.cpp,
lib/Driver/TermPairs14.cpp:11-17, lib/Driver/ConstraintIR.cpp:473-593).

Units are the reference's: nm, ps, amu, kJ/mol, e.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

# include/mdir/Support/PhysicalConstants.h:18-21 (SI defining constants and
# CODATA 2022 permittivity), as exact binary64 values.
COULOMB = float.fromhex("0x1.15def44ac2182p+7")   # kJ nm mol^-1 e^-2
BOLTZMANN = float.fromhex("0x1.1072c483af26dp-7")  # kJ mol^-1 K^-1
KJ_PER_KCAL = 4.184

LJ_NONE, LJ_POTENTIAL_SHIFT, LJ_FORCE_SWITCH, LJ_POTENTIAL_SWITCH = 0, 1, 2, 3

# TIP3P (Jorgensen 1983) as config 2 states it.
TIP3P = dict(q_o=-0.834, q_h=0.417, sigma_o=0.31506, eps_o=0.6364,
             m_o=15.9994, m_h=1.008, d_oh=0.09572, d_hh=0.15139)


def ewald_beta(cutoff: float, tolerance: float = 1e-5) -> float:
    """beta with erfc(beta rc) = tolerance by 100 halvings
    (lib/Driver/TermNonbonded.cpp:778-794)."""
    low, high = 0.0, 1.0
    while math.erfc(high * cutoff) > tolerance:
        high *= 2.0
    for _ in range(100):
        middle = 0.5 * (low + high)
        if math.erfc(middle * cutoff) > tolerance:
            low = middle
        else:
            high = middle
    return 0.5 * (low + high)


def smooth_size(least: int) -> int:
    """Smallest even size >= least with factors 2,3,5,7
    (lib/Driver/TermNonbonded.cpp:756-767)."""
    size = least
    while True:
        if size % 2 == 0:
            rest = size
            for f in (2, 3, 5, 7):
                while rest % f == 0:
                    rest //= f
            if rest == 1:
                return size
        size += 1


def pme_grid(box, spacing: float = 0.12):
    """Grid no wider than `spacing` (TermNonbonded.cpp:1108-1127)."""
    return tuple(smooth_size(int(math.ceil(L / spacing - 1e-9))) for L in box)


@dataclasses.dataclass
class System:
    positions: np.ndarray            # (n, 3) f64
    masses: np.ndarray               # (n,) f64
    charges: np.ndarray              # (n,) f64
    types: np.ndarray                # (n,) i32
    sigma: np.ndarray                # (ntypes, ntypes) f64
    epsilon: np.ndarray              # (ntypes, ntypes) f64
    box: np.ndarray                  # (3,) f64
    cutoff: float = 0.9
    switch_from: float = 0.0
    lj_modifier: int = LJ_NONE
    coulomb_shift: bool = False
    dispersion: bool = False
    beta: float = 0.0                # 0: no electrostatics
    pme_grid: tuple = (0, 0, 0)
    pme_order: int = 4
    exclusions: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros((0, 2), np.int32))   # (m, 2) i<j
    pairs14: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros((0, 2), np.int32))
    pairs14_sigma: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros(0))
    pairs14_eps: np.ndarray = dataclasses.field(          # eps * s_LJ
        default_factory=lambda: np.zeros(0))
    pairs14_qq: np.ndarray = dataclasses.field(           # f q_i q_j s_C
        default_factory=lambda: np.zeros(0))
    waters: np.ndarray = dataclasses.field(               # first atom of each
        default_factory=lambda: np.zeros(0, np.int32))
    water_geometry: tuple = (TIP3P["d_oh"], TIP3P["d_hh"])
    velocities: Optional[np.ndarray] = None
    # bonded tuples (energies in kJ/mol, lengths nm, angles rad):
    # bonds and Urey-Bradley 1-3 pairs (k/2)(r - r0)^2, angles (apex the
    # middle atom) (k/2)(theta - theta0)^2, dihedrals k (1 + cos(n phi -
    # phase)) (Amber's impropers among them), harmonic impropers
    # (k/2) wrap(phi - phi0)^2
    bonds: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 2), np.int32))
    bond_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    bond_r0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    urey_bradley: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros((0, 2), np.int32))
    ub_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    ub_r0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    angles: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 3), np.int32))
    angle_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    angle_theta0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    dihedrals: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros((0, 4), np.int32))
    dihedral_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    dihedral_n: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    dihedral_phase: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    impropers: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros((0, 4), np.int32))
    improper_k: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    improper_phi0: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    # bond constraints (SHAKE clusters: one centre atom and the atoms bonded
    # to it, contiguous in the atom order)
    constraints: np.ndarray = dataclasses.field(
        default_factory=lambda: np.zeros((0, 2), np.int32))
    constraint_length: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    @property
    def ntypes(self) -> int:
        return int(self.sigma.shape[0])

    def degrees_of_freedom(self) -> int:
        """3N - 3 - 3 per rigid water - 1 per bond constraint
        (System.cpp getDegreesOfFreedom)."""
        return 3 * self.n - 3 - 3 * len(self.waters) - len(self.constraints)

    # tuple kinds, in the order the kernels sort them
    TUPLE_BOND, TUPLE_UREY_BRADLEY, TUPLE_ANGLE, TUPLE_DIHEDRAL, TUPLE_IMPROPER = range(5)

    def tuples(self):
        """All bonded tuples as (atoms (T, 4) i32, kind (T,) i32, params
        (T, 4) f64), sorted by kind: bond (k, r0), Urey-Bradley (k, r0),
        angle (k, theta0), dihedral (k, n, phase), improper (k, phi0)."""
        rows, kinds, params = [], [], []

        def add(atoms, kind, cols):
            atoms = np.asarray(atoms, np.int32).reshape(-1, atoms.shape[1] if np.ndim(atoms) == 2
                                                        else 1)
            m = atoms.shape[0]
            if m == 0:
                return
            a = np.zeros((m, 4), np.int32)
            a[:, :atoms.shape[1]] = atoms
            pr = np.zeros((m, 4))
            for c, v in enumerate(cols):
                pr[:, c] = np.asarray(v, np.float64)
            rows.append(a)
            kinds.append(np.full(m, kind, np.int32))
            params.append(pr)

        add(self.bonds, self.TUPLE_BOND, (self.bond_k, self.bond_r0))
        add(self.urey_bradley, self.TUPLE_UREY_BRADLEY, (self.ub_k, self.ub_r0))
        add(self.angles, self.TUPLE_ANGLE, (self.angle_k, self.angle_theta0))
        add(self.dihedrals, self.TUPLE_DIHEDRAL,
            (self.dihedral_k, self.dihedral_n, self.dihedral_phase))
        add(self.impropers, self.TUPLE_IMPROPER, (self.improper_k, self.improper_phi0))
        if not rows:
            return np.zeros((0, 4), np.int32), np.zeros(0, np.int32), np.zeros((0, 4))
        return np.vstack(rows), np.concatenate(kinds), np.vstack(params)

    def clusters(self):
        """The constraint clusters: (first, size, centre offset) (m, 3) i32
        and each atom's constraint length to its centre (n,; 0 for a centre
        and for atoms in no cluster).  A cluster is an atom and the atoms
        constrained to it, each of which has no other constraint; they must
        be consecutive in the atom order (as in Amber topologies, where
        hydrogens follow their heavy atom), at most 5."""
        n = self.n
        length = np.zeros(n)
        if len(self.constraints) == 0:
            return np.zeros((0, 3), np.int32), length
        partners = {}
        for (i, j), d in zip(np.asarray(self.constraints), np.asarray(self.constraint_length)):
            partners.setdefault(int(i), []).append((int(j), float(d)))
            partners.setdefault(int(j), []).append((int(i), float(d)))
        centres = sorted(a for a, ps in partners.items() if len(ps) > 1 or
                         (len(ps) == 1 and len(partners[ps[0][0]]) == 1 and a < ps[0][0]))
        out, seen = [], set()
        for c in centres:
            if c in seen:
                continue
            members = [c] + [j for j, _ in partners[c]]
            for j, d in partners[c]:
                if len(partners[j]) != 1:
                    raise ValueError(f"constraint of atoms {c} and {j}: clusters are an atom "
                                     "and atoms bonded to it alone")
                length[j] = d
            lo, hi = min(members), max(members)
            if hi - lo + 1 != len(members) or len(members) > 5:
                raise ValueError(f"the constraint cluster of atom {c} is not {len(members)} "
                                 "consecutive atoms (at most 5)")
            seen.update(members)
            out.append((lo, len(members), c - lo))
        return np.array(out, np.int32).reshape(-1, 3), length

    def exclusion_csr(self):
        """Symmetric CSR of excluded partners per atom."""
        n = self.n
        rows = [[] for _ in range(n)]
        for i, j in self.exclusions:
            rows[int(i)].append(int(j))
            rows[int(j)].append(int(i))
        start = np.zeros(n + 1, np.int32)
        for i in range(n):
            start[i + 1] = start[i] + len(rows[i])
        atoms = np.array([j for r in rows for j in sorted(r)], np.int32)
        return start, atoms

    def kinetic_energy(self) -> float:
        v = self.velocities
        return 0.5 * float(np.sum(self.masses[:, None] * v * v))

    def copy(self) -> "System":
        return dataclasses.replace(
            self, positions=self.positions.copy(),
            velocities=None if self.velocities is None else self.velocities.copy())


def argon_fcc(cells: int = 6, lattice: float = 0.57964333, sigma: float = 0.34,
              epsilon: float = 0.99773551, mass: float = 39.95,
              cutoff: float = 0.85, switch_from: float = 0.75,
              lj_modifier: int = LJ_POTENTIAL_SWITCH,
              dispersion: bool = False) -> System:
    """Argon on a face-centred cubic lattice, atoms ordered as
    examples/argon/argon.mlir:102-134 places them (corner, then the face
    centres normal to z, y, x), offset a quarter cell."""
    n = 4 * cells ** 3
    x = np.zeros((n, 3))
    for i in range(n):
        cell, basis = divmod(i, 4)
        rest, cx = divmod(cell, cells)
        cz, cy = divmod(rest, cells)
        c = (cx, cy, cz)
        normal = 3 - basis
        for k in range(3):
            offset = 0.0 if (basis == 0 or k == normal) else 0.5
            x[i, k] = (c[k] + offset + 0.25) * lattice
    box = np.full(3, cells * lattice)
    return System(positions=x, masses=np.full(n, mass), charges=np.zeros(n),
                  types=np.zeros(n, np.int32), sigma=np.array([[sigma]]),
                  epsilon=np.array([[epsilon]]), box=box, cutoff=cutoff,
                  switch_from=switch_from, lj_modifier=lj_modifier,
                  dispersion=dispersion)


def argon_config1() -> System:
    """Liquid argon, 864 atoms on an FCC lattice (6x6x6 cells of 4), cubic
    box 3.47 nm, sigma 0.3405 nm, epsilon 0.996 kJ/mol, cutoff 1.0 nm with
    the dispersion correction."""
    return argon_fcc(cells=6, lattice=3.47 / 6, sigma=0.3405, epsilon=0.996,
                     cutoff=1.0, switch_from=0.0, lj_modifier=LJ_NONE,
                     dispersion=True)


def _random_rotations(rng: np.random.Generator, count: int) -> np.ndarray:
    """Uniform rotations from unit quaternions (Shoemake 1992)."""
    u = rng.random((count, 3))
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]),
                  np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]),
                  np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], axis=1)
    a, b, c, d = q.T
    return np.stack([
        np.stack([1 - 2 * (c * c + d * d), 2 * (b * c - a * d), 2 * (b * d + a * c)], 1),
        np.stack([2 * (b * c + a * d), 1 - 2 * (b * b + d * d), 2 * (c * d - a * b)], 1),
        np.stack([2 * (b * d - a * c), 2 * (c * d + a * b), 1 - 2 * (b * b + c * c)], 1),
    ], axis=1)


def tip3p_box(nwaters: int, edge: float, seed: int = 2024, cutoff: float = 0.9,
              tolerance: float = 1e-5, grid=None, order: int = 4,
              jitter: float = 0.08) -> System:
    """A TIP3P water box: waters on a seeded jittered cubic lattice of
    ceil(n^(1/3))^3 sites (a seeded choice of n of them) with uniformly
    random orientations; PME with beta from `tolerance`, grid of spacing
    <= 0.12 nm unless given, SETTLE-rigid waters, O-H/H-H excluded."""
    rng = np.random.default_rng(seed)
    m = int(math.ceil(nwaters ** (1.0 / 3.0) - 1e-9))
    spacing = edge / m
    sites = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m),
                                 indexing="ij"), -1).reshape(-1, 3)
    keep = np.sort(rng.choice(len(sites), size=nwaters, replace=False))
    centres = (sites[keep] + 0.5) * spacing
    centres += (rng.random((nwaters, 3)) - 0.5) * 2 * jitter * spacing
    p = TIP3P
    half = 0.5 * p["d_hh"]
    h = math.sqrt(p["d_oh"] ** 2 - half ** 2)
    total = p["m_o"] + 2 * p["m_h"]
    ra = 2 * p["m_h"] * h / total
    body = np.array([[0.0, ra, 0.0], [-half, ra - h, 0.0], [half, ra - h, 0.0]])
    rot = _random_rotations(rng, nwaters)
    x = (centres[:, None, :] + np.einsum("wij,aj->wai", rot, body)).reshape(-1, 3)
    n = 3 * nwaters
    q = np.tile([p["q_o"], p["q_h"], p["q_h"]], nwaters)
    masses = np.tile([p["m_o"], p["m_h"], p["m_h"]], nwaters)
    types = np.tile(np.array([0, 1, 1], np.int32), nwaters)
    sigma = np.array([[p["sigma_o"], 0.0], [0.0, 0.0]])
    eps = np.array([[p["eps_o"], 0.0], [0.0, 0.0]])
    first = np.arange(0, n, 3, dtype=np.int32)
    excl = np.concatenate([np.stack([first, first + 1], 1),
                           np.stack([first, first + 2], 1),
                           np.stack([first + 1, first + 2], 1)]).astype(np.int32)
    excl = excl[np.lexsort((excl[:, 1], excl[:, 0]))]
    box = np.full(3, float(edge))
    if grid is None:
        grid = pme_grid(box)
    return System(positions=x, masses=masses, charges=q, types=types,
                  sigma=sigma, epsilon=eps, box=box, cutoff=cutoff,
                  beta=ewald_beta(cutoff, tolerance), pme_grid=tuple(grid),
                  pme_order=order, exclusions=excl, waters=first,
                  water_geometry=(p["d_oh"], p["d_hh"]))


def tip3p_config2(seed: int = 2024) -> System:
    """7,900 TIP3P waters (23,700 atoms) in a 6.2 nm cube, PME rc 0.9 nm,
    tolerance 1e-5, 64^3 grid, order 4."""
    return tip3p_box(7900, 6.2, seed=seed, cutoff=0.9, grid=(64, 64, 64))


# ---- initial velocities -----------------------------------------------------

_MASK64 = (1 << 64) - 1


def _splitmix64(state: int):
    state = (state + 0x9E3779B97F4A7C15) & _MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    return state, z ^ (z >> 31)


def _xoshiro_uniforms(seed: int, count: int) -> np.ndarray:
    """`count` numbers in (0, 1] of xoshiro256** seeded by splitmix64
    (lib/Driver/InitialVelocityDraws.h:18-48)."""
    s = []
    st = seed & _MASK64
    for _ in range(4):
        st, z = _splitmix64(st)
        s.append(z)
    out = np.empty(count)
    rot = lambda x, k: ((x << k) | (x >> (64 - k))) & _MASK64
    s0, s1, s2, s3 = s
    for i in range(count):
        result = (rot((s1 * 5) & _MASK64, 7) * 9) & _MASK64
        t = (s1 << 17) & _MASK64
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        s3 = rot(s3, 45)
        out[i] = ((result >> 11) + 1.0) * 2.0 ** -53
    return out


def constrain_water_velocities(x, v, waters, m_o, m_h):
    """Removes the bond-relative velocities of rigid waters: impulses along
    the three bonds from a 3x3 solve (ConstraintIR.cpp:392-471,
    System.cpp:1158-1215).  Vectorised over waters; returns a copy."""
    v = v.copy()
    o = np.asarray(waters)
    if len(o) == 0:
        return v
    xa, xb, xc = x[o], x[o + 1], x[o + 2]
    e0 = xb - xa
    e1 = xc - xa
    e2 = xc - xb
    for e in (e0, e1, e2):
        e /= np.linalg.norm(e, axis=1)[:, None]
    va, vb, vc = v[o], v[o + 1], v[o + 2]
    r = np.stack([-np.sum(e0 * (vb - va), 1), -np.sum(e1 * (vc - va), 1),
                  -np.sum(e2 * (vc - vb), 1)], 1)
    io, ih = 1.0 / m_o, 1.0 / m_h
    A = np.zeros((len(o), 3, 3))
    A[:, 0, 0] = A[:, 1, 1] = io + ih
    A[:, 2, 2] = 2 * ih
    A[:, 0, 1] = A[:, 1, 0] = np.sum(e0 * e1, 1) * io
    A[:, 0, 2] = A[:, 2, 0] = -np.sum(e0 * e2, 1) * ih
    A[:, 1, 2] = A[:, 2, 1] = np.sum(e1 * e2, 1) * ih
    t = np.linalg.solve(A, r[..., None])[..., 0]
    i0, i1, i2 = t[:, :1] * e0, t[:, 1:2] * e1, t[:, 2:3] * e2
    v[o] = va - io * (i0 + i1)
    v[o + 1] = vb + ih * (i0 - i2)
    v[o + 2] = vc + ih * (i1 + i2)
    return v


def initial_velocities(system: System, seed: int, temperature: float) -> np.ndarray:
    """Maxwell-Boltzmann velocities as the reference draws them
    (lib/Driver/System.cpp:1298-1350): six uniforms per particle in order,
    Box-Muller normals of width sqrt(kT/m), constrained, centre of mass at
    rest, scaled to the temperature with the constrained degrees of freedom."""
    n = system.n
    u = _xoshiro_uniforms(seed, 6 * n).reshape(n, 3, 2)
    normals = np.sqrt(-2.0 * np.log(u[..., 0])) * np.cos(6.283185307179586 * u[..., 1])
    m = system.masses
    width = np.where(m > 0, np.sqrt(BOLTZMANN * temperature / np.where(m > 0, m, 1.0)), 0.0)
    v = width[:, None] * normals
    if n < 2:
        return v
    if len(system.waters):
        w0 = int(system.waters[0])
        v = constrain_water_velocities(system.positions, v, system.waters,
                                       m[w0], m[w0 + 1])
    v -= (m[:, None] * v).sum(0) / m.sum()
    kin = 0.5 * float(np.sum(m[:, None] * v * v))
    target = 0.5 * system.degrees_of_freedom() * BOLTZMANN * temperature
    if kin > 0:
        v *= np.sqrt(target / kin)
    return v


# ---- parameters per type and the frames of the differentiation workload ----

def lorentz_berthelot(sigma, epsilon, table_sigma=None):
    """Type-pair tables from per-type sigma and epsilon:
    sigma_ab = (sigma_a + sigma_b) / 2, eps_ab = sqrt(eps_a eps_b); where
    eps_ab is 0 the sigma of `table_sigma` is kept (it does not act)."""
    s = np.asarray(sigma, np.float64)
    e = np.asarray(epsilon, np.float64)
    eps = np.sqrt(np.outer(e, e))
    sig = 0.5 * (s[:, None] + s[None, :])
    if table_sigma is not None:
        sig = np.where(eps != 0.0, sig, np.asarray(table_sigma, np.float64))
    return sig, eps


def type_parameters(system: System):
    """Per-type sigma and epsilon: the diagonal of the tables."""
    return np.diag(system.sigma).copy(), np.diag(system.epsilon).copy()


def with_parameters(system: System, charges=None, sigma=None, epsilon=None) -> System:
    """A copy with other charges and per-type sigma / epsilon (tables by the
    Lorentz-Berthelot rules)."""
    s0, e0 = type_parameters(system)
    sig, eps = lorentz_berthelot(s0 if sigma is None else sigma,
                                 e0 if epsilon is None else epsilon, system.sigma)
    return dataclasses.replace(
        system, positions=system.positions.copy(),
        charges=system.charges.copy() if charges is None else np.asarray(charges, np.float64),
        sigma=sig, epsilon=eps)


def solvated_solutes(nwaters: int = 1000, nsolutes: int = 30, seed: int = 4,
                     ntypes: int = 4, cutoff: float = 0.9, grid=(32, 32, 32),
                     tolerance: float = 1e-5, jitter: float = 0.05) -> System:
    """TIP3P water with five-site solutes (a centre and four sites at 0.1 nm
    on a tetrahedron; all ten pairs of a solute excluded): solute atoms of
    `ntypes` Lennard-Jones types with seeded sigma in [0.3, 0.4] nm, epsilon
    in [0.2, 1.0] kJ/mol and charges in [-0.5, 0.5] e; Lorentz-Berthelot
    tables.  The box holds the waters at 1.0 g/cm^3 plus 0.1 nm^3 per
    solute; waters and solute centres sit on a jittered cubic lattice, with
    no water on the six sites next to a solute.
    The default (1,000 waters, 30 solutes: 3,150 atoms, PME 32^3, rc 0.9 nm)
    is the frame-gradient workload."""
    rng = np.random.default_rng(seed)
    p = TIP3P
    nsol_atoms = 5 * nsolutes
    n = 3 * nwaters + nsol_atoms
    m_water = (p["m_o"] + 2 * p["m_h"]) / 6.02214076e23  # g
    vol_w = nwaters * m_water / 1.0 * 1e21                # nm^3 at 1.0 g/cm^3
    edge = (vol_w + 0.1 * nsolutes) ** (1.0 / 3.0)
    m = int(math.ceil((nwaters + 7 * nsolutes) ** (1.0 / 3.0) - 1e-9))
    spacing = edge / m
    ids = np.arange(m ** 3)
    coords = np.stack([ids % m, (ids // m) % m, ids // (m * m)], 1)
    # solute centres: spread out, then the six neighbours blocked
    blocked = np.zeros(m ** 3, bool)
    centres = []
    for s in rng.permutation(m ** 3):
        if blocked[s]:
            continue
        c = coords[s]
        nb = [(c + d) % m for d in ([0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0],
                                      [0, -1, 0], [0, 0, 1], [0, 0, -1])]
        idx = [int(v[0] + m * v[1] + m * m * v[2]) for v in nb]
        if any(blocked[i] for i in idx):
            continue
        centres.append(s)
        for i in idx:
            blocked[i] = True
        if len(centres) == nsolutes:
            break
    if len(centres) < nsolutes:
        raise ValueError("too many solutes for the lattice")
    free = np.flatnonzero(~blocked)
    wsites = np.sort(rng.choice(free, size=nwaters, replace=False))
    half = 0.5 * p["d_hh"]
    h = math.sqrt(p["d_oh"] ** 2 - half ** 2)
    total = p["m_o"] + 2 * p["m_h"]
    ra = 2 * p["m_h"] * h / total
    body = np.array([[0.0, ra, 0.0], [-half, ra - h, 0.0], [half, ra - h, 0.0]])
    wc = (coords[wsites] + 0.5) * spacing + (rng.random((nwaters, 3)) - 0.5) * 2 * jitter * spacing
    rot = _random_rotations(rng, nwaters)
    xw = (wc[:, None, :] + np.einsum("wij,aj->wai", rot, body)).reshape(-1, 3)
    tet = 0.1 / math.sqrt(3.0) * np.array([[0.0, 0.0, 0.0], [1, 1, 1], [1, -1, -1],
                                           [-1, 1, -1], [-1, -1, 1]])
    sc = (coords[np.array(centres)] + 0.5) * spacing
    rot_s = _random_rotations(rng, nsolutes)
    xs = (sc[:, None, :] + np.einsum("wij,aj->wai", rot_s, tet)).reshape(-1, 3)
    x = np.vstack([xw, xs])
    sig_t = np.concatenate([[p["sigma_o"], 0.0], rng.uniform(0.3, 0.4, ntypes)])
    eps_t = np.concatenate([[p["eps_o"], 0.0], rng.uniform(0.2, 1.0, ntypes)])
    table0 = np.zeros((ntypes + 2, ntypes + 2))
    sig, eps = lorentz_berthelot(sig_t, eps_t, table0)
    types = np.concatenate([np.tile(np.array([0, 1, 1]), nwaters),
                            rng.integers(2, 2 + ntypes, nsol_atoms)]).astype(np.int32)
    q = np.concatenate([np.tile([p["q_o"], p["q_h"], p["q_h"]], nwaters),
                        rng.uniform(-0.5, 0.5, nsol_atoms)])
    masses = np.concatenate([np.tile([p["m_o"], p["m_h"], p["m_h"]], nwaters),
                             np.full(nsol_atoms, 12.011)])
    first = np.arange(0, 3 * nwaters, 3, dtype=np.int32)
    excl = [np.stack([first, first + 1], 1), np.stack([first, first + 2], 1),
            np.stack([first + 1, first + 2], 1)]
    o = 3 * nwaters + 5 * np.arange(nsolutes)
    for i in range(5):
        for j in range(i + 1, 5):
            excl.append(np.stack([o + i, o + j], 1))
    excl = np.concatenate(excl).astype(np.int32)
    excl = excl[np.lexsort((excl[:, 1], excl[:, 0]))]
    box = np.full(3, float(edge))
    return System(positions=x, masses=masses, charges=q, types=types, sigma=sig,
                  epsilon=eps, box=box, cutoff=cutoff,
                  beta=ewald_beta(cutoff, tolerance), pme_grid=tuple(grid), pme_order=4,
                  exclusions=excl, waters=first, water_geometry=(p["d_oh"], p["d_hh"]))


def lj_contacts(system: System, x: np.ndarray, fraction: float = 0.8) -> np.ndarray:
    """The non-excluded pairs (i < j) with Lennard-Jones (epsilon > 0)
    closer than `fraction` of their sigma at positions x (periodic box),
    as an (m, 2) array."""
    from scipy.spatial import cKDTree
    L = np.asarray(system.box, np.float64)
    lj = system.epsilon > 0
    none = np.zeros((0, 2), np.int64)
    if not lj.any():
        return none
    reach = fraction * float(system.sigma[lj].max())
    tree = cKDTree(np.mod(x, L), boxsize=L)
    pairs = tree.query_pairs(reach, output_type="ndarray")
    if len(pairs) == 0:
        return none
    i, j = np.minimum(pairs[:, 0], pairs[:, 1]), np.maximum(pairs[:, 0], pairs[:, 1])
    ti, tj = system.types[i], system.types[j]
    d = x[j] - x[i]
    d -= L * np.rint(d / L)
    r = np.linalg.norm(d, axis=1)
    close = lj[ti, tj] & (r < fraction * system.sigma[ti, tj])
    ex = system.exclusions
    if len(ex) and close.any():
        key = lambda a, b: a.astype(np.int64) * system.n + b
        close &= ~np.isin(key(i, j), key(ex[:, 0], ex[:, 1]))
    return np.stack([i[close], j[close]], 1)


def perturbed_frames(system: System, nframes: int, seed: int = 9, scale: float = 0.02,
                     angle: float = 0.1, contact: float = 0.8):
    """nframes positions (nframes, n, 3): the system's positions with seeded
    moves: every rigid water turned by a random rotation of Gaussian angle
    (`angle` rad) about its oxygen and moved as a unit, every other atom
    moved on its own, by Gaussian displacements of `scale` nm per
    component.  Molecules in a non-excluded Lennard-Jones pair closer than
    `contact` sigma draw their moves again (from the same seeded stream)
    until none is left, so no frame carries overlaps that would dominate
    its forces (pairs already that close in the system's own positions
    are left as they are)."""
    rng = np.random.default_rng(seed)
    x0 = system.positions
    n = system.n
    w = np.asarray(system.waters, np.int64)
    # molecule of each atom: a water (its index) or the atom on its own
    mol = np.arange(n) + len(w)
    for k in range(3):
        mol[w + k] = np.arange(len(w))
    is_free = np.ones(n, bool)
    for k in range(3):
        is_free[w + k] = False
    free_atoms = np.flatnonzero(is_free)

    def move(x, waters, atoms):
        x[atoms] = x0[atoms] + rng.normal(0.0, scale, (len(atoms), 3))
        if len(waters) == 0:
            return
        o = w[waters]
        axis = rng.normal(size=(len(o), 3))
        axis /= np.linalg.norm(axis, axis=1)[:, None]
        phi = rng.normal(0.0, angle, len(o))
        shift = rng.normal(0.0, scale, (len(o), 3))
        c, s_ = np.cos(phi)[:, None], np.sin(phi)[:, None]
        x[o] = x0[o] + shift
        for k in (1, 2):   # Rodrigues' rotation of the O-H arms
            r = x0[o + k] - x0[o]
            kr = np.cross(axis, r)
            kd = np.sum(axis * r, axis=1)[:, None]
            x[o + k] = x[o] + r * c + kr * s_ + axis * kd * (1.0 - c)

    # contacts of the system's own positions are its own: only those the
    # moves create are drawn again
    key = lambda p: p[:, 0].astype(np.int64) * n + p[:, 1]
    base = key(lj_contacts(system, x0, contact))
    out = np.empty((nframes,) + x0.shape)
    for f in range(nframes):
        x = x0.copy()
        move(x, np.arange(len(w)), free_atoms)
        for _round in range(200):
            bad = lj_contacts(system, x, contact)
            bad = bad[~np.isin(key(bad), base)]
            if len(bad) == 0:
                break
            m = np.unique(mol[bad.ravel()])
            move(x, m[m < len(w)], m[m >= len(w)] - len(w))
        else:
            raise ValueError("Lennard-Jones contacts left after 200 rounds of moves")
        out[f] = x
    return out


def perturbed_parameters(system: System, seed: int = 13, rel: float = 0.02, dq: float = 0.02):
    """Charges and per-type sigma, epsilon moved by seeded amounts (the
    parameters of the reference forces of the force-matching workload)."""
    rng = np.random.default_rng(seed)
    s0, e0 = type_parameters(system)
    q = system.charges + rng.uniform(-dq, dq, system.n)
    s = s0 * (1.0 + rng.uniform(-rel, rel, s0.shape))
    e = e0 * (1.0 + rng.uniform(-rel, rel, e0.shape))
    return q, s, e


AVOGADRO = 6.02214076e23


def water_edge(nwaters: int, density: float = 1.0) -> float:
    """Edge (nm) of the cube holding `nwaters` TIP3P waters at `density`
    (g/cm^3)."""
    grams = nwaters * (TIP3P["m_o"] + 2 * TIP3P["m_h"]) / AVOGADRO
    return (grams / density) ** (1.0 / 3.0) * 1e7


def replica_sizes(count: int = 256, seed: int = 5, low: int = 500, high: int = 8000):
    """Numbers of waters of the batching workload: uniform in [low, high]."""
    return np.random.default_rng(seed).integers(low, high + 1, count)


def tip3p_replica(nwaters: int, seed: int) -> System:
    """One replica of the batching workload: `nwaters` TIP3P waters in a
    cube at 1.0 g/cm^3, PME with rc 0.9 nm, tolerance 1e-5, a grid no wider
    than 0.12 nm, order 4."""
    return tip3p_box(int(nwaters), water_edge(int(nwaters)), seed=seed, cutoff=0.9)

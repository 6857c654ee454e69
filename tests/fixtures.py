"""Fixtures shared by the tests: the reference's own input data read as text,
and the linear congruential positions of its neighbor tests."""
from __future__ import annotations

import math
import os

import numpy as np

from mdir_amd.systems import System, COULOMB

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KCAL = 4.184


# Amber's topology and coordinate readers are the package's own
# (mdir_amd/amber.py); the charge unit is e x sqrt(332.0522...) = 18.2223.
from mdir_amd.amber import CHARGE_UNIT as AMBER_CHARGE_UNIT  # noqa: E402
from mdir_amd.amber import read_inpcrd, read_prmtop  # noqa: E402,F401


def load_opc3(cutoff=0.9, beta=3.47045919371208278, grid=(64, 64, 64), order=8):
    """The three OPC waters of test/Driver/Inputs/opc/three.{prmtop,inpcrd}
    (data files of the reference's tests, kept as fixtures).  Each residue
    is O, H, H and a massless charge site EP; the restart file gives EP no
    meaningful position, so it is put where the OPC model has it: on the
    bisector of the H-O-H angle, at the O-EP bond length of the topology's
    bond table from the oxygen.  Every pair within a residue is excluded."""
    top = read_prmtop(os.path.join(GOLDEN, "opc3", "three.prmtop"))
    q = np.asarray(top["CHARGE"]) / AMBER_CHARGE_UNIT
    n = q.size
    x, box = read_inpcrd(os.path.join(GOLDEN, "opc3", "three.inpcrd"), n)
    kinds = top["AMBER_ATOM_TYPE"]
    # bond table entries: (3 i, 3 j, type + 1) triplets, lengths in Angstrom
    bonds = np.asarray(top["BONDS_WITHOUT_HYDROGEN"] + top["BONDS_INC_HYDROGEN"]).reshape(-1, 3)
    length = {}
    for a3, b3, t in bonds:
        length[frozenset((a3 // 3, b3 // 3))] = 0.1 * top["BOND_EQUIL_VALUE"][t - 1]
    first = [p - 1 for p in top["RESIDUE_POINTER"]] + [n]
    excl = []
    for lo, hi in zip(first[:-1], first[1:]):
        members = list(range(lo, hi))
        excl += [(i, j) for k, i in enumerate(members) for j in members[k + 1:]]
        oxygen = [i for i in members if kinds[i] == "OW"]
        hydrogens = [i for i in members if kinds[i] == "HW"]
        for site in (i for i in members if kinds[i] == "EP"):
            o = oxygen[0]
            u = sum((x[h] - x[o]) / np.linalg.norm(x[h] - x[o]) for h in hydrogens)
            x[site] = x[o] + length[frozenset((site, o))] * u / np.linalg.norm(u)
    return System(positions=x, masses=np.ones(n), charges=q,
                  types=np.zeros(n, np.int32), sigma=np.zeros((1, 1)),
                  epsilon=np.zeros((1, 1)), box=box, cutoff=cutoff, beta=beta,
                  pme_grid=tuple(grid), pme_order=order,
                  exclusions=np.array(excl, np.int32))


def lcg_positions(count, edge):
    """test/Runtime/Inputs/neighbors_reference.py:29-39 (seed 42)."""
    state = 42
    out = np.zeros((count, 3))
    for i in range(count):
        for k in range(3):
            state = (state * 1103515245 + 12345) % 2147483648
            out[i, k] = state / 2147483648.0 * edge
    return out


def chain_box(nchains=60, edge=3.2, seed=11, beta=3.47045919371208278,
              grid=(32, 32, 32), order=4, lj_modifier=0, cutoff=0.9,
              switch_from=0.7):
    """Four-atom chains (bonds 0.153 nm, bends near 112 degrees) on a
    jittered lattice in random orientations: every pair within a chain is excluded and the
    1-4 pair (0, 3) is scaled (s_LJ = 0.5, s_C = 1/1.2, the Amber factors).
    Two types, charges +-0.25 e."""
    rng = np.random.default_rng(seed)
    d = 0.153
    theta = math.radians(112.0)
    local = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0],
                      [d - d * math.cos(theta), d * math.sin(theta), 0.0]])
    local = np.vstack([local, local[2] + np.array([d, 0.0, 0.02])])
    local -= local.mean(axis=0)
    m = int(math.ceil(nchains ** (1 / 3) - 1e-9))
    spacing = edge / m
    slots = rng.permutation(m ** 3)[:nchains]
    xs, q, t, exc, p14 = [], [], [], [], []
    for c in range(nchains):
        a, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        ix = np.array([slots[c] % m, (slots[c] // m) % m, slots[c] // (m * m)])
        base = (ix + 0.5) * spacing + rng.uniform(-0.03, 0.03, 3)
        xs.append(base + local @ a.T)
        q += [0.25, -0.25, 0.25, -0.25]
        t += [0, 1, 1, 0]
        o = 4 * c
        exc += [(o + i, o + j) for i in range(4) for j in range(i + 1, 4)]
        p14.append((o, o + 3))
    x = np.vstack(xs)
    n = x.shape[0]
    sigma = np.array([[0.34, 0.33], [0.33, 0.32]])
    eps = np.array([[0.36, 0.40], [0.40, 0.45]])
    types = np.array(t, np.int32)
    charges = np.array(q)
    p14 = np.array(p14, np.int32)
    ti, tj = types[p14[:, 0]], types[p14[:, 1]]
    return System(positions=x, masses=np.full(n, 12.011), charges=charges,
                  types=types, sigma=sigma, epsilon=eps, box=np.full(3, edge),
                  cutoff=cutoff, switch_from=switch_from, lj_modifier=lj_modifier,
                  beta=beta, pme_grid=grid, pme_order=order,
                  exclusions=np.array(exc, np.int32), pairs14=p14,
                  pairs14_sigma=sigma[ti, tj], pairs14_eps=0.5 * eps[ti, tj],
                  pairs14_qq=COULOMB * charges[p14[:, 0]] * charges[p14[:, 1]] / 1.2)

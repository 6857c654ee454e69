"""Fixtures shared by the tests: the reference's own input data read as text,
and the linear congruential positions of its neighbor tests."""
from __future__ import annotations

import math
import os

import numpy as np

from mdir_amd.systems import System, COULOMB

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KCAL = 4.184


def read_prmtop(path):
    """Sections of an Amber prmtop, as the reference's
    test/Driver/Inputs/ewald_reference.py:24-39 reads them."""
    sections, flag, fmt = {}, None, None
    for line in open(path):
        if line.startswith("%FLAG"):
            flag = line.split()[1]
            sections[flag] = []
        elif line.startswith("%FORMAT"):
            fmt = line[line.index("(") + 1: line.index(")")]
        elif line.startswith("%"):
            continue
        elif flag:
            spec = fmt.split("a")[-1].split("I")[-1].split("E")[-1].split(".")[0]
            width = int("".join(c for c in spec if c.isdigit()))
            body = line.rstrip("\n")
            for i in range(0, len(body), width):
                field = body[i: i + width]
                if field.strip():
                    sections[flag].append(field.strip())
    return sections


def load_opc3(cutoff=0.9, beta=3.47045919371208278, grid=(64, 64, 64), order=8):
    """The three OPC waters of test/Driver/Inputs/opc/three.{prmtop,inpcrd}
    (data files of the reference's tests, copied as fixtures), extra points
    placed along the bisector at their bond length as
    ewald_reference.py:56-72 places them; every pair of a residue excluded."""
    top = read_prmtop(os.path.join(GOLDEN, "opc3", "three.prmtop"))
    charges = np.array([float(q) / 18.2223 for q in top["CHARGE"]])
    n = len(charges)
    starts = [int(s) - 1 for s in top["RESIDUE_POINTER"]] + [n]
    types = top["AMBER_ATOM_TYPE"]
    lines = open(os.path.join(GOLDEN, "opc3", "three.inpcrd")).read().splitlines()
    numbers = []
    for line in lines[2:]:
        numbers += [float(line[i: i + 12]) for i in range(0, len(line), 12)
                    if line[i: i + 12].strip()]
    x = np.array([[0.1 * numbers[3 * i + k] for k in range(3)] for i in range(n)])
    box = np.array([0.1 * numbers[3 * n + k] for k in range(3)])
    bonds = [int(v) for v in top["BONDS_WITHOUT_HYDROGEN"]]
    lengths = [0.1 * float(v) for v in top["BOND_EQUIL_VALUE"]]
    for b in range(0, len(bonds), 3):
        i, j, t = bonds[b] // 3, bonds[b + 1] // 3, bonds[b + 2] - 1
        site, owner = (i, j) if types[i] == "EP" else (j, i)
        if types[site] != "EP":
            continue
        units = []
        for a in (owner + 1, owner + 2):
            d = x[a] - x[owner]
            units.append(d / np.linalg.norm(d))
        s = units[0] + units[1]
        x[site] = x[owner] + lengths[t] * s / np.linalg.norm(s)
    excl = []
    for r in range(len(starts) - 1):
        for i in range(starts[r], starts[r + 1]):
            for j in range(i + 1, starts[r + 1]):
                excl.append((i, j))
    return System(positions=x, masses=np.ones(n), charges=charges,
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

"""Amber topologies (prmtop) and coordinates (inpcrd) as a System.

The counterpart of the reference's lib/Driver/Amber.cpp (its topology
reader) for the terms this path has: charges, masses, Lennard-Jones type
tables, exclusions, 1-4 pairs with their scale factors, bonds, angles,
proper and improper dihedrals, rigid waters and bonds to hydrogen as
constraints.  Units follow the reference (lib/Driver/Amber.cpp:39-40,
:611-760): charges CHARGE / 18.2223 e; lengths Angstrom x 0.1 nm; energies
kcal x 4.184 kJ; Amber's bond and angle energies K (x - x0)^2 become
(k/2) (x - x0)^2 with k = 2 K; a dihedral PK (1 + cos(n phi - phase)).

The file format: sections "%FLAG NAME", "%FORMAT(<count><kind><width>)"
(a Fortran edit descriptor), then fixed-width fields.
"""
from __future__ import annotations

import math
import re

import numpy as np

from .systems import COULOMB, System

CHARGE_UNIT = 18.2223  # e x sqrt(Amber's Coulomb constant in kcal A / mol)
KCAL = 4.184
NM_PER_ANGSTROM = 0.1
WATER_RESIDUES = ("WAT", "HOH", "TIP3", "SOL")

_FIELD = re.compile(r"(\d*)([AaIiEeFf])(\d+)(?:\.\d+)?")


def read_prmtop(path) -> dict:
    """{FLAG: list of values} of an Amber topology file."""
    sections = {}
    name, kind, width = None, None, None
    with open(path) as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            if line.startswith("%FLAG"):
                name = line[5:].strip()
                sections[name] = []
            elif line.startswith("%FORMAT"):
                m = _FIELD.search(line)
                kind, width = m.group(2).lower(), int(m.group(3))
            elif line.startswith("%") or name is None:
                continue
            else:
                cast = {"a": str.strip, "i": int}.get(kind, float)
                fields = (line[k: k + width] for k in range(0, len(line), width))
                sections[name].extend(cast(f) for f in fields if f.strip())
    return sections


def read_inpcrd(path, n):
    """Positions (nm, (n, 3)) and box edges (nm) of an Amber coordinate file:
    a title, the atom count, then reals in 12-column fields (3n positions in
    Angstrom, then the box lengths and angles)."""
    with open(path) as fh:
        body = fh.read().splitlines()[2:]
    vals = [float(line[k: k + 12]) for line in body for k in range(0, len(line), 12)
            if line[k: k + 12].strip()]
    x = NM_PER_ANGSTROM * np.asarray(vals[: 3 * n]).reshape(n, 3)
    box = NM_PER_ANGSTROM * np.asarray(vals[3 * n: 3 * n + 3])
    return x, box


def snap_phase(phase: float) -> float:
    """A dihedral phase within 1e-3 rad of a multiple of pi/2 taken as that
    multiple, in (-pi, pi] (the files write pi as 3.141594)."""
    q = round(phase / (0.5 * math.pi))
    if abs(phase - q * 0.5 * math.pi) < 1e-3:
        q %= 4
        return (0.0, 0.5 * math.pi, math.pi, -0.5 * math.pi)[q]
    return phase


def _lj_table(top, ntypes):
    """Per type pair (sigma nm, epsilon kJ/mol) from the A and B tables
    (A = 4 eps sigma^12, B = 4 eps sigma^6, kcal/mol with Angstrom)."""
    index = np.asarray(top["NONBONDED_PARM_INDEX"]).reshape(ntypes, ntypes)
    A, B = top["LENNARD_JONES_ACOEF"], top["LENNARD_JONES_BCOEF"]
    sigma, eps = np.zeros((ntypes, ntypes)), np.zeros((ntypes, ntypes))
    for a in range(ntypes):
        for b in range(ntypes):
            k = index[a, b]
            if k < 0:
                raise ValueError("10-12 hydrogen-bond terms are not supported")
            ak, bk = A[k - 1], B[k - 1]
            if ak > 0.0 and bk > 0.0:
                sigma[a, b] = (ak / bk) ** (1.0 / 6.0) * NM_PER_ANGSTROM
                eps[a, b] = bk * bk / (4.0 * ak) * KCAL
    return sigma, eps


def load_amber(prmtop, inpcrd, cutoff: float = 0.9, pme: bool = True,
               rigid_water: bool = True, hydrogen_bonds: bool = False,
               ewald_tolerance: float = 1e-5, dispersion: bool = True) -> System:
    """A System from Amber files.  rigid_water: waters (residues WAT, HOH,
    TIP3, SOL of an oxygen and two hydrogens) are SETTLE units and their
    bonds and angles leave the potential; hydrogen_bonds: the other bonds to
    hydrogen are constraints (SHAKE) and leave the potential.  pme: Ewald
    electrostatics with beta from erfc(beta rc) = tolerance and a grid of
    <= 0.12 nm spacing; otherwise no electrostatics beyond the 1-4 pairs."""
    from .systems import ewald_beta, pme_grid
    top = read_prmtop(prmtop)
    p = top["POINTERS"]
    n, ntypes = p[0], p[1]
    x, box = read_inpcrd(inpcrd, n)
    if box.size < 3 or not np.all(box > 0):
        box = NM_PER_ANGSTROM * np.asarray(top["BOX_DIMENSIONS"][1:4])
    q = np.asarray(top["CHARGE"], np.float64) / CHARGE_UNIT
    masses = np.asarray(top["MASS"], np.float64)
    types = np.asarray(top["ATOM_TYPE_INDEX"], np.int32) - 1
    sigma, eps = _lj_table(top, ntypes)

    # exclusions: NUMBER_EXCLUDED_ATOMS per atom, then that many partners
    # (1-based, 0 = none) in EXCLUDED_ATOMS_LIST
    excl, at = set(), 0
    lst = top["EXCLUDED_ATOMS_LIST"]
    for i, cnt in enumerate(top["NUMBER_EXCLUDED_ATOMS"]):
        for j in lst[at: at + cnt]:
            if j > 0:
                excl.add((min(i, j - 1), max(i, j - 1)))
        at += cnt

    # waters
    names = top["RESIDUE_LABEL"]
    first = [r - 1 for r in top["RESIDUE_POINTER"]] + [n]
    atomic = top.get("ATOMIC_NUMBER")
    is_h = (np.asarray(atomic) == 1) if atomic else np.array(
        [nm.startswith("H") for nm in top["ATOM_NAME"]])
    waters = []
    for name, lo, hi in zip(names, first[:-1], first[1:]):
        if rigid_water and name in WATER_RESIDUES and hi - lo == 3 and \
                not is_h[lo] and is_h[lo + 1] and is_h[lo + 2]:
            waters.append(lo)
    in_water = np.zeros(n, bool)
    for w in waters:
        in_water[w: w + 3] = True

    # bonds: (3i, 3j, type) triplets
    bk, br = top["BOND_FORCE_CONSTANT"], top["BOND_EQUIL_VALUE"]
    bonds, constraints, water_len = [], [], {}
    for flag, with_h in (("BONDS_INC_HYDROGEN", True), ("BONDS_WITHOUT_HYDROGEN", False)):
        t = np.asarray(top.get(flag, []), np.int64).reshape(-1, 3)
        for a3, b3, ty in t:
            i, j = int(a3) // 3, int(b3) // 3
            k = 2.0 * bk[ty - 1] * KCAL / NM_PER_ANGSTROM ** 2
            r0 = br[ty - 1] * NM_PER_ANGSTROM
            if in_water[i] and in_water[j]:
                water_len[(is_h[i] and is_h[j])] = r0
                continue
            if hydrogen_bonds and with_h and (is_h[i] or is_h[j]):
                constraints.append((i, j, r0))
                continue
            bonds.append((i, j, k, r0))

    ak, at0 = top["ANGLE_FORCE_CONSTANT"], top["ANGLE_EQUIL_VALUE"]
    angles = []
    for flag in ("ANGLES_INC_HYDROGEN", "ANGLES_WITHOUT_HYDROGEN"):
        t = np.asarray(top.get(flag, []), np.int64).reshape(-1, 4)
        for a3, b3, c3, ty in t:
            i, j, k = int(a3) // 3, int(b3) // 3, int(c3) // 3
            if in_water[i] and in_water[j] and in_water[k]:
                continue
            angles.append((i, j, k, 2.0 * ak[ty - 1] * KCAL, at0[ty - 1]))

    # dihedrals: (3i, 3j, +-3k, +-3l, type); a negative third index leaves
    # the 1-4 pair out, a negative fourth marks an improper; a negative
    # periodicity continues with the next type's term
    dk, dn, dph = (top["DIHEDRAL_FORCE_CONSTANT"], top["DIHEDRAL_PERIODICITY"],
                   top["DIHEDRAL_PHASE"])
    ntors = len(dk)
    scee = top.get("SCEE_SCALE_FACTOR") or [1.2] * ntors
    scnb = top.get("SCNB_SCALE_FACTOR") or [2.0] * ntors
    dihedrals, p14 = [], {}
    for flag in ("DIHEDRALS_INC_HYDROGEN", "DIHEDRALS_WITHOUT_HYDROGEN"):
        t = np.asarray(top.get(flag, []), np.int64).reshape(-1, 5)
        for a3, b3, c3, d3, ty in t:
            i, j, k, l = (abs(int(v)) // 3 for v in (a3, b3, c3, d3))
            term = int(ty)
            while True:
                per = dn[term - 1]
                dihedrals.append((i, j, k, l, dk[term - 1] * KCAL, int(round(abs(per))),
                                  snap_phase(dph[term - 1])))
                if per >= 0.0:
                    break
                term += 1
            if c3 < 0 or d3 < 0 or dn[term - 1] == 0.0:
                continue
            key = (min(i, l), max(i, l))
            if key in p14:
                raise ValueError(f"the 1-4 pair {key} is carried by two dihedrals")
            p14[key] = (1.0 / scee[term - 1], 1.0 / scnb[term - 1])

    # 1-4 pairs: sigma and epsilon of the two types' pair, scaled
    pairs14, s14, e14, qq14 = [], [], [], []
    for (i, l), (sc, sl) in sorted(p14.items()):
        pairs14.append((i, l))
        s14.append(sigma[types[i], types[l]])
        e14.append(eps[types[i], types[l]] * sl)
        qq14.append(COULOMB * q[i] * q[l] * sc)

    if waters:
        d_oh = water_len.get(False, 0.09572)
        d_hh = water_len.get(True, 0.15136)
    else:
        d_oh, d_hh = 0.09572, 0.15136
    kw = {}
    if pme:
        kw = dict(beta=ewald_beta(cutoff, ewald_tolerance), pme_grid=pme_grid(box),
                  pme_order=4)
    ex = np.array(sorted(excl), np.int32).reshape(-1, 2)
    return System(
        positions=x, masses=masses, charges=q, types=types, sigma=sigma, epsilon=eps,
        box=np.asarray(box, np.float64), cutoff=cutoff, dispersion=dispersion,
        exclusions=ex, pairs14=np.array(pairs14, np.int32).reshape(-1, 2),
        pairs14_sigma=np.array(s14), pairs14_eps=np.array(e14), pairs14_qq=np.array(qq14),
        waters=np.array(waters, np.int32), water_geometry=(d_oh, d_hh),
        bonds=np.array([b[:2] for b in bonds], np.int32).reshape(-1, 2),
        bond_k=np.array([b[2] for b in bonds]), bond_r0=np.array([b[3] for b in bonds]),
        angles=np.array([a[:3] for a in angles], np.int32).reshape(-1, 3),
        angle_k=np.array([a[3] for a in angles]), angle_theta0=np.array([a[4] for a in angles]),
        dihedrals=np.array([d[:4] for d in dihedrals], np.int32).reshape(-1, 4),
        dihedral_k=np.array([d[4] for d in dihedrals]),
        dihedral_n=np.array([d[5] for d in dihedrals], np.int32),
        dihedral_phase=np.array([d[6] for d in dihedrals]),
        constraints=np.array([c[:2] for c in constraints], np.int32).reshape(-1, 2),
        constraint_length=np.array([c[2] for c in constraints]), **kw)

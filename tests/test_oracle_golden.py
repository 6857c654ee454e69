"""The oracle against the reference's own golden values (no GPU).

Each expected number is quoted from the reference's tests and cited."""
import numpy as np
import pytest

from fixtures import KCAL, lcg_positions, load_opc3
from mdir_amd import systems as S
from oracle import oracle as O


def test_ewald_beta_bisection():
    """test/Driver/pme.test:8 passes beta = 3.47045919371208278 nm^-1, the
    bisection of TermNonbonded.cpp:778-794 at rc 0.9 nm, tolerance 1e-5."""
    assert O.lib().orc_ewald_beta(0.9, 1e-5) == pytest.approx(3.47045919371208278, rel=1e-15)
    assert S.ewald_beta(0.9) == pytest.approx(3.47045919371208278, rel=1e-15)


def test_opc3_terms():
    """test/Driver/pme.test:20-23, kcal/mol (FileCheck matches the printed
    digits as a prefix, so each value holds to 1e-7)."""
    s = load_opc3()
    _, e = O.Oracle(s).total()
    assert e[O.E_COUL_DIRECT] / KCAL == pytest.approx(-14.3572978, abs=1e-7)
    assert e[O.E_COUL_EXCL] / KCAL == pytest.approx(535.1260682, abs=1e-7)
    assert e[O.E_COUL_SELF] / KCAL == pytest.approx(-539.7928271, abs=1e-7)
    # PME at 64^3 order 8 agrees with the Ewald sum to its six digits
    assert e[O.E_COUL_RECIP] / KCAL == pytest.approx(4.6972413, abs=5e-7)


def test_opc3_ewald_sum():
    """The Ewald sum of test/Driver/Inputs/ewald_reference.py: 4.6972413."""
    s = load_opc3()
    e = O.Oracle(s).ewald_reciprocal(forces=False)
    assert e / KCAL == pytest.approx(4.6972413, abs=1e-7)


def test_opc3_default_grid():
    """pme.test:17-18: 4.693822 kcal/mol with 35 points and order 4."""
    s = load_opc3(grid=(35, 35, 35), order=4)
    _, e = O.Oracle(s).total()
    assert e[O.E_COUL_RECIP] / KCAL == pytest.approx(4.693822, abs=5e-7)


def test_argon_lattice_energy():
    """test/Driver/argon.test:8-12: -5459.64 kJ/mol for the lattice of
    examples/argon/argon.mlir (switch 0.75-0.85 nm)."""
    a = S.argon_fcc()
    F, e = O.Oracle(a).total()
    assert e[O.E_LJ] == pytest.approx(-5459.64, abs=0.005)
    assert np.abs(F).max() < 1e-9


# test/Runtime/neighbors-matrix.mlir:111-160 and neighbors-matrix-gpu.mlir:148-186
MATRIX_CASES = [
    # count, edge, reach, cell width, row width, entries, checksum, largest
    (200, 6.0, 1.5, 1.5, 128, 2666, 27305838, 22),
    (200, 6.0, 1.5, 0.75, 128, 2666, 27305838, 22),
    (200, 6.0, 1.5, 0.5, 128, 2666, 27305838, 22),
    (200, 3.5, 1.5, 0.75, 128, 13056, 132165308, 80),
    (200, 3.5, 1.5, 1.5, 128, 13056, 132165308, 80),
    (200, 2.9, 1.4, 1.5, 128, 18476, 187655900, 107),
    (200, 5.0, 1.5, 1.5, 8, 1600, None, 33),
    (2000, 12.0, 1.5, 1.5, 128, 32704, None, 31),
    (2000, 12.0, 1.5, 0.75, 128, 32704, None, 31),
    (2000, 7.0, 1.0, 3.0, 128, 48778, None, 42),
    (2000, 5.8, 0.8, 3.0, 128, 43958, None, 40),
    (2000, 10.0, 1.5, 1.5, 8, 16000, None, 44),
]


@pytest.mark.parametrize("case", MATRIX_CASES)
def test_neighbor_matrix_golden(case):
    n, edge, reach, width, row, entries, checksum, largest = case
    x = lcg_positions(n, edge)
    counts, index, lg = O.neighbors_matrix(x, [edge] * 3, reach, width, row)
    assert int(counts.sum()) == entries
    assert lg == largest
    if checksum is not None:
        total = sum((i + 1) * (int(index[i, k]) + 1)
                    for i in range(n) for k in range(counts[i]))
        assert total == checksum


def test_neighbor_matrix_brute_force():
    """neighbors_reference.py method: all pairs in f64 within reach are in
    the rows, and every entry is within the margin."""
    n, edge, reach = 300, 4.0, 1.2
    x = lcg_positions(n, edge)
    counts, index, _ = O.neighbors_matrix(x, [edge] * 3, reach)
    d = x[:, None, :] - x[None, :, :]
    d -= edge * np.round(d / edge)
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    limit = reach + 3e-6 * 3 * edge
    for i in range(n):
        row = set(index[i, :counts[i]].tolist())
        assert set(np.nonzero(r[i] < reach)[0].tolist()) <= row
        assert all(r[i, j] < limit + 1e-6 for j in row)


def test_cell_width_choice():
    """@mdrt.cell_width picks the least-cost of reach, reach/2, reach/3."""
    w = O.lib().orc_cell_width(2000, np.array([12.0, 12.0, 12.0]).ctypes.data_as(
        O.PD), 1.5, 0.0)
    assert w in (1.5, 0.75, 0.5)

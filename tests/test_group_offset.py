"""The f32 arithmetic the pair loop relies on in boxes with group origins
(mdir_amd/csrc/md_common.h, group_offset): origins on a 1/64 nm grid, the
box edge split as its f32 value plus the rest.  (ogj - ogi) + k box is
exact in f32 whenever its magnitude stays below the edge, and the j
position it puts in the i group's frame is then within one rounding of the
exact value.  Checked here against f64 arithmetic, in which every one of
these quantities is exact."""
import numpy as np

F = np.float32


def origins(pos):
    """Lower corners on the 1/64 nm grid, as nb_sort computes them in f32."""
    return (np.floor(pos.astype(F) * F(64.0)) * F(0.015625)).astype(F)


def test_origin_difference_plus_shift_is_exact():
    rng = np.random.default_rng(11)
    for edge in (1.7, 2.5, 6.2, 8.2, 21.5, 100.3, 999.9, 12345.6):
        box = F(edge)
        lo = F(edge - float(box))
        n = 200_000
        # pairs of origins within cut-off + buffer + extents (2.5 nm here) of
        # each other across the boundary, the case of every list entry in
        # a dense system
        a = rng.uniform(-edge / 2, edge / 2, n)
        d = rng.uniform(-2.5, 2.5, n) if edge > 5.0 else rng.uniform(-edge / 2, edge / 2, n)
        b = a + d
        k = np.where(b >= edge / 2, -1.0, np.where(b < -edge / 2, 1.0, 0.0))
        b = b + k * edge                      # the j origin's image inside the box
        ogi, ogj = origins(a), origins(b)
        s32 = (ogj - ogi) + F(-k) * box       # the pair loop's order, in f32
        s64 = (ogj.astype(np.float64) - ogi.astype(np.float64)) - k * float(box)
        inside = np.abs(s64) < edge
        assert inside.mean() > 0.99
        assert np.array_equal(s32[inside].astype(np.float64), s64[inside]), edge
        # the box_lo part: one rounding on an offset of the order of the
        # cut-off, so the total is within half an ulp of |s| + |k lo|
        t32 = s32 + F(-k) * lo
        t64 = s64 - k * float(lo)
        err = np.abs(t32.astype(np.float64) - t64)
        ulp = np.spacing(np.abs(t32).astype(F)).astype(np.float64)
        assert (err[inside] <= 0.5 * ulp[inside] + 1e-30).all(), edge


def test_relative_positions_beat_box_centre_spacing():
    """At a 21.5 nm edge, a j atom placed in the i group's frame through the
    origins carries at most ~1.2e-7 nm of rounding per coordinate. About
    the box centre each position alone carries up to 4.8e-7 nm."""
    rng = np.random.default_rng(5)
    edge = 21.5
    n = 100_000
    xi = rng.uniform(-edge / 2, edge / 2 - 3.0, n)
    xj = xi + rng.uniform(-1.0, 1.0, n)
    ogi, ogj = origins(xi - 0.3), origins(xj - 0.3)  # lower corners below the atoms
    ri = (xi - ogi.astype(np.float64)).astype(F)
    rj = (xj - ogj.astype(np.float64)).astype(F)
    dx32 = ri - (rj + (ogj - ogi))
    err_origin = np.abs(dx32.astype(np.float64) - (xi - xj))
    dxc = xi.astype(F) - xj.astype(F)
    err_centre = np.abs(dxc.astype(np.float64) - (xi - xj))
    assert err_origin.max() < 4.0e-7
    assert np.sqrt((err_origin ** 2).mean()) < 0.5 * np.sqrt((err_centre ** 2).mean())

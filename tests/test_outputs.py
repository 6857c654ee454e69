"""Outputs without a GPU: the DCD oracle pinned against the record layout
the reference's own trajectory readers assume and against the published
format, the library's DCD writer against the oracle byte for byte, the
checkpoint container, and the report of a failed staged copy."""
import os
import struct
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import dcd as odcd  # noqa: E402


def frames_of(seed, F=3, n=7):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 4.0, size=(F, n, 3))
    boxes = np.array([[3.1 + 0.01 * k, 2.9, 3.3 - 0.02 * k] for k in range(F)])
    return x, boxes


def test_oracle_matches_the_layout_the_reference_readers_assume():
    """test/Driver/Inputs/dcd_frames.py and dcd_prefix.py of the reference
    read the header numbers at byte 8, skip the second record by its
    length, take the particle count 4 bytes into the third, and step over
    frames of (8 + 48) + 3 (8 + 4 n) bytes: the oracle's file must read
    back through exactly that arithmetic."""
    x, boxes = frames_of(1)
    data = odcd.write(x, boxes, first=10, period=10, timestep_ps=0.002)
    size, = struct.unpack_from("<i", data, 0)
    assert size == 84
    numbers = struct.unpack_from("<20i", data, 8)
    assert numbers[0] == 3 and numbers[1] == 10 and numbers[2] == 10 and numbers[3] == 30
    offset = 4 + size + 4
    size, = struct.unpack_from("<i", data, offset)
    offset += 4 + size + 4
    count, = struct.unpack_from("<i", data, offset + 4)
    offset += 12
    assert count == 7
    frame = (8 + 48) + 3 * (8 + 4 * count)
    assert len(data) == offset + 3 * frame


def read_published_layout(data):
    """A reader written from the CHARMM / NAMD DCD description (as VMD's
    dcdplugin documents it): 84-byte first record 'CORD' + 20 int32 (NSET,
    ISTART, NSAVC, ..., DELTA as f32 in word 10, the unit-cell flag in word
    11, the CHARMM version in word 20), the title block (NTITLE + 80-byte
    lines), NATOM, then per frame the 6 f64 unit cell (A, gamma, B, beta,
    alpha, C) when flagged and the X, Y, Z f32 records."""
    pos = 0

    def rec():
        nonlocal pos
        n, = struct.unpack_from("<i", data, pos)
        body = data[pos + 4: pos + 4 + n]
        m, = struct.unpack_from("<i", data, pos + 4 + n)
        assert m == n
        pos += n + 8
        return body

    h = rec()
    assert h[:4] == b"CORD"
    icntrl = struct.unpack("<20i", h[4:])
    delta = struct.unpack("<f", h[4 + 9 * 4: 4 + 10 * 4])[0]
    t = rec()
    ntitle, = struct.unpack_from("<i", t, 0)
    assert len(t) == 4 + 80 * ntitle
    natom, = struct.unpack("<i", rec())
    frames = []
    while pos < len(data):
        cell = struct.unpack("<6d", rec()) if icntrl[10] else None
        xyz = [np.frombuffer(rec(), "<f4") for _ in range(3)]
        frames.append((cell, np.stack(xyz, 1)))
    return icntrl, delta, t[4:84], natom, frames


def test_oracle_against_the_published_format():
    x, boxes = frames_of(2)
    data = odcd.write(x, boxes, first=20, period=5, timestep_ps=0.002)
    icntrl, delta, title, natom, frames = read_published_layout(data)
    assert icntrl[0] == 3 and icntrl[1] == 20 and icntrl[2] == 5 and icntrl[3] == 15
    assert icntrl[10] == 1 and icntrl[19] == 24
    # 2 fs in AKMA time units (48.88821 fs), rounded once to f32
    assert delta == np.float32(0.002 / 0.04888821)
    assert abs(delta * 48.88821 - 2.0) < 1e-6
    assert title.rstrip().startswith(b"Written by")
    assert natom == 7 and len(frames) == 3
    for k, (cell, xyz) in enumerate(frames):
        a, cg, b, cb, ca, c = cell
        assert (cg, cb, ca) == (0.0, 0.0, 0.0)  # right angles: cosines 0
        np.testing.assert_allclose([a, b, c], boxes[k] * 10.0, rtol=1e-15)
        # Angstrom, rounded once to f32
        np.testing.assert_array_equal(xyz, (x[k] * 10.0).astype(np.float32))


def test_oracle_scales_as_the_reference_does():
    """Positions: value x (1 / 0.1) then f32 (Output.cpp:1701 with the
    scale 1 / units::length); cell edges: edge / 0.1 (:1239).  A cell edge
    where edge / 0.1 and edge x 10 differ in f64 tells the two apart."""
    edge = next(e for e in np.linspace(2.0, 4.0, 2001) if e / 0.1 != e * 10.0)
    data = odcd.write(np.zeros((1, 1, 3)) + 0.1, np.array([[edge, edge, edge]]), 0, 1, 0.001)
    _, _, _, _, frames = read_published_layout(data)
    cell, xyz = frames[0]
    assert cell[0] == edge / 0.1 and cell[0] != edge * 10.0
    assert xyz[0, 0] == np.float32(1.0)


def test_oracle_continuation_keeps_and_cuts():
    x, boxes = frames_of(3, F=5)
    data = odcd.write(x, boxes, first=10, period=10, timestep_ps=0.002)
    cut = odcd.continued(data, 2, 0.002)
    assert cut == odcd.write(x[:2], boxes[:2], first=10, period=10, timestep_ps=0.002)
    # continuing with the rest gives the whole file again
    rest = b"".join(odcd.frame(x[k], boxes[k], True) for k in range(2, 5))
    whole = odcd.continued(data, 2, 0.002)[:0] + odcd.header(7, 5, 10, 10, 0.002, True) + \
        cut[len(odcd.header(7, 2, 10, 10, 0.002, True)):] + rest
    assert whole == data


def lib_or_skip():
    from mdir_amd import runtime as rt
    try:
        return rt.lib()
    except rt.LibraryMissing:
        pytest.skip("libmdrt_hip.so is not built")


def test_library_dcd_writer_matches_the_oracle(tmp_path):
    """mdrtHipDcdWrite (the trajectory writer's file code, called without a
    device) writes the oracle's bytes."""
    import ctypes as C
    from mdir_amd import runtime as rt
    L = lib_or_skip()
    for seed, periodic in ((4, True), (5, False)):
        x, boxes = frames_of(seed, F=4, n=11)
        path = str(tmp_path / f"t{seed}.dcd")
        xs = np.ascontiguousarray(x)
        bs = np.ascontiguousarray(boxes)
        rc = L.mdrtHipDcdWrite(path.encode(), 11, 4, 30, 15, 0.0025,
                               rt.ptr(xs, C.c_double), rt.ptr(bs, C.c_double) if periodic else None)
        assert rc == 0, L.mdrtHipLastError()
        expect = odcd.write(x, boxes, first=30, period=15, timestep_ps=0.0025, periodic=periodic)
        assert open(path, "rb").read() == expect


def test_library_dcd_reader_reads_the_oracle(tmp_path):
    from mdir_amd.output import read_dcd
    x, boxes = frames_of(6, F=2, n=5)
    path = tmp_path / "r.dcd"
    path.write_bytes(odcd.write(x, boxes, first=4, period=2, timestep_ps=0.002))
    numbers, titles, pos, cells = read_dcd(str(path))
    assert numbers[0] == 2 and numbers[1] == 4 and numbers[2] == 2
    assert titles[0].startswith("Written by")
    np.testing.assert_array_equal(pos, (x * 10.0).astype(np.float32))
    np.testing.assert_array_equal(cells[:, [0, 2, 5]], boxes / 0.1)


def test_checkpoint_container_round_trip(tmp_path):
    from mdir_amd.output import Checkpoint, read_checkpoint, write_checkpoint
    rng = np.random.default_rng(7)
    n = 9
    ck = Checkpoint(step=120, time=0.24, positions=rng.normal(size=(n, 3)),
                    velocities=rng.normal(size=(n, 3)), forces=rng.normal(size=(n, 3)),
                    masses=rng.uniform(1, 16, n), species=np.arange(n, dtype=np.int32) % 2,
                    box=np.array([3.0, 3.1, 3.2]), integrator="langevin-baoab",
                    timestep=0.002, seed=2**63 + 5, trajectory="run.dcd", frames=12,
                    temperature=300.0, friction=1.0, list_origin=rng.normal(size=(n, 3)),
                    prune_origin=None, list_kept=True, barostat_dv=0.5,
                    barostat_counts=np.array([10, 4, 0, 0], np.int64))
    path = str(tmp_path / "c.ckpt")
    write_checkpoint(path, ck)
    write_checkpoint(path, ck)
    assert os.path.exists(path + ".prev")
    back = read_checkpoint(path)
    for k in ("positions", "velocities", "forces", "masses", "species", "box", "list_origin",
              "barostat_counts"):
        np.testing.assert_array_equal(getattr(back, k), getattr(ck, k))
    assert back.prune_origin is None
    for k in ("step", "time", "integrator", "timestep", "seed", "trajectory", "frames",
              "temperature", "friction", "list_kept", "barostat_dv", "part"):
        assert getattr(back, k) == getattr(ck, k), k
    # the container holds no pickles: numpy reads it with pickles refused
    with np.load(path, allow_pickle=False) as z:
        assert "positions" in z.files


def test_failed_staged_copy_is_reported_and_aborts():
    """mdrtStagingFail reports a wait that failed on a writer's thread as
    any failed device call (mdrt_cuda.c:3820-3827): the message names the
    call and the frame, the failure hook runs, and the process aborts.
    Forced from the host with an error code; nothing touches a device."""
    from test_abi import run_isolated
    lib_or_skip()
    code = ("HOOK = C.CFUNCTYPE(None)(lambda: sys.stderr.write('hook ran\\n'))\n"
            "L.mdrtSetFailureHook(C.cast(HOOK, C.c_void_p))\n"
            "L.mdrtStagingFail(719, b'hipEventSynchronize')\n")
    rc, err = run_isolated(code)
    assert rc == -6, (rc, err)
    assert "mdrt: hipEventSynchronize of a staged frame" in err, err
    assert "hook ran" in err, err
    # a result of 0 is no failure
    rc, err = run_isolated("L.mdrtStagingFail(0, b'x')\nprint('alive')")
    assert rc == 0, (rc, err)

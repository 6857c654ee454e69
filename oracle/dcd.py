"""The reference's DCD writer restated in Python, byte for byte.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package
mdir_amd.

Follows lib/Driver/Trajectory.cpp of the reference:
  - records: int32 length, bytes, int32 length (writeRecord, :48-52);
  - header (writeHeader, :56-89): "CORD" + 20 int32 numbers, [0] frames,
    [1] first step, [2] period, [3] frames x period, [9] the f32 bits of
    timestep / 0.04888821, [10] 1 with a cell, [19] 24; a title record of
    one 80-character line padded with blanks; the particle count;
  - frame (writeFrame, :180-206): the cell record (a, cos gamma, b,
    cos beta, cos alpha, c; f64) when periodic, then the x, y, z records of
    f32, and the header rewritten with the new frame count;
  - continuation (append, :111-176): the first `keep` frames kept, the
    rest cut, the header rewritten with the step of the first frame the file
    holds.
Positions are scaled as lib/Driver/Output.cpp:1701 scales them, value x
(1 / 0.1) in f64 and then rounded to f32; the cell edges as :1239-1240,
edge / 0.1 in f64.  The title text is this writer's own (the reference
writes its program's name there).
"""
from __future__ import annotations

import struct

import numpy as np

TITLE = b"Written by mdir_amd"


def record(body: bytes) -> bytes:
    n = struct.pack("<i", len(body))
    return n + body + n


def header(n: int, frames: int, first: int, period: int, timestep_ps: float,
           periodic: bool) -> bytes:
    numbers = [0] * 20
    numbers[0] = frames
    numbers[1] = first
    numbers[2] = period
    numbers[3] = frames * period
    akma = np.float32(timestep_ps / 0.04888821)
    numbers[9] = struct.unpack("<i", akma.tobytes())[0]
    numbers[10] = 1 if periodic else 0
    numbers[19] = 24
    out = record(b"CORD" + struct.pack("<20i", *numbers))
    out += record(struct.pack("<i", 1) + TITLE.ljust(80, b" "))
    out += record(struct.pack("<i", n))
    return out


def frame(x_nm: np.ndarray, box_nm, periodic: bool) -> bytes:
    out = b""
    if periodic:
        a, b, c = (float(box_nm[k]) / 0.1 for k in range(3))
        # tilts 0: cos gamma = 0 / b, cos beta = 0 / c, cos alpha = (0 + b 0) / (b c)
        out += record(struct.pack("<6d", a, 0.0, b, 0.0, 0.0, c))
    scale = 1.0 / 0.1
    for k in range(3):
        out += record((scale * np.asarray(x_nm, np.float64)[:, k]).astype("<f4").tobytes())
    return out


def write(frames_nm, boxes_nm, first: int, period: int, timestep_ps: float,
          periodic: bool = True) -> bytes:
    """The bytes of a whole file of the given frames ((F, n, 3) nm)."""
    frames_nm = np.asarray(frames_nm, np.float64)
    F, n = frames_nm.shape[0], frames_nm.shape[1]
    out = header(n, F, first, period, timestep_ps, periodic)
    for k in range(F):
        out += frame(frames_nm[k], boxes_nm[k] if periodic else None, periodic)
    return out


def frame_bytes(n: int, periodic: bool) -> int:
    return (8 + 48 if periodic else 0) + 3 * (8 + 4 * n)


def continued(data: bytes, keep: int, timestep_ps: float, periodic: bool = True) -> bytes:
    """A file continued with its first `keep` frames kept (append, :111-176):
    the header rewritten with `keep` frames, the step of its first frame
    and its period, the frames after `keep` cut."""
    numbers = struct.unpack_from("<20i", data, 8)
    n = struct.unpack_from("<i", data, 8 + 84 + 8 + 84 + 4)[0]
    head = len(header(n, 0, 0, 1, timestep_ps, periodic))
    out = header(n, keep, numbers[1], numbers[2], timestep_ps, periodic)
    return out + data[head: head + keep * frame_bytes(n, periodic)]

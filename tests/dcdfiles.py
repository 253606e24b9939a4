"""DCD files of every flavour erdst's DCD plugin reads, written from scratch
for the trajectory tests: CHARMM or X-PLOR headers, 4- or 8-byte record
markers, either byte order, unit-cell blocks in degrees or cosines (or of an
unrecognised size), fixed atoms, a fourth coordinate record, any number of
title lines, a wrong NSET and truncated files.  The layout is the one
vmdplugins/dcdplugin.c reads (header :155-398, frames :400-626)."""
from __future__ import annotations

import struct

import numpy as np


def write_dcd(path, frames, cells=None, *, charmm=True, endian="<", rec64=False, ntitle=1, four_d=False,
              free=None, cell_block=48, nset=None, delta=0.002, istart=0, nsavc=1, tail=b""):
    """frames: list of [N, 3] arrays; cells: per frame (A, gamma, B, beta,
    alpha, C) or None for no unit-cell block; free: 1-based indices of the
    free atoms (the rest are fixed); tail: bytes appended after the frames."""
    e = endian
    frames = [np.asarray(f, np.float32) for f in frames]
    n = frames[0].shape[0]

    def rec(body: bytes) -> bytes:
        m = struct.pack(e + ("q" if rec64 else "i"), len(body))
        return m + body + m

    ctrl = [0] * 20
    ctrl[0] = len(frames) if nset is None else nset
    ctrl[1], ctrl[2] = istart, nsavc
    ctrl[8] = 0 if free is None else n - len(free)
    head = b"CORD"
    if charmm:
        ctrl[10] = 1 if cells is not None else 0
        ctrl[11] = 1 if four_d else 0
        ctrl[19] = 24
        body = b"".join(struct.pack(e + "i", c) for c in ctrl[:9]) + struct.pack(e + "f", delta) \
            + b"".join(struct.pack(e + "i", c) for c in ctrl[10:])
    else:
        body = b"".join(struct.pack(e + "i", c) for c in ctrl[:9]) + struct.pack(e + "d", delta) \
            + b"".join(struct.pack(e + "i", c) for c in ctrl[11:])
    out = [rec(head + body)]
    titles = b"".join(f"REMARK title line {k}".ljust(80).encode() for k in range(ntitle))
    out.append(rec(struct.pack(e + "i", ntitle) + titles))
    out.append(rec(struct.pack(e + "i", n)))
    if free is not None:
        out.append(rec(np.asarray(free, e + "i4").tobytes()))
    for k, f in enumerate(frames):
        if cells is not None:
            if cell_block == 48:
                out.append(rec(np.asarray(cells[k], e + "f8").tobytes()))
            else:
                out.append(rec(b"\0" * cell_block))
        sel = slice(None) if (free is None or k == 0) else np.asarray(free) - 1
        for ax in range(3):
            out.append(rec(np.ascontiguousarray(f[sel, ax]).astype(e + "f4").tobytes()))
        if four_d:
            out.append(rec(np.zeros(n if (free is None or k == 0) else len(free), e + "f4").tobytes()))
    data = b"".join(out) + tail
    with open(path, "wb") as fh:
        fh.write(data)
    return data


def random_frames(rng, nframe, natom, box=30.0):
    return [rng.uniform(-box, 2 * box, (natom, 3)).astype(np.float32) for _ in range(nframe)]

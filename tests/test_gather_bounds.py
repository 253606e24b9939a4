"""CPU model of the streaming brick gather's index bounds (k_gather_stream in
ermod-openacc_amd/csrc/er_kernels.h): for every atom count a brick can hold,
the passes of an item read only LDS record slots that its stage() copied,
only global records inside the brick, and stencil points inside the staged
region.  The constants are read from the kernel source, so the model follows
a change of them.  The model is also shown to catch the unsigned-wrap loop
bound of a half-width final pass tried once (it ran past the staged records
for 33-63 and 97-127 atoms)."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = open(os.path.join(ROOT, "ermod-openacc_amd", "csrc", "er_kernels.h")).read()


def define(name):
    return int(re.search(rf"#define {name} (\d+)", SRC).group(1))


RCAP, NT, BRICK = define("ER_GS_RCAP"), define("ER_GS_NT"), define("ER_BRICK")
ORD, NV, XL = 4, 4, 4
BZ = 6 if ORD <= 4 else 4
RS, RZ = BRICK + ORD - 1, BZ + ORD - 1
APASS = NT // (XL * NV)
RECU = 4                       # float4 per 64-byte record
NREG = RS * RZ * RS * NV       # float4 of the staged region


def staged_records(a0, a1):
    """Record slots stage() copies: u0 + lane < nrec * RECU."""
    return min(a1 - a0, RCAP) * RECU


def passes(a0, a1, half_pass_bug=False):
    """(base, pos, from_lds, bound) of every live lane of every pass of
    compute(): whole passes over the staged records [a0, split), then over
    the rest [split, a1) from global memory.  half_pass_bug: the dropped
    variant's first loop, `for (base = a0; split - base > APASS / 2; base +=
    APASS)` in unsigned arithmetic, followed by a half-width pass."""
    split = min(a1, a0 + RCAP)
    out = []
    base = a0
    if half_pass_bug:
        guard = 0
        while ((split - base) & 0xffffffff) > APASS // 2 and guard < 8 * RCAP // APASS:
            out += [(base, base + s, True, split) for s in range(APASS) if base + s < split]
            out.append((base, None, True, split))
            base += APASS
            guard += 1
        out += [(base, base + s, True, split) for s in range(APASS // 2) if base + s < split]
    else:
        while base < split:
            out += [(base, base + s, True, split) for s in range(APASS) if base + s < split]
            out.append((base, None, True, split))
            base += APASS
    base = split
    while base < a1:
        out += [(base, base + s, False, a1) for s in range(APASS) if base + s < a1]
        out.append((base, None, False, a1))
        base += APASS
    return out


def violations(a0, a1, **kw):
    """Record reads outside what the item holds, and passes started at or
    beyond their bound (a loop running past the item's end)."""
    bad = 0
    nst = staged_records(a0, a1)
    for base, pos, lds, bound in passes(a0, a1, **kw):
        if pos is None:
            bad += not (base < bound)
        elif lds:
            bad += not (0 <= (pos - a0) * RECU and (pos - a0) * RECU + RECU <= nst)
        else:
            bad += not (a0 <= pos < a1)
    return bad


def test_pass_bounds_for_every_atom_count():
    assert APASS == 64 and RCAP % APASS == 0
    for n in range(0, 4 * RCAP + 3):
        for a0 in (0, 17, 1000):
            assert violations(a0, a0 + n) == 0, n
            seen = sorted(p for _, p, _, _ in passes(a0, a0 + n) if p is not None)
            assert seen == list(range(a0, a0 + n)), n          # every atom exactly once


def test_model_catches_the_unsigned_wrap_bound():
    bad = [n for n in range(0, 2 * RCAP) if violations(0, n, half_pass_bug=True)]
    assert bad and all(33 <= n % 64 <= 63 for n in bad)


def test_stencil_reads_stay_in_the_staged_region():
    """Records hold the stencil corner lx, ly, lz = (grid index - brick
    origin) + ORD - 1 with the grid index inside the brick; lane (cx, j4)
    reads region point (lz - c3, ly - c2, lx - cx) for c2, c3 < ORD."""
    idx = []
    for lz in range(ORD - 1, ORD - 1 + BZ):
        for ly in range(ORD - 1, ORD - 1 + BRICK):
            for lx in range(ORD - 1, ORD - 1 + BRICK):
                for cx in range(XL):
                    for c3 in range(ORD):
                        for c2 in range(ORD):
                            for j4 in range(NV):
                                b0 = (((lz - (ORD - 1)) * RS + (ly - (ORD - 1))) * RS + (lx - cx)) * NV + j4
                                idx.append(b0 + ((ORD - 1 - c3) * RS + (ORD - 1 - c2)) * RS * NV)
    idx = np.array(idx)
    assert idx.min() >= 0 and idx.max() < NREG
    # the corner fields fit their 8-bit slots
    assert ORD - 1 + BRICK - 1 < 256 and ORD - 1 + BZ - 1 < 256

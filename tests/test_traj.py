"""The native DCD reader (include/er_traj.h) against the reference's own DCD
plugin, built from vmdplugins/dcdplugin.c into oracle/_ref by oracle/ref.mk
with the cell construction of erdst/vmdfio.c:355-378: every flavour must give
bit-identical coordinates and cells and the same number of frames.  Where
oracle/_ref is absent (no reference sources), the same files are checked
against what was written and against the closed form of the cell."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from dcdfiles import random_frames, write_dcd
from erdst_gpu import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libdcdref.so")


def plugin_read(path, maxframes=64, cap=4096):
    """Frames read by the reference plugin: (xyz [F, N, 3], cells [F, 3, 3]) or None if it will not open."""
    if not os.path.exists(REF_LIB):
        pytest.skip("reference DCD plugin not built (oracle/_ref)")
    L = C.CDLL(REF_LIB)
    L.ref_dcd_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    xyz = np.zeros((maxframes, cap, 3), np.float32)
    cell = np.zeros((maxframes, 9), np.float32)
    n = C.c_int()
    nf = L.ref_dcd_read(os.fsencode(str(path)), maxframes, cap, xyz.ctypes.data, cell.ctypes.data, C.byref(n))
    if nf == -1:
        return None
    assert nf >= 0
    na = n.value
    flat = xyz.reshape(-1)[: maxframes * na * 3].reshape(maxframes, na, 3)
    return flat[:nf].copy(), np.stack([c.reshape((3, 3), order="F") for c in cell[:nf]]) if nf else np.zeros((0, 3, 3))


def native_read(path):
    try:
        t = capi.Trajectory(str(path))
    except capi.TrajError:
        return None
    with t:
        frames = list(t)
    if not frames:
        return np.zeros((0, t.natoms, 3), np.float32), np.zeros((0, 3, 3), np.float32)
    return np.stack([f[0] for f in frames]), np.stack([f[1] for f in frames])


def cell_closed_form(A, B, C_, alpha, beta, gamma):
    """Cell vectors from lengths and angles in degrees (textbook form)."""
    ca, cb, cg = (math.cos(math.radians(x)) for x in (alpha, beta, gamma))
    sg = math.sin(math.radians(gamma))
    v = (ca - cg * cb) / sg
    return np.array([[A, B * cg, C_ * cb], [0, B * sg, C_ * v], [0, 0, C_ * math.sqrt(1 - cb * cb - v * v)]])


FLAVOURS = {
    "charmm_deg": dict(cells="ortho_deg"),
    "charmm_cos": dict(cells="tri_cos"),
    "charmm_tri_deg": dict(cells="tri_deg"),
    "xplor": dict(charmm=False, cells=None),
    "big_endian": dict(endian=">", cells="tri_deg"),
    "rec64": dict(rec64=True, cells="ortho_deg"),
    "rec64_big": dict(rec64=True, endian=">", cells="tri_cos"),
    "fixed_atoms": dict(cells="ortho_deg", free="some"),
    "fixed_atoms_big": dict(cells="tri_deg", free="some", endian=">"),
    "four_d": dict(cells="ortho_deg", four_d=True),
    "no_titles": dict(cells="ortho_deg", ntitle=0),
    "three_titles": dict(charmm=False, cells=None, ntitle=3),
    "nset_wrong": dict(cells="ortho_deg", nset=100),
    "nset_zero": dict(cells="tri_deg", nset=0),
    "truncated": dict(cells="ortho_deg", tail=b"\x30\x00\x00\x00" + b"\x00" * 20),
    "odd_cell_block": dict(cells="ortho_deg", cell_block=40),
    "charmm_no_cell": dict(cells=None),
}


def _cells(kind, nframe, rng):
    if kind is None:
        return None
    out = []
    for k in range(nframe):
        A, B, C_ = rng.uniform(20, 60, 3)
        if kind == "ortho_deg":
            out.append((A, 90.0, B, 90.0, 90.0, C_))
        elif kind == "tri_deg":
            out.append((A, 75.0 + k, B, 80.0, 85.5, C_))
        else:  # cosines, as CHARMM and NAMD > 2.5 write them
            out.append((A, 0.1, B, -0.2 + 0.01 * k, 0.3, C_))
    return out


def _write(tmp_path, name, nframe=4, natom=37, seed=0):
    spec = dict(FLAVOURS[name])
    rng = np.random.default_rng(seed)
    frames = random_frames(rng, nframe, natom)
    cells = _cells(spec.pop("cells"), nframe, rng)
    if spec.get("free") == "some":
        spec["free"] = np.sort(rng.choice(np.arange(1, natom + 1), natom - 9, replace=False))
    path = tmp_path / f"{name}.dcd"
    write_dcd(path, frames, cells, **spec)
    return path, frames, cells, spec


@pytest.mark.parametrize("name", sorted(FLAVOURS))
def test_native_reader_equals_reference_plugin(tmp_path, name):
    path, frames, cells, spec = _write(tmp_path, name)
    ours = native_read(path)
    ref = plugin_read(path)
    assert ours is not None and ref is not None
    assert ours[0].shape == ref[0].shape, (ours[0].shape, ref[0].shape)
    np.testing.assert_array_equal(ours[0].view(np.uint32), ref[0].view(np.uint32))
    np.testing.assert_array_equal(ours[1].view(np.uint32), ref[1].view(np.uint32))


@pytest.mark.parametrize("name", sorted(FLAVOURS))
def test_native_reader_returns_what_was_written(tmp_path, name):
    path, frames, cells, spec = _write(tmp_path, name)
    xyz, cell = native_read(path)
    # the frame count comes from the file size with 48-byte cell blocks
    # assumed (dcdplugin.c:838-851): 4 frames of 37 atoms with 40-byte blocks
    # fill 4 x 516 bytes, which that count reads as 1 + 1540 // 524 = 3
    nexp = len(frames) - 1 if name == "odd_cell_block" else len(frames)
    assert xyz.shape[0] == nexp
    free = spec.get("free")
    for k, f in enumerate(frames[:nexp]):
        want = f.copy()
        if free is not None and k > 0:       # fixed atoms keep their first-frame positions
            fixed = np.setdiff1d(np.arange(f.shape[0]), np.asarray(free) - 1)
            want[fixed] = frames[0][fixed]
        np.testing.assert_array_equal(xyz[k], want)
        if cells is None or spec.get("cell_block", 48) != 48:
            lengths, angles = (1.0, 1.0, 1.0), (90.0, 90.0, 90.0)
        else:
            A, g, B, b, a, C_ = cells[k]
            if max(abs(g), abs(b), abs(a)) <= 1.0:
                a, b, g = (math.degrees(math.acos(x)) for x in (a, b, g))
            lengths, angles = (A, B, C_), (a, b, g)
        np.testing.assert_allclose(cell[k], cell_closed_form(*lengths, *angles), rtol=2e-6, atol=1e-5)
    info = capi.Trajectory(str(path)).info()
    assert info["nframes"] == nexp
    assert info["charmm"] == spec.get("charmm", True)
    assert info["rec64"] == spec.get("rec64", False)
    assert info["swapped"] == (spec.get("endian", "<") == ">")
    assert info["nfixed"] == (0 if free is None else xyz.shape[1] - len(free))


def test_orthorhombic_cell_is_diagonal(tmp_path):
    path, frames, cells, _ = _write(tmp_path, "charmm_deg")
    _, cell = native_read(path)
    for k, (A, _, B, _, _, C_) in enumerate(cells):
        np.testing.assert_array_equal(np.diag(cell[k]), np.float32([A, B, C_]))
        off = cell[k][~np.eye(3, dtype=bool)]
        assert np.all(np.abs(off) < 1e-5 * max(A, B, C_))


def test_big_endian_four_d(tmp_path):
    """The plugin compares the 4-D flag without swapping it (dcdplugin.c:239)
    and so misreads opposite-endian 4-D files; the native reader swaps it."""
    rng = np.random.default_rng(3)
    frames = random_frames(rng, 3, 11)
    path = tmp_path / "b4.dcd"
    write_dcd(path, frames, _cells("ortho_deg", 3, rng), endian=">", four_d=True)
    xyz, _ = native_read(path)
    assert xyz.shape[0] == 3
    for k in range(3):
        np.testing.assert_array_equal(xyz[k], frames[k])


def test_not_a_dcd_and_empty_files(tmp_path):
    bad = tmp_path / "bad.dcd"
    bad.write_bytes(b"\x54\x00\x00\x00CORX" + b"\0" * 200)
    with pytest.raises(capi.TrajError) as e:
        capi.Trajectory(str(bad))
    assert e.value.status == capi.ER_TRAJ_FORMAT
    # a header with no complete frame fails to open, as the plugin's open does
    rng = np.random.default_rng(5)
    data = write_dcd(tmp_path / "one.dcd", random_frames(rng, 1, 7), None)
    cut = tmp_path / "cut.dcd"
    cut.write_bytes(data[:-5])
    with pytest.raises(capi.TrajError):
        capi.Trajectory(str(cut))
    if os.path.exists(REF_LIB):
        assert plugin_read(cut) is None
    with pytest.raises(capi.TrajError) as e:
        capi.Trajectory(str(tmp_path / "missing.dcd"))
    assert e.value.status == capi.ER_TRAJ_IO


def test_atom_count_must_match(tmp_path):
    path, *_ = _write(tmp_path, "charmm_deg")
    with capi.Trajectory(str(path)) as t:
        with pytest.raises(capi.TrajError) as e:
            t.read(natoms=t.natoms + 1)
        assert e.value.status == capi.ER_E_ARG
        assert t.read() is not None

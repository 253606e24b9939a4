"""Host-side logic and the C ABI library, without a GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from erdst_gpu import capi, synth, topology
from erdst_gpu.params import hot_path_params, temperature_kt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(name):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(er_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)          # loads without touching a device
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["er_capi.h", "er_traj.h"]
    names = header_functions("er_capi.h")
    assert "er_run" in names and "er_create" in names
    for n in names + header_functions("er_traj.h"):
        assert hasattr(lib, n), n
    assert sorted(capi.EXPORTED) == names
    assert sorted(capi.EXPORTED_TRAJ) == header_functions("er_traj.h")


def test_water_and_solute_geometry():
    g = synth.water_geometry()
    r1, r2 = np.linalg.norm(g[1] - g[0]), np.linalg.norm(g[2] - g[0])
    ang = np.degrees(np.arccos(np.dot(g[1] - g[0], g[2] - g[0]) / (r1 * r2)))
    assert abs(r1 - 0.9572) < 1e-12 and abs(r2 - 0.9572) < 1e-12 and abs(ang - 104.52) < 1e-9
    t = synth.tetra_geometry()
    assert np.allclose(np.linalg.norm(t[1:], axis=1), 1.09)
    assert np.allclose(t.sum(0), 0.0)


def test_rotations_are_proper_and_rigid_insertions_keep_shape():
    rng = np.random.default_rng(0)
    R = synth.random_rotations(rng, 100)
    assert np.allclose(np.einsum("nij,nkj->nik", R, R), np.eye(3), atol=1e-12)
    assert np.allclose(np.linalg.det(R), 1.0)
    w = synth.config0()
    ins = w.insertions(3).astype(np.float64)
    d = np.linalg.norm(ins[:, 1:] - ins[:, :1], axis=2)
    assert ins.shape == (100, 5, 3) and np.allclose(d, 1.09, atol=1e-5)


def test_frames_are_seeded_and_distinct():
    w = synth.config0()
    a, b = w.frame(2), synth.config0().frame(2)
    assert a.dtype == np.float32 and a.shape == (3000, 3)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(w.frame(2), w.frame(3))


def test_topology_matches_setparam():
    w = synth.config0()
    t = w.topo
    assert (t.natom, t.nmol, t.numslv) == (3000, 1000, 1)
    # LJ types in order of first appearance: O, H(water), C, H(solute)
    assert t.ntype == 4
    np.testing.assert_array_equal(t.ljtype[:3], [0, 1, 1])
    np.testing.assert_array_equal(t.slt_ljtype, [2, 3, 3, 3, 3])
    ene = t.ljene64.reshape(4, 4, order="F")
    lsq = t.ljlensq64.reshape(4, 4, order="F")
    np.testing.assert_allclose(ene[0, 2], np.sqrt(np.float32(0.152) * np.float32(0.11)), rtol=1e-7)
    np.testing.assert_allclose(lsq[0, 2], ((np.float32(3.151) + np.float32(3.4)) / 2) ** 2, rtol=1e-7)
    np.testing.assert_allclose(t.charge32[:3], np.float32(18.22261721) * np.float32([-0.834, 0.417, 0.417]))
    assert t.mol_begin[-1] == t.natom and np.all(np.diff(t.mol_begin) == 3)


def test_soln_topology_uvspec():
    w = synth.config1()
    t = w.topo
    assert t.tagslt.tolist() == [0]
    assert t.uvspec[0] == 0 and np.all(t.uvspec[1:] == 1) and t.numslv == 1
    # soln with several solute molecules: the solute is also a solvent species
    sp = [synth.water_species(4), synth.tetra_species(), synth.water_species(3)]
    sp[1].nmol = 2
    t2 = topology.build_topology(sp, solute_index=1, refs=False)
    assert t2.uvspec.tolist() == [1] * 4 + [2] * 2 + [3] * 3 and t2.tagslt.tolist() == [4, 5]
    # a single solute in the middle: later species slide down (engproc.F90:112-118)
    sp[1].nmol = 1
    t3 = topology.build_topology(sp, solute_index=1, refs=False)
    assert t3.uvspec.tolist() == [1] * 4 + [0] + [2] * 3 and t3.numslv == 2


def test_rmin_half_format_converts_sigma():
    sp = [synth.water_species(2), synth.tetra_species()]
    t = topology.build_topology(sp, 1, True, ljformat=topology.LJFMT_EPS_RMINH)
    lsq = t.ljlensq32.reshape(t.ntype, t.ntype, order="F")
    s0 = np.float32(topology.SGMCNV) * np.float32(3.151)
    assert np.isclose(lsq[0, 0], s0 * s0, rtol=1e-6)


def test_parameter_defaults():
    p = hot_path_params(3, (72, 72, 72))
    assert (p["upljcut"], p["lwljcut"], p["ljswitch"], p["cltype"]) == (12.0, 10.0, 0, 2)
    assert (p["corrcal"], p["wgtslf"], p["estype"], p["selfcal"]) == (1, 1, 2, 0)
    s = hot_path_params(1, (64, 64, 64))
    assert (s["corrcal"], s["wgtslf"]) == (0, 0)
    # temp = inptemp * 8.314510e-3 / 4.184 with single-precision literals
    assert abs(temperature_kt(300.0) / (300 * 8.314510e-3 / 4.184) - 1) < 1e-7
    assert temperature_kt(300.0) != temperature_kt(300.0, double_build=True)


def test_meshes():
    w = synth.config2()
    um, us, uc = w.uvmax_uvsoft_uvcrd()
    assert um.tolist() == [600] and us.tolist() == [500]
    assert np.all(np.diff(uc) > 0) and uc[0] == -20.0 and abs(uc[-1] - 1e4) < 1e-9
    w1 = synth.config1()
    um, us, uc = w1.uvmax_uvsoft_uvcrd()
    assert um.tolist() == [1200] and us.tolist() == [1200]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_context_without_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(capi, "LIB_PATH", "/nonexistent/liberdst_hip.so")
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(ImportError):
        capi.load()


def test_fortran_host_binary_loads():
    host = os.path.join(os.path.dirname(capi.LIB_PATH), "erdst_host")
    r = subprocess.run([host], capture_output=True, text=True, timeout=60)
    assert "usage: erdst_host" in (r.stdout + r.stderr)


def test_multi_solute_topologies_share_one_lj_table():
    from erdst_gpu import synth
    w3 = synth.config3()
    topos = [w3.topo_for(k) for k in (0, 5, 6, 15)]
    for t in topos[1:]:
        assert t.ntype == topos[0].ntype
        np.testing.assert_array_equal(t.ljene32, topos[0].ljene32)
        np.testing.assert_array_equal(t.ljlensq32, topos[0].ljlensq32)
        np.testing.assert_array_equal(t.ljtype, topos[0].ljtype)
    # each solute's own sites keep their own (sigma, epsilon)
    for k, t in zip((0, 5, 6, 15), topos):
        sp = w3.solutes[k]
        sig = np.sqrt(t.ljlensq32.reshape(t.ntype, t.ntype)[t.slt_ljtype, t.slt_ljtype])
        np.testing.assert_allclose(sig, sp.sigma.astype(np.float32), rtol=1e-6)


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of er_topology and er_params have the C header's
    size and field offsets (force_calculation included)."""
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "er_capi.h"\n'
                   "int main(void) { printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(er_topology), sizeof(er_params),"
                   " offsetof(er_params, temp), offsetof(er_params, force_calculation),"
                   " offsetof(er_topology, ljlensq_mat)); return 0; }\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(capi.ErTopology), ctypes.sizeof(capi.ErParams), capi.ErParams.temp.offset,
            capi.ErParams.force_calculation.offset, capi.ErTopology.ljlensq_mat.offset]
    assert got == want

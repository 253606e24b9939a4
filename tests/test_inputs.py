"""erdst's input files (erdst_gpu/inputs.py, after erdst/setconf.F90): the
topology built from MDinfo / MolPrm / SltInfo files equals the one the
GPU tests build programmatically, in every array and bit for bit; the six
LJ conventions convert back to the same epsilon and sigma; LJTable, getmass,
PermIndex, the weight files, the &ene_param defaults and getscrn."""
import math

import numpy as np
import pytest

from erdst_gpu import inputs as I
from erdst_gpu import synth
from erdst_gpu.topology import build_topology

f32 = np.float32


def fmt(x):
    return np.format_float_positional(f32(x), unique=True, trim="0") if abs(float(x)) >= 1e-4 or x == 0 \
        else np.format_float_scientific(f32(x), unique=True)


def write_species_file(path, sp, new_format=True, rigid=False, types=None, lj=None):
    """One line per site; lj = (column2, column3) overrides (eps, sigma)."""
    c2, c3 = lj if lj is not None else (sp.eps, sp.sigma)
    with open(path, "w") as fh:
        for i in range(sp.nsite):
            typ = (types or ["OW"] * sp.nsite)[i]
            cols = [fmt(sp.charge[i]), fmt(c2[i]), fmt(c3[i])]
            if rigid:
                cols += [fmt(v) for v in sp.geometry[i]]
            if new_format:
                fh.write(f"{i + 1} 15.9994 {typ} {typ}{i + 1} " + " ".join(cols) + "\n")
            else:
                fh.write(f"{i + 1:5d}  {typ:6s}" + "  ".join(cols) + "\n")


def write_config0_inputs(d, new_solvent=True, new_solute=False):
    w = synth.config0()
    topo = w.topo
    water, solute = topo.species[0], topo.species[topo.solute_index]
    (d / "MDinfo").write_text(f"10 1\n{water.nmol}\n{water.nsite}\n")
    write_species_file(d / "MolPrm1", water, new_format=new_solvent, types=["OW", "HW", "HW"])
    rigid = w.slttype == I.SLT_REFS_RIGID
    write_species_file(d / "SltInfo", solute, new_format=new_solute, rigid=rigid,
                       types=["C"] + ["H"] * (solute.nsite - 1))
    return w


@pytest.mark.parametrize("new_solvent,new_solute", [(True, False), (False, True), (True, True)])
def test_files_reproduce_programmatic_topology(tmp_path, new_solvent, new_solute):
    w = write_config0_inputs(tmp_path, new_solvent, new_solute)
    params = dict(slttype=w.slttype, ljformat=0, cmbrule=0)
    sysin = I.load_system(tmp_path, params)
    a, b = sysin.topology, w.topo
    assert (a.natom, a.nmol, a.ntype, a.numslv) == (b.natom, b.nmol, b.ntype, b.numslv)
    for name in ("mol_begin", "ljtype", "charge32", "charge64", "uvspec", "ljene32", "ljlensq32", "ljene64",
                 "ljlensq64", "slt_ljtype", "slt_charge32", "slt_charge64"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)
    assert sysin.mdinfo.nrun == 10
    # masses: new format as written, old format from getmass
    water_mass = f32(15.9994) if new_solvent else np.array([I.getmass("OW"), I.getmass("HW"), I.getmass("HW")])
    np.testing.assert_array_equal(sysin.mass[:3], np.broadcast_to(water_mass, 3))
    if w.slttype == I.SLT_REFS_RIGID:
        np.testing.assert_array_equal(sysin.bfcoord, np.asarray(b.species[b.solute_index].geometry, np.float32))


@pytest.mark.parametrize("ljformat", [1, 2, 3, 4])
def test_lj_conventions_convert_to_epsilon_sigma(tmp_path, ljformat):
    rng = np.random.default_rng(ljformat)
    eps = rng.uniform(0.01, 0.3, 12)
    sig = rng.uniform(2.0, 4.0, 12)
    eps[3] = 0.0                                   # a site without LJ
    if ljformat == 1:
        c2, c3 = eps, sig / float(I.SGMCNV)        # Rmin/2
    elif ljformat == 2:
        c2, c3 = eps * 4.184, sig / 10.0           # kJ/mol, nm
    elif ljformat == 3:
        c2, c3 = 4 * eps * sig ** 12, 4 * eps * sig ** 6
    else:
        c2, c3 = 4 * eps * 4.184 * (sig / 10) ** 12, 4 * eps * 4.184 * (sig / 10) ** 6
    e, s = I.convert_lj(ljformat, np.array(c2, np.float32), np.array(c3, np.float32))
    live = eps > 0
    np.testing.assert_allclose(e, eps, rtol=2e-6)
    np.testing.assert_allclose(s[live], sig[live], rtol=2e-6)
    if ljformat in (3, 4):
        assert e[3] == 0.0                         # C = 0: epsilon 0 (setconf.F90:738-740)


def test_ljtable_format(tmp_path):
    w = synth.config0()
    topo = w.topo
    water, solute = topo.species[0], topo.species[topo.solute_index]
    (tmp_path / "MDinfo").write_text(f"1 1\n{water.nmol}\n{water.nsite}\n")
    # ljformat 5: the second LJ column holds the type number
    write_species_file(tmp_path / "MolPrm1", water, lj=([1, 2, 2], [0, 0, 0]))
    write_species_file(tmp_path / "SltInfo", solute, lj=([3] + [2] * (solute.nsite - 1), [0] * solute.nsite),
                       rigid=w.slttype == I.SLT_REFS_RIGID)
    sig = np.array([[3.1, 2.0, 3.3], [2.0, 1.0, 2.2], [3.3, 2.2, 3.4]])
    eps = np.array([[0.15, 0.0, 0.12], [0.0, 0.0, 0.0], [0.12, 0.0, 0.11]])
    rows = "\n".join(" ".join(str(v) for v in r) for r in sig) + "\n" + "\n".join(
        ", ".join(str(v) for v in r) for r in eps)
    (tmp_path / "LJTable").write_text(f"3\n{rows}\n")
    t = I.load_system(tmp_path, dict(slttype=w.slttype, ljformat=5)).topology
    assert t.ntype == 3
    np.testing.assert_array_equal(t.ljtype[:3], [0, 1, 1])
    np.testing.assert_array_equal(t.slt_ljtype, [2] + [1] * (solute.nsite - 1))
    np.testing.assert_array_equal(t.ljlensq32.reshape(3, 3, order="F"), (sig.astype(np.float32) ** 2))
    np.testing.assert_array_equal(t.ljene32.reshape(3, 3, order="F"), eps.astype(np.float32))


def test_getmass_table():
    C, H = f32(12.0107), f32(1.00794)
    assert I.getmass("CH3") == C + f32(3.0) * H
    assert I.getmass("CH2X") == C + f32(2.0) * H
    assert I.getmass("CH") == C + H
    assert I.getmass("Cl-") == f32(35.4527) and I.getmass("C1") == C
    assert I.getmass("OW") == f32(15.9994) and I.getmass("Na+") == f32(22.989770)
    assert I.getmass("MW") == 0.0 and I.getmass("D") == f32(2.014102)
    with pytest.raises(I.InputError):
        I.getmass("Xe")


def test_mdinfo_list_directed(tmp_path):
    (tmp_path / "MDinfo").write_text("  500, 3   ! frames, species\n2*1000\n 40\n3, 3\n  1\n")
    md = I.read_mdinfo(tmp_path / "MDinfo")
    assert (md.nrun, md.ntype, md.nmol, md.nsite) == (500, 3, [1000, 1000, 40], [3, 3, 1])


def test_list_directed_rules():
    r = I.ListDirected(["1, , 3 /  9", "4 5", "'a b', 7"])
    assert r.read([int, int, int, int], prev=[0, 8, 0, 6]) == [1, 8, 3, 6]     # null keeps, '/' ends
    assert r.read([float, float]) == [4.0, 5.0]
    assert r.read([str, int]) == ["a b", 7]
    with pytest.raises(EOFError):
        r.read([int])
    with pytest.raises(ValueError):
        I.ListDirected(["OW"]).read([float])
    assert I.ListDirected(["1.5d-3"]).read([float]) == [1.5e-3]


def test_permindex(tmp_path):
    (tmp_path / "PermIndex").write_text("1 3\n3 1\n")
    perm = I.read_permindex(tmp_path / "PermIndex", 4)
    np.testing.assert_array_equal(perm, [2, 1, 0, 3])
    xyz = np.arange(12, dtype=np.float32).reshape(4, 3)
    out = I.apply_permutation(xyz, perm)
    np.testing.assert_array_equal(out[2], xyz[0])     # trajectory atom 1 -> program atom 3
    np.testing.assert_array_equal(out[0], xyz[2])
    (tmp_path / "Bad").write_text("1 3\n")                 # atom 3 twice, atom 1 never
    with pytest.raises(I.InputError, match="set_pmt"):
        I.read_permindex(tmp_path / "Bad", 4)
    (tmp_path / "Bad2").write_text("5 1\n")
    with pytest.raises(I.InputError, match="set_pmt"):
        I.read_permindex(tmp_path / "Bad2", 4)


def test_weight_files(tmp_path):
    (tmp_path / "W").write_text("1 0.5\n2 1.5\n")
    sw = I.WeightFile(tmp_path / "W", rewind=False)
    assert [sw.next(), sw.next()] == [f32(0.5), f32(1.5)]
    with pytest.raises(I.InputError):
        sw.next()
    st = I.WeightFile(tmp_path / "W", rewind=True)
    assert [st.next() for _ in range(5)] == [f32(0.5), f32(1.5), f32(0.5), f32(1.5), f32(0.5)]


PARAMS = """!  parameters for a refs run
&ene_param
      iseed = 123,        ! seed
      slttype = 3,
      boxshp = 1,  estype = 2,
      inptemp = 298.15,
      elecut = 10.0,                                        ! coulomb cutoff
      cltype = 2, screen = 0.35,  splodr = 4, ms1max = 96,  ! PME
      ms3max = 80,
      hostspec(2) = 1,
      engdiv = 10
/
&hist
      eclbin = 5.0e-2, ecfbin = 2.0e-3, ec0bin = 2.0e-4, finfac = 10.0,
      ecdmin = -80.0,  ecfmns = -0.20,  ecdcen = 0.0,    eccore = 20.0,
      ecdmax = 1.0e11, pecore = 200
/
"""


def test_parameters_and_iniparam_defaults(tmp_path):
    (tmp_path / "parameters_er").write_text(PARAMS)
    nml = I.read_namelist(tmp_path / "parameters_er", "ene_param")
    p = I.iniparam(nml)
    assert p["slttype"] == 3 and p["engdiv"] == 10 and p["splodr"] == 4
    assert (p["ms1max"], p["ms2max"], p["ms3max"]) == (96, 96, 80)   # ms2max follows ms1max; ms3max explicit
    assert p["corrcal"] == 1 and p["wgtslf"] == 1 and p["maxins"] == 1000
    assert p["elecut"] == 10.0 and p["upljcut"] == 12.0 and p["lwljcut"] == 10.0
    assert p["hostspec"] == {2: 1}
    assert p["temp"] == pytest.approx(298.15 * 8.314510e-3 / 4.184, rel=1e-6)
    hist = I.read_namelist(tmp_path / "parameters_er", "hist")
    assert float(hist["ecdmax"]) == 1.0e11 and int(hist["pecore"]) == 200
    soln = I.iniparam({"slttype": "1"})
    assert soln["corrcal"] == 0 and soln["wgtslf"] == 0 and soln["maxins"] == 1 and soln["splodr"] == 6
    assert soln["screen"] > 0.0                      # from getscrn with the default ewtoler
    with pytest.raises(I.InputError, match="set_prs"):
        I.iniparam({"slttype": "1", "cltype": "1"})
    with pytest.raises(I.InputError, match="set_ins"):
        I.iniparam({"slttype": "1", "wgtins": "1"})


@pytest.mark.parametrize("tol,cut,kind", [(1e-5, 12.0, "relative"), (1e-6, 10.0, "relative"),
                                          (1e-4, 9.0, "distance")])
def test_getscrn_meets_its_tolerance(tol, cut, kind):
    s = I.getscrn(tol, cut, kind)
    val = math.erfc(s * cut) / (cut if kind == "distance" else 1.0)
    assert abs(val - tol) <= 1e-15 + 1e-12 * tol
    assert 0.1 < s < 1.0


def test_soln_solute_species_and_molprm_numbering(tmp_path):
    # three species; the second is the solute with two molecules: MolPrm1 for
    # species 1, SltInfo for species 2 (which still advances the counter),
    # MolPrm3 for species 3 (setconf.F90:698-704)
    sp = synth.water_species(2)
    (tmp_path / "MDinfo").write_text("1 3\n2 2 3\n3 3 3\n")
    for name in ("MolPrm1", "SltInfo", "MolPrm3"):
        write_species_file(tmp_path / name, sp)
    t = I.load_system(tmp_path, dict(slttype=1, ljformat=0, sltspec=2)).topology
    assert t.nmol == 7 and t.natom == 21
    np.testing.assert_array_equal(t.tagslt, [2, 3])
    np.testing.assert_array_equal(t.uvspec, [1, 1, 2, 2, 3, 3, 3])

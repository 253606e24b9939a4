"""erdst over a complete ERmod run directory (erdst_gpu/erdst.py): parameters_er,
MDinfo, MolPrm1, SltInfo, EcdMesh, HISTORY and SltConf, as the reference run
of the 1000-water refs configuration is laid out (slttype 3, insertions from
SltConf as they are).  Every output file equals the one written from the same
inputs fed programmatically through the Python binding."""
import json
import os

import numpy as np
import pytest

import erutil as U
from dcdfiles import write_dcd
from erdst_gpu import capi, erdst, synth
from erdst_gpu.engstore import BlockWriter

pytestmark = pytest.mark.gpu

NINS, FRAMES = 40, 4


def fmt(x):
    return repr(float(np.float32(x)))


def write_run_directory(d, w, mesh_shift=0.0):
    topo = w.topo
    water, solute = topo.species[0], topo.species[topo.solute_index]
    mesh, pecore = w.mesh[0]
    mesh = np.asarray(mesh, np.float64) + mesh_shift
    (d / "parameters_er").write_text(f"""&ene_param
  slttype = 3, insposition = 1, insorient = 1, maxins = {NINS},
  ljformat = 0, inptemp = 300.0,
  elecut = 12.0, cltype = 2, screen = 0.35, splodr = 4, ms1max = {w.params['ms'][0]},
  selfcal = 1, engdiv = 2
/
&hist
  eclbin = 5.0e-2, ecfbin = 2.0e-3, ec0bin = 2.0e-4, finfac = 10.0,
  ecdmin = -20.0, ecfmns = -0.20, ecdcen = 0.0, eccore = 5.0,
  ecdmax = 1.0e4, pecore = {pecore}, meshread = 1
/
""")
    lines = [f"0 {U.SELF_MESH.size}"] + [f"0 {i + 1} {float(v)!r}" for i, v in enumerate(U.SELF_MESH)]
    lines += [f"1 {len(mesh)} {pecore}"] + [f"1 {i + 1} {float(v)!r}" for i, v in enumerate(mesh)]
    (d / "EcdMesh").write_text("\n".join(lines) + "\n")
    (d / "MDinfo").write_text(f"{FRAMES} 1\n{water.nmol}\n{water.nsite}\n")
    with open(d / "MolPrm1", "w") as fh:
        for i, typ in enumerate(["OW", "HW", "HW"]):
            fh.write(f"{i + 1} 15.9994 {typ} {typ}{i + 1} {fmt(water.charge[i])} {fmt(water.eps[i])} "
                     f"{fmt(water.sigma[i])}\n")
    with open(d / "SltInfo", "w") as fh:
        for i in range(solute.nsite):
            typ = "C" if i == 0 else "H"
            fh.write(f"{i + 1:4d}  {typ}  {fmt(solute.charge[i])}  {fmt(solute.eps[i])}  {fmt(solute.sigma[i])}\n")
    L = np.diag(w.cell).astype(float)
    write_dcd(d / "HISTORY", [w.frame(f) for f in range(FRAMES)], [(L[0], 90.0, L[1], 90.0, 90.0, L[2])] * FRAMES)
    confs = [c for f in range(3) for c in w.insertions(f)[:NINS]]
    write_dcd(d / "SltConf", confs, None, charmm=False)


@pytest.mark.parametrize("autotune", [False, True, "cache"])
def test_erdst_driver_over_a_run_directory(tmp_path, autotune):
    """Every output byte-identical to the same inputs through the binding;
    also with tuned launch shapes, and with shapes stored in a tuning cache
    by a first run and read back by a second."""
    w = synth.config0()
    run_dir = tmp_path / "run"
    run_dir.mkdir()
    write_run_directory(run_dir, w)
    cache = str(tmp_path / "shapes.json") if autotune == "cache" else None
    if cache:   # a first run tunes and stores the shapes (its outputs elsewhere)
        (tmp_path / "first").mkdir()
        erdst.run(str(run_dir), out=str(tmp_path / "first"), log=lambda *a: None, autotune=True,
                  tuning_cache=cache)
        assert len(json.load(open(cache))) == 1
    erdst.run(str(run_dir), log=lambda *a: None, autotune=bool(autotune), tuning_cache=cache)
    # the same inputs through the binding, written by the same writers
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    p = U.workload_params(w, selfcal=1)
    ctx = capi.Context(w.topo, p)
    uvmax, uvsoft, uvcrd = w.uvmax_uvsoft_uvcrd()
    ctx.set_mesh(uvmax, uvsoft, uvcrd, escrd=U.SELF_MESH)
    ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    wr = BlockWriter(str(ref_dir), uvmax, uvsoft, uvcrd, p["temp"], soln=False, engdiv=2, escrd=U.SELF_MESH)
    hist = capi.Trajectory(str(run_dir / "HISTORY"))
    for blk in range(2):
        for f in range(2 * blk, 2 * blk + 2):
            xyz, cell = hist.read()
            ctx.set_frame(xyz, cell)
            ctx.set_insertions(w.insertions(f % 3)[:NINS])
            ctx.run(capi.ER_BATCHED)
            wr.frame()
        wr.store(ctx.fetch())
        ctx.clear()
    names = sorted(os.listdir(ref_dir))
    assert names == ["corref.01", "corref.02", "engref.01", "engref.02", "progress.tt", "slfeng.01", "slfeng.02",
                     "uvrange.tt", "weight_refs"]
    for n in names:
        assert (run_dir / n).read_bytes() == (ref_dir / n).read_bytes(), n


def test_erdst_driver_soln_run_directory(tmp_path):
    """A soln run directory (configuration 1: a flexible 20-site solute, species
    1 in SltInfo, in 8000 waters, MolPrm1): engsln, aveuv, weight_soln,
    uvrange and flcuv equal those written from the binding on the same
    inputs."""
    w = synth.config1()
    topo = w.topo
    chain, water = topo.species[0], topo.species[1]
    mesh, pecore = w.mesh[0]
    run_dir = tmp_path / "run"
    run_dir.mkdir()
    (run_dir / "parameters_er").write_text(f"""&ene_param
  slttype = 1, ljformat = 0, inptemp = 300.0,
  elecut = 12.0, cltype = 2, screen = 0.35, splodr = 4, ms1max = {w.params['ms'][0]}, engdiv = 1
/
&hist
  eclbin = 5.0e-2, ecfbin = 2.0e-3, ec0bin = 2.0e-4, finfac = 10.0,
  ecdmin = -20.0, ecfmns = -0.20, ecdcen = 0.0, eccore = 5.0,
  ecdmax = 1.0e4, pecore = 100, meshread = 1
/
""")
    lines = [f"1 {len(mesh)} {pecore}"] + [f"1 {i + 1} {float(v)!r}" for i, v in enumerate(mesh)]
    (run_dir / "EcdMesh").write_text("\n".join(lines) + "\n")
    (run_dir / "MDinfo").write_text(f"2 2\n1 {water.nmol}\n{chain.nsite} {water.nsite}\n")
    with open(run_dir / "SltInfo", "w") as fh:
        for i in range(chain.nsite):
            fh.write(f"{i + 1} 12.0107 C C{i + 1} {fmt(chain.charge[i])} {fmt(chain.eps[i])} {fmt(chain.sigma[i])}\n")
    with open(run_dir / "MolPrm1", "w") as fh:
        for i, typ in enumerate(["OW", "HW", "HW"]):
            fh.write(f"{i + 1} 15.9994 {typ} {typ}{i + 1} {fmt(water.charge[i])} {fmt(water.eps[i])} "
                     f"{fmt(water.sigma[i])}\n")
    L = np.diag(w.cell).astype(float)
    write_dcd(run_dir / "HISTORY", [w.frame(f) for f in (0, 1)], [(L[0], 90.0, L[1], 90.0, 90.0, L[2])] * 2)
    erdst.run(str(run_dir), log=lambda *a: None)
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    p = U.workload_params(w)
    ctx = capi.Context(topo, p)
    uvmax, uvsoft, uvcrd = w.uvmax_uvsoft_uvcrd()
    ctx.set_mesh(uvmax, uvsoft, uvcrd)
    ctx.set_solute_molecules(topo.tagslt)
    wr = BlockWriter(str(ref_dir), uvmax, uvsoft, uvcrd, p["temp"], soln=True, engdiv=1, wgtslf=False,
                     corrcal=False)
    hist = capi.Trajectory(str(run_dir / "HISTORY"))
    for _ in range(2):
        xyz, cell = hist.read()
        ctx.set_frame(xyz, cell)
        ctx.run(capi.ER_BATCHED)
        wr.frame(ctx.flceng(1))
    wr.store(ctx.fetch())
    names = sorted(os.listdir(ref_dir))
    assert names == ["aveuv.tt", "engsln.tt", "flcuv.tt", "uvrange.tt", "weight_soln"]
    for n in names:
        assert (run_dir / n).read_bytes() == (ref_dir / n).read_bytes(), n


@pytest.mark.parametrize("permuted", [False, True])
def test_erdst_driver_soln_sphere_region(tmp_path, permuted):
    """Soln with a sphere region around the water aggregate (insposition 2,
    hostspec = species 2): the bounds accept the solute in one of two frames
    only (distances from the mass-weighted centres); the outputs equal the
    binding's run over the accepted frame, and flcuv.tt has that frame's
    line alone.  Permuted: HISTORY holds the atoms in reverse order and a
    PermIndex file maps them back (setconf.F90:877-919, 971-975), so the
    region test must read the permuted coordinates."""
    from erdst_gpu import inputs as I
    w = synth.config1()
    topo = w.topo
    chain, water = topo.species[0], topo.species[1]
    mesh, pecore = w.mesh[0]
    run_dir = tmp_path / "run"
    run_dir.mkdir()
    mw = np.tile(np.float32([15.9994, 1.008, 1.008]), water.nmol).astype(np.float64)
    mc = np.full(chain.nsite, 12.0107)
    dist = []
    for f in (0, 1):
        x = w.frame(f).astype(np.float64)
        c_slt = (x[:chain.nsite] * mc[:, None]).sum(0) / mc.sum()
        c_agg = (x[chain.nsite:] * mw[:, None]).sum(0) / mw.sum()
        dist.append(float(np.linalg.norm(c_slt - c_agg)))
    keep = int(np.argmin(dist))
    up = 0.5 * (dist[0] + dist[1])
    (run_dir / "parameters_er").write_text(f"""&ene_param
  slttype = 1, ljformat = 0, inptemp = 300.0,
  elecut = 12.0, cltype = 2, screen = 0.35, splodr = 4, ms1max = {w.params['ms'][0]}, engdiv = 1,
  insposition = 2, lwreg = 0.0, upreg = {up!r}, hostspec(1) = 2
/
&hist
  eclbin = 5.0e-2, ecfbin = 2.0e-3, ec0bin = 2.0e-4, finfac = 10.0,
  ecdmin = -20.0, ecfmns = -0.20, ecdcen = 0.0, eccore = 5.0,
  ecdmax = 1.0e4, pecore = 100, meshread = 1
/
""")
    lines = [f"1 {len(mesh)} {pecore}"] + [f"1 {i + 1} {float(v)!r}" for i, v in enumerate(mesh)]
    (run_dir / "EcdMesh").write_text("\n".join(lines) + "\n")
    (run_dir / "MDinfo").write_text(f"2 2\n1 {water.nmol}\n{chain.nsite} {water.nsite}\n")
    with open(run_dir / "SltInfo", "w") as fh:
        for i in range(chain.nsite):
            fh.write(f"{i + 1} 12.0107 C C{i + 1} {fmt(chain.charge[i])} {fmt(chain.eps[i])} {fmt(chain.sigma[i])}\n")
    with open(run_dir / "MolPrm1", "w") as fh:
        for i, (typ, m) in enumerate([("OW", "15.9994"), ("HW", "1.008"), ("HW", "1.008")]):
            fh.write(f"{i + 1} {m} {typ} {typ}{i + 1} {fmt(water.charge[i])} {fmt(water.eps[i])} "
                     f"{fmt(water.sigma[i])}\n")
    L = np.diag(w.cell).astype(float)
    frames = [w.frame(f) for f in (0, 1)]
    if permuted:
        n = topo.natom
        (run_dir / "PermIndex").write_text("".join(f"{i + 1} {n - i}\n" for i in range(n)))
        frames = [x[::-1].copy() for x in frames]
    write_dcd(run_dir / "HISTORY", frames, [(L[0], 90.0, L[1], 90.0, 90.0, L[2])] * 2)
    erdst.run(str(run_dir), log=lambda *a: None)
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    p = U.workload_params(w)
    ctx = capi.Context(topo, p)
    uvmax, uvsoft, uvcrd = w.uvmax_uvsoft_uvcrd()
    ctx.set_mesh(uvmax, uvsoft, uvcrd)
    ctx.set_solute_molecules(topo.tagslt)
    xyz, cell = list(capi.Trajectory(str(run_dir / "HISTORY")))[keep]
    if permuted:
        xyz = xyz[::-1].copy()
    ctx.set_frame(xyz, cell)
    ctx.run(capi.ER_BATCHED)
    wr = BlockWriter(str(ref_dir), uvmax, uvsoft, uvcrd, p["temp"], soln=True, engdiv=1, wgtslf=False,
                     corrcal=False)
    wr.store(ctx.fetch())
    for n in ("aveuv.tt", "engsln.tt", "uvrange.tt", "weight_soln"):
        assert (run_dir / n).read_bytes() == (ref_dir / n).read_bytes(), n
    fl = open(run_dir / "flcuv.tt").read().splitlines()
    assert len(fl) == 1 and int(fl[0][:9]) == keep + 1


def test_erdst_driver_stops_on_a_halt(tmp_path, capsys):
    """A mesh whose lowest point lies above the lowest pair energy: erdst
    halts with eng_min in the first frame (engproc.F90:1142-1147) and writes
    no block file; the command exits with status 1 and names the code."""
    w = synth.config0()
    run_dir = tmp_path / "run"
    run_dir.mkdir()
    write_run_directory(run_dir, w, mesh_shift=19.95)   # ecdmin -0.05 kcal/mol
    rc = erdst.main([str(run_dir)])
    err = capsys.readouterr().err
    assert rc == 1 and "eng_min" in err, err
    names = os.listdir(run_dir)
    assert not any(n.startswith(("engref", "corref", "slfeng", "weight")) for n in names), names

"""The Fortran erdst driver (fortran/erdst_host.f90) over the C ABI: per-block
engref/corref/weight_refs/uvrange files in the reference's formats, equal to
the same run through the Python binding."""
import os
import subprocess

import numpy as np
import pytest

import erutil as U
from erdst_gpu import capi, synth
from erdst_gpu.workload_io import write_workload

HOST = os.path.join(os.path.dirname(capi.LIB_PATH), "erdst_host")


def read_fortran_record(path):
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    assert int(raw[4 + n:8 + n].view(np.int32)[0]) == n and raw.size == n + 8
    return raw[4:4 + n].view(np.float64)


def read_engfile(path):
    lines = open(path).read().splitlines()
    assert lines[0].split() == ["#", "bin_left", "bin_middle", "spec", "ratio", "dens"]
    rows = []
    for ln in lines[1:]:
        assert len(ln) == 85           # (g15.7,g15.7,i5,g25.15,g25.15)
        rows.append([float(ln[0:15]), float(ln[15:30]), int(ln[30:35]), float(ln[35:60]), float(ln[60:85])])
    return np.array(rows)


@pytest.mark.gpu
def test_fortran_host_two_blocks(tmp_path):
    w = synth.config0()
    p = U.workload_params(w)
    frames = [0, 1, 2, 3]
    wl = tmp_path / "workload.bin"
    write_workload(wl, w, p, frames, engdiv=2, nins=40)
    r = subprocess.run([HOST, str(wl)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # the same run through the Python binding
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd())
    ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    mesh = w.uvmax_uvsoft_uvcrd()[2]
    for blk in (1, 2):
        for f in frames[2 * (blk - 1):2 * blk]:
            ctx.set_frame(w.frame(f), w.cell)
            ctx.set_insertions(w.insertions(f)[:40])
            ctx.run(capi.ER_BATCHED)
        g = ctx.fetch()
        ctx.clear()
        eng = read_engfile(tmp_path / f"engref.0{blk}")
        np.testing.assert_allclose(eng[:, 0], mesh, rtol=1e-6)
        np.testing.assert_allclose(eng[:, 3], g["edens"] / g["engnorm"], rtol=1e-13, atol=1e-300)
        cor = read_fortran_record(tmp_path / f"corref.0{blk}").reshape(g["ecorr"].shape)
        np.testing.assert_allclose(cor, g["ecorr"] / g["engnorm"], rtol=1e-13, atol=1e-300)
    wt = open(tmp_path / "weight_refs").read().splitlines()
    assert len(wt) == 2 and all(len(x) == 45 for x in wt)
    uvr = open(tmp_path / "uvrange.tt").read().splitlines()
    assert uvr[0] == " species     minimum        maximum" and len(uvr) == 3


@pytest.mark.gpu
def test_fortran_host_autotune_changes_no_output(tmp_path):
    """"autotune" tunes the launch shapes on the first frame; every output
    file is byte-identical to the untuned run's."""
    w = synth.config0()
    p = U.workload_params(w)
    runs = {}
    for name, extra in (("builtin", []), ("tuned", ["autotune"])):
        d = tmp_path / name
        d.mkdir()
        write_workload(d / "workload.bin", w, p, [0, 1, 2, 3], engdiv=2, nins=40)
        r = subprocess.run([HOST, "workload.bin"] + extra, cwd=d, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        runs[name] = {f.name: f.read_bytes() for f in sorted(d.iterdir()) if f.name != "workload.bin"}
        if extra:
            assert "launch shapes: chunk" in r.stdout
    assert sorted(runs["tuned"]) == sorted(runs["builtin"]) and "engref.02" in runs["tuned"]
    for name, data in runs["builtin"].items():
        assert runs["tuned"][name] == data, name


@pytest.mark.gpu
def test_fortran_host_selfcal_and_progress(tmp_path):
    """selfcal = 1: slfeng.XX per block equals eself/engnorm of the same run
    through the Python binding, written as the Python writer writes it;
    progress.tt holds one frame number per frame."""
    from erdst_gpu import engstore
    w = synth.config0()
    p = U.workload_params(w, selfcal=1)
    frames = [0, 1, 2, 3]
    wl = tmp_path / "workload.bin"
    write_workload(wl, w, p, frames, engdiv=2, nins=40, escrd=U.SELF_MESH)
    r = subprocess.run([HOST, str(wl)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd(), escrd=U.SELF_MESH)
    ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    for blk in (1, 2):
        for f in frames[2 * (blk - 1):2 * blk]:
            ctx.set_frame(w.frame(f), w.cell)
            ctx.set_insertions(w.insertions(f)[:40])
            ctx.run(capi.ER_BATCHED)
        g = ctx.fetch()
        ctx.clear()
        engstore.write_slfeng(tmp_path / f"py_slfeng.0{blk}", g["eself"] / g["engnorm"], U.SELF_MESH)
        assert open(tmp_path / f"slfeng.0{blk}").read() == open(tmp_path / f"py_slfeng.0{blk}").read()
        rows = open(tmp_path / f"slfeng.0{blk}").read().splitlines()
        vals = np.array([float(ln[15:40]) for ln in rows])
        assert len(rows) == U.SELF_MESH.size and abs(vals.sum() - 1.0) < 1e-12
    assert open(tmp_path / "progress.tt").read() == "".join(engstore.progress_line(f) + "\n" for f in (1, 2, 3, 4))


@pytest.mark.gpu
def test_fortran_host_soln_flcuv(tmp_path):
    """A soln run (one flexible solute in 8000 waters): flcuv.tt holds, per
    frame, the solute's summed pair energy with the solvent, equal to the
    Python binding's flceng written by the Python writer and to the FP64
    oracle's sums within the summed energy tolerance."""
    from erdst_gpu import engstore
    w = synth.config1()
    p = U.workload_params(w)
    write_workload(tmp_path / "workload.bin", w, p, [0, 1])
    r = subprocess.run([HOST, str(tmp_path / "workload.bin")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd())
    ctx.set_solute_molecules(w.topo.tagslt)
    want = []
    sp = U.species_of(w.topo, 1, w.topo.tagslt)[0]
    for f in (0, 1):
        ctx.set_frame(w.frame(f), w.cell)
        ctx.run(capi.ER_BATCHED)
        fl = ctx.flceng(1)
        want += engstore.flcuv_lines(f + 1, fl)
        uo, _ = U.oracle_soln(w, f, p)
        ref = np.array([uo[0][sp == s + 1].sum() for s in range(w.topo.numslv)])
        tol = U.energy_tolerance(uo[0][sp > 0], U.nsites(w)).sum()
        assert np.all(np.abs(fl[0] - ref) <= tol), (fl[0], ref, tol)
    got = open(tmp_path / "flcuv.tt").read().splitlines()
    assert got == want and len(got) == 2


@pytest.mark.gpu
def test_fortran_host_reads_history_and_sltconf(tmp_path):
    """Frames from a HISTORY DCD and insertions from a SltConf DCD (one solute
    configuration per DCD frame, 3 frames' worth, so the fourth frame's
    insertions wrap around to the file's start as erdst's OUTconfig does):
    the engref/corref/weight files equal those of the same inputs through the
    Python binding, and the DCD coordinates are the workload's float32
    coordinates exactly."""
    from dcdfiles import write_dcd
    w = synth.config0()
    p = U.workload_params(w)
    frames, nins = [0, 1, 2, 3], 40
    L = np.diag(w.cell).astype(float)
    write_dcd(tmp_path / "HISTORY", [w.frame(f) for f in frames], [(L[0], 90.0, L[1], 90.0, 90.0, L[2])] * 4)
    confs = [c for f in frames[:3] for c in w.insertions(f)[:nins]]
    write_dcd(tmp_path / "SltConf", confs, None, charmm=False)
    write_workload(tmp_path / "workload.bin", w, p, frames, engdiv=2, nins=nins, inline_system=False,
                   inline_insertions=False)
    r = subprocess.run([HOST, "workload.bin", "history=HISTORY", "sltconf=SltConf"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd())
    ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    hist = capi.Trajectory(str(tmp_path / "HISTORY"))
    slt = capi.Trajectory(str(tmp_path / "SltConf"))
    for blk in (1, 2):
        for f in frames[2 * (blk - 1):2 * blk]:
            xyz, cell = hist.read()
            np.testing.assert_array_equal(xyz, w.frame(f))
            ins = []
            for _ in range(nins):
                fr = slt.read()
                if fr is None:                      # wrap around
                    slt.close()
                    slt = capi.Trajectory(str(tmp_path / "SltConf"))
                    fr = slt.read()
                ins.append(fr[0])
            ins = np.stack(ins)
            np.testing.assert_array_equal(ins, w.insertions(f % 3)[:nins])
            ctx.set_frame(xyz, cell)
            ctx.set_insertions(ins)
            ctx.run(capi.ER_BATCHED)
        g = ctx.fetch()
        ctx.clear()
        eng = read_engfile(tmp_path / f"engref.0{blk}")
        np.testing.assert_allclose(eng[:, 3], g["edens"] / g["engnorm"], rtol=1e-13, atol=1e-300)
        cor = read_fortran_record(tmp_path / f"corref.0{blk}").reshape(g["ecorr"].shape)
        np.testing.assert_allclose(cor, g["ecorr"] / g["engnorm"], rtol=1e-13, atol=1e-300)
    assert open(tmp_path / "progress.tt").read() == " 1\n 2\n 3\n 4\n"


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [dict(cltype=3), dict(cltype=0), dict(cltype=0, boxshp=0, estype=1)])
def test_fortran_host_interaction_variants(tmp_path, variant):
    """PPPM, bare Coulomb and an open system through the Fortran host: its
    engref equals the Python binding's on the same inputs."""
    w = synth.config0()
    p = U.workload_params(w, **variant)
    write_workload(tmp_path / "workload.bin", w, p, [0, 1], nins=24)
    r = subprocess.run([HOST, "workload.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd())
    ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    for f in (0, 1):
        ctx.set_frame(w.frame(f), w.cell)
        ctx.set_insertions(w.insertions(f)[:24])
        ctx.run(capi.ER_BATCHED)
    g = ctx.fetch()
    eng = read_engfile(tmp_path / "engref.tt")
    np.testing.assert_allclose(eng[:, 3], g["edens"] / g["engnorm"], rtol=1e-13, atol=1e-300)


@pytest.mark.gpu
def test_fortran_host_files_feed_slvfe(tmp_path):
    """erdst (Fortran host, soln and refs runs) -> files -> slvfe merge: the
    solvation free energy equals the one from the Python binding's blocks."""
    from erdst_gpu import slvfe
    soln, refs = synth.dmu_pair(0)
    for sub, w, engdiv in (("soln", soln, 2), ("refs", refs, 1)):
        d = tmp_path / sub
        d.mkdir()
        p = U.workload_params(w)
        write_workload(d / "workload.bin", w, p, range(10), engdiv=engdiv, nins=100)
        r = subprocess.run([HOST, str(d / "workload.bin")], cwd=d, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "soln" / "aveuv.tt").exists()
    res = slvfe.run_merge(slvfe.FeVars(), str(tmp_path))
    sg, rg = U.gpu_dmu_blocks(soln, refs, 5, 2, 100)
    kT = U.workload_params(refs)["temp"]
    ref = slvfe.dmu_from_blocks(sg, rg, *refs.uvmax_uvsoft_uvcrd(), kT, str(tmp_path / "py"))
    print(f"dmu from host files {res.dmu:.6f}, from the binding {ref.dmu:.6f}")
    np.testing.assert_allclose(res.chmpt, ref.chmpt, rtol=0, atol=1e-9)

"""erdst over an ERmod run directory, with the per-frame work on the GPU.

    python -m erdst_gpu.erdst [DIRECTORY] [--plain] [--autotune [--tuning-cache FILE]] [--out OUTDIR]

Reads what erdst reads (erdst/setconf.F90, erdst/engproc.F90:36-318):
parameters_er (&ene_param, &hist), MDinfo, MolPrm<n>, SltInfo, LJTable,
EcdInfo / EcdMesh, PermIndex, SysWght, SltWght, and the DCD trajectories
HISTORY and SltConf; runs the engdiv x frame loop of trjmain
(erdst/enganal.F90:82-95) through the C ABI; writes engsln|engref.XX,
corsln|corref.XX, slfeng.XX, weight_soln|weight_refs, aveuv.tt, uvrange.tt
and flcuv.tt|progress.tt.

Supported: bare Coulomb, PME or PPPM (cltype 0, 2, 3) in a periodic cell,
and bare Coulomb in an open system (boxshp 0); soln
runs; refs runs whose solute configurations are used as they are
(insposition 1 and insorient 1):
flexible solutes (slttype 3) from SltConf, one DCD frame per insertion,
reopened at its end, and rigid solutes (slttype 2) at their SltInfo
coordinates.  Soln runs apply the sphere and slab region rejection of
check_mol_configuration.  Random insertion (the reference's compiler-specific
random_number stream) and RMSD-based restraints raise InputError.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import capi
from .engstore import BlockWriter
from .inputs import (SLT_REFS_FLEX, SLT_REFS_RIGID, SLT_SOLN, InputError, WeightFile, apply_permutation,
                     iniparam, load_system, read_namelist, read_permindex)
from .mesh import species_meshes
from .tuning import tune_cached, tuning_key


def _hot_params(p: dict) -> dict:
    return dict(elecut=float(p["elecut"]), lwljcut=float(p["lwljcut"]), upljcut=float(p["upljcut"]),
                screen=float(p["screen"]), ljswitch=int(p["ljswitch"]), cltype=int(p["cltype"]),
                splodr=int(p["splodr"]), ms=(int(p["ms1max"]), int(p["ms2max"]), int(p["ms3max"])),
                boxshp=int(p["boxshp"]), slttype=int(p["slttype"]), wgtslf=int(p["wgtslf"]),
                estype=int(p["estype"]), corrcal=int(p["corrcal"]), selfcal=int(p["selfcal"]),
                temp=float(p["temp"]), force_calculation=int(bool(p["force_calculation"])))


def region_distance(p: dict, sysin, xyz, cell, mol: int) -> float:
    """check_mol_configuration's distance for a soln solute molecule
    (erdst/engproc.F90:1197-1245): its centre of mass relative to the
    aggregate of the hostspec species (erdst/insertion.F90:587-614,
    erdst/bestfit.F90:252-266), both mass-weighted means of the raw
    coordinates; the norm for a sphere, the component along c for a slab
    (its magnitude for a symmetric slab)."""
    topo = sysin.topology
    mb, mass = topo.mol_begin, sysin.mass.astype(np.float64)
    x = np.asarray(xyz, np.float64)

    def com(sel):
        m = mass[sel]
        if m.sum() == 0.0:
            raise InputError("zero total mass", "bst_zrw")
        return (x[sel] * m[:, None]).sum(0) / m.sum()
    hosts = {int(v) for v in p["hostspec"].values()} if isinstance(p["hostspec"], dict) else set()
    agg = np.concatenate([np.arange(mb[i], mb[i + 1]) for i in range(topo.nmol) if topo.moltype[i] in hosts])
    dx = com(np.arange(mb[mol], mb[mol + 1])) - com(agg)
    if p["insposition"] == 2:
        return float(np.sqrt(dx @ dx))
    c = np.asarray(cell, np.float64)
    d = float(np.linalg.inv(c)[2] @ dx) * float(np.linalg.norm(c[:, 2]))
    return abs(d) if p["insposition"] == 4 else d


def _check_supported(p: dict):
    if p["cltype"] not in (0, 2, 3):
        raise InputError("bare Coulomb (cltype = 0), PME (2) and PPPM (3) run on the GPU path, not Ewald (1)")
    if p["boxshp"] not in (0, 1):
        raise InputError("boxshp must be 0 (open, bare Coulomb) or 1 (periodic)")
    if p["slttype"] != SLT_SOLN and not (p["insposition"] == 1 and p["insorient"] == 1):
        raise InputError("refs runs need the solute configurations as they are (insposition = 1, insorient = 1)")
    if p["slttype"] != SLT_SOLN and p["insstructure"] != 0:
        raise InputError("structure rejection (insstructure) is not supported")
    if p["slttype"] == SLT_SOLN and p["insposition"] in (5, 6):
        raise InputError("RMSD-based region rejection (insposition 5, 6) is not supported")
    if p["slttype"] == SLT_SOLN and p["insposition"] in (2, 3, 4):
        hosts = [v for v in p["hostspec"].values()] if isinstance(p["hostspec"], dict) else []
        if not any(int(v) >= 1 for v in hosts):
            raise InputError("a region needs a hostspec species", "set_ins")
        if p["lwreg"] > p["upreg"] or (p["insposition"] != 3 and min(p["lwreg"], p["upreg"]) < 0.0):
            raise InputError("region bounds", "set_reg")


def run(directory: str, out: str | None = None, mode: int = capi.ER_BATCHED, log=print,
        autotune: bool = False, tuning_cache: str | None = None) -> dict:
    """One erdst run; returns the last block's fetched accumulators.  With
    autotune the launch shapes are timed on the first frame (er_autotune),
    or taken from tuning_cache when it holds this system's; either way no
    output changes."""
    d = directory
    out = out or d
    path = lambda name: os.path.join(d, name)
    p = iniparam(read_namelist(path("parameters_er"), "ene_param"))
    _check_supported(p)
    sysin = load_system(d, p)
    topo = sysin.topology
    soln = p["slttype"] == SLT_SOLN
    hist = read_namelist(path("parameters_er"), "hist")
    uvmax, uvsoft, uvcrd, escrd = species_meshes(
        path("parameters_er"), topo.numslv,
        path("EcdInfo") if os.path.exists(path("EcdInfo")) else None,
        path("EcdMesh") if int(hist.get("meshread", 0)) == 1 else None)
    ctx = capi.Context(topo, _hot_params(p))
    ctx.set_mesh(uvmax, uvsoft, uvcrd, escrd if p["selfcal"] == 1 else None)
    if soln:
        ctx.set_solute_molecules(topo.tagslt)
    else:
        ctx.set_solute(topo.slt_ljtype, topo.slt_charge32)
    writer = BlockWriter(out, uvmax, uvsoft, uvcrd, p["temp"], soln=soln, engdiv=p["engdiv"],
                         wgtslf=bool(p["wgtslf"]), corrcal=bool(p["corrcal"]),
                         escrd=escrd if p["selfcal"] == 1 else None)
    maxcnf, skpcnf, engdiv = sysin.mdinfo.nrun, p["skpcnf"], p["engdiv"]
    frames_per_div = maxcnf // skpcnf // engdiv
    if frames_per_div <= 0:
        raise InputError("fewer frames than blocks", "eng_par")
    perm = read_permindex(path("PermIndex"), topo.natom) if os.path.exists(path("PermIndex")) else None
    syswgt = WeightFile(path("SysWght"), rewind=False) if p["wgtsys"] == 1 else None
    history = capi.Trajectory(path("HISTORY"))
    if history.natoms != topo.natom:
        raise InputError("# of atoms in HISTORY does not match the configurations")
    nslt = 0 if soln else len(topo.slt_charge32)
    nins = 1 if soln else p["maxins"]
    sltconf = None
    sltwgt = None
    if p["slttype"] == SLT_REFS_FLEX:
        sltconf = capi.Trajectory(path("SltConf"))
        if sltconf.natoms != nslt:
            raise InputError("# of atoms in SltConf does not match the solute")
        if p["wgtins"] == 1:
            sltwgt = WeightFile(path("SltWght"), rewind=True)
    g = None
    tune = autotune and mode == capi.ER_BATCHED

    def apply_tuning():
        if tuning_cache is None:
            log(f"erdst: launch shapes {ctx.autotune()}")
            return
        mb = topo.mol_begin
        sites = [int(mb[m + 1] - mb[m]) for m in topo.tagslt] if soln else [nslt]
        key = tuning_key(topo, _hot_params(p), sites, len(topo.tagslt) if soln else nins)
        shapes, found = tune_cached(ctx, key, tuning_cache)
        log(f"erdst: launch shapes {shapes} ({'from' if found else 'added to'} {tuning_cache})")
    def run_frame(m):
        # engconst halts inside the frame on eng_min / eng_sft, and on the
        # 'mbin' warning unless force_calculation (engproc.F90:1142-1167,
        # mpiproc.F90:238-252): the frame's output lines and the block's
        # files are not written
        ctx.run(m)
        try:
            ctx.sync()
        except capi.ErError as e:
            if e.status != capi.ER_E_HALT:
                raise
            code, sp, en = ctx.last_error()
            raise InputError(f"energy of {en:.4g} for {sp}-th species: {e}", code) from None
    active = list(range(len(topo.tagslt)))
    for division in range(1, engdiv + 1):
        for _ in range(frames_per_div):
            w_sys = np.float32(1.0)
            for _ in range(skpcnf):
                fr = history.read()
                if fr is None:
                    raise InputError("HISTORY has no further frame", "set_trj")
                if syswgt is not None:
                    w_sys = syswgt.next()
            xyz, cell = fr
            if perm is not None:
                xyz = apply_permutation(xyz, perm)
            ctx.set_frame(xyz, cell, float(w_sys))
            if not soln:
                ins = np.empty((nins, nslt, 3), np.float32)
                w_slt = np.ones(nins, np.float32)
                for n in range(nins):
                    if p["slttype"] == SLT_REFS_RIGID:
                        ins[n] = sysin.bfcoord
                        continue
                    c = sltconf.read()
                    if c is None:                       # reopen at the end (setconf.F90:106-116)
                        sltconf.close()
                        sltconf = capi.Trajectory(path("SltConf"))
                        c = sltconf.read()
                        if c is None:
                            raise InputError("Failed to reload solute trajectory")
                    ins[n] = c[0]
                    if sltwgt is not None:
                        w_slt[n] = sltwgt.next()
                ctx.set_insertions(ins, w_slt if sltwgt is not None else None)
            if soln and p["insposition"] in (2, 3, 4):
                # check_mol_configuration per solute molecule (engproc.F90:
                # 750-756): a rejected one adds nothing this frame
                keep = [k for k, m in enumerate(topo.tagslt)
                        if p["lwreg"] <= region_distance(p, sysin, xyz, cell, int(m)) <= p["upreg"]]
                if not keep:
                    writer.frame(np.zeros((0, topo.numslv)), skpcnf, [], len(topo.tagslt))
                    continue
                if keep != active:
                    ctx.set_solute_molecules(topo.tagslt[keep])
                    active = keep
                if tune:
                    apply_tuning()
                    tune = False
                run_frame(mode)
                writer.frame(ctx.flceng(len(keep)), skpcnf, keep, len(topo.tagslt))
                continue
            if tune:
                apply_tuning()
                tune = False
            run_frame(mode)
            writer.frame(ctx.flceng(len(topo.tagslt)) if soln else None, skpcnf)
        g = ctx.fetch()
        writer.store(g)
        ctx.clear()
        log(f"erdst: block {division} of {engdiv} stored")
    history.close()
    if sltconf is not None:
        sltconf.close()
    return g


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m erdst_gpu.erdst", description=__doc__.split("\n\n")[0])
    ap.add_argument("directory", nargs="?", default=".")
    ap.add_argument("--out", default=None, help="directory for the output files (default: the run directory)")
    ap.add_argument("--plain", action="store_true", help="one insertion per launch sequence")
    ap.add_argument("--autotune", action="store_true", help="time the launch shapes on the first frame")
    ap.add_argument("--tuning-cache", default=None,
                    help="with --autotune: take the shapes from this JSON file, or add them to it")
    a = ap.parse_args(argv)
    if a.tuning_cache and not a.autotune:
        ap.error("--tuning-cache needs --autotune")
    try:
        run(a.directory, a.out, capi.ER_PLAIN if a.plain else capi.ER_BATCHED, autotune=a.autotune,
            tuning_cache=a.tuning_cache)
    except (InputError, capi.TrajError, capi.ErError) as e:
        print(f"erdst: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Stream file read by the Fortran host driver (fortran/erdst_host.f90).

Layout (little-endian, no record markers):
  int32[22]  numatm nummol ntype numslv nslt nins nframes ermax slttype
             ms1 ms2 ms3 splodr ljswitch estype corrcal wgtslf selfcal engdiv esmax
             cltype boxshp
  float32[4] elecut lwljcut upljcut screen ; float64 temp
  int32 mol_begin[nummol+1] (1-based), ljtype[numatm] (1-based), uvspec[nummol]
  float32 charge[numatm], ljene[ntype^2], ljlensq[ntype^2]
  int32 uvmax[numslv], uvsoft[numslv] ; float64 uvcrd[ermax]
  selfcal = 1: float64 escrd[esmax] (the self-energy mesh, ercrd(:,0))
  refs: int32 slt_ljtype[nslt] (1-based), float32 slt_charge[nslt]
  soln: int32 tagslt[nins] (1-based)
  float32 cell[9] (column-major)
  per frame: float32 sitepos[3*numatm] ; refs: float32 sltpos[3*nslt*nins]
  (sitepos is left out when the host reads frames from a HISTORY DCD, sltpos
  when it reads insertions from a SltConf DCD)
"""
from __future__ import annotations

import numpy as np


def write_workload(path, w, params: dict, frames, engdiv: int = 1, nins: int | None = None, escrd=None,
                   inline_system: bool = True, inline_insertions: bool = True):
    t = w.topo
    if (escrd is not None) != (params["selfcal"] == 1):
        raise ValueError("a self-energy mesh is written exactly when selfcal = 1")
    esmax = 0 if escrd is None else len(escrd)
    refs = params["slttype"] != 1
    nslt = len(t.slt_charge32) if refs else 0
    nins = (nins or w.nins) if refs else len(t.tagslt)
    um, us, uc = w.uvmax_uvsoft_uvcrd()
    frames = list(frames)
    hdr = np.array([t.natom, t.nmol, t.ntype, t.numslv, nslt, nins, len(frames), int(um.sum()),
                    params["slttype"], *params["ms"], params["splodr"], params["ljswitch"], params["estype"],
                    params["corrcal"], params["wgtslf"], params["selfcal"], engdiv, esmax,
                    params.get("cltype", 2), params.get("boxshp", 1)], np.int32)
    with open(path, "wb") as fh:
        hdr.tofile(fh)
        np.array([params["elecut"], params["lwljcut"], params["upljcut"], params["screen"]], np.float32).tofile(fh)
        np.array([params["temp"]], np.float64).tofile(fh)
        (t.mol_begin + 1).astype(np.int32).tofile(fh)
        (t.ljtype + 1).astype(np.int32).tofile(fh)
        t.uvspec.astype(np.int32).tofile(fh)
        t.charge32.astype(np.float32).tofile(fh)
        t.ljene32.astype(np.float32).tofile(fh)
        t.ljlensq32.astype(np.float32).tofile(fh)
        um.astype(np.int32).tofile(fh)
        us.astype(np.int32).tofile(fh)
        uc.astype(np.float64).tofile(fh)
        if escrd is not None:
            np.asarray(escrd, np.float64).tofile(fh)
        if refs:
            (t.slt_ljtype + 1).astype(np.int32).tofile(fh)
            t.slt_charge32.astype(np.float32).tofile(fh)
        else:
            (t.tagslt + 1).astype(np.int32).tofile(fh)
        w.cell.astype(np.float32).ravel(order="F").tofile(fh)
        for f in frames:
            if inline_system:
                w.frame(f).astype(np.float32).tofile(fh)
            if refs and inline_insertions:
                w.insertions(f)[:nins].astype(np.float32).tofile(fh)
    return nins

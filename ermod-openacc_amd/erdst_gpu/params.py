"""&ene_param values read by the hot path, with iniparam's defaults
(erdst/setconf.F90:133-438)."""
from __future__ import annotations

import numpy as np

SLT_SOLN, SLT_REFS_RIGID, SLT_REFS_FLEX = 1, 2, 3


def temperature_kt(inptemp: float = 300.0, double_build: bool = False) -> float:
    """temp = inptemp * 8.314510e-3 / 4.184 (setconf.F90:304).  In the default
    (single) build both literals are default-kind REAL, i.e. float32."""
    if double_build:
        return inptemp * 8.314510e-3 / 4.184
    return float(np.float64(inptemp) * np.float64(np.float32(8.314510e-3)) / np.float64(np.float32(4.184)))


def hot_path_params(slttype: int, ms, *, elecut=12.0, upljcut=None, lwljcut=None, screen=0.35,
                    splodr=4, ljswitch=0, estype=2, corrcal=None, wgtslf=None, selfcal=0,
                    inptemp=300.0, cltype=2, boxshp=1, force_calculation=False) -> dict:
    """Parameters after iniparam's defaulting: upljcut and lwljcut default to
    12 and 10 A, set from the default elecut before the namelist is read
    (setconf.F90:180-188), so a user's elecut does not move them; corrcal
    and wgtslf default to YES for refs with PME, corrcal NO for soln
    (setconf.F90:219-230)."""
    up = 12.0 if upljcut is None else upljcut
    lw = 10.0 if lwljcut is None else lwljcut
    refs = slttype != SLT_SOLN
    return dict(elecut=float(elecut), lwljcut=float(lw), upljcut=float(up), screen=float(screen),
                ljswitch=int(ljswitch), cltype=int(cltype), splodr=int(splodr), ms=tuple(int(m) for m in ms),
                boxshp=int(boxshp), slttype=int(slttype), wgtslf=int((1 if refs else 0) if wgtslf is None else wgtslf),
                estype=int(estype), corrcal=int((1 if refs else 0) if corrcal is None else corrcal),
                selfcal=int(selfcal), temp=temperature_kt(inptemp), force_calculation=int(bool(force_calculation)))

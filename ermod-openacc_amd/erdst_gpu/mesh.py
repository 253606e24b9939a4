"""The energy-coordinate mesh of enginit (erdst/engproc.F90:127-281).

For each species pti = 0 (the solute self energy) .. numslv the &hist
namelist is read again, optionally overridden by that species' EcdInfo line,
and the mesh is built from seven region boundaries: linear bins of eclbin,
ecfbin, ec0bin, ecfbin, eclbin, then pecore log-spaced bins up to ecdmax.
With meshread an EcdMesh block replaces the built mesh.  As in the reference,
&hist values the namelist does not set keep the values the previous species
left in them (the boundaries are shifted in place), and the EcdMesh line
counter is the mesh build's loop index, so a species without a block keeps
its &hist mesh only when that mesh has a core region.
"""
from __future__ import annotations

import math

import numpy as np

from .inputs import InputError, ListDirected, read_namelist

RGLMAX, LARGE = 5, 10000
_HIST_REAL = ("ecdmin", "ecfmns", "ecdcen", "eccore", "ecdmax", "eclbin", "ecfbin", "ec0bin", "finfac")
_HIST_INT = ("pecore", "ecprread", "meshread", "peread")


def _nint(x: float) -> int:
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def build_mesh(h: dict):
    """Mesh of one species from the &hist values in h (modified in place as
    the reference modifies them).  Returns (mesh [pemax], pemax, pesoft)."""
    return _build_mesh(h)[:3]


def _build_mesh(h: dict):
    """build_mesh, plus the DO index the build's last region loop ends with
    (engproc.F90:222): maxrg of that region, i.e. pemax with a core region
    and pesoft + 1 without."""
    ectmvl = h["finfac"] * h["ecfbin"]
    h["ecdmin"] = h["ecdmin"] - ectmvl
    h["ecfmns"] = h["ecfmns"] - ectmvl
    ecmns0 = h["ecdcen"] - ectmvl
    ecpls0 = 2.0 * h["ecdcen"] - ecmns0
    ecfpls = 2.0 * h["ecdcen"] - h["ecfmns"]
    h["eccore"] = h["eccore"] + ecfpls - h["ecdcen"]
    cdrgvl = [h["ecdmin"], h["ecfmns"], ecmns0, ecpls0, ecfpls, h["eccore"], h["ecdmax"]]
    uprgcd = [1] + [0] * (RGLMAX + 1)
    for regn in range(1, RGLMAX + 1):
        factor = h["eclbin"] if regn in (1, 5) else h["ecfbin"] if regn in (2, 4) else h["ec0bin"]
        incre = cdrgvl[regn] - cdrgvl[regn - 1]
        if incre < factor:
            raise InputError("energy-coordinate region narrower than its bin", "eng_ecd")
        uprgcd[regn] = uprgcd[regn - 1] + _nint(incre / factor)
    pecore = int(h["pecore"])
    pesoft = uprgcd[RGLMAX] - uprgcd[0]
    pemax = pesoft + pecore
    uprgcd[RGLMAX + 1] = pemax
    if pemax > LARGE:
        raise InputError("mesh larger than 10000 points", "eng_siz")
    ercrd = np.zeros(max(LARGE, pemax + 1) + 1)          # 1-based
    for r in range(RGLMAX + 2):
        ercrd[uprgcd[r]] = cdrgvl[r]
    nreg = RGLMAX if pecore == 0 else RGLMAX + 1
    for regn in range(1, nreg + 1):
        minrg, maxrg = uprgcd[regn - 1], uprgcd[regn]
        if regn <= RGLMAX:
            incre = ercrd[maxrg] - ercrd[minrg]
        else:
            incre = math.log(ercrd[maxrg] / ercrd[minrg])
        factor = incre / float(np.float32(maxrg - minrg))
        for iduv in range(minrg, maxrg):
            inc = factor * float(np.float32(iduv - minrg))
            ercrd[iduv] = ercrd[minrg] + inc if regn <= RGLMAX else ercrd[minrg] * math.exp(inc)
    return ercrd[1:pemax + 1].copy(), pemax, pesoft, uprgcd[nreg]


def _ecdinfo_line(path, pti):
    """The EcdInfo line of species pti (first column), after a comment line."""
    with open(path) as fh:
        lines = fh.readlines()
    for ln in lines[1:]:
        try:
            (q,) = ListDirected([ln]).read([int])
        except (EOFError, ValueError):
            continue
        if q == pti:
            kinds = [int] + [float] * 8 + ([] if pti == 0 else [float, int])
            return ListDirected([ln]).read(kinds)
    return None


def species_meshes(parameters_path, numslv: int, ecdinfo_path=None, ecdmesh_path=None):
    """uvmax, uvsoft, uvcrd (species 1..numslv concatenated) and escrd
    (species 0, its soft part), as enginit builds them."""
    h: dict = {}
    uvmax, uvsoft, crd = [], [], []
    escrd = None
    for pti in range(numslv + 1):
        nml = read_namelist(parameters_path, "hist")
        for k in _HIST_REAL:
            if k in nml:
                h[k] = float(nml[k].replace("d", "e").replace("D", "E"))
        for k in _HIST_INT:
            if k in nml:
                h[k] = int(nml[k])
        h.setdefault("ecprread", 0)
        h.setdefault("meshread", 0)
        if h.get("peread", 0) == 1:
            h["ecprread"] = 1
        missing = [k for k in _HIST_REAL + ("pecore",) if k not in h]
        if missing:
            raise InputError(f"&hist lacks {missing}")
        if h["ecprread"] == 1:
            if ecdinfo_path is None:
                raise InputError("EcdInfo requested but absent", "eng_eci")
            v = _ecdinfo_line(ecdinfo_path, pti)
            if v is not None:
                if pti == 0:
                    h["pecore"] = 0
                else:
                    h["ecdmax"], h["pecore"] = v[9], v[10]
                    if h["pecore"] <= 1:
                        raise InputError("EcdInfo pecore", "eng_pcr")
                (h["eclbin"], h["ecfbin"], h["ec0bin"], h["finfac"], h["ecdmin"], h["ecfmns"], h["ecdcen"],
                 h["eccore"]) = v[1:9]
                if h["eccore"] < 1.0e-20:
                    h["pecore"] = 0
                if h["pecore"] == 1:
                    raise InputError("EcdInfo pecore", "eng_pcr")
        # the reference reuses the build loop's index as the EcdMesh line
        # counter: a species without a block passes the count check only when
        # that index ended at pemax (a core region), as erdst's does
        mesh, pemax, pesoft, mesh_iduv = _build_mesh(h)
        if h["meshread"] == 1:
            if ecdmesh_path is None:
                raise InputError("EcdMesh requested but absent", "eng_ecm")
            with open(ecdmesh_path) as fh:
                lines = fh.readlines()
            start = True
            mesh = np.concatenate([mesh, np.zeros(LARGE)])
            for ln in lines:
                try:
                    (q,) = ListDirected([ln]).read([int])
                except (EOFError, ValueError):
                    continue
                if q != pti:
                    continue
                if start:
                    start = False
                    mesh_iduv = 0
                    if pti == 0:
                        _, pemax = ListDirected([ln]).read([int, int])
                        pecore = 0
                    else:
                        _, pemax, pecore = ListDirected([ln]).read([int, int, int])
                    pesoft = pemax - pecore
                else:
                    mesh_iduv += 1
                    mesh[mesh_iduv - 1] = ListDirected([ln]).read([int, int, float])[2]
            ok = mesh_iduv == pemax and pemax > pemax - pesoft
            ok = ok and all(mesh[i] < mesh[i + 1] for i in range(pemax - 1))
            if not ok:
                raise InputError(f"EcdMesh block of species {pti}", "eng_emf")
            mesh = mesh[:pemax]
        if pti == 0:
            escrd = mesh[:pesoft].copy()
        else:
            uvmax.append(pemax)
            uvsoft.append(pesoft)
            crd.append(mesh[:pemax])
    return (np.array(uvmax, np.int32), np.array(uvsoft, np.int32), np.concatenate(crd) if crd else np.zeros(0),
            escrd)

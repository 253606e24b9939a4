"""engstore's normalisation and file writers for runs driven from Python.

The Fortran host (fortran/erdst_host.f90) writes the same files from the
same er_fetch results; this module is its counterpart for the Python
binding and the input side of slvfe.  Formats follow erdst/engproc.F90:
  engsln.XX / engref.XX  '(g15.7,g15.7,i5,g25.15,g25.15)' left, middle,
                         species, ratio, ratio*invwidth   (:683-693)
  corsln.XX / corref.XX  one unformatted record of ermax^2 float64 (:705-709)
  weight_soln / weight_refs  '(i5,e20.8)' or '(i5,e20.8,e20.7)' (:641-656)
  aveuv.tt               '(i5,9999f15.5)' per block (soln only, :631-639)
  uvrange.tt             (:658-669)
"""
from __future__ import annotations

import math

import numpy as np


def fortran_g(x: float, w: int, d: int) -> str:
    """Fortran Gw.d output of one value (F editing with 4 trailing blanks
    inside 0.1 <= |x| < 10**d, E editing outside)."""
    x = float(x)
    if x != 0.0 and math.isfinite(x):
        ax = abs(x)
        # decide on the value rounded to d significant digits
        r = float(f"{ax:.{d - 1}e}")
        if 0.1 <= r < 10.0 ** d:
            k = int(math.floor(math.log10(r))) + 1
            s = f"{x:.{max(d - k, 0)}f}"
            return s.rjust(w - 4) + "    "
    elif x == 0.0:
        return f"{0.0:.{d - 1}f}".rjust(w - 4) + "    "
    return fortran_e(x, w, d)


def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d: 0.ddddE+ee."""
    x = float(x)
    if x == 0.0:
        return (f"0.{'0' * d}E+00").rjust(w)
    if not math.isfinite(x):
        return ("Infinity" if x > 0 else "-Infinity").rjust(w)
    e = int(math.floor(math.log10(abs(x)))) + 1
    m = abs(x) / 10.0 ** e
    digits = round(m * 10 ** d)
    if digits >= 10 ** d:
        digits //= 10
        e += 1
    sign = "-" if x < 0 else ""
    body = f"{sign}0.{digits:0{d}d}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    return body.rjust(w)


def bin_info(uvmax, uvsoft, uvcrd):
    """(left, middle, invwidth, species) of every bin,
    representative_bin_info erdst/engproc.F90:1355-1387."""
    uvmax = np.asarray(uvmax)
    uvsoft = np.asarray(uvsoft)
    uvcrd = np.asarray(uvcrd, float)
    ermax = int(uvmax.sum())
    left = uvcrd.copy()
    mid = np.zeros(ermax)
    inv = np.zeros(ermax)
    spec = np.zeros(ermax, np.int64)
    off = 0
    for p, n in enumerate(uvmax):
        for i in range(n):
            g = off + i
            idpt = i + 1
            spec[g] = p + 1
            if idpt <= uvsoft[p]:
                mid[g] = (uvcrd[g] + uvcrd[g + 1]) / 2.0
                inv[g] = 1.0 / (uvcrd[g + 1] - uvcrd[g])
            elif idpt < n:
                mid[g] = math.sqrt(uvcrd[g] * uvcrd[g + 1])
                inv[g] = 1.0 / (uvcrd[g + 1] - uvcrd[g])
            else:
                mid[g] = uvcrd[g]
                inv[g] = 0.0
        off += n
    return left, mid, inv, spec


def normalised(fetch: dict):
    """engstore's normalisation of one block (erdst/engproc.F90:609-630):
    edens/engnorm, ecorr/engnorm, and the two weight-file columns."""
    en = fetch["engnorm"]
    out = dict(edens=fetch["edens"] / en, ecorr=None if fetch.get("ecorr") is None else fetch["ecorr"] / en,
               avediv1=en / fetch["engsmpl"])
    return out


def self_weight_column(fetch: dict, kT: float, soln: bool) -> float:
    """avediv(:,2) = voffset -/+ kT log(avslf/engnorm), erdst/engproc.F90:622-628."""
    avslf = fetch["avslf"] / fetch["engnorm"]
    return fetch["voffset"] - kT * math.log(avslf) if soln else fetch["voffset"] + kT * math.log(avslf)


def write_engfile(path, edens_norm, uvmax, uvsoft, uvcrd):
    left, mid, inv, spec = bin_info(uvmax, uvsoft, uvcrd)
    with open(path, "w") as fh:
        fh.write("# " + "bin_left".rjust(13) + "bin_middle".rjust(15) + "spec".rjust(5)
                 + "ratio".rjust(25) + "dens".rjust(25) + "\n")
        for i in range(len(edens_norm)):
            fh.write(fortran_g(left[i], 15, 7) + fortran_g(mid[i], 15, 7) + f"{spec[i]:5d}"
                     + fortran_g(edens_norm[i], 25, 15) + fortran_g(edens_norm[i] * inv[i], 25, 15) + "\n")


def write_corfile(path, ecorr_norm):
    """One Fortran unformatted sequential record (4-byte length markers),
    column-major."""
    a = np.asarray(ecorr_norm, np.float64)
    body = np.asfortranarray(a).ravel(order="F").tobytes()
    n = np.array([len(body)], np.int32).tobytes()
    with open(path, "wb") as fh:
        fh.write(n + body + n)


def write_weights(path, avediv1, avediv2=None):
    with open(path, "w") as fh:
        for k, a in enumerate(avediv1, 1):
            line = f"{k:5d}" + fortran_e(a, 20, 8)
            if avediv2 is not None:
                line += fortran_e(avediv2[k - 1], 20, 7)
            fh.write(line + "\n")


def write_aveuv(path, aveuv_blocks):
    with open(path, "w") as fh:
        for k, row in enumerate(aveuv_blocks, 1):
            fh.write(f"{k:5d}" + "".join(f"{v:15.5f}" for v in np.atleast_1d(row)) + "\n")


def suffix(division: int, engdiv: int) -> str:
    return "tt" if engdiv == 1 else f"{division:02d}"


class BlockWriter:
    """Accumulates per-block results and writes the erdst output set into a
    directory, as engstore does at each block end."""

    def __init__(self, directory, uvmax, uvsoft, uvcrd, kT: float, soln: bool, engdiv: int,
                 wgtslf: bool = True, corrcal: bool = True):
        self.dir, self.kT, self.soln, self.engdiv = directory, kT, soln, engdiv
        self.mesh = (np.asarray(uvmax), np.asarray(uvsoft), np.asarray(uvcrd, float))
        self.wgtslf, self.corrcal = wgtslf, corrcal
        self.avediv1, self.avediv2, self.aveuv = [], [], []

    def store(self, fetch: dict):
        import os
        division = len(self.avediv1) + 1
        nrm = normalised(fetch)
        self.avediv1.append(nrm["avediv1"])
        self.avediv2.append(self_weight_column(fetch, self.kT, self.soln))
        if self.soln:
            self.aveuv.append(np.asarray(fetch["slnuv"]) / fetch["engnorm"])
        sfx = suffix(division, self.engdiv)
        stem = "engsln" if self.soln else "engref"
        write_engfile(os.path.join(self.dir, f"{stem}.{sfx}"), nrm["edens"], *self.mesh)
        if self.corrcal and nrm["ecorr"] is not None:
            write_corfile(os.path.join(self.dir, f"{'corsln' if self.soln else 'corref'}.{sfx}"), nrm["ecorr"])
        if division == self.engdiv:
            write_weights(os.path.join(self.dir, "weight_soln" if self.soln else "weight_refs"),
                          self.avediv1, self.avediv2 if self.wgtslf else None)
            if self.soln:
                write_aveuv(os.path.join(self.dir, "aveuv.tt"), self.aveuv)

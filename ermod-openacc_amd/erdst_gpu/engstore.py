"""engstore's normalisation and file writers for runs driven from Python.

The Fortran host (fortran/erdst_host.f90) writes the same files from the
same er_fetch results; this module is its counterpart for the Python
binding and the input side of slvfe.  Formats follow erdst/engproc.F90:
  engsln.XX / engref.XX  '(g15.7,g15.7,i5,g25.15,g25.15)' left, middle,
                         species, ratio, ratio*invwidth   (:683-693)
  corsln.XX / corref.XX  one unformatted record of ermax^2 float64 (:705-709)
  weight_soln / weight_refs  '(i5,e20.8)' or '(i5,e20.8,e20.7)' (:641-656)
  aveuv.tt               '(i5,9999f15.5)' per block (soln only, :631-639)
  uvrange.tt             '(i5,2f15.5)' or '(i5,f15.5,g18.5)' per species (:658-669)
  slfeng.XX              '(g15.7,g25.15)' self-energy distribution (selfcal, :712-722)
  flcuv.tt               '(i9,9999f15.5)' or '(2i9,9999f15.5)' per frame and
                         solute molecule (soln, :485-499); progress.tt, the
                         frame number list-directed (refs, :504-505)
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
            s = f"{x:#.{max(d - k, 0)}f}"   # F editing keeps the point with no decimals
            return s.rjust(w - 4) + "    "
    elif x == 0.0:
        return f"{x:.{d - 1}f}".rjust(w - 4) + "    "     # keeps the sign of -0.0, as G editing does
    return fortran_e(x, w, d)


def fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d: w asterisks when the value does not fit."""
    x = float(x)
    if not math.isfinite(x):
        s = "NaN" if x != x else ("Infinity" if x > 0 else "-Infinity")
        if len(s) > w:
            s = "Inf" if x > 0 else "-Inf"
        return s.rjust(w) if len(s) <= w else "*" * w
    s = f"{x:.{d}f}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d: 0.ddddE+ee."""
    x = float(x)
    if x == 0.0:
        return (("-" if math.copysign(1.0, x) < 0 else "") + f"0.{'0' * d}E+00").rjust(w)
    if not math.isfinite(x):
        return ("Infinity" if x > 0 else "-Infinity").rjust(w)
    # d significant digits of the exact binary value, correctly rounded
    mant, ex = f"{abs(x):.{d - 1}e}".split("e")
    digits = int(mant.replace(".", ""))
    e = int(ex) + 1
    sign = "-" if x < 0 else ""
    if abs(e) <= 99:
        body = f"{sign}0.{digits:0{d}d}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:                                # a three-digit exponent replaces the E
        body = f"{sign}0.{digits:0{d}d}{'+' if e >= 0 else '-'}{abs(e):03d}"
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
            if idpt <= uvsoft[p] and g + 1 < ermax:
                # (for the last bin of a species without a core region this is
                # the next species' first point, as the reference reads it)
                mid[g] = (uvcrd[g] + uvcrd[g + 1]) / 2.0
                inv[g] = 1.0 / (uvcrd[g + 1] - uvcrd[g])
            elif idpt < n:
                mid[g] = math.sqrt(uvcrd[g] * uvcrd[g + 1])
                inv[g] = 1.0 / (uvcrd[g + 1] - uvcrd[g])
            else:
                # the open top bin; also the last bin overall when its species
                # has no core region, where the reference reads past the mesh
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
    """aveuv.tt; aveuv is real (float32) in the reference (erdst/engmain.F90:339)."""
    with open(path, "w") as fh:
        for k, row in enumerate(aveuv_blocks, 1):
            row = np.atleast_1d(np.asarray(row, np.float64)).astype(np.float32).astype(np.float64)
            fh.write(f"{k:5d}" + "".join(fortran_f(v, 15, 5) for v in row) + "\n")


def write_uvrange(path, minuv, maxuv):
    """uvrange.tt: species 0 (the solute self energy) to numslv, minimum and
    maximum of the energies binned over the run (erdst/engproc.F90:658-669);
    minuv/maxuv are real (float32) in the reference."""
    mn = np.asarray(minuv, np.float32).astype(np.float64)
    mx = np.asarray(maxuv, np.float32).astype(np.float64)
    with open(path, "w") as fh:
        fh.write(" species     minimum        maximum\n")
        for p in range(mn.size):
            if mx[p] < 1.0e5:
                fh.write(f"{p:5d}" + fortran_f(mn[p], 15, 5) + fortran_f(mx[p], 15, 5) + "\n")
            else:
                fh.write(f"{p:5d}" + fortran_f(mn[p], 15, 5) + fortran_g(mx[p], 18, 5) + "\n")


def selfenergy_bin_coords(escrd):
    """representative_bin_selfenergy (erdst/engproc.F90:1344-1353): the bin
    middle (escrd(i) + escrd(i+1))/2 below the top bin, escrd(esmax) for the
    top bin, formed in float64 and stored as real (float32)."""
    e = np.asarray(escrd, np.float64)
    out = np.empty_like(e)
    out[:-1] = (e[:-1] + e[1:]) / 2.0
    out[-1] = e[-1]
    return out.astype(np.float32)


def write_slfeng(path, eself_norm, escrd):
    """slfeng.XX: the solute self-energy distribution eself/engnorm per bin
    (erdst/engproc.F90:712-722)."""
    coord = selfenergy_bin_coords(escrd).astype(np.float64)
    with open(path, "w") as fh:
        for i, v in enumerate(np.asarray(eself_norm, np.float64)):
            fh.write(fortran_g(coord[i], 15, 7) + fortran_g(v, 25, 15) + "\n")


def flcuv_lines(frame_label: int, flceng, indices=None, maxdst: int | None = None) -> list[str]:
    """flcuv.tt lines of one frame of a soln run (erdst/engproc.F90:485-499):
    per solute molecule the sum of its pair energies with each solvent
    species; the solute index is written only when the run has more than
    one solute molecule.  flceng [rows, numslv] is stored as real (float32)
    like the reference's accumulator; indices (0-based, default all) name
    the solute molecule of each row when some were rejected this frame, and
    maxdst is the run's solute-molecule count (default: the rows)."""
    fl = np.atleast_2d(np.asarray(flceng, np.float64)).astype(np.float32).astype(np.float64)
    idx = list(range(fl.shape[0])) if indices is None else list(indices)
    maxdst = fl.shape[0] if maxdst is None else maxdst
    vals = lambda row: "".join(fortran_f(v, 15, 5) for v in row)
    if maxdst == 1:
        return [f"{frame_label:9d}" + vals(fl[r]) for r in range(len(idx))]
    return [f"{k + 1:9d}{frame_label:9d}" + vals(fl[r]) for r, k in enumerate(idx)]


def progress_line(frame_label: int) -> str:
    """progress.tt line of one frame of a refs run (erdst/engproc.F90:504-505):
    a list-directed write of the frame number, laid out as the Fortran
    runtime of the host driver writes it (flang: one blank, then the digits)."""
    return f" {frame_label:d}"


def suffix(division: int, engdiv: int) -> str:
    return "tt" if engdiv == 1 else f"{division:02d}"


class BlockWriter:
    """Accumulates per-block results and writes the erdst output set into a
    directory, as engstore does at each block end."""

    def __init__(self, directory, uvmax, uvsoft, uvcrd, kT: float, soln: bool, engdiv: int,
                 wgtslf: bool = True, corrcal: bool = True, escrd=None):
        self.dir, self.kT, self.soln, self.engdiv = directory, kT, soln, engdiv
        self.mesh = (np.asarray(uvmax), np.asarray(uvsoft), np.asarray(uvcrd, float))
        self.wgtslf, self.corrcal = wgtslf, corrcal
        self.escrd = None if escrd is None else np.asarray(escrd, float)
        self.avediv1, self.avediv2, self.aveuv = [], [], []
        self.nframe = 0

    def frame(self, flceng=None, skpcnf: int = 1, indices=None, maxdst=None):
        """After each frame's er_run, as engconst does (erdst/engproc.F90:
        308-315, 485-509): soln runs append flceng [nsolute, numslv] to
        flcuv.tt, refs runs the frame number to progress.tt.  The reference
        opens both files with status='new' at its first frame."""
        import os
        self.nframe += 1
        label = self.nframe * skpcnf
        name = os.path.join(self.dir, "flcuv.tt" if self.soln else "progress.tt")
        if self.nframe == 1 and os.path.exists(name):
            raise FileExistsError(name)
        lines = flcuv_lines(label, flceng, indices, maxdst) if self.soln else [progress_line(label)]
        with open(name, "a") as fh:
            fh.write("".join(ln + "\n" for ln in lines))

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
        if self.escrd is not None:
            write_slfeng(os.path.join(self.dir, f"slfeng.{sfx}"), np.asarray(fetch["eself"]) / fetch["engnorm"],
                         self.escrd)
        if division == self.engdiv:
            if fetch.get("minuv") is not None:
                write_uvrange(os.path.join(self.dir, "uvrange.tt"), fetch["minuv"], fetch["maxuv"])
            write_weights(os.path.join(self.dir, "weight_soln" if self.soln else "weight_refs"),
                          self.avediv1, self.avediv2 if self.wgtslf else None)
            if self.soln:
                write_aveuv(os.path.join(self.dir, "aveuv.tt"), self.aveuv)

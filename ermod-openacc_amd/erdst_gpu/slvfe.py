"""slvfe: solvation free energy from the energy-representation distributions.

A numpy restatement of ERmod's slvfe program (slvfe/slvfe.F90,
slvfe/sfemain.F90), the step after erdst: it reads the solution
distribution (engsln), the reference-solvent density and correlation
matrix (engref, corref), optionally merges blocks, and evaluates the
energy-representation free-energy functional for each mesh coarsening
("group").  slvfe is built in double precision (the real-as-real(8) flag
plus -DDP of slvfe/configure.ac:46-81), so everything here is float64.

The default &fevars settings (slvfe/sfemain.F90:24-88) are those of
FeVars below: invmtrx='reg', zerosft='eczr', functional='py-hnc',
normalize='yes', wgtfnform='harm', extsln='lin', infchk='not' (five groups,
pickgr=3).  ljlrc='yes' (the LJ long-range correction of
slvfe/sfecorrect.F90) is not restated and raises.

Entry points:
  chmpot(dist, fv, group, inft)  one free-energy evaluation (one block, one
                                 mesh grouping), slvfe/slvfe.F90:586-873
  run_range(dist, fv)            all groupings of one block ('range')
  run_merge(soln_dir, refs_dir, fv)  the block loop of sfemain with
                                 datread's file handling ('merge')
  main()                         `python -m erdst_gpu.slvfe`, reading
                                 parameters_fe like the reference
"""
from __future__ import annotations

import math
import os
import re
import sys
from dataclasses import dataclass, field, fields, replace

import numpy as np


class SlvfeError(RuntimeError):
    """A condition on which the reference program stops."""


@dataclass
class FeVars:
    """The &fevars namelist with its defaults (slvfe/sfemain.F90:24-88)."""
    clcond: str = "merge"
    uvread: str = "yes"
    slfslt: str = "yes"
    ljlrc: str = "not"
    infchk: str = "not"
    meshread: str = "not"
    cumuint: str = "not"
    write_mesherror: str = "cnd"
    extsln: str = "lin"
    slncor: str = "not"
    refmerge: str = "yes"
    readwgtfl: str = "yes"
    invmtrx: str = "reg"
    zerosft: str = "eczr"
    wrtzrsft: str = "not"
    wgtfnform: str = "harm"
    wgtf2smpl: str = "yes"
    normalize: str = "yes"
    showdst: str = "not"
    functional: str = "py-hnc"
    numprm: int = 0
    numsln: int = 0
    numref: int = 0
    numdiv: int = 0
    inptemp: float = 300.0
    avevolume: float = 0.0
    et: float = 0.0
    pickgr: int = 3
    msemin: int = 1
    msemax: int = 5
    mesherr: float = 0.1
    extthres_soln: int = 1
    extthres_refs: int = 1
    minthres_soln: int = 0
    minthres_refs: int = 0
    norm_error: float = 1.0e-8
    itrmax: int = 100
    ermax_limit: int = 15000
    tiny: float = 1.0e-8
    large: int = 500000
    solndirec: str = "soln"
    refsdirec: str = "refs"
    wgtslnfl: str = "weight_soln"
    wgtreffl: str = "weight_refs"
    slndnspf: str = "engsln"
    slncorpf: str = "corsln"
    refdnspf: str = "engref"
    refcorpf: str = "corref"
    aveuvfile: str = "aveuv.tt"
    engmeshfile: str = "EngMesh"
    cumuintfl: str = "cumsfe"

    def resolved(self) -> "FeVars":
        """Apply init_sysvars' derived defaults (slvfe/sfemain.F90:151-167)."""
        fv = replace(self)
        if fv.numprm <= 0:
            fv.numprm = 11 if fv.infchk == "yes" else 5
        if fv.pickgr < fv.msemin:
            raise SlvfeError("pickgr must not be below msemin")
        if fv.pickgr > fv.msemax:
            raise SlvfeError("pickgr must not exceed msemax")
        if fv.pickgr > fv.numprm:
            raise SlvfeError("pickgr must not exceed numprm")
        if fv.et != 0.0:
            fv.functional = "thnc"
        if fv.ljlrc == "yes":
            raise SlvfeError("ljlrc='yes' (LJ long-range correction) is not supported")
        if fv.invmtrx not in ("gce", "reg", "evd"):
            fv.invmtrx = "reg"       # getinscv's first-call default, slvfe/slvfe.F90:1014-1026
        return fv


def kt_of(temp_kelvin: float) -> float:
    """kT in kcal/mol, slvfe/slvfe.F90:199."""
    return temp_kelvin * 8.314510e-3 / 4.184


def group_table(numprm: int, infchk: str):
    """(group, inft) for each parameter set of 'range'/'merge',
    slvfe/slvfe.F90:168-196."""
    out = []
    for prmcnt in range(1, numprm + 1):
        if infchk == "yes":
            tab = {1: (1, 0), 2: (1, 60), 3: (1, 80), 4: (1, 100), 5: (2, 0), 6: (3, 0), 7: (4, 0),
                   8: (5, 0), 9: (5, 60), 10: (5, 80), 11: (5, 100), 12: (8, 0)}
            out.append(tab[prmcnt] if prmcnt <= 12 else (10 + (prmcnt - 13) * 5, 0))
        else:
            out.append((prmcnt if prmcnt <= 10 else 10 + (prmcnt - 10) * 2, 0))
    return out


def mesh_layout(rdcrd: np.ndarray, rdspec: np.ndarray, tiny: float):
    """Per-species bin counts and log-core lengths from the bin_middle
    column, slvfe/slvfe.F90:104-139 (a run of equal spacings resets the core
    count: the last non-linear stretch is the core)."""
    numslv = int(rdspec.max())
    rduvmax = np.zeros(numslv, np.int64)
    rduvcore = np.zeros(numslv, np.int64)
    k = 0
    crdprev = crddif_prev = 0.0
    for iduv in range(len(rdcrd)):
        pti = int(rdspec[iduv])
        crdnow = float(rdcrd[iduv])
        if pti != k:
            rduvmax[pti - 1] = 1
            rduvcore[pti - 1] = 0
            k = pti
        else:
            rduvmax[pti - 1] += 1
            crddif_now = crdnow - crdprev
            if rduvmax[pti - 1] > 2:
                if abs(crddif_now - crddif_prev) < tiny:
                    rduvcore[pti - 1] = 0
                else:
                    rduvcore[pti - 1] += 1
            crddif_prev = crddif_now
        crdprev = crdnow
    return rduvmax, rduvcore


@dataclass
class Distributions:
    """What datread collects for one block (slvfe/slvfe.F90:292-446).

    rdcrd: bin_middle of every bin; rddst: solution distribution (engsln
    'ratio' column); rddns: reference density (engref 'ratio'); rdcor:
    reference correlation matrix (corref), rdslc: solution correlation
    (corsln, only with slncor='yes'); rdspec: 1-based solvent species of
    each bin; aveuv: average solute-solvent energy per species; slfeng:
    solute self energy; nummol: molecules per species (taken from the
    rddst sums when None)."""
    rdcrd: np.ndarray
    rddst: np.ndarray
    rddns: np.ndarray
    rdcor: np.ndarray
    rdspec: np.ndarray
    aveuv: np.ndarray | None = None
    slfeng: float = 0.0
    rdslc: np.ndarray | None = None
    nummol: np.ndarray | None = None
    leftbin: np.ndarray | None = None

    @property
    def numslv(self) -> int:
        return int(self.rdspec.max())

    def checked_nummol(self, tiny: float, previous: np.ndarray | None = None) -> np.ndarray:
        """The normalisation check of datread (slvfe/slvfe.F90:416-433)."""
        nm = np.zeros(self.numslv)
        for pti in range(1, self.numslv + 1):
            m = self.rdspec == pti
            factor = float(np.sum(self.rddst[m]))
            ampl = float(np.sum(self.rddns[m]))
            bad = abs(factor - ampl) > tiny
            if previous is not None:
                bad |= _nint(factor) != _nint(previous[pti - 1]) or _nint(ampl) != _nint(previous[pti - 1])
            if bad:
                raise SlvfeError(f"species {pti}: the normalised distribution does not sum to its molecule count")
            nm[pti - 1] = float(_nint(factor))
        return nm if previous is None else previous


def _nint(x: float) -> int:
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


@dataclass
class Grouped:
    """The coarsened mesh and distributions that chmpot works on."""
    uvcrd: np.ndarray
    edist: np.ndarray
    edens: np.ndarray
    ecorr: np.ndarray
    edscr: np.ndarray | None
    uvspec: np.ndarray
    uvmax: np.ndarray
    idrduv: np.ndarray

    @property
    def gemax(self) -> int:
        return len(self.uvcrd)


def regroup(d: Distributions, rduvmax, rduvcore, group: int, inft: int, slncor: bool) -> Grouped:
    """Merge `group` consecutive bins (and inft% of the log core into the
    last bin) per species, slvfe/slvfe.F90:614-703."""
    numslv = len(rduvmax)
    ermax = len(d.rdcrd)
    uvmax = np.zeros(numslv, np.int64)
    idrduv = np.zeros(ermax, np.int64)     # 0-based coarse index
    gemax = 0
    for pti in range(numslv):
        k = int(rduvcore[pti]) * inft // 100
        m = (int(rduvmax[pti]) - k) // group
        if m < 1:
            raise SlvfeError(f"group {group} leaves no bins for species {pti + 1}")
        uvmax[pti] = m
        cnt = int(np.sum(rduvmax[:pti]))
        j = int(np.sum(uvmax[:pti]))
        for iduv in range(cnt, cnt + int(rduvmax[pti])):
            kk = (iduv - cnt) // group + 1
            if kk >= m:
                kk = m
            idrduv[iduv] = kk + j - 1
        gemax += m
    if int(np.sum(rduvmax)) != ermax:
        raise SlvfeError("the bin count read from the distribution files differs from the one given")
    uvcrd = np.zeros(gemax)
    gpnum = np.zeros(gemax, np.int64)
    uvspec = np.zeros(gemax, np.int64)
    for iduv in range(ermax):
        k = idrduv[iduv]
        uvcrd[k] += d.rdcrd[iduv]
        gpnum[k] += 1
        uvspec[k] = d.rdspec[iduv]
    nz = gpnum > 0
    uvcrd[nz] /= gpnum[nz]
    # the last coarse bin of a species keeps the last fine coordinate
    cnt = np.cumsum(rduvmax) - 1
    kk = np.cumsum(uvmax) - 1
    uvcrd[kk] = d.rdcrd[cnt]
    edist = np.zeros(gemax)
    edens = np.zeros(gemax)
    np.add.at(edist, idrduv, d.rddst)
    np.add.at(edens, idrduv, d.rddns)
    P = np.zeros((ermax, gemax))
    P[np.arange(ermax), idrduv] = 1.0
    ecorr = P.T @ d.rdcor @ P
    edscr = P.T @ d.rdslc @ P if slncor else None
    return Grouped(uvcrd, edist, edens, ecorr, edscr, uvspec, uvmax, idrduv)


def distnorm(g: Grouped, nummol: np.ndarray, fv: FeVars, slncor: bool) -> None:
    """Scale each species' distribution to its molecule count and iterate
    the correlation matrix to the matching marginals, slvfe/slvfe.F90:1467-1542
    (a Gauss-Seidel sweep: each correc(iduv) uses the ones updated before it
    in the same sweep)."""
    numslv = len(nummol)
    for cnt in (1, 2):
        edhst = (g.edist if cnt == 1 else g.edens).copy()
        edmcr = None
        if cnt == 2:
            edmcr = g.ecorr.copy()
        elif slncor:
            edmcr = g.edscr.copy()
        for pti in range(1, numslv + 1):
            m = g.uvspec == pti
            factor = float(np.sum(edhst[m]))
            factor = nummol[pti - 1] / factor if factor > 0.0 else 0.0
            edhst[m] = factor * edhst[m]
        if edmcr is not None:
            masks = [g.uvspec == pti for pti in range(1, numslv + 1)]
            errtmp = fv.norm_error + 1.0
            itrcnt = 0
            correc = np.ones(g.gemax)
            while errtmp > fv.norm_error and itrcnt <= fv.itrmax:
                for iduv in range(g.gemax):
                    lcsln = 0.0
                    col = edmcr[:, iduv]
                    for pti in range(numslv):
                        m = masks[pti]
                        ampl = float(np.dot(correc[m], col[m]))
                        if ampl > 0.0:
                            lcsln += nummol[pti] / ampl
                    lcsln /= numslv
                    correc[iduv] = lcsln * edhst[iduv]
                edmcr = np.outer(correc, correc) * edmcr
                pos = edhst > 0.0
                errtmp = float(np.max(np.abs(correc[pos] - 1.0))) if pos.any() else 0.0
                itrcnt += 1
                if itrcnt >= fv.itrmax:
                    raise SlvfeError(f"distnorm stopped after {itrcnt} sweeps with the marginals still "
                                     f"off by {errtmp}")
        if cnt == 1:
            g.edist = edhst
            if slncor:
                g.edscr = edmcr
        else:
            g.edens = edhst
            g.ecorr = edmcr


class _State:
    """Module-level arrays of sfecalc that the helper functions share."""

    def __init__(self, g: Grouped, nummol, fv: FeVars, kT: float):
        self.g, self.nummol, self.fv, self.kT = g, nummol, fv, kT
        n = g.gemax
        numslv = len(nummol)
        self.slncv = np.zeros(n)
        self.inscv = np.zeros(n)
        self.sdrcv = np.zeros(n)
        self.zrsln = np.zeros(numslv)
        self.zrref = np.zeros(numslv)
        self.zrsdr = np.zeros(numslv)

    # wgtdst, slvfe/slvfe.F90:1381-1419 (0-based iduv)
    def wgtdst(self, iduv, cnt, systype, wgttype):
        fsln, fref = self.g.edist[iduv], self.g.edens[iduv]
        jdg = 1
        if cnt == 2:
            if systype in ("slncv", "extsl"):
                jdg = 9
            if wgttype == "smpl":
                jdg = 2
            if self.fv.wgtf2smpl == "yes":
                jdg = 2
        if jdg not in (1, 2):
            raise SlvfeError("internal error: weight function asked for an unknown system or weight type")
        if jdg == 2:
            return fref
        if wgttype == "smpl":
            if cnt == 2:
                raise SlvfeError("internal error: weight function asked for an unknown system or weight type")
            return fsln
        if wgttype == "geom":
            f = fsln * fref
            return math.sqrt(f) if f > 0.0 else 0.0
        f = fsln + fref
        return fsln * fref / f if f > 0.0 else 0.0

    # getwght, slvfe/slvfe.F90:1281-1316
    def getwght(self, pti, cnt, systype, wgttype, engtype):
        g = self.g
        weight = np.zeros(g.gemax)
        sp = g.uvspec == pti
        for iduv in np.nonzero(sp)[0]:
            weight[iduv] = self.wgtdst(iduv, cnt, systype, wgttype)
        if engtype == "yes":
            m = sp & (weight > 0.0)
            minuv = float(np.min(np.abs(g.uvcrd[m])))
            ampl = self.nummol[pti - 1] * (np.abs(g.uvcrd[sp]) - minuv)
            weight[sp] = np.exp(-ampl / self.kT) * weight[sp]
        ampl = float(np.sum(weight[sp]))
        if not ampl > 0.0:
            raise SlvfeError(f"species {pti}: every bin has zero weight")
        return weight / ampl

    # cvfcen, slvfe/slvfe.F90:1243-1279
    def cvfcen(self, pti, cnt, systype, wgttype, engtype, inscv_now=None):
        weight = self.getwght(pti, cnt, systype, wgttype, engtype)
        sp = self.g.uvspec == pti
        if systype == "slncv":
            if cnt == 2:
                raise SlvfeError("internal error: central-value average asked for an unknown kind")
            cv = self.slncv
        elif systype == "inscv":
            cv = inscv_now if inscv_now is not None else (self.sdrcv if cnt == 1 else self.inscv)
        elif systype == "uvcrd":
            cv = self.g.uvcrd
        else:
            raise SlvfeError("internal error: central-value average asked for an unknown kind")
        return float(np.sum(cv[sp] * weight[sp]))

    # zeroec, slvfe/slvfe.F90:1318-1355: bin nearest to zero energy
    def zeroec(self, pti, cnt):
        g, tiny = self.g, self.fv.tiny
        dist = g.edist if cnt == 1 else g.edens
        k = None
        for iduv in range(g.gemax - 1):
            if g.uvspec[iduv] != pti:
                continue
            if g.uvcrd[iduv] <= 0.0 and g.uvcrd[iduv + 1] >= 0.0:
                factor = abs(g.uvcrd[iduv])
                ampl = g.uvcrd[iduv + 1]
                if ampl > factor + tiny:
                    k = iduv
                elif ampl < factor - tiny:
                    k = iduv + 1
                elif abs(ampl - factor) <= tiny:
                    k = iduv if dist[iduv] >= dist[iduv + 1] else iduv + 1
                else:
                    raise SlvfeError("internal error: zero-energy bin search asked for an unknown system")
        if k is None:
            raise SlvfeError(f"no energy-coordinate bin brackets zero for species {pti}")
        return k

    def wgtmxco(self, pti):
        return 1.0 / self.nummol[pti - 1]


def getslncv(st: _State, d: Distributions) -> None:
    """Solution-side indirect potential with extrapolation into poorly
    sampled bins, slvfe/slvfe.F90:875-1001."""
    g, fv, kT = st.g, st.fv, st.kT
    n = g.gemax
    min_rddst = float(np.min(d.rddst[d.rddst > 0.0]))
    min_rddns = float(np.min(d.rddns[d.rddns > 0.0]))
    ext_target = np.ones(n, bool)
    ofdmp = 10
    for iduv in range(n):
        factor = g.edist[iduv] / min_rddst
        m = ofdmp * fv.extthres_soln
        j = m if factor > m else _nint(factor)
        factor = g.edens[iduv] / min_rddns
        m = ofdmp * fv.extthres_refs
        k = m if factor > m else _nint(factor)
        if j < fv.extthres_soln or k < fv.extthres_refs:
            ext_target[iduv] = False
    with np.errstate(divide="ignore", invalid="ignore"):
        t = ext_target
        st.slncv[t] = -kT * np.log(g.edist[t] / g.edens[t]) - g.uvcrd[t]
    for iduv in range(n):
        if ext_target[iduv]:
            continue
        if g.edist[iduv] <= 0.0:
            st.slncv[iduv] = 0.0
            continue
        pti = g.uvspec[iduv]
        same = (g.uvspec == pti) & ext_target
        # nearest extrapolation targets below/above (defaults: first/last bin)
        below = np.nonzero(same[:iduv])[0]
        above = np.nonzero(same[iduv + 1:])[0]
        m = int(below[-1]) if below.size else 0
        k = iduv + 1 + int(above[0]) if above.size else n - 1
        if fv.extsln == "sim":
            if abs(m - iduv) < abs(k - iduv):
                factor = st.slncv[m]
            elif abs(m - iduv) > abs(k - iduv):
                factor = st.slncv[k]
            else:
                factor = (st.slncv[m] + st.slncv[k]) / 2.0
        else:
            if ext_target[m] and ext_target[k]:
                j = m if abs(m - iduv) < abs(k - iduv) else k
            elif ext_target[m]:
                j = m
            elif ext_target[k]:
                j = k
            else:
                raise SlvfeError(f"no sampled bin to extrapolate from at energy {g.uvcrd[iduv]}")
            work = np.zeros(n)
            for iduvp in np.nonzero(same)[0]:
                if iduvp == iduv:
                    raise SlvfeError("extrapolation found no usable partner bin (inconsistent data)")
                factor = g.uvcrd[iduvp] - g.uvcrd[j]
                if iduvp < iduv:
                    factor = -factor - 2.0 * (g.uvcrd[j] - g.uvcrd[iduv])
                with np.errstate(over="ignore"):
                    work[iduvp] = math.exp(-factor / kT) if -factor / kT < 709.0 else math.inf
                work[iduvp] *= st.wgtdst(iduvp, 1, "extsl", fv.wgtfnform)
            with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
                factor = float(np.sum(work[work > 0.0]))
                work = work / factor
                pos = work > 0.0
                mat11 = float(np.sum(work[pos] * g.uvcrd[pos]))
                mat22 = float(np.sum(work[pos] * g.uvcrd[pos] * g.uvcrd[pos]))
                mat12 = float(np.sum(work[pos] * st.slncv[pos]))
                mat21 = float(np.sum(work[pos] * g.uvcrd[pos] * st.slncv[pos]))
                det = mat22 - mat11 ** 2
                a0 = (mat22 * mat12 - mat11 * mat21) / det
                a1 = (mat21 - mat11 * mat12) / det
            factor = a0 + a1 * g.uvcrd[iduv]
        st.slncv[iduv] = factor
    for pti in range(1, len(st.nummol) + 1):
        zs = fv.zerosft
        if zs in ("eczr", "orig"):
            cvzero = 0.0
        elif zs == "mxco":
            factor = st.wgtmxco(pti)
            cvzero = factor * st.cvfcen(pti, 1, "slncv", fv.wgtfnform, "not") \
                - (1.0 - factor) * st.cvfcen(pti, 2, "uvcrd", "smpl", "yes")
        elif zs == "zero":
            cvzero = st.cvfcen(pti, 1, "slncv", fv.wgtfnform, "yes")
        elif zs == "cntr":
            cvzero = st.cvfcen(pti, 1, "slncv", fv.wgtfnform, "not")
        else:
            raise SlvfeError("unknown zerosft choice")
        st.slncv[g.uvspec == pti] -= cvzero
        st.zrsln[pti - 1] = cvzero


def posv(mat: np.ndarray, vec: np.ndarray):
    """Cholesky solve from the upper triangle with posv_wrap's acceptance
    tests (slvfe/slvfe.F90:464-534): max |A x - b| <= 1e-6 and max |x| <= 1e6.
    Returns (x, info)."""
    from scipy.linalg import cho_factor, cho_solve, LinAlgError
    try:
        c = cho_factor(mat, lower=False, check_finite=False)
    except (LinAlgError, ValueError):
        return None, 1
    x = cho_solve(c, vec, check_finite=False)
    if not np.all(np.isfinite(x)):
        return None, 1
    residual = np.abs(mat @ x - vec)
    if residual.max() > 1.0e-6 or np.abs(x).max() > 1.0e6:
        return x, 1
    return x, 0


def getinscv(st: _State) -> dict:
    """Indirect potentials from the inverse of the density-fluctuation
    matrix, slvfe/slvfe.F90:1003-1232.  Returns which solver each system
    used."""
    g, fv, kT = st.g, st.fv, st.kT
    numslv = len(st.nummol)
    used = {}
    for cnt in (1, 2):
        if cnt == 1 and fv.slncor != "yes":
            continue
        mode = fv.invmtrx
        edvec = (g.edist if cnt == 1 else g.edens).copy()
        edmcr = (g.edscr if cnt == 1 else g.ecorr).copy()
        pos = edvec > 0.0
        both = np.outer(pos, pos)
        edmcr = np.where(both, edmcr - np.outer(edvec, edvec), 0.0)
        idx = np.nonzero(~pos)[0]
        edmcr[idx, idx] = 1.0
        ddiff = -kT * (g.edist - g.edens)
        x = None
        if mode in ("gce", "reg"):
            A = edmcr.copy()
            work = ddiff.copy()
            work[~pos] = 0.0
            if mode == "gce":
                for pti in range(1, numslv + 1):
                    k = st.zeroec(pti, cnt)
                    A[k, :] = 0.0
                    A[:, k] = 0.0
                    A[k, k] = 1.0
                    work[k] = 0.0
            else:
                for pti in range(1, numslv + 1):
                    sel = pos & (g.uvspec == pti)
                    regcnt = float(np.count_nonzero(sel))
                    regfac = float(np.sum(np.diag(edmcr)[sel])) / regcnt / regcnt
                    A[np.ix_(sel, sel)] += regfac
            x, info = posv(A, work)
            if info != 0:
                print(" the regularised correlation matrix is not positive definite; "
                      "solving by eigendecomposition instead", file=sys.stderr)
                mode = "evd"
                x = None
        if mode == "evd":
            w, v = np.linalg.eigh(edmcr, UPLO="U")
            p0 = numslv
            coef = np.zeros(g.gemax)
            for iduv in range(p0, g.gemax):
                coef[iduv] = float(np.sum(ddiff[pos] * v[pos, iduv])) / w[iduv]
            x = v[:, p0:] @ coef[p0:]
        used["soln" if cnt == 1 else "refs"] = mode
        # reference value for the additive constant
        zerouv = np.zeros(numslv)
        for pti in range(1, numslv + 1):
            if fv.zerosft in ("eczr", "orig"):
                zerouv[pti - 1] = x[st.zeroec(pti, cnt)]
            elif fv.zerosft == "mxco":
                zerouv[pti - 1] = st.cvfcen(pti, cnt, "inscv", "smpl", "yes", inscv_now=x)
        # solute-solvent potential -> indirect PMF
        with np.errstate(divide="ignore", invalid="ignore"):
            cv = np.where(pos, ddiff / np.where(pos, edvec, 1.0) - x, 0.0)
        if cnt == 1:
            st.sdrcv = cv
        else:
            st.inscv = cv
        for pti in range(1, numslv + 1):
            zs = fv.zerosft
            if zs in ("eczr", "orig"):
                cvzero = -zerouv[pti - 1]
            elif zs == "mxco":
                factor = st.wgtmxco(pti)
                cvzero = factor * st.cvfcen(pti, cnt, "inscv", fv.wgtfnform, "not") \
                    - (1.0 - factor) * zerouv[pti - 1]
            elif zs == "zero":
                cvzero = st.cvfcen(pti, cnt, "inscv", fv.wgtfnform, "yes")
            elif zs == "cntr":
                cvzero = st.cvfcen(pti, cnt, "inscv", fv.wgtfnform, "not")
            else:
                raise SlvfeError("unknown zerosft choice")
            sp = g.uvspec == pti
            if cnt == 1:
                st.sdrcv[sp] -= cvzero
                st.zrsdr[pti - 1] = cvzero
            else:
                st.inscv[sp] -= cvzero
                st.zrref[pti - 1] = cvzero
    return used


def sfewgt(fsln: float, fref: float) -> float:
    """Mixing weight of the solution and reference functionals,
    slvfe/slvfe.F90:1422-1434."""
    if fsln >= fref:
        return 1.0
    factor = (fsln - fref) / (fsln + fref)
    return 1.0 - factor ** 2


def pyhnc(indpmf: float, cnt: int, kT: float) -> float:
    """PY below zero, HNC above, slvfe/slvfe.F90:1438-1465."""
    factor = indpmf / kT
    if cnt in (1, 2):
        if factor < 0.0:
            if cnt == 1:
                intg = factor + factor / (math.exp(-factor) - 1.0)
            else:
                intg = math.log(1.0 - factor) * (1.0 / factor - 1.0)
            return intg + 1.0
        return factor / 2.0
    if cnt == 3:
        if factor > 0.0:
            return 1.0 - math.log(1.0 + factor) * (1.0 / factor + 1.0)
        return -factor / 2.0
    raise SlvfeError("py-hnc switch called with a system index other than 1 or 2")


def thnc(indpmf: float, et: float, cnt: int, kT: float) -> float:
    """Truncated HNC, slvfe/slvfe.F90:1414-1436 (et = -huge gives HNC)."""
    factor = indpmf / kT
    if cnt == 1:
        return factor / 2.0
    if cnt == 2:
        return (0.0 if indpmf < et else factor) / 2.0
    raise SlvfeError("hnc switch called with a system index other than 1 or 2")


@dataclass
class ChmResult:
    """chmpot's output for one block and one grouping."""
    chmpt: np.ndarray            # [0] total, [1:] per solvent species (kcal/mol)
    aveuv: np.ndarray
    zrsln: np.ndarray
    zrref: np.ndarray
    zrsdr: np.ndarray
    solver: dict
    grouped: Grouped = field(repr=False)
    slncv: np.ndarray = field(repr=False)
    inscv: np.ndarray = field(repr=False)
    cumsfe: np.ndarray | None = field(default=None, repr=False)


def chmpot(d: Distributions, fv: FeVars, group: int, inft: int = 0, *, kT: float | None = None,
           nummol: np.ndarray | None = None, rduvmax=None, rduvcore=None) -> ChmResult:
    """One solvation-free-energy evaluation, slvfe/slvfe.F90:586-873."""
    fv = fv.resolved()
    kT = kt_of(fv.inptemp) if kT is None else kT
    slncor = fv.slncor == "yes"
    if slncor and d.rdslc is None:
        raise SlvfeError("slncor='yes' needs the solution correlation matrix")
    if rduvmax is None:
        rduvmax, rduvcore = mesh_layout(d.rdcrd, d.rdspec, fv.tiny)
    if nummol is None:
        nummol = d.nummol if d.nummol is not None else d.checked_nummol(fv.tiny)
    numslv = len(rduvmax)
    g = regroup(d, rduvmax, rduvcore, group, inft, slncor)
    if fv.normalize == "yes":
        distnorm(g, nummol, fv, slncor)
    st = _State(g, nummol, fv, kT)
    getslncv(st, d)
    solver = getinscv(st)
    if fv.uvread != "not" and d.aveuv is None:
        raise SlvfeError("uvread='yes' needs the average solute-solvent energy (aveuv)")
    if fv.uvread == "not":
        aveuv = np.array([float(np.sum((g.uvcrd * g.edist)[g.uvspec == p])) for p in range(1, numslv + 1)])
    else:
        aveuv = np.asarray(d.aveuv, float)
    soln_zero = fv.minthres_soln * float(np.min(d.rddst[d.rddst > 0.0]))
    refs_zero = fv.minthres_refs * float(np.min(d.rddns[d.rddns > 0.0]))
    func = fv.functional.lower()
    if func not in ("py-hnc", "hnc", "none", "thnc"):
        raise SlvfeError("unknown functional")
    chmpt = np.zeros(numslv + 1)
    cumsfe = np.zeros(g.gemax)
    for pti in range(1, numslv + 1):
        uvpot = 0.0
        slvfe = 0.0
        for iduv in range(g.gemax):
            if g.uvspec[iduv] != pti:
                continue
            ed, en = g.edist[iduv], g.edens[iduv]
            if not (ed <= soln_zero and en <= refs_zero):
                uvpot += g.uvcrd[iduv] * ed
                slvfe -= kT * (ed - en)
                lcent = -(st.slncv[iduv] + st.zrsln[pti - 1] + g.uvcrd[iduv])
                special = slncor and ed > soln_zero and en <= refs_zero
                if special:
                    ampl = lcent * en / ed
                    lcent = ampl - (st.zrsln[pti - 1] + g.uvcrd[iduv]) * (1.0 - en / ed)
                slvfe += lcent * ed
                if func == "py-hnc":
                    lcsln = pyhnc(st.slncv[iduv], 1, kT)
                    lcref = pyhnc(st.inscv[iduv], 2, kT)
                elif func == "hnc":
                    lcsln = thnc(st.slncv[iduv], -sys.float_info.max, 1, kT)
                    lcref = thnc(st.inscv[iduv], -sys.float_info.max, 2, kT)
                elif func == "none":
                    lcsln = lcref = 0.0
                else:
                    lcsln = thnc(st.slncv[iduv], fv.et, 1, kT)
                    lcref = thnc(st.inscv[iduv], fv.et, 2, kT)
                if special:
                    lcsln = pyhnc(st.sdrcv[iduv] + st.zrsdr[pti - 1], 3, kT)
                ampl = sfewgt(ed, en)
                factor = ampl * lcsln + (1.0 - ampl) * lcref
                slvfe += kT * factor * (ed - en)
            cumsfe[iduv] = uvpot + slvfe
        chmpt[pti] = slvfe + aveuv[pti - 1]
    chmpt[0] = float(np.sum(chmpt[1:]))
    if fv.slfslt == "yes":
        chmpt[0] += d.slfeng
    return ChmResult(chmpt, aveuv, st.zrsln, st.zrref, st.zrsdr, solver, g, st.slncv, st.inscv, cumsfe)


def run_range(d: Distributions, fv: FeVars, *, kT: float | None = None):
    """All groupings of one block (clcond='range'), slvfe/slvfe.F90:160-197.
    Returns [(group, inft, ChmResult)]."""
    fv = fv.resolved()
    rduvmax, rduvcore = mesh_layout(d.rdcrd, d.rdspec, fv.tiny)
    nummol = d.nummol if d.nummol is not None else d.checked_nummol(fv.tiny)
    return [(grp, inf, chmpot(d, fv, grp, inf, kT=kT, nummol=nummol, rduvmax=rduvmax, rduvcore=rduvcore))
            for grp, inf in group_table(fv.numprm, fv.infchk)]


# ---------------------------------------------------------------------------
# files

def read_engfile(path):
    """engsln/engref text (header + left, middle, spec, ratio, dens):
    returns (leftbin, middle, spec, ratio) as datread reads them
    (slvfe/slvfe.F90:369-391)."""
    rows = []
    with open(path) as fh:
        fh.readline()
        for ln in fh:
            if ln.strip():
                p = ln.split()
                rows.append((float(p[0]), float(p[1]), int(p[2]), float(p[3])))
    a = np.array(rows, dtype=object)
    return (np.array(a[:, 0], float), np.array(a[:, 1], float), np.array(a[:, 2], np.int64),
            np.array(a[:, 3], float))


def read_corfile(path, ermax: int) -> np.ndarray:
    """corref/corsln: one unformatted sequential record of ermax^2 float64,
    column-major (slvfe/slvfe.F90:392-401)."""
    raw = np.fromfile(path, dtype=np.uint8)
    n = ermax * ermax * 8
    if raw.size < 8:
        raise SlvfeError(f"{path}: empty correlation file")
    head = int(raw[:4].view(np.int32)[0])
    if head == n and raw.size == n + 8:
        body = raw[4:4 + n]
    else:
        # records longer than 2 GiB are split in subrecords (negative markers)
        parts, off = [], 0
        while off < raw.size:
            m = int(raw[off:off + 4].view(np.int32)[0])
            ln = abs(m)
            parts.append(raw[off + 4:off + 4 + ln])
            off += ln + 8
        body = np.concatenate(parts)
        if body.size != n:
            raise SlvfeError(f"{path}: record length {body.size} != {n}")
    return body.view(np.float64).reshape((ermax, ermax), order="F").copy()


def _suffix(i: int, tt: bool) -> str:
    return "tt" if tt else f"{i:02d}"


def _count_files(direc: str, trunk: str):
    """check_tt_or_numeric_files, slvfe/sfemain.F90:186-226."""
    if os.path.exists(os.path.join(direc, f"{trunk}.tt")):
        for i in range(1, 100):
            if os.path.exists(os.path.join(direc, f"{trunk}.{i:02d}")):
                raise SlvfeError(f"{trunk}.tt and numbered {trunk}.NN files are both present; keep one kind")
        return 1, True
    n = 0
    for i in range(1, 100):
        if os.path.exists(os.path.join(direc, f"{trunk}.{i:02d}")):
            n = i
        else:
            if i == 1:
                raise SlvfeError(f"{direc} holds no {trunk}.tt and no {trunk}.01")
            break
    return n, False


def _read_weights(path, n):
    out = []
    with open(path) as fh:
        for _ in range(n):
            p = fh.readline().split()
            out.append(float(p[1]))
    return np.array(out)


@dataclass
class MergeResult:
    """Per-block, per-grouping free energies and the summary statistics of
    wrtmerge (slvfe/slvfe.F90:1733-1868)."""
    groups: list                 # [(group, inft)]
    chmpt: np.ndarray            # [numslv+1, prmmax, numrun]
    blockuv: np.ndarray | None   # [numslv+1, numrun]
    slfeng: float
    grref: int

    @property
    def numrun(self) -> int:
        return self.chmpt.shape[2]

    def mean(self, prm: int | None = None, pti: int = 0) -> float:
        prm = self.grref if prm is None else prm
        return float(np.mean(self.chmpt[pti, prm, :]))

    def error95(self, prm: int | None = None, pti: int = 0) -> float:
        """2 x standard error of the block values, as wrtmerge computes it."""
        prm = self.grref if prm is None else prm
        v = self.chmpt[pti, prm, :]
        r = float(len(v))
        if len(v) < 2:
            return 0.0
        avecp = float(np.sum(v)) / r
        var = float(np.sum(v ** 2)) / r - avecp ** 2
        std = math.sqrt(r / (r - 1.0)) * math.sqrt(max(var, 0.0))
        return 2.0 * std / math.sqrt(r)

    @property
    def dmu(self) -> float:
        """Total solvation free energy at the picked grouping, averaged over blocks."""
        return self.mean()


def run_merge(fv: FeVars, base: str = ".", log=None) -> MergeResult:
    """clcond='merge': block loop over engsln/engref/corref files with the
    weight files, slvfe/slvfe.F90:62-73,229-286,292-446 and sfemain's loop."""
    fv = fv.resolved()
    log = log or (lambda s: None)
    soln = os.path.join(base, fv.solndirec)
    refs = os.path.join(base, fv.refsdirec)
    nsln_files, sln_tt = _count_files(soln, fv.slndnspf)
    nref_files, ref_tt = _count_files(refs, fv.refdnspf)
    numsln = fv.numsln if 0 < fv.numsln <= nsln_files else nsln_files
    numref = fv.numref if 0 < fv.numref <= nref_files else nref_files
    numdiv = fv.numdiv if 0 < fv.numdiv < numsln else numsln
    if numsln % numdiv != 0:
        i = numdiv + 1
        while i <= numsln and numsln % i != 0:
            i += 1
        numdiv = i
    if fv.refmerge == "not":
        if numdiv > numref:
            raise SlvfeError("refmerge='not' pairs solution blocks with reference blocks one to one, so numdiv may not exceed numref")
        numref = numdiv * (numref // numdiv)
    maxsln, maxref, numrun, prmmax = numsln, numref, numdiv, fv.numprm
    groups = group_table(prmmax, fv.infchk)

    path1 = os.path.join(soln, f"{fv.slndnspf}.{_suffix(1, sln_tt)}")
    _, rdcrd, rdspec, _ = read_engfile(path1)
    ermax = len(rdcrd)
    if ermax > fv.ermax_limit:
        raise SlvfeError("ermax is above the supported bin count")
    rduvmax, rduvcore = mesh_layout(rdcrd, rdspec, fv.tiny)
    numslv = len(rduvmax)
    kT = kt_of(fv.inptemp)

    uvread = fv.uvread
    uvene = None
    if uvread != "not":
        p = os.path.join(soln, fv.aveuvfile)
        if os.path.exists(p):
            uvene = np.zeros((numslv, maxsln))
            with open(p) as fh:
                for i in range(maxsln):
                    parts = fh.readline().split()
                    uvene[:, i] = [float(x) for x in parts[1:1 + numslv]]
        else:
            uvread = "not"
            log(f" Warning: Although the uvread parameter was set to 'yes', it is changed into 'not' "
                f"since the {p} file was not found")
    fv = replace(fv, uvread=uvread)

    wgtsln = np.ones(maxsln)
    wgtref = np.ones(maxref)
    if fv.readwgtfl == "yes":
        wgtsln = _read_weights(os.path.join(soln, fv.wgtslnfl), maxsln)
        wgtref = _read_weights(os.path.join(refs, fv.wgtreffl), maxref)
    wgtsln = wgtsln / np.sum(wgtsln)
    wgtref = wgtref / np.sum(wgtref)

    slfslt = fv.slfslt
    slfeng = 0.0
    if slfslt == "yes":
        if fv.readwgtfl == "not":
            raise SlvfeError("slfslt='yes' needs readwgtfl='yes' (the self energy comes from the weight file)")
        p = os.path.join(refs, fv.wgtreffl)
        if not os.path.exists(p):
            raise SlvfeError("slfslt='yes' but the weight_refs file is missing")
        with open(p) as fh:
            for i in range(maxref):
                parts = fh.readline().split()
                if len(parts) < 3:
                    slfslt = "not"
                    log(" Warning: Although the slfslt parameter was set to 'yes', it is changed into 'not'.")
                    break
                slfeng += wgtref[i] * float(parts[2])
    fv = replace(fv, slfslt=slfslt)
    slncor = fv.slncor == "yes"

    chmpt = np.zeros((numslv + 1, prmmax, numrun))
    blockuv = np.zeros((numslv + 1, numrun)) if uvread != "not" else None
    rddns = rdcor = None
    nummol = None
    for cntrun in range(1, numrun + 1):
        if maxsln >= numrun:
            k = maxsln // numrun
            slnini, slnfin = (cntrun - 1) * k + 1, cntrun * k
        else:
            slnini = slnfin = (cntrun - 1) % maxsln + 1
        if fv.refmerge == "not":
            if maxref >= numrun:
                m = maxref // numrun
                refini, reffin = (cntrun - 1) * m + 1, cntrun * m
            else:
                refini = reffin = (cntrun - 1) % maxref + 1
        else:
            refini, reffin = 1, maxref
        rddst = np.zeros(ermax)
        rdslc = np.zeros((ermax, ermax)) if slncor else None
        if cntrun == 1 or fv.refmerge == "not":
            rddns = np.zeros(ermax)
            rdcor = np.zeros((ermax, ermax))
        s = slice(slnini - 1, slnfin)
        wgtsln[s] = wgtsln[s] / np.sum(wgtsln[s])
        leftsln = None
        for i in range(slnini, slnfin + 1):
            lb, crd, spec, ratio = read_engfile(os.path.join(soln, f"{fv.slndnspf}.{_suffix(i, sln_tt)}"))
            leftsln = lb
            rddst += wgtsln[i - 1] * ratio
            rdspec = _species_run_index(spec)
            rdcrd = crd
            if slncor:
                rdslc += wgtsln[i - 1] * read_corfile(
                    os.path.join(soln, f"{fv.slncorpf}.{_suffix(i, sln_tt)}"), ermax)
        if cntrun == 1 or fv.refmerge == "not":
            s = slice(refini - 1, reffin)
            wgtref[s] = wgtref[s] / np.sum(wgtref[s])
            for i in range(refini, reffin + 1):
                lb, crd, spec, ratio = read_engfile(os.path.join(refs, f"{fv.refdnspf}.{_suffix(i, ref_tt)}"))
                if not np.array_equal(lb, leftsln):
                    raise SlvfeError("the solution and reference meshes differ")
                rddns += wgtref[i - 1] * ratio
                rdspec = _species_run_index(spec)
                rdcrd = crd
                rdcor += wgtref[i - 1] * read_corfile(
                    os.path.join(refs, f"{fv.refcorpf}.{_suffix(i, ref_tt)}"), ermax)
        d = Distributions(rdcrd, rddst, rddns, rdcor, rdspec, None, slfeng, rdslc)
        nummol = d.checked_nummol(fv.tiny, nummol)
        if cntrun == 1:
            for pti in range(numslv):
                log(f"  Number of the {pti + 1:3d}-th solvent  = {int(nummol[pti]):12d}")
        if uvread != "not":
            d.aveuv = np.array([float(np.sum(wgtsln[slnini - 1:slnfin] * uvene[p, slnini - 1:slnfin]))
                                for p in range(numslv)])
            blockuv[1:, cntrun - 1] = d.aveuv
            blockuv[0, cntrun - 1] = float(np.sum(d.aveuv)) + (slfeng if slfslt == "yes" else 0.0)
        for prmcnt, (grp, inf) in enumerate(groups):
            r = chmpot(d, fv, grp, inf, kT=kT, nummol=nummol, rduvmax=rduvmax, rduvcore=rduvcore)
            chmpt[:, prmcnt, cntrun - 1] = r.chmpt
    grref = next(i for i, (grp, _) in enumerate(groups) if grp == fv.pickgr)
    return MergeResult(groups, chmpt, blockuv, slfeng, grref)


def dmu_from_blocks(soln_blocks, refs_blocks, uvmax, uvsoft, uvcrd, kT_erdst: float, directory: str,
                    fv: FeVars | None = None) -> MergeResult:
    """The erdst -> slvfe hand-over for runs driven from Python: write both
    runs' block results in engstore's formats (engsln.XX + weight_soln +
    aveuv.tt under directory/soln, engref.XX + corref.XX + weight_refs
    under directory/refs) and run slvfe's merge over them.  Each block is
    a dict with edens, ecorr (refs), engnorm, engsmpl, avslf, voffset and
    slnuv (soln), as er_fetch returns them.  kT_erdst is erdst's temp."""
    from . import engstore
    fv = fv or FeVars()
    for sub, blocks, soln in ((fv.solndirec, soln_blocks, True), (fv.refsdirec, refs_blocks, False)):
        d = os.path.join(directory, sub)
        os.makedirs(d, exist_ok=True)
        w = engstore.BlockWriter(d, uvmax, uvsoft, uvcrd, kT_erdst, soln=soln, engdiv=len(blocks),
                                 wgtslf=not soln, corrcal=not soln)
        for b in blocks:
            w.store(b)
    return run_merge(fv, directory)


def _species_run_index(spec: np.ndarray) -> np.ndarray:
    """rdspec: the species column renumbered by runs (slvfe/slvfe.F90:383-390)."""
    out = np.zeros(len(spec), np.int64)
    k = m = 0
    for i, pti in enumerate(spec):
        if pti != k:
            k = pti
            m += 1
        out[i] = m
    return out


def format_merge(res: MergeResult, fv: FeVars) -> str:
    """The tables wrtmerge prints (slvfe/slvfe.F90:1733-1868) for infchk='not'."""
    fv = fv.resolved()
    numslv = res.chmpt.shape[0] - 1
    lines = []
    if res.blockuv is not None:
        lines += ["", "", " cumulative average & 95% error for solvation energy"]
        lines += _cumu_lines(res.blockuv, numslv)
        lines.append("")
    numrun = res.numrun
    for pti in range(numslv + 1):
        if numslv == 1 and pti != 0:
            continue
        avcp0 = res.mean(res.grref, pti)
        for prm, (grp, inf) in enumerate(res.groups):
            avecp = res.mean(prm, pti)
            if prm == 0:
                lines.append("")
                if pti == 0:
                    lines.append(" group    solvation free energy     difference" if numrun == 1 else
                                 " group    solvation free energy     error          difference")
                if numslv > 1:
                    lines.append("  total solvation free energy" if pti == 0 else
                                 f"  contribution from {pti:2d}-th solvent component")
            if numrun == 1:
                lines.append(f"{grp:4d}{avecp:20.5f}{avecp - avcp0:21.5f}")
            else:
                lines.append(f"{grp:4d}{avecp:20.5f}{res.error95(prm, pti):18.5f}{avecp - avcp0:18.5f}")
    lines += ["", "", " cumulative average & 95% error for solvation free energy"]
    lines += _cumu_lines(res.chmpt[:, res.grref, :], numslv)
    return "\n".join(lines)


def _cumu_lines(data: np.ndarray, numslv: int):
    """wrtcumu, slvfe/slvfe.F90:1870-1972."""
    out = []
    numrun = data.shape[1]
    runcp = np.zeros(numslv + 1)
    runer = np.zeros(numslv + 1)
    for cntrun in range(1, numrun + 1):
        r = float(cntrun)
        wr = np.zeros(2 * numslv + 2)
        for pti in range(numslv + 1):
            v = data[pti, cntrun - 1]
            runcp[pti] += v
            runer[pti] += v ** 2
            avecp = runcp[pti] / r
            wr[2 * pti] = avecp
            if cntrun > 1:
                f = runer[pti] / r - avecp ** 2
                wr[2 * pti + 1] = 0.0 if f <= 0.0 else (2.0 / math.sqrt(r)) * math.sqrt(r / (r - 1.0)) * math.sqrt(f)
        if cntrun == 1:
            if numslv == 1:
                out.append(f"{cntrun:3d}{wr[0]:11.4f}")
            else:
                out.append(f"{cntrun:3d}{wr[0]:11.4f}" + "".join(f"{wr[2 * p]:22.4f}" for p in range(1, numslv + 1)))
        else:
            if numslv == 1:
                out.append(f"{cntrun:3d}{wr[0]:11.4f}{wr[1]:11.4f}")
            else:
                out.append(f"{cntrun:3d}" + "".join(f"{x:11.4f}" for x in wr))
    return out


_NML_RE = re.compile(r"&\s*fevars(.*?)(?:/|&end)", re.IGNORECASE | re.DOTALL)


def read_parameters_fe(path: str) -> FeVars:
    """The &fevars namelist of parameters_fe (slvfe/sfemain.F90:98-104)."""
    fv = FeVars()
    if not os.path.exists(path):
        return fv
    text = open(path).read()
    m = _NML_RE.search(text)
    if not m:
        return fv
    body = re.sub(r"!.*", "", m.group(1))
    types = {f.name: f.type for f in fields(FeVars)}
    for key, val in re.findall(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\s]+)", body):
        k = key.lower()
        if k not in types:
            continue
        v = val.strip().strip("'\"")
        t = types[k]
        if t in ("int", int):
            setattr(fv, k, int(v))
        elif t in ("float", float):
            setattr(fv, k, float(v.lower().replace("d", "e")))
        elif t in ("bool", bool):
            setattr(fv, k, v.lower().strip(".") in ("true", "t"))
        else:
            setattr(fv, k, v)
    return fv


def main(argv=None) -> int:
    """`python -m erdst_gpu.slvfe` in a directory holding parameters_fe and
    the soln/ and refs/ folders (clcond='merge')."""
    fv = read_parameters_fe("parameters_fe")
    if fv.clcond != "merge":
        raise SlvfeError("only clcond='merge' runs from the command line; call chmpot()/run_range() "
                         "for 'basic' and 'range'")
    res = run_merge(fv, ".", log=print)
    if fv.slfslt == "yes" and res.slfeng != 0.0:
        print(f"  Self-energy of the solute   =   {res.slfeng:12.4f}  kcal/mol")
    print(format_merge(res, fv))
    return 0


if __name__ == "__main__":
    sys.exit(main())

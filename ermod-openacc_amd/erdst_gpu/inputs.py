"""erdst's input files, read the way erdst/setconf.F90 reads them.

* ``parameters_er``: the ``&ene_param`` and ``&hist`` namelists, with
  iniparam's defaults and checks (setconf.F90:133-438) and getscrn
  (:492-512).
* ``MDinfo``: frame count, species count, molecules and sites per species
  (:64-73).
* ``MolPrm<n>`` / ``SltInfo``: one line per site, in the new form
  ``index mass type name charge eps sigma [x y z]`` or the old form
  ``index type charge eps sigma [x y z]`` with the mass from the type name
  (:705-727, getmass :1016-1076); the LJ columns in any of the six ljformat
  conventions (:728-746); ``LJTable`` for ljformat 5 (:796-808).
* ``PermIndex``: trajectory-to-program atom order (:877-919).
* ``SysWght`` / ``SltWght``: per-frame and per-configuration weights
  (:988-1013; erdst/insertion.F90:480-490).

Values the reference holds as default ``real`` are float32 here and are
converted in float32; list-directed reads follow the Fortran rules the
reference relies on (values across records, commas or blanks, ``r*c``
repeats, null values, ``/`` ending a read).
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from .params import temperature_kt
from .topology import LJCMB_ARITH, Species, Topology, build_topology

f32 = np.float32

LJFMT_EPS_CAL_SGM_NM, LJFMT_EPS_RMINH, LJFMT_EPS_J_SGM_A, LJFMT_A_C, LJFMT_C12_C6, LJFMT_TABLE = range(6)
SLT_SOLN, SLT_REFS_RIGID, SLT_REFS_FLEX = 1, 2, 3
SGMCNV = f32(1.7817974362806784)     # Rmin/2 -> sigma (setconf.F90:536)
LENCNV = f32(10.0)                   # nm -> Angstrom
ENGCNV = f32(1.0) / f32(4.184)       # kJ/mol -> kcal/mol, a default-real constant


class InputError(ValueError):
    """A condition under which erdst stops; ``code`` is its halt code, if any."""

    def __init__(self, message, code=""):
        super().__init__(f"{message} ({code})" if code else message)
        self.code = code


# ---------------------------------------------------------------------------
# Fortran list-directed input
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"""\s*(?:'((?:[^']|'')*)'|"((?:[^"]|"")*)"|(/)|(,)|([^\s,/'"]+))""")
_NULL = object()


def _record_values(line: str):
    """Values of one record: strings (unquoted or quoted), _NULL for empty
    values between separators, and '/' as its own marker."""
    out, pos, last_was_value = [], 0, False
    line = line.rstrip("\n")
    while pos < len(line):
        m = _TOKEN.match(line, pos)
        if not m or m.end() == pos:
            break
        pos = m.end()
        q1, q2, slash, comma, word = m.groups()
        if slash:
            out.append("/")
            return out
        if comma:
            if not last_was_value:
                out.append(_NULL)
            last_was_value = False
            continue
        val = q1.replace("''", "'") if q1 is not None else q2.replace('""', '"') if q2 is not None else word
        if q1 is None and q2 is None:
            rep = re.fullmatch(r"(\d+)\*(.*)", word)
            if rep:
                n, v = int(rep.group(1)), rep.group(2)
                out.extend([_NULL if v == "" else v] * n)
                last_was_value = True
                continue
        out.append(val)
        last_was_value = True
    return out


def _convert(tok, kind):
    if kind is int:
        if not re.fullmatch(r"[+-]?\d+", tok):
            raise ValueError(f"not an integer: {tok!r}")
        return int(tok)
    if kind in (float, f32, np.float64):
        t = tok.replace("d", "e").replace("D", "E")
        if not re.fullmatch(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?", t):
            raise ValueError(f"not a real: {tok!r}")
        return f32(float(t)) if kind is f32 else float(t)
    if isinstance(kind, tuple) and kind[0] is str:     # (str, length)
        return tok[: kind[1]]
    if kind is str:
        return tok
    raise TypeError(kind)


class ListDirected:
    """Successive list-directed READ statements over the records of a file:
    each read starts at the next record and takes its values across records;
    a '/' ends it, leaving the remaining items null (None)."""

    def __init__(self, lines):
        self.lines = list(lines)
        self.pos = 0

    @classmethod
    def from_file(cls, path):
        with open(path) as fh:
            return cls(fh.readlines())

    def read(self, kinds, prev=None):
        """Values for ``kinds`` (int, float, f32, str or (str, length));
        prev supplies the values null items keep.  EOFError at end of file,
        ValueError on a value that does not convert."""
        out = list(prev) if prev is not None else [None] * len(kinds)
        k = 0
        while k < len(kinds):
            if self.pos >= len(self.lines):
                raise EOFError("end of file")
            vals = _record_values(self.lines[self.pos])
            self.pos += 1
            for v in vals:
                if v == "/":
                    return out
                if k >= len(kinds):
                    break
                if v is not _NULL:
                    out[k] = _convert(v, kinds[k])
                k += 1
        return out


# ---------------------------------------------------------------------------
# parameters_er (&ene_param, &hist)
# ---------------------------------------------------------------------------
def read_namelist(path_or_text, group: str) -> dict:
    """Name -> value(s) of one namelist group; values are strings or lists of
    strings as written (converted by the caller); array elements written as
    name(i) = v are returned under 'name' as {i: v}."""
    text = open(path_or_text).read() if os.path.exists(str(path_or_text)) else str(path_or_text)
    lines = [ln.split("!", 1)[0] if "'" not in ln and '"' not in ln else ln for ln in text.splitlines()]
    body = "\n".join(lines)
    m = re.search(r"[&$]" + group + r"\b(.*?)(?:^|[\s,])(?:/|[&$]end)", body, flags=re.S | re.I | re.M)
    if not m:
        raise InputError(f"namelist group &{group} not found")
    out: dict = {}
    items = re.split(r"([A-Za-z_][A-Za-z0-9_]*(?:\(\s*\d+\s*\))?)\s*=", m.group(1))
    for i in range(1, len(items), 2):
        name = items[i].strip().lower()
        vals = [v for v in _record_values(items[i + 1].replace("\n", " ")) if v is not _NULL]
        am = re.fullmatch(r"([a-z_][a-z0-9_]*)\(\s*(\d+)\s*\)", name)
        if am:
            out.setdefault(am.group(1), {})[int(am.group(2))] = vals[0] if len(vals) == 1 else vals
        else:
            out[name] = vals[0] if len(vals) == 1 else vals
    return out


def getscrn(ewtoler: float, elecut: float, scrtype: str = "relative") -> float:
    """Ewald screening parameter from the real-space tolerance by bisection
    (setconf.F90:492-512): erfc(screen*elecut) (divided by elecut for
    'distance') equals ewtoler to 1e-15.  The bar is below float32
    resolution, so the bisection runs in float64, as the reference's
    double-precision build does; a single-precision build never meets it."""
    error, factor, small, large = 1.0e-15, 1.0 + 1.0e-15, 0.0, 1.0e3
    scr = 0.0
    for _ in range(400):
        if factor <= error:
            break
        scr = (small + large) / 2.0
        factor = math.erfc(scr * elecut)
        if scrtype == "distance":
            factor /= elecut
        if factor > ewtoler:
            small = scr
        else:
            large = scr
        factor = abs(factor - ewtoler)
    return scr


_INT_KEYS = {"iseed", "skpcnf", "corrcal", "selfcal", "slttype", "wgtslf", "wgtsys", "wgtins", "boxshp", "estype",
             "sltspec", "ljformat", "ljswitch", "insorigin", "insposition", "insorient", "insstructure", "sltpick",
             "refpick", "inscnd", "inscfg", "engdiv", "maxins", "intprm", "cmbrule", "cltype", "splodr",
             "ew1max", "ew2max", "ew3max", "ms1max", "ms2max", "ms3max", "ermax_limit", "pecore"}
_STR_KEYS = {"scrtype"}
_LOGICAL_KEYS = {"force_calculation"}


def _typed(name, v):
    if isinstance(v, dict):
        return {k: _typed(name, x) for k, x in v.items()}
    if isinstance(v, list):
        return [_typed(name, x) for x in v]
    if name in _INT_KEYS:
        return _convert(v, int)
    if name in _STR_KEYS:
        return v
    if name in _LOGICAL_KEYS:
        return v.strip(".").lower().startswith("t")
    return float(_convert(v, f32))


def iniparam(ene_param: dict) -> dict:
    """&ene_param with iniparam's defaults (setconf.F90:133-438): the
    trajectory-reading defaults, then the namelist, then the defaults that
    depend on slttype and cltype, then the namelist again for what the user
    set explicitly; the temperature in kcal/mol; getscrn when screen <= 1e-20;
    the parameter checks, raising InputError with the halt code."""
    user = {k: _typed(k, v) for k, v in ene_param.items()}
    if "slttype" not in user:
        raise InputError("slttype must be given")
    p = dict(boxshp=1, estype=2, cltype=2, ms1max=64, inptemp=300.0, engdiv=1, elecut=12.0, screen=0.0,
             ewtoler=1.0e-5, splodr=6, scrtype="relative", sltpick=0, refpick=0, inscnd=0, inscfg=0,
             insposition=0, insorient=0, insstructure=0, ermax_limit=15000, force_calculation=False,
             skpcnf=1, wgtslf=0, wgtsys=0, wgtins=0, selfcal=0, sltspec=1, ljformat=1, ljswitch=0,
             lwreg=-1.0, upreg=-2.0, lwstr=-1.0, upstr=-2.0, cmbrule=LJCMB_ARITH, iseed=0, hostspec={},
             refspec={}, ew1max=0)
    # upljcut / lwljcut follow the default elecut, before the namelist
    p["upljcut"], p["lwljcut"] = 12.0, 10.0
    p.update(user)
    slttype, cltype = p["slttype"], p["cltype"]
    if slttype == SLT_SOLN:
        p["corrcal"], p["maxins"] = 0, 1
    elif slttype in (SLT_REFS_RIGID, SLT_REFS_FLEX):
        p["corrcal"], p["maxins"] = 1, 1000
        if cltype in (1, 2, 3):
            p["wgtslf"] = 1
    else:
        raise InputError("Unknown slttype")
    p["ew2max"] = p["ew3max"] = p["ew1max"]
    p["ms2max"] = p["ms3max"] = p["ms1max"]
    if p["sltpick"] > 0:
        p["sltspec"] = p["sltpick"]
    if p["refpick"] > 0:
        p["refspec"] = {1: p["refpick"]}
    p["insposition"] = {0: 0, 1: 2, 2: 4, 3: 5}.get(p["inscnd"], None)
    if p["insposition"] is None:
        raise InputError("Unknown inscnd")
    if p["inscfg"] == 0:
        p["insorient"] = 0
    elif p["inscfg"] == 1:
        p["insorient"] = 1
    elif p["inscfg"] == 2:
        p["insorient"], p["insposition"] = 1, 1
    else:
        raise InputError("Unknown inscfg")
    p.update(user)           # the second init_params: explicit values win
    if p["slttype"] == SLT_SOLN:
        p["maxins"] = 1
    p["temp"] = temperature_kt(p["inptemp"])
    if cltype in (1, 2, 3) and p["screen"] <= 1.0e-20:
        if p["ewtoler"] <= 1.0e-20:
            raise InputError("Ewald tolerance", "set_ewa")
        p["screen"] = float(f32(getscrn(p["ewtoler"], p["elecut"], p["scrtype"])))
    if cltype == 1:
        raise InputError("Ewald is not effective in the current version", "set_prs")
    if cltype in (2, 3) and p["ms1max"] * p["ms2max"] * p["ms3max"] == 0:
        raise InputError("PME mesh", "set_ewa")
    if cltype not in (0, 1, 2, 3):
        raise InputError("Unknown cltype")
    if p["ljformat"] not in range(6):
        raise InputError("Unknown ljformat")
    if p["ljswitch"] not in range(4):
        raise InputError("Unknown ljswitch")
    if p["slttype"] in (SLT_REFS_RIGID, SLT_REFS_FLEX) and p["maxins"] < 1:
        raise InputError("maxins", "set_ins")
    if (p["slttype"] == SLT_REFS_RIGID and p["insposition"] == 1 and p["insorient"] == 1
            and p["maxins"] != 1):
        raise InputError("a rigid solute used as is takes maxins = 1", "set_ins")
    if p["boxshp"] == 0 and (p["estype"] == 2 or p["cltype"] != 0):
        raise InputError("non-periodic system", "set_prs")
    if p["wgtins"] == 1 and p["slttype"] != SLT_REFS_FLEX:
        raise InputError("wgtins", "set_ins")
    return p


# ---------------------------------------------------------------------------
# MDinfo, molecule files, LJTable
# ---------------------------------------------------------------------------
@dataclass
class MDInfo:
    nrun: int
    ntype: int
    nmol: list
    nsite: list


def read_mdinfo(path) -> MDInfo:
    """MDinfo (setconf.F90:64-73): nrun ntype / nmol(1:ntype) / nsite(1:ntype)."""
    r = ListDirected.from_file(path)
    nrun, ntype = r.read([int, int])
    nmol = r.read([int] * ntype)
    nsite = r.read([int] * ntype)
    return MDInfo(nrun, ntype, nmol, nsite)


# getmass (setconf.F90:1016-1076): atomic weights by the leading letters of
# the type name; later, longer matches override earlier ones
_MASS1 = {"M": 0.0, "H": 1.00794, "D": 2.014102, "C": 12.0107, "O": 15.9994, "N": 14.00674, "S": 32.066,
          "P": 30.973761, "B": 10.811, "K": 39.0983, "F": 18.9984032, "I": 126.90447}
_MASS2 = {"He": 4.0026, "Si": 28.0855, "Li": 6.941, "Na": 22.989770, "Cl": 35.4527, "Br": 79.904,
          "Ca": 40.078, "Zn": 65.409, "Fe": 55.845, "Cu": 63.546}


def getmass(atmtype: str) -> np.float32:
    t = atmtype[:5]
    m = None
    if t[:1] in _MASS1:
        m = f32(_MASS1[t[:1]])
    if t[:2] in _MASS2:
        m = f32(_MASS2[t[:2]])
    if t[:2] == "CH":
        m = f32(_MASS1["C"]) + f32(_MASS1["H"])
    if t[:3] == "CH2":
        m = f32(_MASS1["C"]) + f32(2.0) * f32(_MASS1["H"])
    if t[:3] == "CH3":
        m = f32(_MASS1["C"]) + f32(3.0) * f32(_MASS1["H"])
    if m is None:
        raise InputError(f"no mass for atom type {atmtype!r} in the old line format")
    return m


@dataclass
class SiteRecords:
    mass: np.ndarray
    atmtype: list
    atmname: list
    charge: np.ndarray      # e, float32
    lj2: np.ndarray         # second LJ column as read, float32
    lj3: np.ndarray         # third LJ column as read, float32
    coord: np.ndarray | None  # rigid solute reference coordinates


def read_molecule_file(path, nsite: int, rigid: bool) -> SiteRecords:
    """nsite lines of MolPrm<n> / SltInfo (setconf.F90:705-727): each line
    (its first 120 characters) is read in the new form, and in the old form
    when that read fails; a rigid test particle carries three more values,
    its reference coordinates."""
    with open(path) as fh:
        lines = fh.readlines()
    if len(lines) < nsite:
        raise InputError(f"{path}: {nsite} site lines expected, {len(lines)} found")
    extra = [f32] * 3 if rigid else []
    mass, types, names, q, e, s, xyz = [], [], [], [], [], [], []
    prev_new = [None, None, None, None, f32(0), f32(0), f32(0)] + [f32(0)] * len(extra)
    for line in lines[:nsite]:
        buf = line.rstrip("\n")[:120]
        try:
            v = ListDirected([buf]).read([int, f32, (str, 12), (str, 8), f32, f32, f32] + extra, prev_new)
            if any(x is None for x in v[:7]):
                raise EOFError
            m, typ, name, rest = v[1], v[2], v[3], v[4:]
        except (EOFError, ValueError):
            v = ListDirected([buf]).read([int, (str, 12), f32, f32, f32] + extra)
            if any(x is None for x in v):
                raise InputError(f"{path}: cannot read site line {buf!r}")
            typ, name, rest = v[1], "", v[2:]
            m = getmass(typ)
        mass.append(m)
        types.append(typ)
        names.append(name)
        q.append(rest[0])
        e.append(rest[1])
        s.append(rest[2])
        if rigid:
            xyz.append(rest[3:6])
    return SiteRecords(np.array(mass, np.float32), types, names, np.array(q, np.float32),
                       np.array(e, np.float32), np.array(s, np.float32),
                       np.array(xyz, np.float32) if rigid else None)


def convert_lj(ljformat: int, lj2, lj3):
    """(epsilon kcal/mol, sigma Angstrom) from the two LJ columns of a
    molecule file (setconf.F90:728-746), in float32 as the reference computes
    them.  ljformat 5 returns the columns unchanged (the type index is the
    second one)."""
    e = np.array(lj2, np.float32)
    s = np.array(lj3, np.float32)
    if ljformat == LJFMT_EPS_RMINH:
        s = (SGMCNV * s).astype(np.float32)
    if ljformat in (LJFMT_A_C, LJFMT_C12_C6):
        for i in range(e.size):
            if s[i] != f32(0.0):
                factor = f32(np.power(f32(e[i] / s[i]), f32(1.0) / f32(6.0)))
                f6 = f32(factor * factor)
                f6 = f32(f32(f6 * f6) * f6)
                e[i], s[i] = f32(s[i] / f32(f32(4.0) * f6)), factor
            else:
                e[i] = f32(0.0)
    if ljformat in (LJFMT_EPS_J_SGM_A, LJFMT_C12_C6):
        e = (ENGCNV * e).astype(np.float32)
        s = (LENCNV * s).astype(np.float32)
    return e, s


def read_ljtable(path):
    """LJTable (setconf.F90:796-808): the type count, then ntype rows of
    sigma (squared on reading) and ntype rows of epsilon."""
    r = ListDirected.from_file(path)
    (n,) = r.read([int])
    sig = np.array([r.read([f32] * n) for _ in range(n)], np.float32)
    eps = np.array([r.read([f32] * n) for _ in range(n)], np.float32)
    sig64 = np.array([[float(x) for x in row] for row in sig])
    return n, (sig * sig).astype(np.float32), eps, sig64 * sig64, eps.astype(np.float64)


@dataclass
class SystemInput:
    """What setparam builds from the input files."""
    params: dict
    mdinfo: MDInfo
    topology: Topology
    mass: np.ndarray                 # per site of the physical system, float32
    solute_mass: np.ndarray          # per site of the solute / test particle
    bfcoord: np.ndarray | None       # rigid test particle reference coordinates
    species: list = field(default_factory=list)


def load_system(directory, params: dict) -> SystemInput:
    """setparam (setconf.F90:516-844) over MDinfo, MolPrm<n>, SltInfo and
    LJTable in ``directory``: species in MDinfo order (the refs test
    particle last, its site count from SltInfo's lines), molecule files read
    and converted, LJ types numbered in order of first appearance (or from
    LJTable), the combining rule applied, charges scaled by 18.22261721."""
    md = read_mdinfo(os.path.join(directory, "MDinfo"))
    slttype, ljformat = params["slttype"], params["ljformat"]
    refs = slttype != SLT_SOLN
    numtype = md.ntype + (1 if refs else 0)
    cnt, nsite = list(md.nmol), list(md.nsite)
    kind = [0] * md.ntype                      # PT_SOLVENT
    slt_path = os.path.join(directory, "SltInfo")
    if not refs:
        solute_index = params.get("sltspec", 1) - 1 if 1 <= params.get("sltspec", 1) <= numtype else 0
        kind[solute_index] = SLT_SOLN
    else:
        n = 0
        with open(slt_path) as fh:
            r = ListDirected(fh.readlines())
        while True:
            try:
                r.read([int])
            except (EOFError, ValueError):
                break
            n += 1
        cnt.append(1)
        nsite.append(n)
        kind.append(slttype)
        solute_index = numtype - 1
    species, records = [], []
    cur_solvent = 0
    for pti in range(numtype):
        if kind[pti] == 0:
            cur_solvent += 1
            path = os.path.join(directory, f"MolPrm{cur_solvent}")
        else:
            if cnt[pti] > 1:
                cur_solvent += 1
            path = slt_path
        rec = read_molecule_file(path, nsite[pti], rigid=(kind[pti] == SLT_REFS_RIGID))
        records.append(rec)
        e, s = convert_lj(ljformat, rec.lj2, rec.lj3)
        species.append(Species(name=os.path.basename(path), nmol=cnt[pti], charge=rec.charge, eps=e, sigma=s,
                               geometry=rec.coord))
    table = None
    if ljformat == LJFMT_TABLE:
        n, lensq, ene, lensq64, ene64 = read_ljtable(os.path.join(directory, "LJTable"))
        types = [np.trunc(rec.lj2).astype(np.int32) for rec in records]
        table = (types, lensq, ene, lensq64, ene64)
    topo = build_topology(species, solute_index, refs, ljformat=0, cmbrule=params.get("cmbrule", LJCMB_ARITH),
                          lj_table=table)
    phys = [k for k in range(numtype) if not (refs and k == solute_index)]
    mass = np.concatenate([np.tile(records[k].mass, cnt[k]) for k in phys]).astype(np.float32)
    return SystemInput(params=params, mdinfo=md, topology=topo, mass=mass, solute_mass=records[solute_index].mass,
                       bfcoord=records[solute_index].coord if kind[solute_index] == SLT_REFS_RIGID else None,
                       species=species)


# ---------------------------------------------------------------------------
# PermIndex, SysWght, SltWght
# ---------------------------------------------------------------------------
def read_permindex(path, natom: int) -> np.ndarray:
    """PermIndex (setconf.F90:877-919): lines 'i p' map the i-th atom of the
    trajectory to the p-th atom of the program (identity elsewhere); both in
    [1, natom], and the result must be a permutation.  Returns the 0-based
    map perm, with program position perm[i] for trajectory atom i."""
    perm = np.arange(1, natom + 1)
    r = ListDirected.from_file(path)
    while True:
        try:
            i, p = r.read([int, int])
        except EOFError:
            break
        if i is None or p is None:
            break
        if not 1 <= i <= natom:
            raise InputError(f"Incorrect range of particle number at {i}", "set_pmt")
        if not 1 <= p <= natom:
            raise InputError(f"{i}-th permutation points to {p}", "set_pmt")
        perm[i - 1] = p
    counts = np.bincount(perm - 1, minlength=natom)
    if np.any(counts != 1):
        raise InputError("atoms with other than one entry: " + str(np.flatnonzero(counts != 1)[:8] + 1), "set_pmt")
    return perm - 1


def apply_permutation(traj_xyz: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """sitepos(:, permutation(1:natom)) = OUTpos (setconf.F90:971-975)."""
    out = np.empty_like(traj_xyz)
    out[perm] = traj_xyz
    return out


class WeightFile:
    """SysWght (per trajectory frame, setconf.F90:988-1013) or SltWght (per
    solute configuration, rewound at its end like SltConf,
    erdst/insertion.F90:480-490): lines 'index weight', read in turn."""

    def __init__(self, path, rewind: bool):
        with open(path) as fh:
            self.lines = fh.readlines()
        self.rewind, self.path = rewind, path
        self.r = ListDirected(self.lines)

    def next(self) -> np.float32:
        for attempt in (0, 1):
            try:
                _, w = self.r.read([int, f32])
                if w is None:
                    raise ValueError
                return w
            except (EOFError, ValueError):
                if not self.rewind or attempt == 1:
                    raise InputError(f"The weight file ({os.path.basename(self.path)}) is ill formed "
                                     "or its length does not match")
                self.r = ListDirected(self.lines)
        raise AssertionError

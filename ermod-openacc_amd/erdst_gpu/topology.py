"""Topology arrays of module engmain, built the way setparam builds them.

Restates erdst/setconf.F90:516-844: sites of each species in order, LJ
types deduplicated by (sigma, epsilon) in order of first appearance
(setconf.F90:756-777), combining rules (setconf.F90:813-827; arithmetic
sigma by default, geometric epsilon), charges scaled by 18.22261721
(setconf.F90:834) and molecule charges as plain sums (setconf.F90:838-840).

Two precisions are produced from the same description: float32 arrays as the
reference's default (single) build holds them -- these feed the GPU library
-- and float64 arrays as its --enable-double build holds them.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

CHARGE_SCALE = 18.22261721          # setconf.F90:834
SGMCNV = 1.7817974362806784         # Rmin/2 -> sigma, setconf.F90:536

LJFMT_EPS_CAL_SGM_A = 0             # epsilon kcal/mol, sigma Angstrom (no conversion)
LJFMT_EPS_RMINH = 1                 # epsilon kcal/mol, Rmin/2 Angstrom
LJCMB_ARITH, LJCMB_GEOM = 0, 1


@dataclass
class Species:
    """One molecular species: per-site charge (e), epsilon, sigma/Rmin2."""
    name: str
    nmol: int
    charge: np.ndarray
    eps: np.ndarray
    sigma: np.ndarray
    geometry: np.ndarray | None = None   # reference site coordinates (rigid)

    @property
    def nsite(self) -> int:
        return int(len(self.charge))


@dataclass
class Topology:
    """engmain arrays.  Index arrays are 0-based here; the C ABI gets 1-based."""
    species: list[Species]
    solute_index: int                   # 0-based index into species
    refs: bool                          # True: solute is a test particle (last)
    # physical system (molecules present in the trajectory)
    natom: int = 0
    nmol: int = 0
    mol_begin: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    ljtype: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    charge32: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    charge64: np.ndarray = field(default_factory=lambda: np.zeros(0))
    uvspec: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    moltype: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    numslv: int = 0
    ntype: int = 0
    ljene32: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    ljlensq32: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    ljene64: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ljlensq64: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # test particle (refs)
    slt_ljtype: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    slt_charge32: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    slt_charge64: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # soln: solute molecule indices (0-based) in the physical system
    tagslt: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))


def build_topology(species: list[Species], solute_index: int, refs: bool,
                   ljformat: int = LJFMT_EPS_CAL_SGM_A, cmbrule: int = LJCMB_ARITH,
                   extra_lj_species: list[Species] | None = None, lj_table=None) -> Topology:
    """setparam for species listed in MDinfo order.  For refs runs the solute
    (test particle) must be the last species (setconf.F90:598-600).

    extra_lj_species: further species whose LJ types are appended to the type
    table after those of `species` (not part of the system), so that several
    test particles analysed in one context share one table.

    lj_table: ljformat 5 (setconf.F90:751-753, 794-808): (per-species 1-based
    type numbers, sigma^2 and epsilon matrices in float32 and float64); the
    types and matrices are taken as given."""
    if refs and solute_index != len(species) - 1:
        raise ValueError("the test particle must be the last species")
    t = Topology(species=species, solute_index=solute_index, refs=refs)
    # LJ type table in order of first appearance (setconf.F90:756-777)
    # (with extra species: the system's solvent species, then the extras in
    # their order, so the table does not depend on which extra is the solute)
    table: list[tuple[float, float]] = []

    def site_keys(sp):
        sig = np.asarray(sp.sigma, np.float32).copy()
        eps = np.asarray(sp.eps, np.float32)
        if ljformat == LJFMT_EPS_RMINH:
            sig = (np.float32(SGMCNV) * sig).astype(np.float32)
        return list(zip(sig.tolist(), eps.tolist()))

    if lj_table is not None:
        types, lensq32, ene32, lensq64, ene64 = lj_table
        ntype = int(np.asarray(lensq32).shape[0])
        site_types = [np.asarray(ty, np.int32) - 1 for ty in types]
        for ty in site_types:
            if ty.size and (ty.min() < 0 or ty.max() >= ntype):
                raise ValueError("LJ type outside the LJTable")
        t.ntype = ntype
        t.ljene32 = np.asarray(ene32, np.float32).ravel(order="F")
        t.ljlensq32 = np.asarray(lensq32, np.float32).ravel(order="F")
        t.ljene64 = np.asarray(ene64, np.float64).ravel(order="F")
        t.ljlensq64 = np.asarray(lensq64, np.float64).ravel(order="F")
        return _assemble(t, species, solute_index, refs, site_types)
    order = list(species)
    if extra_lj_species:
        order = [sp for i, sp in enumerate(species) if not (refs and i == solute_index)] + list(extra_lj_species)
        order += [species[solute_index]] if refs else []
    for sp in order:
        for key in site_keys(sp):
            if key not in table:
                table.append(key)
    site_types = [np.array([table.index(key) for key in site_keys(sp)], np.int32) for sp in species]
    ntype = len(table)
    sig = np.array([k[0] for k in table], np.float32)
    eps = np.array([k[1] for k in table], np.float32)
    if cmbrule == LJCMB_ARITH:
        lensq32 = (((sig[:, None] + sig[None, :]) / np.float32(2.0)) ** 2).astype(np.float32)
        s64 = sig.astype(np.float64)
        lensq64 = ((s64[:, None] + s64[None, :]) / 2.0) ** 2
    else:
        lensq32 = (sig[:, None] * sig[None, :]).astype(np.float32)
        lensq64 = sig.astype(np.float64)[:, None] * sig.astype(np.float64)[None, :]
    ene32 = np.sqrt(eps[:, None] * eps[None, :]).astype(np.float32)
    e64 = eps.astype(np.float64)
    ene64 = np.sqrt(e64[:, None] * e64[None, :])
    t.ntype = ntype
    # column-major [ntype, ntype]; symmetric, so ravel order is immaterial
    t.ljene32, t.ljlensq32 = ene32.ravel(order="F"), lensq32.ravel(order="F")
    t.ljene64, t.ljlensq64 = ene64.ravel(order="F"), lensq64.ravel(order="F")
    return _assemble(t, species, solute_index, refs, site_types)


def _assemble(t: Topology, species, solute_index, refs, site_types) -> Topology:
    """Per-atom and per-molecule arrays of the physical system (setparam's
    fill loop, setconf.F90:780-787, and engproc.F90:112-118 for uvspec)."""
    phys = [k for k in range(len(species)) if not (refs and k == solute_index)]
    mol_begin = [0]
    ljt, q32, q64, uvspec, moltype = [], [], [], [], []
    # uvspec: engproc.F90:112-118 (solute -> 0, later species slide down by one)
    numslt = 1 if refs else species[solute_index].nmol
    for k in phys:
        sp = species[k]
        if refs or numslt == 1:
            if k == solute_index:
                spec = 0
            else:
                spec = k + 1 - (1 if k > solute_index else 0)
        else:
            spec = k + 1
        qq32 = (np.float32(CHARGE_SCALE) * np.asarray(sp.charge, np.float32)).astype(np.float32)
        qq64 = CHARGE_SCALE * np.asarray(np.asarray(sp.charge, np.float32), np.float64)
        for _ in range(sp.nmol):
            mol_begin.append(mol_begin[-1] + sp.nsite)
            ljt.append(site_types[k])
            q32.append(qq32)
            q64.append(qq64)
            uvspec.append(spec)
            moltype.append(k + 1)
    t.mol_begin = np.array(mol_begin, np.int32)
    t.nmol = len(mol_begin) - 1
    t.natom = int(mol_begin[-1])
    t.ljtype = np.concatenate(ljt).astype(np.int32)
    t.charge32 = np.concatenate(q32).astype(np.float32)
    t.charge64 = np.concatenate(q64)
    t.uvspec = np.array(uvspec, np.int32)
    t.moltype = np.array(moltype, np.int32)
    t.numslv = int(t.uvspec.max())
    if refs:
        sp = species[solute_index]
        t.slt_ljtype = site_types[solute_index]
        t.slt_charge32 = (np.float32(CHARGE_SCALE) * np.asarray(sp.charge, np.float32)).astype(np.float32)
        t.slt_charge64 = CHARGE_SCALE * np.asarray(np.asarray(sp.charge, np.float32), np.float64)
    else:
        first = sum(species[k].nmol for k in phys if k < solute_index)
        t.tagslt = np.arange(first, first + species[solute_index].nmol, dtype=np.int32)
    return t

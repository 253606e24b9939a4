"""Shared helpers of the test-suite: build the oracle's view of a workload,
run the oracle, run the HIP library, and compare."""
from __future__ import annotations

import numpy as np

import oracle as O
from erdst_gpu import synth
from erdst_gpu.params import hot_path_params

# Tolerance of per-(insertion, molecule) energies of the FP32 GPU path
# against the FP64 oracle: |E_gpu - E_64| <= RTOL*|E_64| + ATOL.
# RTOL = 1e-5 is the relative bar for pair energies; ATOL = 2e-4 kcal/mol
# (= 1e-5 x a 20 kcal/mol floor) covers the absolute FP32 rounding of the
# ~200-term PME gather and of sums whose terms cancel.
RTOL = 1e-5
ATOL = 2e-4
# Above E_HIGH (the top of every mesh used here, so all such energies share
# the open-ended last bin) near-overlapping sites dominate: the FP32 distance
# of two sites at coordinates ~|x| has relative error ~eps32*|x|/r, and the
# r^-12 term multiplies it by 12, so the bar there is RTOL_HIGH.
E_HIGH = 1.0e4
RTOL_HIGH = 1e-3


def oracle_system(topo) -> O.OracleSystem:
    return O.OracleSystem(topo.mol_begin, topo.charge64, topo.ljtype,
                          topo.ljene64.reshape(topo.ntype, topo.ntype, order="F"),
                          topo.ljlensq64.reshape(topo.ntype, topo.ntype, order="F"))


def oracle_params(p: dict) -> O.Params:
    return O.make_params(elecut=p["elecut"], lwljcut=p["lwljcut"], upljcut=p["upljcut"],
                         screen=p["screen"], ljswitch=p["ljswitch"], splodr=p["splodr"], ms=p["ms"])


def workload_params(w: synth.Workload, **over) -> dict:
    kw = dict(splodr=4)
    kw.update(over)
    return hot_path_params(w.slttype, w.params["ms"], **kw)


def oracle_refs(w, f, params, nins=None, nthreads=0):
    topo = w.topo
    pos = w.frame(f).astype(np.float64)
    ins = w.insertions(f).astype(np.float64)
    if nins is not None:
        ins = ins[:nins]
    return O.energies_refs(oracle_system(topo), oracle_params(params), w.cell.astype(np.float64), pos,
                           topo.slt_charge64, topo.slt_ljtype, ins, rigid=(params["slttype"] == 2),
                           nthreads=nthreads)


def oracle_soln(w, f, params, nthreads=0):
    topo = w.topo
    pos = w.frame(f).astype(np.float64)
    return O.energies_soln(oracle_system(topo), oracle_params(params), w.cell.astype(np.float64), pos,
                           topo.tagslt, nthreads=nthreads)


def oracle_distribution(w, params) -> O.Distribution:
    uvmax, uvsoft, uvcrd = w.uvmax_uvsoft_uvcrd()
    invcl, vol = O.inverted_cell(w.cell.astype(np.float64))
    qtot = float(np.float64(w.topo.charge32.astype(np.float64).sum()))
    return O.Distribution(params["slttype"], uvmax, uvsoft, uvcrd, params["temp"], vol, params["screen"],
                          qtot, wgtslf=params["wgtslf"], estype=params["estype"], selfcal=params["selfcal"],
                          corrcal=params["corrcal"])


def species_of(topo, nins, soln_tag=None):
    """uvspec per insertion row; the soln solute's own entry is skipped (0)."""
    sp = np.tile(topo.uvspec, (nins, 1))
    if soln_tag is not None:
        for n, t in enumerate(np.atleast_1d(soln_tag)):
            sp[n, t] = 0
    return sp


def energy_tolerance(e_ref, rtol=RTOL, atol=ATOL):
    e_ref = np.asarray(e_ref, np.float64)
    return np.where(np.abs(e_ref) > E_HIGH, RTOL_HIGH, rtol) * np.abs(e_ref) + atol


def energy_mismatch(e_gpu, e_ref, rtol=RTOL, atol=ATOL):
    e_gpu = np.asarray(e_gpu, np.float64)
    e_ref = np.asarray(e_ref, np.float64)
    excess = np.abs(e_gpu - e_ref) - energy_tolerance(e_ref, rtol, atol)
    worst = np.unravel_index(np.argmax(excess), excess.shape)
    return int((excess > 0).sum()), float(excess.max()), worst

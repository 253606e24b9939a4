"""ctypes front end of the FP64 oracle (oracle/er_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by the
benchmark's cpu_baseline leg, always as the checker or as the timed host-core
baseline -- never by the product path.

Array conventions follow er_oracle.h (0-based, Fortran memory order).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class Topo(C.Structure):
    _fields_ = [("natom", C.c_int), ("nmol", C.c_int), ("mol_begin", C.c_void_p),
                ("charge", C.c_void_p), ("ljtype", C.c_void_p), ("ntype", C.c_int),
                ("ljene", C.c_void_p), ("ljlensq", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("elecut", C.c_double), ("lwljcut", C.c_double), ("upljcut", C.c_double),
                ("screen", C.c_double), ("ljswitch", C.c_int), ("splodr", C.c_int),
                ("ms", C.c_int * 3), ("cltype", C.c_int), ("boxshp", C.c_int)]


class Box(C.Structure):
    _fields_ = [("cell_normal", C.c_double * 9), ("cell_len", C.c_double * 3),
                ("invcell", C.c_double * 3), ("is_cuboid", C.c_int), ("rot", C.c_double * 9),
                ("open", C.c_int)]


class Dist(C.Structure):
    _fields_ = [("slttype", C.c_int), ("wgtslf", C.c_int), ("estype", C.c_int),
                ("selfcal", C.c_int), ("corrcal", C.c_int), ("numslv", C.c_int),
                ("ermax", C.c_int), ("esmax", C.c_int), ("uvmax", C.c_void_p),
                ("uvsoft", C.c_void_p), ("uvcrd", C.c_void_p), ("escrd", C.c_void_p),
                ("temp", C.c_double), ("volume", C.c_double), ("screen", C.c_double),
                ("total_charge", C.c_double), ("edens", C.c_void_p), ("ecorr", C.c_void_p),
                ("eself", C.c_void_p), ("minuv", C.c_void_p), ("maxuv", C.c_void_p),
                ("slnuv", C.c_void_p), ("flceng", C.c_void_p), ("engnorm", C.c_double),
                ("engsmpl", C.c_double), ("avslf", C.c_double), ("voffset", C.c_double),
                ("voffset_initialized", C.c_int), ("insdst", C.c_void_p),
                ("errcode", C.c_char * 8), ("err_species", C.c_int), ("err_energy", C.c_double),
                ("warn_mbin", C.c_int), ("cltype", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.orc_spline_value.restype = C.c_double
        L.orc_spline_value.argtypes = [C.c_int, C.c_double]
        L.orc_init_spline_axis.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_engfac.argtypes = [C.POINTER(Params), _dp, C.c_double, _dp]
        L.orc_inverted_cell.argtypes = [_dp, _dp, C.POINTER(C.c_double)]
        L.orc_normalize_periodic.argtypes = [_dp, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                             C.POINTER(Box)]
        L.orc_pair_energy.restype = C.c_double
        L.orc_pair_energy.argtypes = [C.POINTER(Params), C.POINTER(Box), _dp, _dp, C.c_double,
                                      C.c_double, C.c_double, C.c_double]
        L.orc_fft_r2c.argtypes = [_ip, _dp, _dp]
        L.orc_fft_c2r.argtypes = [_ip, _dp, _dp]
        L.orc_fft1d.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_spline_atom.argtypes = [C.POINTER(Params), _dp, _dp, _dp, _ip]
        L.orc_pme_solute_grid.restype = C.c_double
        L.orc_pme_solute_grid.argtypes = [C.POINTER(Params), _dp, _dp, C.c_int, _dp, _dp,
                                          _dp, _dp, _dp]
        L.orc_binsearch.restype = C.c_int
        L.orc_binsearch.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_build_mesh.restype = C.c_int
        L.orc_build_mesh.argtypes = [C.c_double] * 9 + [C.c_int, _dp, C.POINTER(C.c_int),
                                                        C.POINTER(C.c_int)]
        L.orc_realcal_self.restype = C.c_double
        L.orc_realcal_self.argtypes = [C.POINTER(Params), C.POINTER(Box), C.c_int, _dp, _dp]
        L.orc_energies_refs.restype = C.c_int
        L.orc_energies_refs.argtypes = [C.POINTER(Topo), C.POINTER(Params), _dp, _dp, C.c_int,
                                        _dp, _ip, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int]
        L.orc_energies_soln.restype = C.c_int
        L.orc_energies_soln.argtypes = [C.POINTER(Topo), C.POINTER(Params), _dp, _dp, C.c_int,
                                        _ip, _dp, _dp, C.c_int]
        L.orc_engclear.argtypes = [C.POINTER(Dist)]
        L.orc_pppm_denom.restype = C.c_double
        L.orc_pppm_denom.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double]
        L.orc_pppm_gfb.argtypes = [C.c_int, _dp]
        L.orc_pppm_engfac.argtypes = [C.POINTER(Params), _dp, _dp, C.c_double, _dp]
        L.orc_pppm_engfac.restype = C.c_int
        L.orc_update_histogram.restype = C.c_int
        L.orc_update_histogram.argtypes = [C.POINTER(Dist), C.c_int, _ip, _dp, C.c_double,
                                           C.c_double, C.c_double]
        _lib = L
    return _lib


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def make_params(elecut=12.0, lwljcut=10.0, upljcut=12.0, screen=0.35, ljswitch=0, splodr=4,
                ms=(32, 32, 32), cltype=2, boxshp=1) -> Params:
    p = Params()
    p.elecut, p.lwljcut, p.upljcut, p.screen = elecut, lwljcut, upljcut, screen
    p.ljswitch, p.splodr = ljswitch, splodr
    p.ms[0], p.ms[1], p.ms[2] = ms
    p.cltype = cltype
    p.boxshp = boxshp
    return p


def pppm_gfb(splodr):
    g = np.zeros(16)
    lib().orc_pppm_gfb(int(splodr), g)
    return g[:splodr].copy()


def pppm_denom(splodr, gfb, x, y, z):
    g = f64(np.concatenate([gfb, np.zeros(16 - len(gfb))]))
    return lib().orc_pppm_denom(int(splodr), g, float(x), float(y), float(z))


class OracleSystem:
    """Holds the arrays an orc_topo points at, so they outlive the struct."""

    def __init__(self, mol_begin, charge, ljtype, ljene, ljlensq):
        self.mol_begin = i32(mol_begin)
        self.charge = f64(charge)
        self.ljtype = i32(ljtype)
        self.ljene = f64(ljene).ravel()
        self.ljlensq = f64(ljlensq).ravel()
        self.ntype = int(round(np.sqrt(self.ljene.size)))
        t = Topo()
        t.natom = int(self.charge.size)
        t.nmol = int(self.mol_begin.size - 1)
        t.mol_begin = self.mol_begin.ctypes.data
        t.charge = self.charge.ctypes.data
        t.ljtype = self.ljtype.ctypes.data
        t.ntype = self.ntype
        t.ljene = self.ljene.ctypes.data
        t.ljlensq = self.ljlensq.ctypes.data
        self.topo = t


def spline_value(order, x):
    return lib().orc_spline_value(order, float(x))


def init_spline_axis(order, K):
    out = np.zeros(K)
    lib().orc_init_spline_axis(order, K, out)
    return out


def inverted_cell(cell):
    invcl = np.zeros(9)
    vol = C.c_double()
    lib().orc_inverted_cell(f64(np.asarray(cell).ravel(order="F")), invcl, C.byref(vol))
    return invcl.reshape(3, 3, order="F"), vol.value


def engfac(params: Params, cell):
    """The influence function of params.cltype (2 PME, 3 PPPM) on the half
    spectrum, [K3][K2][K1/2+1]."""
    invcl, vol = inverted_cell(cell)
    K1, K2, K3 = params.ms
    out = np.zeros((K1 // 2 + 1) * K2 * K3)
    if params.cltype == 3:
        if lib().orc_pppm_engfac(C.byref(params), f64(np.asarray(cell).ravel(order="F")),
                                 f64(invcl.ravel(order="F")), vol, out):
            raise ValueError("PPPM needs an even order")
    else:
        lib().orc_engfac(C.byref(params), f64(invcl.ravel(order="F")), vol, out)
    return out.reshape((K3, K2, K1 // 2 + 1))


def fft_r2c(grid):
    """grid[x3,x2,x1] (C order = Fortran (x1,x2,x3)) -> spec[m3,m2,m1] complex"""
    K3, K2, K1 = grid.shape
    K = i32([K1, K2, K3])
    spec = np.zeros(2 * (K1 // 2 + 1) * K2 * K3)
    lib().orc_fft_r2c(K, f64(grid).ravel(), spec)
    return spec.view(np.complex128).reshape(K3, K2, K1 // 2 + 1)


def fft_c2r(spec, K1):
    K3, K2, H = spec.shape
    K = i32([K1, K2, K3])
    out = np.zeros(K1 * K2 * K3)
    lib().orc_fft_c2r(K, np.ascontiguousarray(spec, dtype=np.complex128).view(np.float64).ravel(), out)
    return out.reshape(K3, K2, K1)


def fft1d(x, sign):
    d = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(d.size, sign, d.view(np.float64))
    return d


def spline_atom(params: Params, cell, x):
    invcl, _ = inverted_cell(cell)
    w = np.zeros(3 * params.splodr)
    g = np.zeros(3, dtype=np.int32)
    lib().orc_spline_atom(C.byref(params), f64(invcl.ravel(order="F")), f64(x), w, g)
    return w.reshape(3, params.splodr), g


def pme_solute_grid(params: Params, cell, charges, pos):
    """-> (phi[x3,x2,x1], self_energy) for one solute (raw coordinates)."""
    invcl, vol = inverted_cell(cell)
    ef = engfac(params, cell).ravel()
    K1, K2, K3 = params.ms
    G = K1 * K2 * K3
    Gc = (K1 // 2 + 1) * K2 * K3
    phi, rho, spec = np.zeros(G), np.zeros(G), np.zeros(2 * Gc)
    self_e = lib().orc_pme_solute_grid(C.byref(params), f64(invcl.ravel(order="F")), ef,
                                       len(charges), f64(charges), f64(pos).ravel(), phi, spec, rho)
    return phi.reshape(K3, K2, K1), self_e


def binsearch(coord, v):
    c = f64(coord)
    return lib().orc_binsearch(c, c.size, float(v))


def build_mesh(ecdmin, ecfmns, ecdcen, eccore, ecdmax, eclbin, ecfbin, ec0bin, finfac, pecore):
    mesh = np.zeros(10000)
    pesoft = C.c_int()
    err = C.c_int()
    pemax = lib().orc_build_mesh(ecdmin, ecfmns, ecdcen, eccore, ecdmax, eclbin, ecfbin, ec0bin,
                                 finfac, int(pecore), mesh, C.byref(pesoft), C.byref(err))
    if err.value:
        raise ValueError({1: "eng_ecd", 2: "eng_siz"}[err.value])
    return mesh[:pemax].copy(), pesoft.value


def normalize_periodic(cell, pos, elecut=12.0, upljcut=12.0):
    pos = f64(pos)
    out = np.zeros_like(pos)
    box = Box()
    rc = lib().orc_normalize_periodic(f64(np.asarray(cell).ravel(order="F")), elecut, upljcut,
                                      pos.shape[0], pos.ravel(), out.ravel(), C.byref(box))
    if rc:
        raise ValueError({2: "PPPM needs an even order", 3: "an open system needs bare Coulomb"}.get(
            rc, "cutoff exceeds half the box"))
    return out, box


def pair_energy(params: Params, box: Box, xa, xb, qa, qb, eps, sgm2):
    return lib().orc_pair_energy(C.byref(params), C.byref(box), f64(xa), f64(xb), qa, qb, eps, sgm2)


def realcal_self(params: Params, box: Box, q, x):
    return lib().orc_realcal_self(C.byref(params), C.byref(box), len(q), f64(q), f64(x).ravel())


def energies_refs(sysm: OracleSystem, params: Params, cell, pos, slt_charge, slt_ljtype,
                  slt_pos, rigid=False, nthreads=0):
    """Per-frame refs energies: uv[nins, nmol], uv0[nins]."""
    nins, nslt, _ = np.asarray(slt_pos).shape
    nmol = sysm.topo.nmol
    uv = np.zeros(nins * nmol)
    uv0 = np.zeros(nins)
    rc = lib().orc_energies_refs(C.byref(sysm.topo), C.byref(params),
                                 f64(np.asarray(cell).ravel(order="F")), f64(pos).ravel(), nslt,
                                 f64(slt_charge), i32(slt_ljtype), nins, f64(slt_pos).ravel(),
                                 int(rigid), uv, uv0, int(nthreads))
    if rc:
        raise ValueError({2: "PPPM needs an even order", 3: "an open system needs bare Coulomb"}.get(
            rc, "cutoff exceeds half the box"))
    return uv.reshape(nins, nmol), uv0


def energies_soln(sysm: OracleSystem, params: Params, cell, pos, tagslt, nthreads=0):
    tag = i32(np.atleast_1d(tagslt))
    nmol = sysm.topo.nmol
    uv = np.zeros(tag.size * nmol)
    uv0 = np.zeros(tag.size)
    rc = lib().orc_energies_soln(C.byref(sysm.topo), C.byref(params),
                                 f64(np.asarray(cell).ravel(order="F")), f64(pos).ravel(), tag.size,
                                 tag, uv, uv0, int(nthreads))
    if rc:
        raise ValueError({2: "PPPM needs an even order", 3: "an open system needs bare Coulomb"}.get(
            rc, "cutoff exceeds half the box"))
    return uv.reshape(tag.size, nmol), uv0


class Distribution:
    """orc_dist plus the numpy arrays it points at (engproc.F90:858-964 state)."""

    def __init__(self, slttype, uvmax, uvsoft, uvcrd, temp, volume, screen, total_charge,
                 wgtslf=1, estype=2, selfcal=0, corrcal=1, escrd=None, cltype=2):
        self.uvmax, self.uvsoft, self.uvcrd = i32(uvmax), i32(uvsoft), f64(uvcrd)
        self.numslv = self.uvmax.size
        self.ermax = int(self.uvmax.sum())
        self.escrd = f64(escrd if escrd is not None else np.zeros(1))
        self.esmax = int(self.escrd.size) if escrd is not None else 0
        self.edens = np.zeros(self.ermax)
        self.ecorr = np.zeros(self.ermax * self.ermax)
        self.eself = np.zeros(max(self.esmax, 1))
        self.minuv = np.full(self.numslv + 1, np.inf)
        self.maxuv = np.full(self.numslv + 1, -np.inf)
        self.slnuv = np.zeros(self.numslv)
        self.flceng = np.zeros(self.numslv)
        self.insdst = np.zeros(self.ermax, dtype=np.int32)
        d = Dist()
        d.slttype, d.wgtslf, d.estype, d.selfcal, d.corrcal = slttype, wgtslf, estype, selfcal, corrcal
        d.numslv, d.ermax, d.esmax = self.numslv, self.ermax, self.esmax
        d.uvmax, d.uvsoft = self.uvmax.ctypes.data, self.uvsoft.ctypes.data
        d.uvcrd, d.escrd = self.uvcrd.ctypes.data, self.escrd.ctypes.data
        d.temp, d.volume, d.screen, d.total_charge = temp, volume, screen, total_charge
        for name in ("edens", "ecorr", "eself", "minuv", "maxuv", "slnuv", "flceng", "insdst"):
            setattr(d, name, getattr(self, name).ctypes.data)
        d.voffset = -np.inf
        d.voffset_initialized = 0
        d.cltype = cltype
        self.d = d

    def clear(self):
        lib().orc_engclear(C.byref(self.d))

    def update(self, species, uv_row, uv0, w_sys=1.0, w_slt=1.0):
        rc = lib().orc_update_histogram(C.byref(self.d), len(uv_row), i32(species), f64(uv_row),
                                        float(uv0), float(w_sys), float(w_slt))
        if rc:
            raise RuntimeError(self.d.errcode.decode())

    @property
    def scalars(self):
        d = self.d
        return dict(engnorm=d.engnorm, engsmpl=d.engsmpl, avslf=d.avslf, voffset=d.voffset)

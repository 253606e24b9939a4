"""ctypes binding of the C ABI in include/er_capi.h and include/er_traj.h
(lib/liberdst_hip.so).

The library is loaded from this package's lib/ directory; if it is missing
the import of this module fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "liberdst_hip.so")

ER_OK, ER_E_ARG, ER_E_UNSUPPORTED, ER_E_HIP, ER_E_HALT, ER_E_STATE = range(6)
ER_PLAIN, ER_BATCHED, ER_PLAIN_ASYNC = 0, 1, 2

EXPORTED = ["er_create", "er_set_mesh", "er_set_solute", "er_set_solutes", "er_select_solute",
            "er_set_solute_molecules", "er_set_frame",
            "er_set_frame_device", "er_set_insertions", "er_set_insertions_device", "er_run", "er_run_noalloc",
            "er_sync",
            "er_get_uvengy", "er_get_nins", "er_fetch", "er_get_flceng", "er_clear", "er_last_error", "er_error_string",
            "er_destroy", "er_profile", "er_profile_read", "er_autotune", "er_set_tuning", "er_get_tuning",
            "er_graph_begin", "er_graph_end", "er_graph_launch", "er_graph_destroy", "er_set_cell_width",
            "er_get_cell_width"]
EXPORTED_TRAJ = ["er_traj_open", "er_traj_read", "er_traj_info", "er_traj_close"]
ER_TRAJ_EOF, ER_TRAJ_IO, ER_TRAJ_FORMAT = 100, 101, 102
TRAJ_FLAGS = {"charmm": 1, "cell": 2, "4d": 4, "rec64": 8, "swapped": 16}
KERNELS = ["prep", "spec2d", "axis1", "gather", "pairs", "hist", "weights", "ecorr"]


class ErTopology(C.Structure):
    _fields_ = [("numatm", C.c_int32), ("nummol", C.c_int32), ("ljtype_max", C.c_int32),
                ("numslv", C.c_int32), ("mol_begin_index", C.c_void_p), ("ljtype", C.c_void_p),
                ("uvspec", C.c_void_p), ("charge", C.c_void_p), ("ljene_mat", C.c_void_p),
                ("ljlensq_mat", C.c_void_p)]


class ErParams(C.Structure):
    _fields_ = [("elecut", C.c_float), ("lwljcut", C.c_float), ("upljcut", C.c_float),
                ("screen", C.c_float), ("ljswitch", C.c_int32), ("cltype", C.c_int32),
                ("splodr", C.c_int32), ("ms", C.c_int32 * 3), ("boxshp", C.c_int32),
                ("slttype", C.c_int32), ("wgtslf", C.c_int32), ("estype", C.c_int32),
                ("corrcal", C.c_int32), ("selfcal", C.c_int32), ("temp", C.c_double),
                ("force_calculation", C.c_int32)]


class ErError(RuntimeError):
    def __init__(self, status, message, code=""):
        super().__init__(f"[{status}] {message}" + (f" ({code})" if code else ""))
        self.status, self.code = status, code


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i32, f32, f64 = C.c_void_p, C.c_int32, C.c_float, C.c_double
    L.er_create.argtypes = [C.POINTER(ErTopology), C.POINTER(ErParams), C.POINTER(vp)]
    L.er_set_mesh.argtypes = [vp, vp, vp, vp, i32, vp]
    L.er_set_solute.argtypes = [vp, i32, vp, vp]
    L.er_set_solutes.argtypes = [vp, i32, vp, vp, vp]
    L.er_select_solute.argtypes = [vp, i32]
    L.er_set_solute_molecules.argtypes = [vp, i32, vp]
    L.er_set_frame.argtypes = [vp, vp, vp, f32]
    L.er_set_frame_device.argtypes = [vp, vp, vp, f32]
    L.er_set_insertions.argtypes = [vp, i32, vp, vp]
    L.er_set_insertions_device.argtypes = [vp, i32, vp, vp]
    L.er_run.argtypes = [vp, i32]
    L.er_run_noalloc.argtypes = [vp, i32]
    L.er_sync.argtypes = [vp]
    L.er_get_uvengy.argtypes = [vp, vp, vp]
    L.er_get_nins.argtypes = [vp, C.POINTER(i32)]
    L.er_fetch.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.er_get_flceng.argtypes = [vp, vp]
    L.er_clear.argtypes = [vp]
    L.er_last_error.argtypes = [vp, C.c_char_p, C.POINTER(i32), C.POINTER(f32)]
    L.er_error_string.argtypes = [vp]
    L.er_error_string.restype = C.c_char_p
    L.er_destroy.argtypes = [vp]
    L.er_destroy.restype = None
    L.er_profile.argtypes = [vp, i32]
    L.er_profile_read.argtypes = [vp, i32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.er_autotune.argtypes = [vp, i32]
    L.er_set_tuning.argtypes = [vp, i32, i32, i32]
    L.er_get_tuning.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.er_graph_begin.argtypes = [vp]
    L.er_set_cell_width.argtypes = [vp, f32]
    L.er_get_cell_width.argtypes = [vp, C.POINTER(f32)]
    L.er_graph_end.argtypes = [vp, C.POINTER(i32)]
    L.er_graph_launch.argtypes = [vp, i32]
    L.er_graph_destroy.argtypes = [vp, i32]
    for name in EXPORTED:
        if name not in ("er_error_string", "er_destroy"):
            getattr(L, name).restype = C.c_int
    L.er_traj_open.argtypes = [C.c_char_p, C.POINTER(vp), C.POINTER(i32)]
    L.er_traj_read.argtypes = [vp, i32, vp, vp]
    L.er_traj_info.argtypes = [vp] + [C.POINTER(i32)] * 3 + [C.POINTER(C.c_double)] + [C.POINTER(i32)] * 2
    L.er_traj_close.argtypes = [vp]
    L.er_traj_close.restype = None
    for name in ("er_traj_open", "er_traj_read", "er_traj_info"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


class TrajError(RuntimeError):
    def __init__(self, status, what):
        kind = {ER_E_ARG: "argument", ER_TRAJ_IO: "read", ER_TRAJ_FORMAT: "format"}.get(status, "status")
        super().__init__(f"{what}: {kind} error [{status}]")
        self.status = status


class Trajectory:
    """A DCD trajectory (HISTORY or SltConf) through er_traj_* (include/er_traj.h),
    erdst's open_trajectory / read_trajectory / close_trajectory
    (erdst/trajectory.F90:41-96).  read() returns (coordinates [natoms, 3]
    float32, cell [3, 3] float32 with the cell vectors a, b, c as columns), or
    None after the last frame."""

    def __init__(self, path):
        L = load()
        h, n = C.c_void_p(), C.c_int32()
        rc = L.er_traj_open(os.fsencode(path), C.byref(h), C.byref(n))
        if rc:
            raise TrajError(rc, f"er_traj_open({path})")
        self.h, self.natoms, self.path = h, int(n.value), path

    def info(self) -> dict:
        v = [C.c_int32() for _ in range(3)] + [C.c_double()] + [C.c_int32() for _ in range(2)]
        rc = load().er_traj_info(self.h, *[C.byref(x) for x in v])
        if rc:
            raise TrajError(rc, "er_traj_info")
        flags = v[5].value
        return dict(nframes=v[0].value, istart=v[1].value, nsavc=v[2].value, delta=v[3].value,
                    nfixed=v[4].value, **{k: bool(flags & b) for k, b in TRAJ_FLAGS.items()})

    def read(self, natoms=None):
        n = self.natoms if natoms is None else int(natoms)
        xyz = np.empty((max(n, 0), 3), np.float32)
        cell = np.empty(9, np.float32)
        rc = load().er_traj_read(self.h, n, xyz.ctypes.data, cell.ctypes.data)
        if rc == ER_TRAJ_EOF:
            return None
        if rc:
            raise TrajError(rc, f"er_traj_read({self.path})")
        return xyz, cell.reshape((3, 3), order="F")

    def __iter__(self):
        while (fr := self.read()) is not None:
            yield fr

    def close(self):
        if self.h:
            load().er_traj_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(a):
    return None if a is None else a.ctypes.data


class Context:
    """One er_ctx.  Mirrors the entry points 1:1; raises ErError on failure."""

    def __init__(self, topo, params: dict):
        L = load()
        self._keep = []
        self.nsolute = 1
        t = ErTopology()
        t.numatm, t.nummol = topo.natom, topo.nmol
        t.ljtype_max, t.numslv = topo.ntype, topo.numslv
        arrs = dict(mol_begin_index=np.ascontiguousarray(topo.mol_begin + 1, np.int32),
                    ljtype=np.ascontiguousarray(topo.ljtype + 1, np.int32),
                    uvspec=np.ascontiguousarray(topo.uvspec, np.int32),
                    charge=np.ascontiguousarray(topo.charge32, np.float32),
                    ljene_mat=np.ascontiguousarray(topo.ljene32, np.float32),
                    ljlensq_mat=np.ascontiguousarray(topo.ljlensq32, np.float32))
        for k, v in arrs.items():
            setattr(t, k, v.ctypes.data)
        self._keep.append(arrs)
        p = ErParams()
        p.elecut, p.lwljcut, p.upljcut = params["elecut"], params["lwljcut"], params["upljcut"]
        p.screen, p.ljswitch, p.cltype = params["screen"], params["ljswitch"], params["cltype"]
        p.splodr = params["splodr"]
        for k in range(3):
            p.ms[k] = params["ms"][k]
        p.boxshp, p.slttype, p.wgtslf = params["boxshp"], params["slttype"], params["wgtslf"]
        p.estype, p.corrcal, p.selfcal = params["estype"], params["corrcal"], params["selfcal"]
        p.temp = params["temp"]
        p.force_calculation = int(bool(params.get("force_calculation", 0)))
        h = C.c_void_p()
        rc = L.er_create(C.byref(t), C.byref(p), C.byref(h))
        self.h = h
        self.L = L
        if rc:
            msg = L.er_error_string(h).decode() if h.value else "er_create failed"
            if h.value:
                L.er_destroy(h)
            self.h = C.c_void_p()
            raise ErError(rc, msg)
        self.topo = topo
        self.params = dict(params)

    def _check(self, rc):
        if rc:
            code = C.create_string_buffer(8)
            sp, en = C.c_int32(), C.c_float()
            self.L.er_last_error(self.h, code, C.byref(sp), C.byref(en))
            raise ErError(rc, self.L.er_error_string(self.h).decode(), code.value.decode())

    def set_mesh(self, uvmax, uvsoft, uvcrd, escrd=None):
        self._mesh = (np.ascontiguousarray(uvmax, np.int32), np.ascontiguousarray(uvsoft, np.int32),
                      np.ascontiguousarray(uvcrd, np.float64),
                      None if escrd is None else np.ascontiguousarray(escrd, np.float64))
        um, us, uc, es = self._mesh
        self.ermax = int(um.sum())
        self.esmax = 0 if es is None else int(es.size)
        self._check(self.L.er_set_mesh(self.h, _ptr(um), _ptr(us), _ptr(uc), self.esmax, _ptr(es)))

    def set_solute(self, ljtype0, charge32):
        lj = np.ascontiguousarray(np.asarray(ljtype0) + 1, np.int32)
        q = np.ascontiguousarray(charge32, np.float32)
        self._check(self.L.er_set_solute(self.h, lj.size, _ptr(lj), _ptr(q)))

    def set_solutes(self, solutes):
        """solutes: [(ljtype0, charge32), ...], one per test particle."""
        ns = np.array([len(q) for _, q in solutes], np.int32)
        lj = np.ascontiguousarray(np.concatenate([np.asarray(t) + 1 for t, _ in solutes]), np.int32)
        q = np.ascontiguousarray(np.concatenate([np.asarray(c) for _, c in solutes]), np.float32)
        self._check(self.L.er_set_solutes(self.h, ns.size, _ptr(ns), _ptr(lj), _ptr(q)))
        self.nsolute = int(ns.size)

    def select_solute(self, k):
        """Select solute k (0-based here, 1-based in the C ABI)."""
        self._check(self.L.er_select_solute(self.h, int(k) + 1))

    def set_solute_molecules(self, tagslt0):
        t = np.ascontiguousarray(np.asarray(tagslt0) + 1, np.int32)
        self._check(self.L.er_set_solute_molecules(self.h, t.size, _ptr(t)))

    def set_frame(self, pos, cell, w_sys=1.0):
        pos = np.ascontiguousarray(pos, np.float32)
        cell = np.ascontiguousarray(np.asarray(cell, np.float32).ravel(order="F"))
        self._check(self.L.er_set_frame(self.h, _ptr(pos), _ptr(cell), float(w_sys)))

    def set_frame_device(self, dev_ptr: int, cell, w_sys=1.0):
        cell = np.ascontiguousarray(np.asarray(cell, np.float32).ravel(order="F"))
        self._check(self.L.er_set_frame_device(self.h, C.c_void_p(dev_ptr), _ptr(cell), float(w_sys)))

    def set_insertions(self, sltpos, w_slt=None):
        x = np.ascontiguousarray(sltpos, np.float32)
        w = None if w_slt is None else np.ascontiguousarray(w_slt, np.float32)
        self._check(self.L.er_set_insertions(self.h, x.shape[0], _ptr(x), _ptr(w)))

    def set_insertions_device(self, nins: int, dev_ptr: int, w_slt=None):
        w = None if w_slt is None else np.ascontiguousarray(w_slt, np.float32)
        self._check(self.L.er_set_insertions_device(self.h, int(nins), C.c_void_p(dev_ptr), _ptr(w)))

    def run(self, mode=ER_BATCHED):
        self._check(self.L.er_run(self.h, int(mode)))

    def run_noalloc(self, mode=ER_BATCHED):
        """er_run_noalloc: the same run with no device allocation; an earlier
        er_run of at least as many insertions must have sized the buffers."""
        self._check(self.L.er_run_noalloc(self.h, int(mode)))

    def sync(self):
        self._check(self.L.er_sync(self.h))

    def nins(self):
        """Insertions of the selected solute in the last run (er_get_nins)."""
        n = C.c_int32()
        self._check(self.L.er_get_nins(self.h, C.byref(n)))
        return n.value

    def uvengy(self, nins=None):
        """(uvengy [nins, nmol], uvengy0 [nins]) of the last run; nins, when
        given, must be the run's insertion count."""
        have = self.nins()
        if nins is None:
            nins = have
        elif nins != have:
            raise ValueError(f"uvengy: the last run has {have} insertions of this solute, not {nins}")
        u = np.zeros((nins, self.topo.nmol), np.float32)
        u0 = np.zeros(nins, np.float32)
        self._check(self.L.er_get_uvengy(self.h, _ptr(u), _ptr(u0)))
        return u, u0

    def fetch(self, want_ecorr=True):
        ns = self.topo.numslv
        edens = np.zeros(self.ermax)
        ecorr = np.zeros((self.ermax, self.ermax)) if want_ecorr else None
        eself = np.zeros(max(self.esmax, 1))
        scal = np.zeros(4)
        mn, mx = np.zeros(ns + 1, np.float32), np.zeros(ns + 1, np.float32)
        slnuv = np.zeros(ns)
        self._check(self.L.er_fetch(self.h, _ptr(edens), _ptr(ecorr), _ptr(eself), _ptr(scal), _ptr(mn),
                                    _ptr(mx), _ptr(slnuv)))
        return dict(edens=edens, ecorr=ecorr, eself=eself[:self.esmax], engnorm=scal[0], engsmpl=scal[1],
                    avslf=scal[2], voffset=scal[3], minuv=mn, maxuv=mx, slnuv=slnuv)

    def flceng(self, nslt):
        f = np.zeros((nslt, self.topo.numslv))
        self._check(self.L.er_get_flceng(self.h, _ptr(f)))
        return f

    def clear(self):
        self._check(self.L.er_clear(self.h))

    def last_error(self):
        code = C.create_string_buffer(8)
        sp, en = C.c_int32(), C.c_float()
        self.L.er_last_error(self.h, code, C.byref(sp), C.byref(en))
        return code.value.decode(), sp.value, en.value

    def profile(self, enable=True):
        self._check(self.L.er_profile(self.h, int(enable)))

    def profile_read(self):
        """{kernel: (total_ms, launches)} accumulated since the last read."""
        out = {}
        for k, name in enumerate(KERNELS):
            ms, n = C.c_double(), C.c_int64()
            self._check(self.L.er_profile_read(self.h, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def graph_begin(self):
        """er_graph_begin: record the following device-side setters and
        batched runs into a HIP graph."""
        self._check(self.L.er_graph_begin(self.h))

    def graph_end(self):
        """er_graph_end: the recorded graph's handle."""
        g = C.c_int32()
        self._check(self.L.er_graph_end(self.h, C.byref(g)))
        return int(g.value)

    def graph_launch(self, graph):
        self._check(self.L.er_graph_launch(self.h, int(graph)))

    def graph_destroy(self, graph):
        self._check(self.L.er_graph_destroy(self.h, int(graph)))

    def autotune(self, reps=0):
        """Times launch shapes on the current frame and insertions and keeps
        the fastest (er_autotune); returns tuning()."""
        self._check(self.L.er_autotune(self.h, int(reps)))
        return self.tuning()

    def tuning(self):
        """{"chunk": insertions per launch sequence, "hist_slices": histogram
        workgroups per 16 insertions, "pair_split": [per solute],
        "cell_width": least cell width of the pair search in Angstrom} (0:
        the built-in choice); leaves solute 0 selected."""
        ch, sp, hs = C.c_int32(), C.c_int32(), C.c_int32()
        splits = []
        for k in range(self.nsolute):
            if self.nsolute > 1:
                self.select_solute(k)
            self._check(self.L.er_get_tuning(self.h, C.byref(ch), C.byref(sp), C.byref(hs)))
            splits.append(sp.value)
        if self.nsolute > 1:
            self.select_solute(0)
        cw = C.c_float()
        self._check(self.L.er_get_cell_width(self.h, C.byref(cw)))
        return {"chunk": ch.value, "hist_slices": hs.value, "pair_split": splits, "cell_width": cw.value}

    def set_tuning(self, tuning):
        """Applies a tuning() dictionary (for instance one kept in a cache
        file); leaves solute 0 selected."""
        splits = list(tuning["pair_split"])
        for k, sp in enumerate(splits):
            if len(splits) > 1:
                self.select_solute(k)
            self._check(self.L.er_set_tuning(self.h, int(tuning["chunk"]), int(sp),
                                             int(tuning.get("hist_slices", 0))))
        if len(splits) > 1:
            self.select_solute(0)
        self._check(self.L.er_set_cell_width(self.h, float(tuning.get("cell_width", 0.0))))

    def close(self):
        if self.h is not None and self.h.value:
            self.L.er_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Seeded synthetic erdst workloads (the five benchmark configurations).

There are no trajectories to download, so every input is generated here:
rigid 3-site water on a jittered lattice with uniform random orientations,
a rigid 5-site tetrahedral test particle inserted at uniform random
positions/orientations, a flexible 20-atom chain solute for the solution
run, and a two-species solvent with uneven random solutes.  Coordinates are
float32, as the reference reads them from DCD (erdst/trajectory.F90:75-89).

Random streams: numpy Philox keyed by 20261202 + config, one jumped
substream per frame.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .topology import Species, Topology, build_topology

WATER_Q = np.array([-0.834, 0.417, 0.417])
WATER_EPS = np.array([0.152, 0.0, 0.0])
WATER_SIG = np.array([3.151, 1.0, 1.0])


def water_geometry() -> np.ndarray:
    r, th = 0.9572, np.deg2rad(104.52)
    return np.array([[0.0, 0.0, 0.0],
                     [r * np.sin(th / 2), r * np.cos(th / 2), 0.0],
                     [-r * np.sin(th / 2), r * np.cos(th / 2), 0.0]])


def tetra_geometry() -> np.ndarray:
    d = 1.09 / np.sqrt(3.0)
    return np.array([[0, 0, 0], [d, d, d], [d, -d, -d], [-d, d, -d], [-d, -d, d]], dtype=float)


def water_species(nmol: int) -> Species:
    return Species("water", nmol, WATER_Q.copy(), WATER_EPS.copy(), WATER_SIG.copy(), water_geometry())


def tetra_species() -> Species:
    return Species("tetra", 1, np.array([-0.24, 0.06, 0.06, 0.06, 0.06]),
                   np.array([0.11, 0.03, 0.03, 0.03, 0.03]), np.array([3.4, 2.5, 2.5, 2.5, 2.5]),
                   tetra_geometry())


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    w, x = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    y, z = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def lattice_centres(rng: np.random.Generator, n: int, box: np.ndarray, jitter: float = 0.5,
                    avoid: np.ndarray | None = None, avoid_r: float = 0.0, spare: int = 0) -> np.ndarray:
    """n molecule centres on a jittered cubic-ish lattice (first n of a seeded
    permutation), optionally skipping sites within avoid_r of `avoid`."""
    m = int(np.ceil((n + spare) ** (1.0 / 3.0)))
    while m ** 3 < n + spare:
        m += 1
    idx = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    pts = (idx + 0.5) * (box / m)
    pts = pts[rng.permutation(len(pts))]
    pts = pts + rng.uniform(-jitter, jitter, pts.shape)
    if avoid is not None and avoid_r > 0:
        d = pts[:, None, :] - avoid[None, :, :]
        d -= box * np.round(d / box)
        keep = (np.sqrt((d ** 2).sum(-1)).min(1) > avoid_r)
        pts = pts[keep]
    if len(pts) < n:
        raise ValueError("lattice too small")
    return pts[:n]


def place_rigid(rng: np.random.Generator, geom: np.ndarray, centres: np.ndarray) -> np.ndarray:
    R = random_rotations(rng, len(centres))
    return (centres[:, None, :] + np.einsum("nij,sj->nsi", R, geom)).reshape(-1, 3)


def random_chain(rng: np.random.Generator, n: int, bond: float = 1.5, min_sep: float = 2.2) -> np.ndarray:
    """Self-avoiding random chain with fixed bond length."""
    x = [np.zeros(3)]
    while len(x) < n:
        for _ in range(1000):
            v = rng.normal(size=3)
            v *= bond / np.linalg.norm(v)
            cand = x[-1] + v
            if len(x) < 2 or np.min(np.linalg.norm(np.array(x[:-1]) - cand, axis=1)) > min_sep:
                x.append(cand)
                break
        else:
            raise RuntimeError("chain growth failed")
    x = np.array(x)
    return x - x.mean(0)


def random_solute(rng: np.random.Generator, name: str, nsite: int, qsd: float = 0.3,
                  sig=(2.5, 3.8), eps=(0.03, 0.2)) -> Species:
    q = rng.normal(0.0, qsd, nsite)
    q -= q.mean()
    geom = random_chain(rng, nsite) if nsite > 1 else np.zeros((1, 3))
    return Species(name, 1, q, rng.uniform(*eps, nsite), rng.uniform(*sig, nsite), geom)


def linear_log_mesh(lin_lo: float, lin_hi: float, dlin: float, log_hi: float, nlog: int):
    """Energy mesh: linear points lin_lo, lin_lo+d, ... < lin_hi, then nlog
    log-spaced points from lin_hi to log_hi.  Returns (mesh, pecore)."""
    nlin = int(round((lin_hi - lin_lo) / dlin))
    lin = lin_lo + dlin * np.arange(nlin)
    if nlog == 0:
        return lin, 0
    log = lin_hi * (log_hi / lin_hi) ** (np.arange(nlog) / (nlog - 1))
    return np.concatenate([lin, log]), nlog


@dataclass
class Workload:
    """A generated run: topology, frames, insertions, mesh and parameters."""
    name: str
    topo: Topology
    box: np.ndarray                      # (3,) orthorhombic lengths
    frames: list = field(default_factory=list)   # callables/arrays [natom,3] float32
    nframes: int = 0
    nins: int = 0
    params: dict = field(default_factory=dict)
    mesh: list = field(default_factory=list)     # per solvent species: (mesh, pecore)
    seed: int = 0
    slttype: int = 3
    _frame_fn: object = None
    _ins_fn: object = None

    @property
    def cell(self) -> np.ndarray:
        return np.diag(self.box).astype(np.float32)

    def frame(self, f: int) -> np.ndarray:
        return self._frame_fn(f)

    def insertions(self, f: int) -> np.ndarray:
        """[nins, nslt, 3] float32 absolute insertion coordinates of frame f."""
        return self._ins_fn(f)

    def uvmax_uvsoft_uvcrd(self):
        uvmax, uvsoft, crd = [], [], []
        for mesh, pecore in self.mesh:
            uvmax.append(len(mesh))
            uvsoft.append(len(mesh) - pecore)
            crd.append(np.asarray(mesh, np.float64))
        return np.array(uvmax, np.int32), np.array(uvsoft, np.int32), np.concatenate(crd)


def _stream(seed: int, f: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=seed).jumped(f + 1))


def refs_water(config: int, nwater: int, L: float, nins: int, nframes: int, grid: int,
               solute: Species | None = None, name: str = "") -> Workload:
    """Water + rigid test particle, refs mode (configurations 0 and 2)."""
    seed = 20261202 + config
    solute = solute or tetra_species()
    topo = build_topology([water_species(nwater), solute], solute_index=1, refs=True)
    box = np.array([L, L, L])
    geom_w = water_geometry()

    def frame(f):
        rng = _stream(seed, 2 * f)
        return place_rigid(rng, geom_w, lattice_centres(rng, nwater, box)).astype(np.float32)

    def ins(f):
        rng = _stream(seed, 2 * f + 1)
        centres = rng.uniform(0.0, 1.0, (nins, 3)) * box
        x = place_rigid(rng, solute.geometry, centres)
        return x.reshape(nins, solute.nsite, 3).astype(np.float32)

    mesh = linear_log_mesh(-20.0, 5.0, 0.05, 1.0e4, 100)
    w = Workload(name or f"config{config}", topo, box, nframes=nframes, nins=nins,
                 params=dict(ms=(grid, grid, grid)), mesh=[mesh], seed=seed, slttype=3)
    w._frame_fn, w._ins_fn = frame, ins
    return w


def config0() -> Workload:
    """1000 waters, 31.04 A cube, 100 insertions x 10 frames, PME 32^3."""
    return refs_water(0, 1000, 31.04, 100, 10, 32)


def config2() -> Workload:
    """30 000 waters (90 000 atoms), 67.0 A cube, 1000 insertions x 200 frames, PME 72^3."""
    return refs_water(2, 30000, 67.0, 1000, 200, 72)


def config1() -> Workload:
    """Solution run: 8000 waters, 62.1 A cube, one flexible 20-atom chain
    solute (species 1), 100 frames, PME 64^3."""
    seed = 20261202 + 1
    rng0 = np.random.Generator(np.random.Philox(key=seed))
    chain = random_solute(rng0, "chain20", 20)
    L = 62.1
    box = np.array([L, L, L])
    nw = 8000
    topo = build_topology([chain, water_species(nw)], solute_index=0, refs=False)
    geom_w = water_geometry()
    centre = box / 2

    def frame(f):
        rng = _stream(seed, f)
        sx = centre + chain.geometry + rng.normal(0.0, 0.1, chain.geometry.shape)
        wc = lattice_centres(rng, nw, box, avoid=sx, avoid_r=3.2, spare=600)
        wx = place_rigid(rng, geom_w, wc)
        return np.concatenate([sx, wx]).astype(np.float32)

    mesh = linear_log_mesh(-60.0, 60.0, 0.1, 0, 0)
    w = Workload("config1", topo, box, nframes=100, nins=1, params=dict(ms=(64, 64, 64)),
                 mesh=[mesh], seed=seed, slttype=1)
    w._frame_fn = frame
    return w


def soln_water(config: int, nwater: int, L: float, nframes: int, grid: int,
               solute: Species | None = None, name: str = "") -> Workload:
    """Solution run of the rigid test particle of refs_water in the same
    water box: the solute sits at the box centre with a random orientation
    per frame, and the waters avoid its sites (3.2 A from any solute site),
    so every solute-water energy stays inside the linear part of the mesh
    (erdst halts with eng_sft beyond it in solution runs,
    erdst/engproc.F90:1150-1156).  With refs_water(config, ...) it forms the
    (engsln, engref/corref) pair from which slvfe computes the solvation
    free energy."""
    seed = 20261202 + 100 + config
    solute = solute or tetra_species()
    topo = build_topology([solute, water_species(nwater)], solute_index=0, refs=False)
    box = np.array([L, L, L])
    geom_w = water_geometry()
    centre = box / 2

    def frame(f):
        rng = _stream(seed, f)
        R = random_rotations(rng, 1)[0]
        sx = centre + solute.geometry @ R.T
        wc = lattice_centres(rng, nwater, box, avoid=sx, avoid_r=3.2, spare=max(64, nwater // 50))
        wx = place_rigid(rng, geom_w, wc)
        return np.concatenate([sx, wx]).astype(np.float32)

    mesh = linear_log_mesh(-20.0, 5.0, 0.05, 1.0e4, 100)
    w = Workload(name or f"config{config}-soln", topo, box, nframes=nframes, nins=1,
                 params=dict(ms=(grid, grid, grid)), mesh=[mesh], seed=seed, slttype=1)
    w._frame_fn = frame
    return w


def dmu_pair(config: int = 0, nframes: int | None = None):
    """(solution, reference) workloads of the tetrahedral solute in water at
    the sizes of configuration 0 (1000 waters) or 2 (30 000 waters)."""
    if config == 0:
        refs = config0()
        soln = soln_water(0, 1000, 31.04, nframes or 10, 32)
    elif config == 2:
        refs = config2()
        soln = soln_water(2, 30000, 67.0, nframes or 200, 72)
    else:
        raise ValueError("dmu_pair covers configurations 0 and 2")
    return soln, refs


def region_lattice(rng: np.random.Generator, n: int, lo, hi, jitter: float = 0.5) -> np.ndarray:
    """n centres on a jittered lattice filling the box region [lo, hi) (first
    n of a seeded permutation of the lattice sites)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    ext = hi - lo
    # lattice counts per axis proportional to the extents
    h = (ext.prod() / n) ** (1.0 / 3.0)
    m = np.maximum(1, np.floor(ext / h).astype(int))
    while m.prod() < n:
        m[np.argmax(ext / m)] += 1
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in m], indexing="ij"), -1).reshape(-1, 3)
    pts = lo + (idx + 0.5) * (ext / m)
    pts = pts[rng.permutation(len(pts))[:n]]
    return pts + rng.uniform(-jitter, jitter, pts.shape)


def ethanol_like(rng: np.random.Generator, nmol: int) -> Species:
    """Rigid 9-site molecule: a random 9-site chain (1.4 A bonds), random
    charges summing to zero, sigma U[2.5, 3.8] A, epsilon U[0.03, 0.2]."""
    q = rng.normal(0.0, 0.3, 9)
    q -= q.mean()
    geom = random_chain(rng, 9, bond=1.4, min_sep=1.9)
    return Species("ethanol-like", nmol, q, rng.uniform(0.03, 0.2, 9), rng.uniform(2.5, 3.8, 9), geom)


@dataclass
class MultiSoluteWorkload:
    """One solvent trajectory and several test particles inserted into it
    (configuration 3).  topo_for(k) is the topology of the run of solute k;
    insertions(f, k) its insertions of frame f."""
    name: str
    solvent: list
    solutes: list
    box: np.ndarray
    nframes: int
    nins: int
    params: dict
    mesh: list
    seed: int
    slttype: int = 2
    _frame_fn: object = None
    _ins_fn: object = None

    @property
    def cell(self) -> np.ndarray:
        return np.diag(self.box).astype(np.float32)

    def topo_for(self, k: int) -> Topology:
        """Topology of solute k's run; the LJ type table covers every solute,
        so all runs (and one multi-solute context) share it."""
        return build_topology(list(self.solvent) + [self.solutes[k]], solute_index=len(self.solvent), refs=True,
                              extra_lj_species=self.solutes)

    def frame(self, f: int) -> np.ndarray:
        return self._frame_fn(f)

    def insertions(self, f: int, k: int) -> np.ndarray:
        return self._ins_fn(f, k)

    def workload(self, k: int) -> Workload:
        """The single-solute view of solute k (a Workload)."""
        w = Workload(f"{self.name}-solute{k}", self.topo_for(k), self.box, nframes=self.nframes, nins=self.nins,
                     params=dict(self.params), mesh=self.mesh, seed=self.seed, slttype=self.slttype)
        w._frame_fn = self._frame_fn
        w._ins_fn = lambda f, _k=k: self._ins_fn(f, _k)
        return w

    def uvmax_uvsoft_uvcrd(self):
        return Workload.uvmax_uvsoft_uvcrd(self)


def config3(nins: int = 500, nframes: int = 100) -> MultiSoluteWorkload:
    """Two-species solvent (12 000 waters + 2000 rigid 9-site ethanol-like
    molecules) in an 80 A cube; 16 rigid solutes with log-uniform site counts
    in [1, 60] and zero-sum random charges; 500 insertions per solute per
    frame; 800 bins per species (500 linear from -60 kcal/mol, 300
    log-spaced up to 1e4); PME 80^3."""
    seed = 20261202 + 3
    rng0 = np.random.Generator(np.random.Philox(key=seed))
    eth = ethanol_like(rng0, 2000)
    nw = 12000
    solutes = []
    for k in range(16):
        ns = int(np.clip(np.round(np.exp(rng0.uniform(0.0, np.log(60.0)))), 1, 60))
        sp = random_solute(rng0, f"solute{k}", ns)
        if ns == 1:
            sp.charge[:] = 0.0
        solutes.append(sp)
    solvent = [water_species(nw), eth]
    L = 80.0
    box = np.array([L, L, L])
    geom_w = water_geometry()

    def frame(f):
        rng = _stream(seed, 2 * f)
        centres = lattice_centres(rng, nw + eth.nmol, box)
        wx = place_rigid(rng, geom_w, centres[:nw])
        ex = place_rigid(rng, eth.geometry, centres[nw:])
        return np.concatenate([wx, ex]).astype(np.float32)

    def ins(f, k):
        rng = _stream(seed, 2 * f + 1 + 1000 * (k + 1))
        sp = solutes[k]
        centres = rng.uniform(0.0, 1.0, (nins, 3)) * box
        return place_rigid(rng, sp.geometry, centres).reshape(nins, sp.nsite, 3).astype(np.float32)

    # the 45-site solute reaches below -20 kcal/mol against ethanol-like molecules
    mesh = linear_log_mesh(-60.0, 5.0, 0.13, 1.0e4, 300)
    w = MultiSoluteWorkload("config3", solvent, solutes, box, nframes, nins, dict(ms=(80, 80, 80)),
                            [mesh, mesh], seed)
    w._frame_fn, w._ins_fn = frame, ins
    return w


def config4(nins: int = 500, nframes: int = 50) -> Workload:
    """Orthorhombic 120 x 120 x 100 A box: a central 40 A slab (z in [30, 70))
    of 3000 rigid random 50-site chains (zero-sum charges N(0, 0.2 e), sigma
    U[3, 4] A) flanked by 16 667 rigid waters; a flexible 150-site solute
    (each insertion a jittered conformation) inserted at 500 positions per
    frame inside the slab; PME 128 x 128 x 108 order 4; 600 bins per species
    (500 linear from -150 kcal/mol, 100 log-spaced up to 1e4)."""
    seed = 20261202 + 4
    rng0 = np.random.Generator(np.random.Philox(key=seed))
    q = rng0.normal(0.0, 0.2, 50)
    q -= q.mean()
    chain = Species("chain50", 3000, q, rng0.uniform(0.03, 0.2, 50), rng0.uniform(3.0, 4.0, 50),
                    random_chain(rng0, 50))
    solute = random_solute(rng0, "chain150", 150)
    nw = 16667
    box = np.array([120.0, 120.0, 100.0])
    topo = build_topology([chain, water_species(nw), solute], solute_index=2, refs=True)
    geom_w = water_geometry()
    z0, z1 = 30.0, 70.0

    def frame(f):
        rng = _stream(seed, 2 * f)
        cc = region_lattice(rng, chain.nmol, (0.0, 0.0, z0), (box[0], box[1], z1))
        cx = place_rigid(rng, chain.geometry, cc)
        nlo = nw // 2
        wlo = region_lattice(rng, nlo, (0.0, 0.0, 0.0), (box[0], box[1], z0))
        whi = region_lattice(rng, nw - nlo, (0.0, 0.0, z1), (box[0], box[1], box[2]))
        wx = place_rigid(rng, geom_w, np.concatenate([wlo, whi]))
        return np.concatenate([cx, wx]).astype(np.float32)

    def ins(f):
        rng = _stream(seed, 2 * f + 1)
        centres = np.column_stack([rng.uniform(0.0, box[0], nins), rng.uniform(0.0, box[1], nins),
                                   rng.uniform(z0, z1, nins)])
        R = random_rotations(rng, nins)
        conf = solute.geometry[None, :, :] + rng.normal(0.0, 0.1, (nins, solute.nsite, 3))
        conf -= conf.mean(axis=1, keepdims=True)
        x = centres[:, None, :] + np.einsum("nij,nsj->nsi", R, conf)
        return x.astype(np.float32)

    # chain-solute energies reach -61 kcal/mol within 5 frames in the slab
    mesh = linear_log_mesh(-150.0, 5.0, 0.31, 1.0e4, 100)
    w = Workload("config4", topo, box, nframes=nframes, nins=nins, params=dict(ms=(128, 128, 108)),
                 mesh=[mesh, mesh], seed=seed, slttype=3)
    w._frame_fn, w._ins_fn = frame, ins
    return w

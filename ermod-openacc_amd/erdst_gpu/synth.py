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

"""Shared helpers of the test-suite: build the oracle's view of a workload,
run the oracle, run the HIP library, and compare."""
from __future__ import annotations

import numpy as np

import oracle as O
from erdst_gpu import synth
from erdst_gpu.params import hot_path_params

# Tolerance of per-(insertion, molecule) energies of the FP32 GPU path
# against the FP64 oracle, for a solute of n sites:
#     |E_gpu - E_64| <= ATOL_SITE sqrt(n) + RTOL |E_64|
#  - RTOL = 1e-5: the relative bar for pair energies.  Pair distances come
#    from lattice coordinates (exact periodic wrap), so a pair energy carries
#    only FP32 rounding of its own short difference vector and arithmetic:
#    at most 7.4e-6 relative above 10 kcal/mol and 1.6e-6 above 100 kcal/mol
#    over every benchmark workload.
#  - ATOL_SITE sqrt(n), ATOL_SITE = 2e-5 kcal/mol: below a few kcal/mol the
#    energy is a sum of terms that cancel, dominated by the FP32 PME potential
#    gathered at the molecule's sites, one such sum per solute site, so its
#    error grows like the square root of the site count: 8.6e-6 at 5 sites
#    (configuration 2), 3.0e-5 at 12, 4.8e-5 at 25, 6.6e-5 at 45
#    (configuration 3); the bar is at least twice each.
# The solute self energy uvengy0 sums the FP32 reciprocal-space self term (up
# to ~1e2 kcal/mol for 25 charged sites, from the FP32 solute spectra) with
# the real-space (FP64) and residual terms, which cancel it to below
# 1 kcal/mol, so its error grows with the site count: 5.7e-5 at 25 sites,
# 1.1e-4 at 45 (configuration 3), 1.6e-4 at 150 (configuration 4); its bar
# adds SELF_ATOL_SITE n, twice the largest 2.4e-6 per site.
ATOL_SITE = 2.0e-5
RTOL = 1.0e-5
SELF_ATOL_SITE = 5.0e-6


def oracle_system(topo) -> O.OracleSystem:
    return O.OracleSystem(topo.mol_begin, topo.charge64, topo.ljtype,
                          topo.ljene64.reshape(topo.ntype, topo.ntype, order="F"),
                          topo.ljlensq64.reshape(topo.ntype, topo.ntype, order="F"))


def oracle_params(p: dict) -> O.Params:
    return O.make_params(elecut=p["elecut"], lwljcut=p["lwljcut"], upljcut=p["upljcut"],
                         screen=p["screen"], ljswitch=p["ljswitch"], splodr=p["splodr"], ms=p["ms"],
                         cltype=p.get("cltype", 2), boxshp=p.get("boxshp", 1))


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


def oracle_distribution(w, params, escrd=None) -> O.Distribution:
    uvmax, uvsoft, uvcrd = w.uvmax_uvsoft_uvcrd()
    invcl, vol = O.inverted_cell(w.cell.astype(np.float64))
    qtot = float(np.float64(w.topo.charge32.astype(np.float64).sum()))
    return O.Distribution(params["slttype"], uvmax, uvsoft, uvcrd, params["temp"], vol, params["screen"],
                          qtot, wgtslf=params["wgtslf"], estype=params["estype"], selfcal=params["selfcal"],
                          corrcal=params["corrcal"], escrd=escrd, cltype=params.get("cltype", 2))


# Self-energy mesh for the small test solutes: their self energies (solute
# intramolecular + Ewald self terms) lie within a few 1e-4 kcal/mol of zero,
# so 2e-5-wide bins spread 100 insertions over tens of bins.
SELF_MESH = np.linspace(-2.0e-3, 2.0e-3, 201)


def species_of(topo, nins, soln_tag=None):
    """uvspec per insertion row; the soln solute's own entry is skipped (0)."""
    sp = np.tile(topo.uvspec, (nins, 1))
    if soln_tag is not None:
        for n, t in enumerate(np.atleast_1d(soln_tag)):
            sp[n, t] = 0
    return sp


def energy_tolerance(e_ref, nslt: int, rtol: float = RTOL):
    """The energy bar of a solute of nslt sites at the FP64 energies e_ref."""
    e_ref = np.asarray(e_ref, np.float64)
    return ATOL_SITE * np.sqrt(max(int(nslt), 1)) + rtol * np.abs(e_ref)


def energy_mismatch(e_gpu, e_ref, nslt: int, rtol: float = RTOL):
    """(count beyond the bar, largest excess over it, its index)."""
    e_gpu = np.asarray(e_gpu, np.float64)
    e_ref = np.asarray(e_ref, np.float64)
    excess = np.abs(e_gpu - e_ref) - energy_tolerance(e_ref, nslt, rtol)
    worst = np.unravel_index(np.argmax(excess), excess.shape)
    return int((excess > 0).sum()), float(excess.max()), worst


def self_mismatch(u0_gpu, u0_ref, nslt: int):
    """energy_mismatch for the solute self energies uvengy0."""
    u0_gpu = np.asarray(u0_gpu, np.float64)
    u0_ref = np.asarray(u0_ref, np.float64)
    tol = energy_tolerance(u0_ref, nslt) + SELF_ATOL_SITE * max(int(nslt), 1)
    excess = np.abs(u0_gpu - u0_ref) - tol
    worst = np.unravel_index(np.argmax(excess), excess.shape)
    return int((excess > 0).sum()), float(excess.max()), worst


def nsites(w) -> int:
    """Site count of the workload's solute (the first solute molecule of a
    solution run)."""
    t = w.topo
    if t.slt_charge32.size:
        return int(t.slt_charge32.size)
    j = int(np.atleast_1d(t.tagslt)[0])
    return int(t.mol_begin[j + 1] - t.mol_begin[j])


def bar_text(nslt: int) -> str:
    return (f"|dE| <= {ATOL_SITE * np.sqrt(max(nslt, 1)):.2e} + {RTOL} |E| kcal/mol "
            f"({ATOL_SITE} sqrt({nslt} sites) + {RTOL} relative)")


REL_FLOOR = 1.0  # kcal/mol: relative errors are reported for |E| at or above it
E_WALL = 20.0    # kcal/mol: the report's separate LJ-wall regime


def energy_report(e_gpu, e_ref) -> dict:
    """Error summary of FP32 energies against the FP64 oracle, by regime:
    the largest absolute error where |E| < REL_FLOOR (sums that cancel, where
    only an absolute bar means anything), the largest relative error for
    REL_FLOOR <= |E| < E_WALL, and the largest relative error on the LJ wall
    (|E| >= E_WALL)."""
    e_gpu = np.asarray(e_gpu, np.float64)
    e_ref = np.asarray(e_ref, np.float64)
    d = np.abs(e_gpu - e_ref)
    a = np.abs(e_ref)
    small = a < REL_FLOOR
    mid = (a >= REL_FLOOR) & (a < E_WALL)
    wall = a >= E_WALL
    return dict(max_abs_small=float(d[small].max()) if small.any() else 0.0,
                max_rel_mid=float((d[mid] / a[mid]).max()) if mid.any() else 0.0,
                max_rel_wall=float((d[wall] / a[wall]).max()) if wall.any() else 0.0,
                max_abs=float(d[~wall].max()) if (~wall).any() else 0.0,
                n=int(e_ref.size), by_decade=error_by_decade(e_gpu, e_ref))


def error_by_decade(e_gpu, e_ref, lo=-3, hi=6) -> list:
    """[decade, count, largest absolute error, largest relative error] of
    the energies with 10^decade <= |E| < 10^(decade+1), for decades lo..hi
    (the first also holds |E| below 10^lo, the last everything above)."""
    e_gpu = np.asarray(e_gpu, np.float64)
    e_ref = np.asarray(e_ref, np.float64)
    d = np.abs(e_gpu - e_ref).ravel()
    a = np.abs(e_ref).ravel()
    dec = np.clip(np.floor(np.log10(np.maximum(a, 1e-300))), lo, hi).astype(int)
    out = []
    for k in range(lo, hi + 1):
        m = dec == k
        if m.any():
            rel = d[m] / np.maximum(a[m], 1e-300)
            out.append([k, int(m.sum()), float(f"{d[m].max():.3e}"),
                        float(f"{rel.max():.3e}") if k > lo else None])
    return out


def flip_report(e_gpu, e_ref, species, uvmax, uvcrd, nslt: int) -> dict:
    """Molecules whose bin differs between the GPU's FP32 energies and the
    oracle's FP64 energies.  e_*: [nins, nmol]; species: [nmol] (0 = the
    solute itself, skipped).  Every flip must have a bin edge between the two
    energies within the energy tolerance.  Reported: the count, the largest
    distance of the FP64 energy from that edge in kcal/mol and as a fraction
    of the energy tolerance, and the largest such distance in float32 ulps of
    the larger of |E| and |edge| (edges at exactly 0 included)."""
    e_gpu = np.asarray(e_gpu, np.float64)
    e_ref = np.asarray(e_ref, np.float64)
    idmin = np.concatenate([[0], np.cumsum(uvmax)[:-1]])
    nflip, unexplained = 0, 0
    max_dist, max_frac, max_ulp = 0.0, 0.0, 0.0
    for sp in range(1, len(uvmax) + 1):
        cols = np.nonzero(np.asarray(species) == sp)[0]
        if cols.size == 0:
            continue
        mesh = np.asarray(uvcrd[idmin[sp - 1]:idmin[sp - 1] + uvmax[sp - 1]], np.float64)
        g, r = e_gpu[:, cols], e_ref[:, cols]
        bg = np.searchsorted(mesh, g, side="left")
        br = np.searchsorted(mesh, r, side="left")
        diff = bg != br
        if not diff.any():
            continue
        nflip += int(diff.sum())
        lo = np.minimum(g, r)[diff]
        hi = np.maximum(g, r)[diff]
        rd = r[diff]
        tol = energy_tolerance(rd, nslt)
        ilo = np.searchsorted(mesh, lo, side="left")
        edge = mesh[np.minimum(ilo, mesh.size - 1)]
        ok = (np.searchsorted(mesh, hi, side="left") > ilo) & (hi - lo <= tol)
        unexplained += int((~ok).sum())
        dist = np.abs(rd - edge)
        ulp = np.spacing(np.maximum(np.abs(edge), np.abs(rd)).astype(np.float32)).astype(np.float64)
        max_dist = max(max_dist, float(dist.max()))
        max_frac = max(max_frac, float((dist / tol).max()))
        max_ulp = max(max_ulp, float((dist / ulp).max()))
    return dict(flips=nflip, max_dist_to_edge=max_dist, max_dist_over_tol=max_frac,
                max_ulp_to_edge=max_ulp, unexplained=unexplained, n=int(e_ref.size))


def oracle_block(d) -> dict:
    """An oracle Distribution in er_fetch's dict layout."""
    g = dict(edens=d.edens.copy(), ecorr=d.ecorr.reshape(d.ermax, d.ermax).copy(), slnuv=d.slnuv.copy())
    g.update(d.scalars)
    return g


def oracle_dmu_blocks(soln, refs, frames_per_block, nblocks_soln, nins, f32_energies=False, nthreads=0,
                      refs_frames=None):
    """Oracle soln and refs block results for slvfe (nblocks_soln blocks of
    frames_per_block frames each; refs in one block over refs_frames frames,
    by default as many as the soln run has, with nins insertions each; a
    side with no frames or insertions is skipped).  With f32_energies the
    FP64 energies are rounded to float32 before binning."""
    def cast(x):
        return x.astype(np.float32).astype(np.float64) if f32_energies else x
    ps, pr = workload_params(soln), workload_params(refs)
    sblocks = []
    ds = oracle_distribution(soln, ps)
    sp = species_of(soln.topo, 1, soln.topo.tagslt)[0]
    for b in range(nblocks_soln):
        ds.clear()
        for f in range(b * frames_per_block, (b + 1) * frames_per_block):
            uv, uv0 = oracle_soln(soln, f, ps, nthreads=nthreads)
            ds.update(sp, cast(uv[0]), float(cast(uv0)[0]))
        sblocks.append(oracle_block(ds))
    nref = nblocks_soln * frames_per_block if refs_frames is None else refs_frames
    if nref == 0 or nins == 0:
        return sblocks, []
    dr = oracle_distribution(refs, pr)
    for f in range(nref):
        uv, uv0 = oracle_refs(refs, f, pr, nins=nins, nthreads=nthreads)
        uv, uv0 = cast(uv), cast(uv0)
        for n in range(uv.shape[0]):
            dr.update(refs.topo.uvspec, uv[n], float(uv0[n]))
    return sblocks, [oracle_block(dr)]


def gpu_dmu_blocks(soln, refs, frames_per_block, nblocks_soln, nins, mode=None):
    """The GPU path's soln and refs block results for slvfe, on the same
    frames and insertions as oracle_dmu_blocks."""
    from erdst_gpu import capi
    mode = capi.ER_BATCHED if mode is None else mode
    ps, pr = workload_params(soln), workload_params(refs)
    ctx = capi.Context(soln.topo, ps)
    ctx.set_mesh(*soln.uvmax_uvsoft_uvcrd())
    ctx.set_solute_molecules(soln.topo.tagslt)
    sblocks = []
    for b in range(nblocks_soln):
        for f in range(b * frames_per_block, (b + 1) * frames_per_block):
            ctx.set_frame(soln.frame(f), soln.cell)
            ctx.run(mode)
        sblocks.append(ctx.fetch(want_ecorr=False))
        ctx.clear()
    ctx.close()
    ctx = capi.Context(refs.topo, pr)
    ctx.set_mesh(*refs.uvmax_uvsoft_uvcrd())
    ctx.set_solute(refs.topo.slt_ljtype, refs.topo.slt_charge32)
    for f in range(nblocks_soln * frames_per_block):
        ctx.set_frame(refs.frame(f), refs.cell)
        ctx.set_insertions(refs.insertions(f)[:nins])
        ctx.run(mode)
    rblocks = [ctx.fetch()]
    ctx.close()
    return sblocks, rblocks


def plain_and_batched(w, params, f, nins, capi):
    """Run frame f's first nins insertions in both modes, each in a fresh
    context.  Returns [(uvengy, uvengy0, fetch, oracle Distribution of that
    mode's own energies)] for (batched, plain)."""
    out = []
    for mode in (capi.ER_BATCHED, capi.ER_PLAIN):
        ctx = capi.Context(w.topo, params)
        ctx.set_mesh(*w.uvmax_uvsoft_uvcrd())
        ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
        ctx.set_frame(w.frame(f), w.cell)
        ctx.set_insertions(w.insertions(f)[:nins])
        ctx.run(mode)
        ctx.sync()
        u, u0 = ctx.uvengy(nins)
        g = ctx.fetch()
        ctx.close()
        d = oracle_distribution(w, params)
        for n in range(nins):
            d.update(w.topo.uvspec, u[n].astype(np.float64), float(u0[n]))
        out.append((u, u0, g, d))
    return out


def assert_plain_matches_batched(res, nslt: int):
    """The plain path (one insertion per launch sequence) and the batched
    path add the same terms in different orders (per-atom products stored
    before the molecule sum, different FMA contraction), so their FP32
    energies agree to rounding; each path's distributions
    are exact for its own energies, so the two differ only where an energy
    sits within that rounding of a bin edge."""
    (ub, u0b, gb, db), (up, u0p, gp, dp) = res
    # each is held to the energy bar against the FP64 oracle, so they are held
    # to it against each other (sums over 50-site molecules cancel to ~1e-2
    # from terms ~1e1, so the agreement is absolute there, not relative)
    assert energy_mismatch(ub, up, nslt)[0] == 0
    assert self_mismatch(u0b, u0p, nslt)[0] == 0
    for g, d in ((gb, db), (gp, dp)):
        np.testing.assert_allclose(g["edens"] / g["engnorm"], d.edens / d.d.engnorm, rtol=1e-12, atol=0)
        shape = g["ecorr"].shape
        np.testing.assert_allclose(g["ecorr"] / g["engnorm"], d.ecorr.reshape(shape) / d.d.engnorm,
                                   rtol=1e-12, atol=0)

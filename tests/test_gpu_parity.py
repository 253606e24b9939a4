"""Parity of the HIP path (through the C ABI) against the FP64 oracle.

Energies: per (insertion, molecule) within |dE| <= ATOL_SITE sqrt(n) +
RTOL |E| for a solute of n sites (erutil).  Histograms: the GPU's counts must equal the oracle's binning of
the GPU's own FP32 energies exactly, so edens/ecorr/engnorm agree to FP64
summation-order rounding; bins that differ when the oracle bins its FP64
energies are counted and must all lie within the energy tolerance of an
edge.
"""
import os

import numpy as np
import pytest

import erutil as U
from erdst_gpu import capi, synth

pytestmark = pytest.mark.gpu


def _ctx(w, params, chunk=None, escrd=None):
    if chunk is not None:
        os.environ["ER_CHUNK"] = str(chunk)
    try:
        ctx = capi.Context(w.topo, params)
    finally:
        os.environ.pop("ER_CHUNK", None)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd(), escrd=escrd)
    if params["slttype"] == 1:
        ctx.set_solute_molecules(w.topo.tagslt)
    else:
        ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    return ctx


def _run_refs(ctx, w, f, mode, nins=None):
    ins = w.insertions(f)
    if nins is not None:
        ins = ins[:nins]
    ctx.set_frame(w.frame(f), w.cell)
    ctx.set_insertions(ins)
    ctx.run(mode)
    ctx.sync()
    return ctx.uvengy(ins.shape[0])


@pytest.fixture(scope="module")
def cfg0():
    return synth.config0()


@pytest.mark.parametrize("mode", [capi.ER_BATCHED, capi.ER_PLAIN])
def test_refs_energies_cfg0(cfg0, mode):
    p = U.workload_params(cfg0)
    ctx = _ctx(cfg0, p)
    u, u0 = _run_refs(ctx, cfg0, 0, mode)
    uo, u0o = U.oracle_refs(cfg0, 0, p)
    nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(cfg0))
    print(f"mode {mode}: max |dE| {np.abs(u - uo).max():.3e}, worst excess {worst:.3e} at {at}, "
          f"E_gpu {u[at]:.6f} E_64 {uo[at]:.6f}")
    assert nbad == 0
    nbad0, _, at0 = U.self_mismatch(u0, u0o, U.nsites(cfg0))
    assert nbad0 == 0, (u0[at0], u0o[at0])


def test_uvengy_row_count(cfg0):
    """er_get_nins gives the rows er_get_uvengy writes (the selected solute's
    insertions in the last run); the wrapper refuses another count instead of
    letting the library write past the caller's array."""
    p = U.workload_params(cfg0)
    ctx = _ctx(cfg0, p)
    assert ctx.nins() == 0
    ctx.set_frame(cfg0.frame(0), cfg0.cell)
    ctx.set_insertions(cfg0.insertions(0)[:24])
    ctx.run(capi.ER_BATCHED)
    ctx.sync()
    assert ctx.nins() == 24
    with pytest.raises(ValueError):
        ctx.uvengy(16)
    u, u0 = ctx.uvengy()
    assert u.shape == (24, cfg0.topo.nmol) and u0.shape == (24,)
    np.testing.assert_array_equal(u, ctx.uvengy(24)[0])


def test_plain_matches_batched(cfg0):
    p = U.workload_params(cfg0)
    ub, u0b = _run_refs(_ctx(cfg0, p), cfg0, 1, capi.ER_BATCHED, nins=40)
    up, u0p = _run_refs(_ctx(cfg0, p), cfg0, 1, capi.ER_PLAIN, nins=40)
    np.testing.assert_allclose(ub, up, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(u0b, u0p, rtol=1e-6, atol=1e-6)


def test_plain_async_equals_plain(cfg0):
    """ER_PLAIN_ASYNC issues ER_PLAIN's launches without the host waiting
    after each insertion: energies and distributions bit for bit equal."""
    p = U.workload_params(cfg0)
    out = []
    for mode in (capi.ER_PLAIN, capi.ER_PLAIN_ASYNC):
        ctx = _ctx(cfg0, p)
        u, u0 = _run_refs(ctx, cfg0, 0, mode, nins=20)
        out.append((u, u0, ctx.fetch()))
        ctx.close()
    (ua, u0a, ga), (ub, u0b, gb) = out
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(u0a, u0b)
    for k in ("edens", "ecorr", "minuv", "maxuv"):
        np.testing.assert_array_equal(ga[k], gb[k])
    for k in ("engnorm", "engsmpl", "avslf", "voffset"):
        assert ga[k] == gb[k]


def test_multi_chunk_equals_single_chunk(cfg0):
    p = U.workload_params(cfg0)
    u1, _ = _run_refs(_ctx(cfg0, p, chunk=32), cfg0, 2, capi.ER_BATCHED)
    u2, _ = _run_refs(_ctx(cfg0, p, chunk=128), cfg0, 2, capi.ER_BATCHED)
    np.testing.assert_array_equal(u1, u2)


def _distribution_check(ctx, w, p, frames, mode):
    """Accumulate over frames on the GPU; feed the GPU energies to the oracle's
    update_histogram and compare every accumulator."""
    d_gpu_e = U.oracle_distribution(w, p)   # oracle binning of GPU energies
    d_64 = U.oracle_distribution(w, p)      # oracle binning of FP64 energies
    flips = 0
    for f in frames:
        u, u0 = _run_refs(ctx, w, f, mode)
        uo, u0o = U.oracle_refs(w, f, p)
        for n in range(u.shape[0]):
            d_gpu_e.update(w.topo.uvspec, u[n].astype(np.float64), float(u0[n]))
            d_64.update(w.topo.uvspec, uo[n], u0o[n])
        # bins of GPU vs FP64 energies
        mesh = w.uvmax_uvsoft_uvcrd()[2]
        b_gpu = np.searchsorted(mesh, u.astype(np.float64), side="left")
        b_64 = np.searchsorted(mesh, uo, side="left")
        diff = b_gpu != b_64
        if diff.any():
            # every flip must be explained by the energy tolerance straddling an edge
            lo = np.minimum(u.astype(np.float64), uo)[diff]
            hi = np.maximum(u.astype(np.float64), uo)[diff]
            tol = U.energy_tolerance(uo[diff], U.nsites(w))
            edge_between = np.searchsorted(mesh, hi, side="left") > np.searchsorted(mesh, lo, side="left")
            assert edge_between.all()
            assert (hi - lo <= tol).all()
        flips += int(diff.sum())
    g = ctx.fetch()
    return g, d_gpu_e, d_64, flips


def _assert_same_distribution(g, d):
    shape = g["ecorr"].shape if g["ecorr"] is not None else None
    np.testing.assert_allclose(g["edens"] / g["engnorm"], d.edens / d.d.engnorm, rtol=1e-12, atol=0)
    if shape is not None:
        np.testing.assert_allclose(g["ecorr"] / g["engnorm"], d.ecorr.reshape(shape) / d.d.engnorm,
                                   rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["engnorm"], d.d.engnorm, rtol=1e-7)
    np.testing.assert_allclose(g["avslf"], d.d.avslf, rtol=1e-7)


def test_histograms_cfg0(cfg0):
    p = U.workload_params(cfg0)
    ctx = _ctx(cfg0, p)
    g, dg, d64, flips = _distribution_check(ctx, cfg0, p, [0, 1], capi.ER_BATCHED)
    print(f"bin flips vs FP64 energies: {flips} of {2 * 100 * cfg0.topo.nmol}")
    # Exact binning of the GPU's own energies.  The GPU forms total_weight in
    # FP32 like the reference's single build (engproc.F90:873,901-903,
    # 1026-1036), the oracle in FP64: one common factor (1 + O(1e-8)) scales
    # every weight here (w_sys, w_slt, V and Q are constant), so it cancels
    # in the normalised distributions, which must then agree to FP64
    # summation order -- a single miscounted molecule would show at >= 1e-6.
    _assert_same_distribution(g, dg)
    assert g["engsmpl"] == dg.d.engsmpl == 200
    assert g["voffset"] == dg.d.voffset
    np.testing.assert_array_equal(g["minuv"].astype(np.float64), dg.minuv)
    np.testing.assert_array_equal(g["maxuv"].astype(np.float64), dg.maxuv)
    # symmetric ecorr and per-species row sums N_p * edens (sum_j c_j = N_p)
    np.testing.assert_array_equal(g["ecorr"], g["ecorr"].T)
    np.testing.assert_allclose(g["ecorr"].sum(1), cfg0.topo.nmol * g["edens"], rtol=1e-12)
    # against the FP64 oracle end to end: weights agree to FP32-energy rounding
    np.testing.assert_allclose(g["engnorm"], d64.d.engnorm, rtol=1e-4)


def test_plain_histograms_equal_batched(cfg0):
    p = U.workload_params(cfg0)
    U.assert_plain_matches_batched(U.plain_and_batched(cfg0, p, 3, 50, capi), U.nsites(cfg0))


@pytest.mark.parametrize("ljswitch", [1, 2, 3])
def test_lj_switch_forms(cfg0, ljswitch):
    p = U.workload_params(cfg0, ljswitch=ljswitch)
    u, _ = _run_refs(_ctx(cfg0, p), cfg0, 0, capi.ER_BATCHED, nins=20)
    uo, _ = U.oracle_refs(cfg0, 0, p, nins=20)
    assert U.energy_mismatch(u, uo, U.nsites(cfg0))[0] == 0


def test_noncubic_grid_and_order6():
    w = synth.refs_water(0, 1000, 31.04, 24, 1, 32)
    w.params["ms"] = (40, 32, 32)
    p = U.workload_params(w, splodr=6)
    u, u0 = _run_refs(_ctx(w, p), w, 0, capi.ER_BATCHED)
    uo, u0o = U.oracle_refs(w, 0, p)
    assert U.energy_mismatch(u, uo, U.nsites(w))[0] == 0
    assert U.self_mismatch(u0, u0o, U.nsites(w))[0] == 0


@pytest.mark.parametrize("ms,order", [((60, 60, 60), 5), ((50, 45, 36), 6), ((100, 100, 100), 4),
                                      ((32, 32, 32), 3), ((32, 32, 32), 7), ((36, 36, 36), 8),
                                      ((45, 40, 40), 10)])
def test_any_grid_and_order(ms, order):
    """PME grids without a compiled four-step shape (the direct-DFT plane and
    axis-1 kernels: 60^3, 50 x 45 x 36 with K2 != K3, 100^3, 45 x 40 x 40)
    and spline orders 3 to 10, the brick gather's 3 to 8 and the
    per-molecule gather's 10 (erdst/fft_iface.F90:38-50 plans any size,
    erdst/spline.F90:25-58 any order; odd orders zero engfac at m = K/2,
    erdst/recpcal.F90:142): batched and plain energies against the oracle."""
    w = synth.refs_water(0, 1000, 31.04, 24, 1, 32)
    w.params["ms"] = ms
    p = U.workload_params(w, splodr=order)
    uo, u0o = U.oracle_refs(w, 0, p)
    for mode in (capi.ER_BATCHED, capi.ER_PLAIN):
        ctx = _ctx(w, p)
        u, u0 = _run_refs(ctx, w, 0, mode)
        nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(w))
        print(f"grid {ms} order {order} mode {mode}: {U.energy_report(u, uo)}, uvengy0 "
              f"{np.abs(u0 - u0o).max():.2e}, worst excess {worst:.3e}", flush=True)
        assert nbad == 0, (mode, worst, at, u[at], uo[at])
        assert U.self_mismatch(u0, u0o, U.nsites(w))[0] == 0
        ctx.close()


def test_grid_and_order_limits():
    w = synth.refs_water(0, 1000, 31.04, 4, 1, 32)
    for ms, order in (((32, 32, 32), 2), ((32, 32, 32), 11), ((12, 32, 32), 4), ((32, 200, 200), 4)):
        w.params["ms"] = ms
        with pytest.raises(capi.ErError) as e:
            capi.Context(w.topo, U.workload_params(w, splodr=order))
        assert e.value.status == capi.ER_E_UNSUPPORTED


def test_dense_bricks_read_overflow_records():
    """8000 waters in a 62.1 A cube on a 32^3 grid: about 250 atoms per 8x8x6
    brick, so the batched gather runs atoms beyond its 128 LDS-staged records
    with records read from global memory; energies to the oracle bar."""
    w = synth.refs_water(9, 8000, 62.1, 24, 1, 32)
    p = U.workload_params(w)
    u, u0 = _run_refs(_ctx(w, p), w, 0, capi.ER_BATCHED)
    uo, u0o = U.oracle_refs(w, 0, p)
    assert U.energy_mismatch(u, uo, U.nsites(w))[0] == 0
    assert U.self_mismatch(u0, u0o, U.nsites(w))[0] == 0


def test_soln_config1_two_frames():
    w = synth.config1()
    p = U.workload_params(w, corrcal=1)
    ctx = _ctx(w, p)
    d_gpu_e = U.oracle_distribution(w, p)
    for f in (0, 1):
        ctx.set_frame(w.frame(f), w.cell)
        ctx.run(capi.ER_BATCHED)
        ctx.sync()
        u, u0 = ctx.uvengy(1)
        uo, u0o = U.oracle_soln(w, f, p)
        assert u[0, w.topo.tagslt[0]] == 0.0
        nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(w))
        assert nbad == 0, (worst, at, u[at], uo[at])
        assert U.self_mismatch(u0, u0o, U.nsites(w))[0] == 0
        sp = U.species_of(w.topo, 1, w.topo.tagslt)
        d_gpu_e.update(sp[0], u[0].astype(np.float64), float(u0[0]))
        fl = ctx.flceng(1)
        np.testing.assert_allclose(fl[0], d_gpu_e.flceng, rtol=1e-6)
    g = ctx.fetch()
    _assert_same_distribution(g, d_gpu_e)
    np.testing.assert_allclose(g["slnuv"], d_gpu_e.slnuv, rtol=1e-6)


def test_selfcal_self_energy_distribution(cfg0):
    """selfcal = 1: eself, binned on the GPU from its own self energies, equals
    the oracle's binning of those energies; every sample lands in one bin
    (sum eself = engnorm); the GPU's self energies agree with the FP64
    oracle's, and the solvent distributions are those of a selfcal = 0 run."""
    p = U.workload_params(cfg0, selfcal=1)
    ctx = _ctx(cfg0, p, escrd=U.SELF_MESH)
    ctx0 = _ctx(cfg0, U.workload_params(cfg0))
    d = U.oracle_distribution(cfg0, p, escrd=U.SELF_MESH)
    for f in (0, 1):
        u, u0 = _run_refs(ctx, cfg0, f, capi.ER_BATCHED)
        _run_refs(ctx0, cfg0, f, capi.ER_BATCHED)
        uo, u0o = U.oracle_refs(cfg0, f, p)
        assert U.self_mismatch(u0, u0o, U.nsites(cfg0))[0] == 0
        for n in range(u.shape[0]):
            d.update(cfg0.topo.uvspec, u[n].astype(np.float64), float(u0[n]))
    g, g0 = ctx.fetch(), ctx0.fetch()
    assert g["eself"].shape == (U.SELF_MESH.size,)
    assert np.count_nonzero(g["eself"]) > 10        # the mesh resolves the self energies
    np.testing.assert_allclose(g["eself"] / g["engnorm"], d.eself / d.d.engnorm, rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["eself"].sum(), g["engnorm"], rtol=1e-12)
    assert g["minuv"][0] == d.minuv[0] and g["maxuv"][0] == d.maxuv[0]
    np.testing.assert_array_equal(g["edens"], g0["edens"])
    np.testing.assert_array_equal(g["ecorr"], g0["ecorr"])


@pytest.mark.parametrize("order", [4, 6])
def test_pppm_energies_cfg0(cfg0, order):
    """cltype 3 (PPPM): the GPU's influence function kernel and the unchanged
    PME pipeline against the FP64 oracle's PPPM."""
    p = U.workload_params(cfg0, cltype=3, splodr=order)
    ctx = _ctx(cfg0, p)
    for mode in (capi.ER_BATCHED, capi.ER_PLAIN):
        u, u0 = _run_refs(ctx, cfg0, 0, mode, nins=24)
        uo, u0o = U.oracle_refs(cfg0, 0, p, nins=24)
        nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(cfg0))
        assert nbad == 0, (mode, worst, at, u[at], uo[at])
        assert U.self_mismatch(u0, u0o, U.nsites(cfg0))[0] == 0
    # the GPU's PPPM - PME difference follows the oracle's (both near
    # convergence at order 6, where they differ by a few 1e-5 kcal/mol)
    pme = U.workload_params(cfg0, splodr=order)
    up, _ = _run_refs(_ctx(cfg0, pme), cfg0, 0, capi.ER_BATCHED, nins=24)
    uop, _ = U.oracle_refs(cfg0, 0, pme, nins=24)
    # per entry the two methods differ by less than the energy tolerance
    # here; over all entries the GPU's PPPM energies sit on the oracle's
    # PPPM values and well away from its PME values
    # (entries below the LJ wall, |E| < 20 kcal/mol, where the methods
    # differ by more than 1e-5 kcal/mol).  At order 6 both methods are
    # converged to within the FP32 noise of the gather, and only the parity
    # above applies.
    if order == 4:
        ug = u.astype(np.float64)
        sel = (np.abs(uo) < U.E_WALL) & (np.abs(uo - uop) > 1e-5)
        assert sel.sum() >= 20
        rms = lambda a: float(np.sqrt(np.mean(a[sel] ** 2)))
        assert rms(ug - uo) < 0.3 * rms(ug - uop)


def test_bare_coulomb_cfg0(cfg0):
    """cltype 0: real-space pairs with plain q q / r within elecut, no
    reciprocal or residual part, uvengy0 = 0, volcorrect by the volume only;
    energies against the oracle and the distribution against the oracle's
    binning of the GPU's energies."""
    p = U.workload_params(cfg0, cltype=0)
    ctx = _ctx(cfg0, p)
    d = U.oracle_distribution(cfg0, p)
    for mode in (capi.ER_BATCHED, capi.ER_PLAIN):
        u, u0 = _run_refs(ctx, cfg0, 0, mode, nins=32)
        uo, u0o = U.oracle_refs(cfg0, 0, p, nins=32)
        nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(cfg0))
        assert nbad == 0, (mode, worst, at, u[at], uo[at])
        assert np.all(u0 == 0.0) and np.all(u0o == 0.0)
        for n in range(u.shape[0]):
            d.update(cfg0.topo.uvspec, u[n].astype(np.float64), float(u0[n]))
    g = ctx.fetch()
    _assert_same_distribution(g, d)
    assert g["engsmpl"] == 64
    np.testing.assert_allclose(g["engnorm"], 64 * np.float32(np.prod(np.diag(cfg0.cell))), rtol=1e-6)


@pytest.mark.parametrize("mode", [capi.ER_BATCHED, capi.ER_PLAIN])
def test_open_system_bare_coulomb_cfg0(cfg0, mode):
    """boxshp 0 (an open cluster: the configuration-0 coordinates taken as
    they are), bare Coulomb between every pair, constant volume: energies
    against the oracle; uvengy0 = 0; the distribution against the oracle's
    binning of the GPU's energies with no volume factor."""
    p = U.workload_params(cfg0, cltype=0, boxshp=0, estype=1)
    ctx = _ctx(cfg0, p)
    d = U.oracle_distribution(cfg0, p)
    u, u0 = _run_refs(ctx, cfg0, 0, mode, nins=24)
    uo, u0o = U.oracle_refs(cfg0, 0, p, nins=24)
    nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(cfg0))
    assert nbad == 0, (worst, at, u[at], uo[at])
    assert np.all(u0 == 0.0)
    # far molecules still interact (no Coulomb cutoff in an open system)
    assert np.count_nonzero(u) > 0.9 * u.size
    for n in range(u.shape[0]):
        d.update(cfg0.topo.uvspec, u[n].astype(np.float64), float(u0[n]))
    g = ctx.fetch()
    _assert_same_distribution(g, d)
    np.testing.assert_allclose(g["engnorm"], 24.0, rtol=1e-6)


def test_open_system_needs_bare_coulomb(cfg0):
    with pytest.raises(capi.ErError):
        capi.Context(cfg0.topo, U.workload_params(cfg0, boxshp=0, estype=1))


def test_rotated_cell_cfg0(cfg0):
    """The whole configuration-0 system rotated, so the cell is not upper
    triangular (make_cell_uppertriangular, realcal.F90:776-817): energies
    against the oracle on the same inputs, and against the unrotated
    system's oracle energies (rotation only changes float32 rounding)."""
    rng = np.random.default_rng(21)
    qm, rm = np.linalg.qr(rng.normal(size=(3, 3)))
    R = qm * np.sign(np.diag(rm))
    if np.linalg.det(R) < 0:
        R = -R
    p = U.workload_params(cfg0)
    pos = (cfg0.frame(0).astype(np.float64) @ R.T).astype(np.float32)
    cell = (R @ cfg0.cell.astype(np.float64)).astype(np.float32)
    ins = (cfg0.insertions(0)[:32].astype(np.float64) @ R.T).astype(np.float32)
    assert abs(cell[1, 0]) + abs(cell[2, 0]) + abs(cell[2, 1]) > 1.0
    ctx = _ctx(cfg0, p)
    topo = cfg0.topo
    orc = lambda ps, c, x: U.O.energies_refs(U.oracle_system(topo), U.oracle_params(p), c.astype(np.float64),
                                             ps.astype(np.float64), topo.slt_charge64, topo.slt_ljtype,
                                             x.astype(np.float64), rigid=(p["slttype"] == 2))
    uo, u0o = orc(pos, cell, ins)
    uo_plain, u0o_plain = orc(cfg0.frame(0), cfg0.cell, cfg0.insertions(0)[:32])
    for mode in (capi.ER_BATCHED, capi.ER_PLAIN):
        ctx.set_frame(pos, cell)
        ctx.set_insertions(ins)
        ctx.run(mode)
        ctx.sync()
        u, u0 = ctx.uvengy(32)
        nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(cfg0))
        assert nbad == 0, (mode, worst, at, u[at], uo[at])
        assert U.self_mismatch(u0, u0o, U.nsites(cfg0))[0] == 0
        # against the unrotated system: the rotated FP32 inputs differ from it by their own
        # rounding (half an ulp of 31 A, 1e-6 A), which the LJ wall multiplies to ~1e-5
        assert U.energy_mismatch(u, uo_plain, U.nsites(cfg0), rtol=1e-4)[0] == 0
        assert U.self_mismatch(u0, u0o_plain, U.nsites(cfg0))[0] == 0


def test_lattice_images_cfg0(cfg0):
    """Whole molecules and insertions moved by several cell vectors (up to
    three box lengths out, as an unwrapped trajectory holds them): the
    positions are not wrapped into the cell, every minimum image comes from
    the lattice coordinates, so the energies are the oracle's on the same
    inputs (batched and plain) and those of the unmoved system up to the FP32
    rounding of the moved inputs."""
    rng = np.random.default_rng(5)
    p = U.workload_params(cfg0)
    topo = cfg0.topo
    L = np.diag(cfg0.cell).astype(np.float64)
    pos0 = cfg0.frame(0).astype(np.float64)
    ins0 = cfg0.insertions(0)[:32].astype(np.float64)
    mb = topo.mol_begin
    shift_mol = rng.integers(-3, 4, size=(len(mb) - 1, 3)).astype(np.float64) * L
    pos = pos0 + np.repeat(shift_mol, np.diff(mb), axis=0)[: pos0.shape[0]]
    ins = ins0 + (rng.integers(-3, 4, size=(ins0.shape[0], 1, 3)) * L)
    pos32, ins32 = pos.astype(np.float32), ins.astype(np.float32)
    orc = lambda ps, x: U.O.energies_refs(U.oracle_system(topo), U.oracle_params(p), cfg0.cell.astype(np.float64),
                                          ps.astype(np.float64), topo.slt_charge64, topo.slt_ljtype,
                                          x.astype(np.float64), rigid=(p["slttype"] == 2))
    uo, u0o = orc(pos32, ins32)
    uo_in, _ = orc(cfg0.frame(0), cfg0.insertions(0)[:32])
    ns = U.nsites(cfg0)
    for mode in (capi.ER_BATCHED, capi.ER_PLAIN):
        ctx = _ctx(cfg0, p)
        ctx.set_frame(pos32, cfg0.cell)
        ctx.set_insertions(ins32)
        ctx.run(mode)
        ctx.sync()
        u, u0 = ctx.uvengy(32)
        nbad, worst, at = U.energy_mismatch(u, uo, ns)
        print(f"lattice images mode {mode}: {U.energy_report(u, uo)}, worst excess {worst:.3e}")
        assert nbad == 0, (mode, worst, at, u[at], uo[at])
        assert U.self_mismatch(u0, u0o, ns)[0] == 0
        # against the unmoved system: the moved FP32 inputs carry their own
        # rounding (half an ulp of up to 4 box lengths, 4e-6 A)
        assert U.energy_mismatch(u, uo_in, ns, rtol=2e-4)[0] == 0
        ctx.close()


def test_triclinic_cell_cfg0(cfg0):
    """The configuration-0 system sheared into a skewed cell whose (1,2)
    and (2,3) elements exceed half the diagonal: normalize_cell_vector's
    reduction (realcal.F90:832-847), the non-cuboid minimum image
    (realcal.F90:272-282), the cell list in the reduced cell and PME of the
    skewed lattice, against the oracle on the same inputs (batched and
    plain)."""
    L = float(cfg0.cell[0, 0])
    tri = np.array([[L, 0.7 * L, 0.2 * L], [0.0, L, -0.6 * L], [0.0, 0.0, L]])
    S = tri / L   # maps the cube onto the skewed cell: a periodic configuration stays periodic
    p = U.workload_params(cfg0)
    pos = (cfg0.frame(0).astype(np.float64) @ S.T).astype(np.float32)
    ins = (cfg0.insertions(0)[:32].astype(np.float64) @ S.T).astype(np.float32)
    cell = tri.astype(np.float32)
    topo = cfg0.topo
    uo, u0o = U.O.energies_refs(U.oracle_system(topo), U.oracle_params(p), cell.astype(np.float64),
                                pos.astype(np.float64), topo.slt_charge64, topo.slt_ljtype, ins.astype(np.float64),
                                rigid=(p["slttype"] == 2))
    for mode in (capi.ER_BATCHED, capi.ER_PLAIN):
        ctx = _ctx(cfg0, p)
        ctx.set_frame(pos, cell)
        ctx.set_insertions(ins)
        ctx.run(mode)
        u, u0 = ctx.uvengy(32)
        nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(cfg0))
        print(f"triclinic mode {mode}: {U.energy_report(u, uo)}, worst excess {worst:.3e}")
        assert nbad == 0, (mode, worst, at, u[at], uo[at])
        assert U.self_mismatch(u0, u0o, U.nsites(cfg0))[0] == 0
        ctx.close()


def test_pppm_rejects_odd_order(cfg0):
    with pytest.raises(capi.ErError):
        capi.Context(cfg0.topo, U.workload_params(cfg0, cltype=3, splodr=4) | {"splodr": 5})


def test_eng_min_halt_is_reported(cfg0):
    p = U.workload_params(cfg0)
    ctx = capi.Context(cfg0.topo, p)
    uvmax, uvsoft, uvcrd = cfg0.uvmax_uvsoft_uvcrd()
    ctx.set_mesh(uvmax, uvsoft, uvcrd + 1.0e3)   # every energy below the mesh
    ctx.set_solute(cfg0.topo.slt_ljtype, cfg0.topo.slt_charge32)
    ctx.set_frame(cfg0.frame(0), cfg0.cell)
    ctx.set_insertions(cfg0.insertions(0)[:4])
    ctx.run(capi.ER_BATCHED)
    with pytest.raises(capi.ErError) as ei:
        ctx.sync()
    assert ei.value.status == capi.ER_E_HALT
    code, species, energy = ctx.last_error()
    assert code == "eng_min" and species == 1


def test_halt_reports_the_first_offending_energy_across_runs(cfg0):
    """Several frames run before er_sync: the halt reported is the earliest
    (frame, insertion, molecule) below the mesh, with its own energy."""
    p = U.workload_params(cfg0)
    uvmax, uvsoft, uvcrd = cfg0.uvmax_uvsoft_uvcrd()
    ref = _ctx(cfg0, p)
    u0_frames = []
    for f in (0, 1, 2):
        u, _ = _run_refs(ref, cfg0, f, capi.ER_BATCHED, nins=30)
        u0_frames.append(u)
    # lowest mesh point = the largest of the frames' minimum energies, so that
    # every frame has at least one energy at or below it
    t = max(float(u.min()) for u in u0_frames)
    shifted = uvcrd + (t - uvcrd[0])
    shifted[0] = t
    first = None
    for f, u in enumerate(u0_frames):
        bad = np.argwhere(u.astype(np.float64) <= shifted[0])
        if bad.size:
            n, j = bad[0]                          # row-major: smallest insertion, then molecule
            first = (f, int(n), int(j), float(u[n, j]))
            break
    assert first is not None
    ctx = capi.Context(cfg0.topo, p)
    ctx.set_mesh(uvmax, uvsoft, shifted)
    ctx.set_solute(cfg0.topo.slt_ljtype, cfg0.topo.slt_charge32)
    for f in (0, 1, 2):
        ctx.set_frame(cfg0.frame(f), cfg0.cell)
        ctx.set_insertions(cfg0.insertions(f)[:30])
        ctx.run(capi.ER_BATCHED)
    with pytest.raises(capi.ErError) as ei:
        ctx.sync()
    assert ei.value.status == capi.ER_E_HALT
    code, species, energy = ctx.last_error()
    print("first offending (frame, insertion, molecule, energy):", first, "reported", code, species, energy)
    assert code == "eng_min" and species == 1
    assert np.float32(energy) == np.float32(first[3])


def test_eng_sft_halt_in_soln():
    """Solution run with a soft region ending below the frame's highest
    solute-solvent energy: getiduv halts with eng_sft at the first molecule
    (in molecule order) beyond it (engproc.F90:1150-1156)."""
    w = synth.config1()
    p = U.workload_params(w, corrcal=1)
    uvmax, uvsoft, uvcrd = w.uvmax_uvsoft_uvcrd()
    ref = _ctx(w, p)
    ref.set_frame(w.frame(0), w.cell)
    ref.run(capi.ER_BATCHED)
    ref.sync()
    u, _ = ref.uvengy(1)
    row = u[0].astype(np.float64)
    sp = U.species_of(w.topo, 1, w.topo.tagslt)[0]
    live = np.nonzero(sp > 0)[0]
    top = np.sort(row[live])[-3]                 # three energies lie at or above it
    soft = int(np.searchsorted(uvcrd, top, side="left"))   # mesh points below `top`
    assert 0 < soft < uvmax[0]
    beyond = [j for j in live if np.searchsorted(uvcrd, row[j], side="left") > soft]
    assert beyond
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(uvmax, np.array([soft], np.int32), uvcrd)
    ctx.set_solute_molecules(w.topo.tagslt)
    ctx.set_frame(w.frame(0), w.cell)
    ctx.run(capi.ER_BATCHED)
    with pytest.raises(capi.ErError) as ei:
        ctx.sync()
    assert ei.value.status == capi.ER_E_HALT
    code, species, energy = ctx.last_error()
    print("eng_sft: first molecule beyond the soft region", beyond[0], row[beyond[0]], "reported", code, energy)
    assert code == "eng_sft" and species == 1
    assert np.float32(energy) == np.float32(row[beyond[0]])


def _top_bin_mesh(w):
    """The linear part of the mesh only (no log-spaced core): energies above
    its last point fall in the open top bin."""
    mesh, pecore = w.mesh[0]
    lin = np.asarray(mesh[:len(mesh) - pecore], np.float64)
    return np.array([lin.size], np.int32), np.array([lin.size], np.int32), lin


def test_mbin_warning_stops_the_run_unless_forced(cfg0):
    """An energy below 1e3 kcal/mol in the open top bin raises the 'mbin'
    warning (engproc.F90:1158-1167), on which the reference stops unless
    force_calculation is set (mpiproc.F90:238-252).  Without it er_sync
    reports the first such (insertion, molecule) with its energy; with it
    the run goes on and its distributions equal the oracle's binning of the
    GPU's energies on the same mesh."""
    nins = 40
    uvmax, uvsoft, lin = _top_bin_mesh(cfg0)
    ref = _ctx(cfg0, U.workload_params(cfg0))
    u, u0 = _run_refs(ref, cfg0, 0, capi.ER_BATCHED, nins=nins)
    u64 = u.astype(np.float64)
    hit = np.argwhere((u64 > lin[-1]) & (u64 < 1.0e3))
    assert hit.size, "no energy in the top bin"
    n, j = hit[0]
    for force in (0, 1):
        p = U.workload_params(cfg0) | {"force_calculation": force}
        ctx = capi.Context(cfg0.topo, p)
        ctx.set_mesh(uvmax, uvsoft, lin)
        ctx.set_solute(cfg0.topo.slt_ljtype, cfg0.topo.slt_charge32)
        ctx.set_frame(cfg0.frame(0), cfg0.cell)
        ctx.set_insertions(cfg0.insertions(0)[:nins])
        ctx.run(capi.ER_BATCHED)
        if not force:
            with pytest.raises(capi.ErError) as ei:
                ctx.sync()
            assert ei.value.status == capi.ER_E_HALT
            code, species, energy = ctx.last_error()
            print("mbin: first (insertion, molecule)", (int(n), int(j)), u[n, j], "reported", code, energy)
            assert code == "mbin" and species == 1
            assert np.float32(energy) == u[n, j]
            continue
        ctx.sync()
        g = ctx.fetch()
        import oracle as O
        _, vol = O.inverted_cell(cfg0.cell.astype(np.float64))
        qtot = float(cfg0.topo.charge32.astype(np.float64).sum())
        d = O.Distribution(p["slttype"], uvmax, uvsoft, lin, p["temp"], vol, p["screen"], qtot)
        for k in range(nins):
            d.update(cfg0.topo.uvspec, u64[k], float(u0[k]))
        assert d.d.warn_mbin == 1
        _assert_same_distribution(g, d)

"""The other benchmark configurations through the C ABI against the oracle.

Configuration 3: a two-species solvent (waters + 9-site ethanol-like
molecules, 54 000 atoms, 80 A cube, PME 80^3) with 16 rigid solutes of 1 to
60 sites; each solute is its own run over the same frames.  Energies per
(insertion, molecule) to the erutil bar; the distributions (two species,
800 bins each, 1600^2 ecorr) exactly against the oracle's binning of the
GPU's own energies."""
import numpy as np
import pytest

import erutil as U
from erdst_gpu import capi, synth

pytestmark = pytest.mark.gpu


def _ctx(w, p):
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd())
    ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    return ctx


@pytest.fixture(scope="module")
def cfg3():
    return synth.config3()


@pytest.mark.parametrize("k", range(16))
def test_config3_solute_energies(cfg3, k):
    """Every one of configuration 3's 16 solutes (1 to 45 sites) against the
    oracle: 8 insertions of frame 0."""
    w = cfg3.workload(k)
    p = U.workload_params(w)
    ctx = _ctx(w, p)
    ins = w.insertions(0)[:8]
    ctx.set_frame(w.frame(0), w.cell)
    ctx.set_insertions(ins)
    ctx.run(capi.ER_BATCHED)
    ctx.sync()
    u, u0 = ctx.uvengy(8)
    uo, u0o = U.oracle_refs(w, 0, p, nins=8)
    ns = U.nsites(w)
    nbad, worst, at = U.energy_mismatch(u, uo, ns)
    print(f"ENERGY config3 solute {k} ({w.topo.slt_charge32.size} sites): {U.energy_report(u, uo)}, "
          f"uvengy0 {U.energy_report(u0, u0o)}, worst excess {worst:.3e}")
    assert nbad == 0, (worst, at, u[at], uo[at])
    assert U.self_mismatch(u0, u0o, ns)[0] == 0


@pytest.fixture(scope="module")
def cfg2():
    return synth.config2()


def test_config2_energies_and_distributions(cfg2):
    """Configuration 2 at full size (30 000 waters, 67 A cube, PME 72^3,
    600 bins, ecorr on): frames 0 and 1, 64 insertions each, batched.
    Energies and uvengy0 at the bar against the oracle; edens, ecorr,
    engnorm, avslf, voffset and minuv/maxuv exactly equal to the oracle's
    binning of the GPU's own energies; every bin that differs from the
    binning of the FP64 energies has an edge within the energy tolerance."""
    w = cfg2
    p = U.workload_params(w)
    ctx = _ctx(w, p)
    nins = 64
    d_gpu = U.oracle_distribution(w, p)
    uvmax, _, uvcrd = w.uvmax_uvsoft_uvcrd()
    for f in (0, 1):
        ctx.set_frame(w.frame(f), w.cell)
        ctx.set_insertions(w.insertions(f)[:nins])
        ctx.run(capi.ER_BATCHED)
        ctx.sync()
        u, u0 = ctx.uvengy(nins)
        uo, u0o = U.oracle_refs(w, f, p, nins=nins)
        ns = U.nsites(w)
        nbad, worst, at = U.energy_mismatch(u, uo, ns)
        fl = U.flip_report(u, uo, w.topo.uvspec, uvmax, uvcrd, ns)
        print(f"ENERGY config2 frame {f}: {U.energy_report(u, uo)}, uvengy0 {U.energy_report(u0, u0o)}, "
              f"worst excess {worst:.3e}; flips {fl}")
        assert nbad == 0, (worst, at, u[at], uo[at])
        assert U.self_mismatch(u0, u0o, ns)[0] == 0
        assert fl["unexplained"] == 0
        for n in range(nins):
            d_gpu.update(w.topo.uvspec, u[n].astype(np.float64), float(u0[n]))
    g = ctx.fetch()
    assert g["edens"].size == 600 and g["engsmpl"] == 2 * nins
    np.testing.assert_allclose(g["edens"] / g["engnorm"], d_gpu.edens / d_gpu.d.engnorm, rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["ecorr"] / g["engnorm"], d_gpu.ecorr.reshape(600, 600) / d_gpu.d.engnorm,
                               rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["engnorm"], d_gpu.d.engnorm, rtol=1e-7)
    np.testing.assert_allclose(g["avslf"], d_gpu.d.avslf, rtol=1e-7)
    assert g["voffset"] == d_gpu.d.voffset
    np.testing.assert_array_equal(g["minuv"].astype(np.float64), d_gpu.minuv)
    np.testing.assert_array_equal(g["maxuv"].astype(np.float64), d_gpu.maxuv)
    # invariants: symmetric ecorr, rows sum to N * edens, edens sums to N
    np.testing.assert_array_equal(g["ecorr"], g["ecorr"].T)
    np.testing.assert_allclose(g["ecorr"].sum(1), w.topo.nmol * g["edens"], rtol=1e-12)
    np.testing.assert_allclose(g["edens"].sum() / g["engnorm"], w.topo.nmol, rtol=1e-12)


def test_config3_two_species_distribution(cfg3):
    w = cfg3.workload(0)
    p = U.workload_params(w)
    ctx = _ctx(w, p)
    d = U.oracle_distribution(w, p)
    for f in (0, 1):
        ins = w.insertions(f)[:40]
        ctx.set_frame(w.frame(f), w.cell)
        ctx.set_insertions(ins)
        ctx.run(capi.ER_BATCHED)
        ctx.sync()
        u, u0 = ctx.uvengy(40)
        for n in range(40):
            d.update(w.topo.uvspec, u[n].astype(np.float64), float(u0[n]))
    g = ctx.fetch()
    assert g["edens"].size == 1600
    np.testing.assert_allclose(g["edens"] / g["engnorm"], d.edens / d.d.engnorm, rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["ecorr"] / g["engnorm"], d.ecorr.reshape(1600, 1600) / d.d.engnorm,
                               rtol=1e-12, atol=0)
    # per species: bins of species p sum to its molecule count
    nm = np.bincount(w.topo.uvspec, minlength=3)[1:]
    en = g["edens"] / g["engnorm"]
    np.testing.assert_allclose([en[:800].sum(), en[800:].sum()], nm, rtol=1e-12)


@pytest.fixture(scope="module")
def cfg4():
    return synth.config4()


def test_config4_slab_energies(cfg4):
    """Configuration 4: 200 001 atoms in a 120 x 120 x 100 A box, PME
    128 x 128 x 108 (K2 != K3), a flexible 150-site solute inserted in the
    slab; per-insertion conformations change the solute self energies."""
    w = cfg4
    p = U.workload_params(w)
    ctx = _ctx(w, p)
    ins = w.insertions(0)[:6]
    ctx.set_frame(w.frame(0), w.cell)
    ctx.set_insertions(ins)
    ctx.run(capi.ER_BATCHED)
    ctx.sync()
    u, u0 = ctx.uvengy(6)
    uo, u0o = U.oracle_refs(w, 0, p, nins=6)
    nbad, worst, at = U.energy_mismatch(u, uo, U.nsites(w))
    print(f"ENERGY config 4: {U.energy_report(u, uo)}, worst excess {worst:.3e}, "
          f"uvengy0 {u0.tolist()} vs {u0o.tolist()}")
    assert nbad == 0, (worst, at, u[at], uo[at])
    assert U.self_mismatch(u0, u0o, U.nsites(w))[0] == 0


def test_config4_distribution_plain_equals_batched(cfg4):
    w = cfg4
    p = U.workload_params(w)
    res = U.plain_and_batched(w, p, 1, 20, capi)
    U.assert_plain_matches_batched(res, U.nsites(w))
    nm = np.bincount(w.topo.uvspec, minlength=3)[1:]
    g = res[0][2]
    en = g["edens"] / g["engnorm"]
    np.testing.assert_allclose([en[:600].sum(), en[600:].sum()], nm, rtol=1e-12)


def test_config3_all_solutes_one_context(cfg3):
    """All 16 solutes of configuration 3 in one context (er_set_solutes):
    one er_run per frame processes every solute; each solute's energies and
    distributions equal those of its own single-solute context."""
    nins = 24
    ws = [cfg3.workload(k) for k in range(16)]
    p = U.workload_params(ws[0])
    multi = capi.Context(ws[0].topo, p)
    multi.set_mesh(*cfg3.uvmax_uvsoft_uvcrd())
    multi.set_solutes([(w.topo.slt_ljtype, w.topo.slt_charge32) for w in ws])
    frames = (0, 1)
    singles = []
    for k, w in enumerate(ws):
        ctx = _ctx(w, p)
        singles.append(ctx)
    for f in frames:
        x = cfg3.frame(f)
        multi.set_frame(x, cfg3.cell)
        for k, w in enumerate(ws):
            multi.select_solute(k)
            multi.set_insertions(w.insertions(f)[:nins])
        multi.run(capi.ER_BATCHED)
        multi.sync()
        for k, w in enumerate(ws):
            singles[k].set_frame(x, cfg3.cell)
            singles[k].set_insertions(w.insertions(f)[:nins])
            singles[k].run(capi.ER_BATCHED)
            singles[k].sync()
            multi.select_solute(k)
            um, u0m = multi.uvengy(nins)
            us, u0s = singles[k].uvengy(nins)
            np.testing.assert_array_equal(um, us)
            np.testing.assert_array_equal(u0m, u0s)
    for k in range(16):
        multi.select_solute(k)
        gm, gs = multi.fetch(), singles[k].fetch()
        for key in ("edens", "ecorr"):
            np.testing.assert_array_equal(gm[key], gs[key])
        for key in ("engnorm", "engsmpl", "avslf", "voffset"):
            assert gm[key] == gs[key]

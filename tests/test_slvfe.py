"""slvfe restatement (erdst_gpu/slvfe.py) pinned by closed forms and
invariants of the energy-representation functional.

Independent solvent molecules give exactly solvable inputs: with N
molecules of one species, each drawing its solute-solvent energy from P0
in the reference system, the reference density is N*P0, the per-insertion
count correlation is N(N-1) P0 P0^T + N diag(P0), and the solution
density is N*P0*exp(-e/kT)/Z.  Then
  * the 'none' functional returns the exact -N kT log Z,
  * slncv is the constant kT log Z and, with the 'eczr' shift, inscv is
    the constant kT (1 - 1/Z), so py-hnc adds kT (A - B) sum_i s_i
    (rho_i - rho0_i) with A, B the PY-HNC values of those constants and s_i
    the mixing weights,
  * the 'reg', 'gce' and 'evd' inversions agree (the null space of the
    fluctuation matrix is the all-ones vector, which eczr removes).
"""
import math
import os

import numpy as np
import pytest

from erdst_gpu import engstore, slvfe

KT = slvfe.kt_of(300.0)


def ideal_inputs(n_mol=40, nbin=36, seed=1, lo=-3.0, hi=4.0, species=1):
    rng = np.random.default_rng(seed)
    crds, dst, dns, spec, blocks = [], [], [], [], []
    for p in range(species):
        e = np.linspace(lo, hi, nbin)
        e[np.argmin(np.abs(e))] = 0.0                 # an exact zero-energy bin
        p0 = rng.dirichlet(np.full(nbin, 3.0))
        z = float(np.sum(p0 * np.exp(-e / KT)))
        rho0 = n_mol * p0
        rho = n_mol * p0 * np.exp(-e / KT) / z
        cor = n_mol * (n_mol - 1) * np.outer(p0, p0) + n_mol * np.diag(p0)
        crds.append(e); dst.append(rho); dns.append(rho0); spec.append(np.full(nbin, p + 1))
        blocks.append((cor, z, rho, rho0))
    ermax = nbin * species
    cor = np.zeros((ermax, ermax))
    for p, (c, *_rest) in enumerate(blocks):
        s = slice(p * nbin, (p + 1) * nbin)
        cor[s, s] = c
    # cross-species blocks of independent species: <n_i m_j> = <n_i><m_j>
    for p in range(species):
        for q in range(species):
            if p != q:
                cor[p * nbin:(p + 1) * nbin, q * nbin:(q + 1) * nbin] = np.outer(dns[p], dns[q])
    d = slvfe.Distributions(np.concatenate(crds), np.concatenate(dst), np.concatenate(dns), cor,
                            np.concatenate(spec).astype(np.int64))
    return d, blocks


FV = slvfe.FeVars(uvread="not", slfslt="not")


def test_none_functional_is_exact_for_independent_solvent():
    d, blocks = ideal_inputs()
    r = slvfe.chmpot(d, slvfe.replace(FV, functional="none"), 1)
    cor, z, rho, rho0 = blocks[0]
    exact = -40 * KT * math.log(z)
    assert r.chmpt[1] == pytest.approx(exact, abs=1e-10)
    assert r.chmpt[0] == pytest.approx(exact, abs=1e-10)


def test_pyhnc_closed_form_for_independent_solvent():
    d, blocks = ideal_inputs(seed=3)
    r = slvfe.chmpot(d, FV, 1)
    cor, z, rho, rho0 = blocks[0]
    # slncv = kT log Z; inscv = kT (1 - 1/Z) after the eczr shift
    np.testing.assert_allclose(r.slncv, KT * math.log(z), rtol=0, atol=1e-11)
    np.testing.assert_allclose(r.inscv, KT * (1.0 - 1.0 / z), rtol=0, atol=1e-9)
    a = slvfe.pyhnc(KT * math.log(z), 1, KT)
    b = slvfe.pyhnc(KT * (1.0 - 1.0 / z), 2, KT)
    s = np.array([slvfe.sfewgt(x, y) for x, y in zip(rho, rho0)])
    expect = -40 * KT * math.log(z) + KT * (a - b) * float(np.sum(s * (rho - rho0)))
    assert r.chmpt[1] == pytest.approx(expect, abs=1e-8)
    assert r.solver == {"refs": "reg"}


def test_two_independent_species_add():
    d, blocks = ideal_inputs(species=2, nbin=30, seed=5)
    r = slvfe.chmpot(d, slvfe.replace(FV, functional="none"), 1)
    for p, (_c, z, _r, _r0) in enumerate(blocks):
        assert r.chmpt[p + 1] == pytest.approx(-40 * KT * math.log(z), abs=1e-9)
    assert r.chmpt[0] == pytest.approx(r.chmpt[1] + r.chmpt[2], abs=1e-12)


@pytest.mark.parametrize("seed", [7, 8])
def test_inversion_methods_agree(seed):
    d, _ = ideal_inputs(seed=seed)
    # perturb the solution side so inscv is not constant
    rng = np.random.default_rng(seed)
    d.rddst = d.rddst * np.exp(0.3 * rng.standard_normal(d.rddst.size))
    d.rddst *= d.rddns.sum() / d.rddst.sum()
    vals = {m: slvfe.chmpot(d, slvfe.replace(FV, invmtrx=m), 1) for m in ("reg", "gce", "evd")}
    for m in ("gce", "evd"):
        np.testing.assert_allclose(vals[m].inscv, vals["reg"].inscv, rtol=0, atol=1e-7)
        assert vals[m].chmpt[0] == pytest.approx(vals["reg"].chmpt[0], abs=1e-8)


def test_identical_systems_give_the_average_energy():
    d, _ = ideal_inputs(seed=11)
    d.rddst = d.rddns.copy()
    d.aveuv = np.array([-3.25])
    d.slfeng = 1.5
    r = slvfe.chmpot(d, slvfe.FeVars(), 1)
    assert r.chmpt[1] == pytest.approx(-3.25, abs=1e-12)
    assert r.chmpt[0] == pytest.approx(-3.25 + 1.5, abs=1e-12)


def test_pyhnc_branches_meet_at_zero():
    for cnt in (1, 2):
        for h in (1e-4, -1e-4):
            v = slvfe.pyhnc(h * KT, cnt, KT)
            assert v == pytest.approx(h / 2.0, rel=1e-3)      # value and slope 1/2 from both sides
    for h in (1e-4, -1e-4):
        assert slvfe.pyhnc(h * KT, 3, KT) == pytest.approx(-h / 2.0, rel=1e-3)
    # PY solution branch: f + f/(e^{-f}-1) + 1 at f = -1
    assert slvfe.pyhnc(-KT, 1, KT) == pytest.approx(-1.0 - 1.0 / (math.e - 1.0) + 1.0, rel=1e-14)
    assert slvfe.pyhnc(-KT, 2, KT) == pytest.approx(math.log(2.0) * (-2.0) + 1.0, rel=1e-14)
    with pytest.raises(slvfe.SlvfeError):
        slvfe.pyhnc(0.0, 4, KT)


def test_sfewgt():
    assert slvfe.sfewgt(2.0, 1.0) == 1.0
    assert slvfe.sfewgt(1.0, 1.0) == 1.0
    assert slvfe.sfewgt(0.0, 1.0) == 0.0
    assert slvfe.sfewgt(1.0, 3.0) == pytest.approx(1.0 - 0.25)


def test_linear_extrapolation_reproduces_a_linear_potential():
    """A bin with too few reference samples takes the weighted-least-squares
    line through the well-sampled bins; for an exactly linear indirect
    potential that is the line itself, whatever the weights."""
    nbin = 30
    e = np.linspace(-3.0, 3.0, nbin)
    rng = np.random.default_rng(2)
    p0 = rng.dirichlet(np.full(nbin, 4.0))
    a0, a1 = 0.37, -0.21
    rho0 = 20 * p0
    rho = rho0 * np.exp(-(a0 + a1 * e + e) / KT)
    rho *= rho0.sum() / rho.sum()
    # with extthres_refs = 3, bins below 2.5x the smallest reference value are extrapolated
    weak = [3, 17]
    rho0[weak] = rho0.min() * 0.9
    rho0 *= 20 / rho0.sum()
    rho = rho0 * np.exp(-(a0 + a1 * e + e) / KT)
    rho *= rho0.sum() / rho.sum()
    cor = 20 * 19 * np.outer(rho0 / 20, rho0 / 20) + np.diag(rho0)
    line = -KT * np.log(rho / rho0) - e                # exactly a0' + a1 e
    np.testing.assert_allclose(np.diff(line) / np.diff(e), a1, atol=1e-9)
    # the weak bins' own solution values are spoiled: extrapolation must ignore them
    spoiled = rho.copy()
    spoiled[weak] *= 1.7
    d = slvfe.Distributions(e, spoiled, rho0, cor, np.ones(nbin, np.int64), nummol=np.array([20.0]))
    fv = slvfe.replace(FV, extthres_refs=3, normalize="not")
    r = slvfe.chmpot(d, fv, 1)
    np.testing.assert_allclose(r.slncv, line, atol=1e-9)
    # without the threshold the spoiled bins are taken as they are
    r2 = slvfe.chmpot(d, slvfe.replace(FV, normalize="not"), 1)
    assert abs(r2.slncv[weak[0]] - line[weak[0]]) > 0.1


def test_distnorm_restores_marginals():
    d, _ = ideal_inputs(seed=13)
    rng = np.random.default_rng(13)
    noise = 1.0 + 0.02 * rng.standard_normal(d.rdcor.shape)
    d.rdcor = d.rdcor * (noise + noise.T) / 2.0
    g = slvfe.regroup(d, *slvfe.mesh_layout(d.rdcrd, d.rdspec, 1e-8), 1, 0, False)
    slvfe.distnorm(g, np.array([40.0]), slvfe.FeVars(), False)
    # per species: sum_j ecorr(i, j) = N * edens(i)
    np.testing.assert_allclose(g.ecorr.sum(axis=0), 40.0 * g.edens, rtol=1e-7)
    assert g.edens.sum() == pytest.approx(40.0, rel=1e-14)


def test_regroup_preserves_totals_and_marginals():
    d, _ = ideal_inputs(nbin=37, seed=17)
    lay = slvfe.mesh_layout(d.rdcrd, d.rdspec, 1e-8)
    for grp in (1, 2, 3, 5):
        g = slvfe.regroup(d, *lay, grp, 0, False)
        assert g.gemax == (37 // grp)
        assert g.edist.sum() == pytest.approx(d.rddst.sum(), rel=1e-14)
        np.testing.assert_allclose(g.ecorr.sum(), d.rdcor.sum(), rtol=1e-13)
        assert g.uvcrd[-1] == d.rdcrd[-1]                  # last bin keeps the last coordinate
        assert np.all(np.diff(g.idrduv) >= 0)


def test_mesh_layout_counts_log_core():
    lin = -2.0 + 0.1 * np.arange(20)
    log = np.geomspace(0.2, 100.0, 7) + lin[-1]
    crd = np.concatenate([lin, log])
    mx, core = slvfe.mesh_layout(crd, np.ones(crd.size, np.int64), 1e-8)
    assert mx.tolist() == [27] and core.tolist() == [7]


def test_posv_rejects_singular_and_reports_residual():
    a = np.array([[2.0, 1.0], [1.0, 2.0]])
    x, info = slvfe.posv(a, np.array([3.0, 3.0]))
    assert info == 0 and np.allclose(x, [1.0, 1.0])
    x, info = slvfe.posv(np.array([[1.0, 1.0], [1.0, 1.0]]), np.array([1.0, 0.0]))
    assert info != 0


def test_group_table_defaults():
    assert slvfe.group_table(5, "not") == [(1, 0), (2, 0), (3, 0), (4, 0), (5, 0)]
    assert slvfe.group_table(12, "not")[10:] == [(12, 0), (14, 0)]
    assert slvfe.group_table(11, "yes")[:4] == [(1, 0), (1, 60), (1, 80), (1, 100)]


def test_fortran_formats():
    assert engstore.fortran_g(1.5, 15, 7) == "   1.500000    "
    assert engstore.fortran_g(-20.0, 15, 7) == "  -20.00000    "
    assert engstore.fortran_g(0.0125, 15, 7) == "  0.1250000E-01"
    assert engstore.fortran_g(12345678.0, 15, 7) == "  0.1234568E+08"
    assert engstore.fortran_e(1.0, 20, 8) == "      0.10000000E+01"


def test_merge_reads_what_engstore_writes(tmp_path):
    """Four solution blocks, one reference block: run_merge over the written
    files equals chmpot over the same (file-rounded) distributions."""
    rng = np.random.default_rng(21)
    nbin = 40
    mesh = np.concatenate([np.linspace(-4.0, 2.0, 31), 2.0 * (50.0 / 2.0) ** (np.arange(1, 10) / 9.0)])
    uvmax, uvsoft = np.array([nbin]), np.array([30])
    (tmp_path / "soln").mkdir()
    (tmp_path / "refs").mkdir()
    n = 25
    p0 = rng.dirichlet(np.full(nbin, 2.0))
    e = engstore.bin_info(uvmax, uvsoft, mesh)[1]
    fetch_ref = dict(edens=n * p0 * 7.0, ecorr=(n * (n - 1) * np.outer(p0, p0) + n * np.diag(p0)) * 7.0,
                     engnorm=7.0, engsmpl=9.0, avslf=3.5, voffset=1.25)
    wr = engstore.BlockWriter(str(tmp_path / "refs"), uvmax, uvsoft, mesh, KT, soln=False, engdiv=1)
    wr.store(fetch_ref)
    ws = engstore.BlockWriter(str(tmp_path / "soln"), uvmax, uvsoft, mesh, KT, soln=True, engdiv=4,
                              corrcal=False)
    rho_blocks, uv_blocks = [], []
    for b in range(4):
        q = p0 * np.exp(-(e * (0.6 + 0.1 * b)) / KT)
        q /= q.sum()
        en = 2.0 + b
        uv = float(np.sum(e * n * q))
        ws.store(dict(edens=n * q * en, engnorm=en, engsmpl=3.0, avslf=en, voffset=0.0,
                      slnuv=np.array([uv * en])))
        rho_blocks.append(n * q)
        uv_blocks.append(uv)
    res = slvfe.run_merge(slvfe.FeVars(), str(tmp_path))
    assert res.numrun == 4 and res.chmpt.shape == (2, 5, 4)
    slfeng = 1.25 + KT * math.log(0.5)
    assert res.slfeng == pytest.approx(slfeng, rel=1e-7)
    # direct evaluation on the values as the files hold them
    _, mid, spec, ref_ratio = slvfe.read_engfile(str(tmp_path / "refs" / "engref.tt"))
    cor = slvfe.read_corfile(str(tmp_path / "refs" / "corref.tt"), nbin)
    for b in range(4):
        _, _, _, ratio = slvfe.read_engfile(str(tmp_path / "soln" / f"engsln.{b + 1:02d}"))
        aveuv = float(open(tmp_path / "soln" / "aveuv.tt").read().splitlines()[b].split()[1])
        d = slvfe.Distributions(mid, ratio, ref_ratio, cor, spec, np.array([aveuv]), res.slfeng)
        r = slvfe.chmpot(d, slvfe.FeVars(), 3)
        assert res.chmpt[0, 2, b] == pytest.approx(r.chmpt[0], abs=1e-12)
        np.testing.assert_allclose(ratio, rho_blocks[b], rtol=1e-13)
        assert aveuv == pytest.approx(uv_blocks[b], abs=1e-5)
    text = slvfe.format_merge(res, slvfe.FeVars())
    assert "cumulative average & 95% error for solvation free energy" in text
    assert res.error95() > 0.0
    assert os.path.exists(tmp_path / "soln" / "weight_soln")


def test_parameters_fe_namelist(tmp_path):
    p = tmp_path / "parameters_fe"
    p.write_text("&fevars\n  clcond = 'range', numprm = 7,\n  invmtrx='evd' ! comment\n  inptemp=310.0d0\n/\n")
    fv = slvfe.read_parameters_fe(str(p))
    assert fv.clcond == "range" and fv.numprm == 7 and fv.invmtrx == "evd" and fv.inptemp == 310.0

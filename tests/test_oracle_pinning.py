"""Pins the FP64 oracle (oracle/er_oracle.c) to independent references.

The reference repository holds no tests, golden vectors or fixtures for this
path, and its Fortran+OpenACC+cuFFT build cannot be made here, so every
oracle function is checked against what the mathematics fixes: closed forms,
a brute-force Ewald sum (which PME converges to), numpy's FFT, brute-force
counting, and the invariants of the distributions.  Each check uses a
formula or routine different from the oracle's own.
"""
import numpy as np
import pytest
from scipy.special import erfc

import oracle as O

CH = 18.22261721


# ---------------------------------------------------------------- B-spline
def cox_de_boor(n, x):
    """Cardinal B-spline by recursion (not the explicit sum the oracle uses)."""
    if n == 1:
        return 1.0 if 0.0 <= x < 1.0 else 0.0
    return (x * cox_de_boor(n - 1, x) + (n - x) * cox_de_boor(n - 1, x - 1.0)) / (n - 1)


@pytest.mark.parametrize("t", [0.0, 0.13, 0.5, 0.77, 0.999])
def test_cubic_bspline_closed_form(t):
    ref = [t ** 3 / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (1 - t) ** 3 / 6]
    got = [O.spline_value(4, t + k) for k in range(4)]
    if t == 0.0:
        ref[0] = 0.0  # spline.F90:51 returns 0 at the open end x = 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize("order", [3, 4, 5, 6, 8, 10])
def test_bspline_recursion_partition_symmetry(order):
    rng = np.random.default_rng(order)
    for x in rng.uniform(0.01, order - 0.01, 20):
        assert abs(O.spline_value(order, x) - cox_de_boor(order, x)) < 1e-15 * 10 ** (order / 2)
        assert abs(O.spline_value(order, x) - O.spline_value(order, order - x)) < 1e-12
    for t in rng.uniform(0.001, 0.999, 10):
        # the explicit alternating sum of spline.F90 cancels ~10^(order/2) ulps
        assert abs(sum(O.spline_value(order, t + k) for k in range(order)) - 1.0) < 1e-15 * 10 ** (order / 2)


def test_spline_axis_order4_closed_form():
    # |sum_k M4(k+1) e^{2 pi i k m/K}|^2 = (2/3 + cos(2 pi m/K)/3)^2
    K = 36
    th = 2 * np.pi * np.arange(K) / K
    np.testing.assert_allclose(O.init_spline_axis(4, K), (2 / 3 + np.cos(th) / 3) ** 2, atol=1e-14)


def test_spline_atom_weights_and_base_index():
    p = O.make_params(splodr=6, ms=(40, 32, 36))
    cell = np.diag([30.0, 27.0, 33.0])
    x = np.array([-4.2, 13.7, 40.1])          # raw, outside the cell on two axes
    w, g = O.spline_atom(p, cell, x)
    u = (x / np.diag(cell)) % 1.0
    f = u * np.array(p.ms)
    np.testing.assert_array_equal(g, np.floor(f).astype(int))
    for k in range(3):
        t = f[k] - np.floor(f[k])
        np.testing.assert_allclose(w[k], [cox_de_boor(6, t + c) for c in range(6)], atol=1e-12)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(6, 10, 12), (5, 9, 8), (7, 4, 10), (8, 8, 8)])
def test_fft_against_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    g = rng.standard_normal(shape)                       # [x3, x2, x1]
    K1 = shape[2]
    spec = O.fft_r2c(g)
    np.testing.assert_allclose(spec, np.fft.fftn(g)[:, :, :K1 // 2 + 1], rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(O.fft_c2r(spec, K1), g * g.size, rtol=1e-12, atol=1e-9)
    for n in (7, 12, 45, 72, 128):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        np.testing.assert_allclose(O.fft1d(x, -1), np.fft.fft(x), atol=1e-11)
        np.testing.assert_allclose(O.fft1d(x, +1), np.fft.ifft(x) * n, atol=1e-11)


# ---------------------------------------------------------------- Ewald
def ewald_pair(qa, xa, qb, xb, L, alpha, nmax=18, nimg=2):
    """Ewald interaction energy of charge sets a and b in a cubic box
    (real part over images, reciprocal sum over |m| <= nmax, and the
    neutralising background -pi Q_a Q_b / (alpha^2 V))."""
    V = L ** 3
    E = 0.0
    shifts = np.array([[i, j, k] for i in range(-nimg, nimg + 1) for j in range(-nimg, nimg + 1)
                       for k in range(-nimg, nimg + 1)]) * L
    for i in range(len(qa)):
        for j in range(len(qb)):
            r = np.linalg.norm(xa[i] - xb[j] + shifts, axis=1)
            E += qa[i] * qb[j] * np.sum(erfc(alpha * r) / r)
    m = np.arange(-nmax, nmax + 1)
    M = np.stack(np.meshgrid(m, m, m, indexing="ij"), -1).reshape(-1, 3)
    M = M[np.any(M != 0, axis=1)]
    k = 2 * np.pi * M / L
    k2 = (k ** 2).sum(1)
    Sa = (qa[None, :] * np.exp(1j * k @ xa.T)).sum(1)
    Sb = (qb[None, :] * np.exp(1j * k @ xb.T)).sum(1)
    E += (4 * np.pi / V * np.exp(-k2 / (4 * alpha ** 2)) / k2 * np.real(Sa * np.conj(Sb))).sum()
    return E - np.pi * qa.sum() * qb.sum() / (alpha ** 2 * V)


def ewald_self(q, x, L, alpha, nmax=18, nimg=2):
    """Ewald energy of one charge set with its own images, minus its direct
    minimum-image Coulomb pairs: what uvengy0 restates (engproc.F90:828-855)."""
    V = L ** 3
    E = 0.0
    shifts = np.array([[i, j, k] for i in range(-nimg, nimg + 1) for j in range(-nimg, nimg + 1)
                       for k in range(-nimg, nimg + 1)]) * L
    for i in range(len(q)):
        for j in range(len(q)):
            r = np.linalg.norm(x[i] - x[j] + shifts, axis=1)
            if i == j:
                r = r[r > 0]
            E += 0.5 * q[i] * q[j] * np.sum(erfc(alpha * r) / r)
    m = np.arange(-nmax, nmax + 1)
    M = np.stack(np.meshgrid(m, m, m, indexing="ij"), -1).reshape(-1, 3)
    M = M[np.any(M != 0, axis=1)]
    k = 2 * np.pi * M / L
    k2 = (k ** 2).sum(1)
    S = (q[None, :] * np.exp(1j * k @ x.T)).sum(1)
    E += (2 * np.pi / V * np.exp(-k2 / (4 * alpha ** 2)) / k2 * np.abs(S) ** 2).sum()
    E -= alpha / np.sqrt(np.pi) * (q ** 2).sum()
    E -= np.pi * q.sum() ** 2 / (2 * alpha ** 2 * V)
    for i in range(len(q)):
        for j in range(i + 1, len(q)):
            d = x[i] - x[j]
            d -= L * np.round(d / L)
            E -= q[i] * q[j] / np.linalg.norm(d)
    return E


@pytest.fixture(scope="module")
def ewald_system():
    rng = np.random.default_rng(7)
    L = 25.0
    qa = np.array([0.5, -0.3, 0.2]) * CH
    xa = rng.uniform(0, L, (3, 3))
    xa[1:] = xa[0] + rng.normal(0, 0.8, (2, 3))        # a compact molecule
    qb = np.array([-0.4, 0.1]) * CH
    xb = rng.uniform(0, L, (2, 3))
    xb[1] = xb[0] + np.array([0.9, 0.3, -0.2])
    return L, qa, xa, qb, xb


@pytest.mark.parametrize("order,K,rtol", [(4, 32, 2e-4), (6, 48, 2e-6), (8, 64, 1e-8)])
def test_pme_pair_energy_converges_to_ewald(ewald_system, order, K, rtol):
    L, qa, xa, qb, xb = ewald_system
    alpha = 0.35
    ref = ewald_pair(qa, xa, qb, xb, L, alpha)
    p = O.make_params(screen=alpha, splodr=order, ms=(K, K, K))
    # solvent molecule b (LJ off: one type with epsilon 0)
    sysm = O.OracleSystem([0, 2], qb, [0, 0], np.array([[0.0]]), np.array([[1.0]]))
    uv, uv0 = O.energies_refs(sysm, p, np.diag([L] * 3), xb, qa, np.zeros(3, np.int32), xa[None], nthreads=2)
    assert abs(uv[0, 0] - ref) <= rtol * abs(ref)
    refself = ewald_self(qa, xa, L, alpha)
    assert abs(uv0[0] - refself) <= max(rtol * 30, 1e-9) * abs(refself) + 1e-9


def test_pme_noncubic_grid_converges_to_ewald(ewald_system):
    L, qa, xa, qb, xb = ewald_system
    ref = ewald_pair(qa, xa, qb, xb, L, 0.35)
    p = O.make_params(screen=0.35, splodr=8, ms=(72, 64, 80))
    sysm = O.OracleSystem([0, 2], qb, [0, 0], np.array([[0.0]]), np.array([[1.0]]))
    uv, _ = O.energies_refs(sysm, p, np.diag([L] * 3), xb, qa, np.zeros(3, np.int32), xa[None], nthreads=2)
    assert abs(uv[0, 0] - ref) <= 1e-8 * abs(ref)


def test_engfac_zero_and_symmetry():
    p = O.make_params(splodr=4, ms=(32, 32, 32))
    ef = O.engfac(p, np.diag([31.0, 31.0, 31.0]))
    assert ef[0, 0, 0] == 0.0
    # engfac(m) = engfac(-m): m2 -> K2 - m2, m3 -> K3 - m3 along the kept half
    np.testing.assert_allclose(ef[1:, 1:, :], ef[:0:-1, :0:-1, :], rtol=1e-13)


def test_odd_order_skips_nyquist():
    p = O.make_params(splodr=5, ms=(32, 32, 32))
    ef = O.engfac(p, np.diag([31.0, 31.0, 31.0]))
    assert np.all(ef[:, :, 16] == 0.0) and np.all(ef[16, :, :] == 0.0)


# ---------------------------------------------------------------- pairs
def test_lj_minimum_and_switch_endpoints():
    p = O.make_params()                      # CHARMM potential switch 10..12
    _, box = O.normalize_periodic(np.diag([40.0] * 3), np.zeros((1, 3)))
    sig, eps = 3.0, 0.2
    rmin = 2 ** (1 / 6) * sig
    e = O.pair_energy(p, box, [0, 0, 0], [rmin, 0, 0], 0.0, 0.0, eps, sig * sig)
    assert abs(e + eps) < 1e-12
    lj = lambda r: 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
    assert abs(O.pair_energy(p, box, [0, 0, 0], [10.0, 0, 0], 0, 0, eps, sig * sig) - lj(10.0)) < 1e-15
    assert abs(O.pair_energy(p, box, [0, 0, 0], [11.9999999, 0, 0], 0, 0, eps, sig * sig)) < 1e-12
    assert O.pair_energy(p, box, [0, 0, 0], [12.5, 0, 0], 0, 0, eps, sig * sig) == 0.0


@pytest.mark.parametrize("ljswitch", [1, 2, 3])
def test_switched_lj_is_continuous_and_vanishes(ljswitch):
    p = O.make_params(ljswitch=ljswitch)
    _, box = O.normalize_periodic(np.diag([40.0] * 3), np.zeros((1, 3)))
    f = lambda r: O.pair_energy(p, box, [0, 0, 0], [r, 0, 0], 0, 0, 0.2, 9.0)
    for r in (10.0, 12.0):
        assert abs(f(r - 1e-7) - f(r + 1e-7 if r < 12 else r)) < 1e-8
    assert abs(f(12.0)) < 1e-12
    assert f(12.0001) == 0.0


def test_coulomb_erfc_and_min_image():
    p = O.make_params()
    L = 30.0
    _, box = O.normalize_periodic(np.diag([L] * 3), np.zeros((1, 3)))
    qa, qb = 0.3 * CH, -0.7 * CH
    for r in (0.7, 3.0, 11.9):
        # through the periodic boundary: raw separation L - r
        e = O.pair_energy(p, box, [0.2, 0, 0], [0.2 + L - r, 0, 0], qa, qb, 0.0, 1.0)
        assert abs(e - qa * qb * erfc(0.35 * r) / r) < 1e-12 * abs(qa * qb)
    assert O.pair_energy(p, box, [0, 0, 0], [12.0, 0, 0], qa, qb, 0.0, 1.0) == 0.0


def test_realcal_refs_matches_direct_sum():
    """Brute-force numpy restatement of the per-molecule real-space sum."""
    rng = np.random.default_rng(3)
    L, nmol = 26.0, 40
    pos = rng.uniform(0, L, (nmol, 3))[:, None, :] + np.array([[0, 0, 0], [0.6, 0.7, 0], [-0.6, 0.7, 0]])
    pos = pos.reshape(-1, 3)
    q = np.tile([-0.8, 0.4, 0.4], nmol) * CH
    ljt = np.tile([0, 1, 1], nmol).astype(np.int32)
    sig = np.array([3.1, 1.0, 2.6])
    eps = np.array([0.15, 0.0, 0.05])
    ene = np.sqrt(eps[:, None] * eps[None, :])
    lsq = ((sig[:, None] + sig[None, :]) / 2) ** 2
    sysm = O.OracleSystem(np.arange(0, 3 * nmol + 1, 3), q, ljt, ene, lsq)
    ins = rng.uniform(0, L, (3, 1, 3)) + np.array([[[0, 0, 0], [1.0, 0, 0]]])
    sq = np.array([0.3, -0.3]) * CH
    slj = np.array([2, 2], np.int32)
    p = O.make_params(ms=(32, 32, 32))
    # real-space part only: the difference of two runs that share PME
    uv_full, _ = O.energies_refs(sysm, p, np.diag([L] * 3), pos, sq, slj, ins)
    sys0 = O.OracleSystem(np.arange(0, 3 * nmol + 1, 3), q, ljt, ene * 0, lsq)
    p_far = O.make_params(ms=(32, 32, 32), elecut=12.0)
    uv_nolj, _ = O.energies_refs(sys0, p_far, np.diag([L] * 3), pos, sq, slj, ins)
    for n in range(3):
        for j in range(nmol):
            elj = 0.0
            for s in range(2):
                for a in range(3):
                    d = ins[n, s] - pos[3 * j + a]
                    d -= L * np.round(d / L)
                    r = np.linalg.norm(d)
                    if r <= 12.0 and eps[ljt[3 * j + a]] > 0:
                        e = 4 * ene[2, ljt[3 * j + a]] * ((lsq[2, ljt[3 * j + a]] / r ** 2) ** 6
                                                         - (lsq[2, ljt[3 * j + a]] / r ** 2) ** 3)
                        if r > 10.0:
                            e *= (2 * r * r + 144 - 300) * (r * r - 144) ** 2 / (144 - 100) ** 3
                        elj += e
            assert abs((uv_full[n, j] - uv_nolj[n, j]) - elj) < 1e-9 * max(1.0, abs(elj))


def test_normalize_periodic_wraps_into_cell():
    rng = np.random.default_rng(5)
    L = np.array([30.0, 27.0, 33.0])
    x = rng.uniform(-100, 100, (500, 3))
    xn, box = O.normalize_periodic(np.diag(L), x)
    assert np.all(xn >= 0) and np.all(xn < L)
    np.testing.assert_allclose(((x - xn) / L) - np.round((x - xn) / L), 0, atol=1e-12)
    assert box.is_cuboid == 1


# ---------------------------------------------------------------- histograms
def test_binsearch_counts_points_strictly_below():
    mesh = np.concatenate([np.linspace(-20, 4.95, 500), 5 * 2000 ** (np.arange(100) / 99)])
    rng = np.random.default_rng(11)
    vals = np.concatenate([rng.uniform(-25, 1.2e4, 4000), mesh, mesh[:50] + 1e-12, [-1e9, 1e30]])
    for v in vals:
        assert O.binsearch(mesh, v) == int(np.sum(mesh < v))


def test_build_mesh_from_hist_example():
    """&hist of parameters_er.example: region widths / bin widths give the
    counts; the mesh rises strictly and starts at ecdmin - finfac*ecfbin."""
    mesh, pesoft = O.build_mesh(ecdmin=-80.0, ecfmns=-0.20, ecdcen=0.0, eccore=20.0, ecdmax=1.0e11,
                                eclbin=5.0e-2, ecfbin=2.0e-3, ec0bin=2.0e-4, finfac=10.0, pecore=200)
    sh = 10.0 * 2.0e-3
    edges = [-80.0 - sh, -0.20 - sh, 0.0 - sh, 0.0 + sh, 0.20 + sh, 20.0 + 0.20 + sh]
    widths = [5e-2, 2e-3, 2e-4, 2e-3, 5e-2]
    counts = [int(round((edges[k + 1] - edges[k]) / widths[k])) for k in range(5)]
    assert pesoft == sum(counts)
    assert mesh.size == pesoft + 200
    assert np.all(np.diff(mesh) > 0)
    assert abs(mesh[0] - edges[0]) < 1e-12
    start = np.cumsum([0] + counts)
    for k in range(5):
        np.testing.assert_allclose(mesh[start[k]], edges[k], atol=1e-9)
        np.testing.assert_allclose(np.diff(mesh[start[k]:start[k + 1]]), widths[k], rtol=1e-6)
    logpart = mesh[pesoft:]
    np.testing.assert_allclose(logpart[0], edges[5], rtol=1e-12)
    # pecore points from eccore' to ecdmax inclusive: pecore - 1 equal log steps
    np.testing.assert_allclose(np.diff(np.log(logpart)), np.log(1.0e11 / edges[5]) / 199, rtol=1e-9)
    assert logpart[-1] == 1.0e11
    with pytest.raises(ValueError):
        O.build_mesh(-80.0, -79.99, 0.0, 20.0, 1e11, 5e-2, 2e-3, 2e-4, 10.0, 200)   # eng_ecd


def _dist(rng, nmol=50, wgtslf=1):
    mesh = np.linspace(-5, 5, 41)
    d = O.Distribution(3, [41], [41], mesh, temp=0.6, volume=1000.0, screen=0.35, total_charge=0.0,
                       wgtslf=wgtslf, estype=2)
    return d, mesh


def test_update_histogram_invariants():
    rng = np.random.default_rng(2)
    d, mesh = _dist(rng)
    nmol = 50
    for n in range(30):
        d.update(np.ones(nmol, np.int32), rng.uniform(-4.9, 6, nmol), rng.normal(0, 0.3))
        assert d.insdst.sum() == nmol
    ec = d.ecorr.reshape(d.ermax, d.ermax)
    # (w*k)*q and (w*q)*k round differently, as in engproc.F90:957-958
    np.testing.assert_allclose(ec, ec.T, rtol=1e-15)
    np.testing.assert_allclose(d.edens.sum() / d.d.engnorm, nmol, rtol=1e-13)
    np.testing.assert_allclose(ec.sum(1), nmol * d.edens, rtol=1e-13)
    assert d.d.engsmpl == 30


def test_normalised_distribution_is_voffset_invariant():
    rng = np.random.default_rng(9)
    rows = [rng.uniform(-4.9, 6, 40) for _ in range(20)]
    u0 = rng.normal(0, 0.5, 20)
    out = []
    for order in (range(20), list(range(10, 20)) + list(range(10))):   # different first insertion
        d, _ = _dist(rng)
        for n in order:
            d.update(np.ones(40, np.int32), rows[n], u0[n])
        out.append((d.edens / d.d.engnorm, d.ecorr / d.d.engnorm))
    np.testing.assert_allclose(out[0][0], out[1][0], rtol=1e-12)
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-12)


def _weighted_log_mean_exp(x, w):
    """ln( sum w e^x / sum w ), shifted by max(x) (a different route from the
    oracle's voffset shift)."""
    m = float(np.max(x))
    return m + float(np.log(np.sum(w * np.exp(x - m)) / np.sum(w)))


@pytest.mark.parametrize("slttype,estype,qtot", [(2, 1, 0.0), (3, 1, 0.0), (3, 2, 7.3), (1, 1, 0.0), (1, 2, -4.1)])
def test_insertion_weight_gives_the_self_free_energy(slttype, estype, qtot):
    """The weight update_histogram gives each sample (engproc.F90:883-903,
    with volcorrect :1022-1038) and engstore's normalisation (:609-628) must
    yield the solute's self free energy that statistical mechanics fixes:
    refs runs  avediv2 = voffset + kT ln(avslf/engnorm) = -kT ln<exp(-u0/kT)>_w,
    soln runs  avediv2 = voffset - kT ln(avslf/engnorm) = +kT ln<exp(+u0/kT)>_w,
    where u0 is the solute self energy and w = w_sys w_slt the statistical
    weights, whatever sample sets voffset.  A flipped exponent sign, a
    missing voffset shift or weights applied on one side only fail it; the
    volume and net-charge factor of volcorrect (NPT) multiplies every sample
    alike and must drop out.  The written weight file (the product's writer)
    carries the same value in its third column."""
    import tempfile
    from erdst_gpu.engstore import BlockWriter
    rng = np.random.default_rng(11 + slttype + 7 * estype)
    kT = 0.5961612775922495
    mesh = np.linspace(-5, 5, 41)
    d = O.Distribution(slttype, [41], [41], mesh, temp=kT, volume=2.7e4, screen=0.35, total_charge=qtot,
                       wgtslf=1, estype=estype)
    nmol, nsample = 6, 400
    u0 = rng.uniform(-2.5, 2.5, nsample) + rng.choice([0.0, 1.5], nsample)   # spread over several kT
    w_sys = rng.uniform(0.5, 2.0, nsample)
    w_slt = rng.uniform(0.2, 3.0, nsample)
    for n in range(nsample):
        d.update(np.ones(nmol, np.int32), rng.uniform(-4.9, 4.9, nmol), u0[n], w_sys[n], w_slt[n])
    sc = d.scalars
    w = w_sys * w_slt
    sign = +1.0 if slttype == 1 else -1.0
    exact = sign * kT * _weighted_log_mean_exp(sign * u0 / kT, w)
    got = sc["voffset"] - sign * kT * np.log(sc["avslf"] / sc["engnorm"])
    assert abs(got - exact) <= 1e-12 * max(1.0, abs(exact)), (got, exact)
    assert sc["voffset"] == u0[0]
    # the same through the product's writer (weight_refs / weight_soln, e20.7)
    with tempfile.TemporaryDirectory() as td:
        wr = BlockWriter(td, np.array([41]), np.array([41]), mesh, kT, soln=slttype == 1, engdiv=1)
        wr.store(dict(edens=d.edens, ecorr=d.ecorr.reshape(41, 41), eself=np.zeros(0), slnuv=d.slnuv,
                      minuv=None, maxuv=None, **sc))
        name = "weight_soln" if slttype == 1 else "weight_refs"
        col3 = float(open(f"{td}/{name}").read().split()[2])
    assert abs(col3 - exact) <= 1e-6 * max(1.0, abs(exact))


def madelung_cubic_point_charge(Q, L):
    """Ewald energy of one point charge with its periodic images and a
    uniform neutralising background in a cubic box: Q^2 xi / (2 L) with the
    simple-cubic Ewald Madelung constant xi = -2.837297479 (Nijboer and De
    Wette 1957)."""
    return Q * Q * -2.837297479 / (2.0 * L)


def test_volcorrect_factor_is_the_background_energy_of_the_net_charge():
    """volcorrect (engproc.F90:1022-1038) multiplies each sample's weight by
    V exp(pi Q^2 / (2 alpha^2 V) / kT) = V exp(-E_bg / kT), E_bg being the
    energy of the net charge Q with the uniform neutralising background.
    E_bg is pinned here: with it, the Ewald energy of a lone charge is the
    alpha-independent Madelung energy; without it, not.  The oracle's weight
    of a single sample then equals V exp(-E_bg/kT), and at constant volume
    the factor cancels from every normalised output."""
    L, Q, kT = 25.0, 3.0 * CH, 0.6
    V = L ** 3
    x = np.zeros((1, 3))
    for alpha in (0.3, 0.35, 0.45):
        full = ewald_self(np.array([Q]), x, L, alpha)
        assert abs(full - madelung_cubic_point_charge(Q, L)) <= 1e-9 * abs(full)
        e_bg = -np.pi * Q * Q / (2 * alpha ** 2 * V)
        assert abs((full - e_bg) - madelung_cubic_point_charge(Q, L)) > 1e-3 * abs(full)   # needed
        mesh = np.linspace(-5, 5, 41)
        d = O.Distribution(3, [41], [41], mesh, temp=kT, volume=V, screen=alpha, total_charge=Q, wgtslf=1,
                           estype=2)
        d.update(np.ones(4, np.int32), np.array([0.1, 0.2, -1.0, 3.0]), 0.0)
        np.testing.assert_allclose(d.d.engnorm, V * np.exp(-e_bg / kT), rtol=1e-13)
    # constant volume: a net charge rescales engnorm and avslf alike
    rng = np.random.default_rng(5)
    rows = [rng.uniform(-4.9, 4.9, 8) for _ in range(30)]
    u0 = rng.normal(0, 0.4, 30)
    out = []
    for q in (0.0, 2.0 * CH):
        d = O.Distribution(3, [41], [41], np.linspace(-5, 5, 41), temp=kT, volume=V, screen=0.35,
                           total_charge=q, wgtslf=1, estype=2)
        for n in range(30):
            d.update(np.ones(8, np.int32), rows[n], u0[n])
        out.append((d.edens / d.d.engnorm, d.ecorr / d.d.engnorm, d.d.avslf / d.d.engnorm))
    for a, b in zip(*out):
        np.testing.assert_allclose(a, b, rtol=1e-12)


def test_energy_below_mesh_halts_eng_min():
    rng = np.random.default_rng(1)
    d, mesh = _dist(rng)
    with pytest.raises(RuntimeError, match="eng_min"):
        d.update(np.ones(3, np.int32), np.array([0.0, -5.0, 1.0]), 0.0)   # v == mesh[0]


def test_self_energy_histogram():
    """selfcal: every update adds its weight to exactly one self-energy bin,
    the one binsearch gives on escrd (erdst/engproc.F90:895-899, 1130-1140);
    values above the top edge land in the top bin, values at or below the
    bottom edge halt with eng_min; the solvent histograms do not change."""
    rng = np.random.default_rng(4)
    escrd = np.linspace(-1.0, 1.0, 21)
    mesh = np.linspace(-5, 5, 41)
    mk = lambda selfcal: O.Distribution(3, [41], [41], mesh, temp=0.6, volume=1000.0, screen=0.35,
                                        total_charge=0.0, wgtslf=0, estype=2, selfcal=selfcal,
                                        escrd=escrd if selfcal else None)
    d, d0 = mk(1), mk(0)
    u0s = np.concatenate([rng.uniform(-0.99, 0.99, 40), [1.0, 1.7, 25.0]])
    for u0 in u0s:
        row = rng.uniform(-4.9, 6, 20)
        before, n0 = d.eself.copy(), d.d.engnorm
        d.update(np.ones(20, np.int32), row, u0)
        d0.update(np.ones(20, np.int32), row, u0)
        changed = np.flatnonzero(d.eself != before)
        assert changed.size == 1
        assert changed[0] == min(O.binsearch(escrd, u0), escrd.size) - 1
        np.testing.assert_allclose(d.eself[changed[0]] - before[changed[0]], d.d.engnorm - n0, rtol=1e-13)
    np.testing.assert_allclose(d.eself.sum(), d.d.engnorm, rtol=1e-13)
    assert d.eself[-1] > 0                      # 1.7 and 25 above the top edge
    np.testing.assert_array_equal(d.edens, d0.edens)
    assert d.minuv[0] == u0s.min() and d.maxuv[0] == 25.0
    with pytest.raises(RuntimeError, match="eng_min"):
        d.update(np.ones(3, np.int32), np.zeros(3), -1.0)   # at the bottom edge


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_pppm_denominator_is_the_alias_sum(order):
    """calc_gfb_pppm / gf_denom (recpcal.F90:179-214): sum_l gf_b(l) x^(l-1)
    at x = sin^2(theta) is the aliased assignment-function sum
    sum_n (sin(theta) / (theta + pi n))^(2 order), checked by brute force."""
    gfb = O.pppm_gfb(order)
    for theta in np.linspace(0.05, np.pi / 2, 9):
        x = np.sin(theta) ** 2
        poly = sum(gfb[l] * x ** l for l in range(order))
        n = np.arange(-4000, 4001)
        brute = np.sum((np.sin(theta) / (theta + np.pi * n)) ** (2 * order))
        assert abs(poly - brute) <= 1e-10 * brute
        assert O.pppm_denom(order, gfb, x, x, x) == pytest.approx(poly ** 6, rel=1e-13)


@pytest.mark.parametrize("order,K,rtol", [(4, 32, 2e-4), (6, 48, 2e-6), (8, 64, 1e-8)])
def test_pppm_pair_energy_converges_to_ewald(ewald_system, order, K, rtol):
    """PPPM (cltype 3) differs from PME only in the influence function; its
    energies must converge to the Ewald sum like PME's."""
    L, qa, xa, qb, xb = ewald_system
    ref = ewald_pair(qa, xa, qb, xb, L, 0.35)
    p = O.make_params(screen=0.35, splodr=order, ms=(K, K, K), cltype=3)
    sysm = O.OracleSystem([0, 2], qb, [0, 0], np.array([[0.0]]), np.array([[1.0]]))
    uv, uv0 = O.energies_refs(sysm, p, np.diag([L] * 3), xb, qa, np.zeros(3, np.int32), xa[None], nthreads=2)
    assert abs(uv[0, 0] - ref) <= rtol * abs(ref)
    refself = ewald_self(qa, xa, L, 0.35)
    assert abs(uv0[0] - refself) <= max(rtol * 30, 1e-9) * abs(refself) + 1e-9


def test_pppm_influence_function_properties():
    """Zero at k = 0; the alias range is empty beyond m = 0 here, so each
    value is exp(-k^2/4a^2) U(k)^2 4 pi / (V k^2 denom); odd orders refused."""
    L, K, P = 25.0, 16, 4
    p = O.make_params(screen=0.35, splodr=P, ms=(K, K, K), cltype=3)
    g = O.engfac(p, np.diag([L] * 3))
    assert g[0, 0, 0] == 0.0 and np.all(g >= 0.0)
    gfb = O.pppm_gfb(P)
    m = np.array([3, -2, 5])                       # (rc1, rc2, rc3) = (3, 14, 5)
    k = 2 * np.pi * m / L
    th = np.pi * m / K
    U2 = np.prod((np.sin(th) / th) ** (2 * P))
    den = O.pppm_denom(P, gfb, *(np.sin(th) ** 2))
    want = np.exp(-(k @ k) / (4 * 0.35 ** 2)) * U2 * 4 * np.pi / (L ** 3 * (k @ k) * den)
    assert g[5, 14, 3] == pytest.approx(want, rel=1e-12)
    with pytest.raises(ValueError):
        O.engfac(O.make_params(splodr=5, ms=(K, K, K), cltype=3), np.diag([L] * 3))


def test_bare_coulomb_is_the_minimum_image_sum():
    """cltype 0 (realcal.F90:367-666, engproc.F90:819-840): each pair adds
    q_i q_j / r for r < elecut under the minimum image, no reciprocal or
    residual term; uvengy0 is 0 (the self terms return at once); volcorrect
    multiplies by the volume only."""
    rng = np.random.default_rng(11)
    L, n = 30.0, 40
    pos = rng.uniform(0, L, (2 * n, 3))
    pos[1::2] = pos[::2] + rng.normal(0, 0.9, (n, 3))       # diatomic molecules
    q = np.tile([0.4, -0.4], n) * CH
    slt = rng.uniform(0, L, (3, 3))
    slt[1:] = slt[0] + rng.normal(0, 0.9, (2, 3))
    qs = np.array([0.3, -0.1, -0.2]) * CH
    p = O.make_params(screen=0.35, splodr=4, ms=(32, 32, 32), cltype=0)
    sysm = O.OracleSystem(np.arange(0, 2 * n + 1, 2), q, np.zeros(2 * n, np.int32), np.array([[0.0]]),
                          np.array([[1.0]]))
    uv, uv0 = O.energies_refs(sysm, p, np.diag([L] * 3), pos, qs, np.zeros(3, np.int32), slt[None], nthreads=2)
    want = np.zeros(n)
    for m in range(n):
        for a in (2 * m, 2 * m + 1):
            for s in range(3):
                d = slt[s] - pos[a]
                d -= L * np.round(d / L)
                r = np.linalg.norm(d)
                if r < 12.0:
                    want[m] += qs[s] * q[a] / r
    np.testing.assert_allclose(uv[0], want, rtol=1e-12, atol=1e-12)
    assert uv0[0] == 0.0
    d = O.Distribution(3, [41], [41], np.linspace(-50, 50, 41), temp=0.6, volume=L ** 3, screen=0.35,
                       total_charge=1.0, wgtslf=1, estype=2, cltype=0)
    d.update(np.ones(n, np.int32), np.clip(uv[0], -49, 49), 0.0)
    assert d.d.engnorm == pytest.approx(L ** 3, rel=1e-15)


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_rotated_system_has_the_same_energies():
    """A cell that is not upper triangular is brought to upper-triangular
    form by a rotation (make_cell_uppertriangular, realcal.F90:776-817):
    rotating frame, cell and insertions together leaves every energy and
    self energy unchanged."""
    rng = np.random.default_rng(5)
    L, n = 26.0, 60
    pos = rng.uniform(0, L, (2 * n, 3))
    pos[1::2] = pos[::2] + rng.normal(0, 0.9, (n, 3))
    q = np.tile([0.4, -0.4], n) * CH
    slt = rng.uniform(0, L, (4, 3, 3))
    slt[:, 1:] = slt[:, :1] + rng.normal(0, 0.9, (4, 2, 3))
    qs = np.array([0.3, -0.1, -0.2]) * CH
    sysm = O.OracleSystem(np.arange(0, 2 * n + 1, 2), q, np.tile([0, 1], n), np.array([[0.2, 0.1], [0.1, 0.05]]),
                          np.array([[9.0, 7.0], [7.0, 5.0]]))
    R = _rotation(rng)
    cell = np.diag([L, L * 1.1, L * 0.95])
    for cltype in (0, 2):
        p = O.make_params(screen=0.35, splodr=6, ms=(48, 48, 48), cltype=cltype)
        u, u0 = O.energies_refs(sysm, p, cell, pos, qs, np.array([0, 1, 1], np.int32), slt, nthreads=2)
        rc = R @ cell
        assert abs(rc[1, 0]) + abs(rc[2, 0]) + abs(rc[2, 1]) > 1.0
        ur, u0r = O.energies_refs(sysm, p, rc, pos @ R.T, qs, np.array([0, 1, 1], np.int32), slt @ R.T,
                                  nthreads=2)
        np.testing.assert_allclose(ur, u, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(u0r, u0, rtol=1e-9, atol=1e-9)


def test_open_system_bare_coulomb_is_the_all_pairs_sum():
    """boxshp 0 (realcal.F90:387-395, 367-666): no periodic images, Coulomb
    between every pair (reelcut infinite), LJ within upljcut; a second
    configuration shifted by a box length changes the energies, which a
    periodic treatment would not."""
    rng = np.random.default_rng(13)
    n = 30
    pos = rng.uniform(0, 20.0, (2 * n, 3))
    pos[1::2] = pos[::2] + rng.normal(0, 0.9, (n, 3))
    q = np.tile([0.4, -0.4], n) * CH
    slt = rng.uniform(0, 20.0, (1, 3, 3))
    slt[0, 1:] = slt[0, 0] + rng.normal(0, 0.9, (2, 3))
    qs = np.array([0.3, -0.1, -0.2]) * CH
    eps, sig2 = np.array([[0.15]]), np.array([[9.0]])
    sysm = O.OracleSystem(np.arange(0, 2 * n + 1, 2), q, np.zeros(2 * n, np.int32), eps, sig2)
    p = O.make_params(cltype=0, boxshp=0, upljcut=12.0, lwljcut=10.0, ljswitch=0)
    uv, uv0 = O.energies_refs(sysm, p, np.zeros((3, 3)), pos, qs, np.zeros(3, np.int32), slt, nthreads=2)
    want = np.zeros(n)
    for m in range(n):
        for a in (2 * m, 2 * m + 1):
            for s in range(3):
                r = np.linalg.norm(slt[0, s] - pos[a])
                e = qs[s] * q[a] / r
                if r <= 12.0:
                    x6 = (sig2[0, 0] / r ** 2) ** 3
                    lj = 4 * eps[0, 0] * x6 * (x6 - 1)
                    if r > 10.0:
                        lw2, up2 = 100.0, 144.0
                        lj *= (2 * r * r + up2 - 3 * lw2) * (r * r - up2) ** 2 / (up2 - lw2) ** 3
                    e += lj
                want[m] += e
    np.testing.assert_allclose(uv[0], want, rtol=1e-12, atol=1e-12)
    assert uv0[0] == 0.0
    far = slt + np.array([60.0, 0.0, 0.0])
    uvf, _ = O.energies_refs(sysm, p, np.zeros((3, 3)), pos, qs, np.zeros(3, np.int32), far, nthreads=2)
    assert np.max(np.abs(uvf - uv)) > 1e-3
    with pytest.raises(ValueError, match="bare Coulomb"):
        O.energies_refs(sysm, O.make_params(cltype=2, boxshp=0), np.zeros((3, 3)), pos, qs,
                        np.zeros(3, np.int32), slt, nthreads=2)


def _triclinic(L):
    """Columns a, b, c of a cell whose (1,2) and (2,3) elements exceed half
    the diagonal, so normalize_cell_vector reduces columns 2 and 3."""
    return np.array([[L, 0.7 * L, 0.2 * L],
                     [0.0, L, -0.6 * L],
                     [0.0, 0.0, L]])


def test_triclinic_bare_coulomb_is_the_image_sum():
    """A skewed periodic cell (realcal.F90:272-282 minimum image after
    normalize_cell_vector's reduction, realcal.F90:832-847): each pair adds
    q_i q_j / r over every periodic image within elecut, computed here by
    brute force over lattice translations; the same lattice given by a
    reduced or an unreduced basis gives the same energies."""
    rng = np.random.default_rng(17)
    L, n = 30.0, 40
    cell = _triclinic(L)
    frac = rng.uniform(0, 1, (n, 3))
    pos = np.empty((2 * n, 3))
    pos[::2] = frac @ cell.T
    pos[1::2] = pos[::2] + rng.normal(0, 0.9, (n, 3))
    q = np.tile([0.4, -0.4], n) * CH
    slt = (rng.uniform(0, 1, 3) @ cell.T) + np.vstack([np.zeros(3), rng.normal(0, 0.9, (2, 3))])
    qs = np.array([0.3, -0.1, -0.2]) * CH
    p = O.make_params(screen=0.35, splodr=4, ms=(32, 32, 32), cltype=0)
    sysm = O.OracleSystem(np.arange(0, 2 * n + 1, 2), q, np.zeros(2 * n, np.int32), np.array([[0.0]]),
                          np.array([[1.0]]))
    uv, _ = O.energies_refs(sysm, p, cell, pos, qs, np.zeros(3, np.int32), slt[None], nthreads=2)
    shifts = np.array([[i, j, k] for i in (-2, -1, 0, 1, 2) for j in (-2, -1, 0, 1, 2) for k in (-2, -1, 0, 1, 2)])
    lat = shifts @ cell.T
    want = np.zeros(n)
    for m in range(n):
        for a in (2 * m, 2 * m + 1):
            for s in range(3):
                r = np.linalg.norm(slt[s] - pos[a] + lat, axis=1)
                want[m] += np.sum(qs[s] * q[a] / r[r < 12.0])
    np.testing.assert_allclose(uv[0], want, rtol=1e-12, atol=1e-12)
    # the reduced basis of the same lattice: b - a, then c + (b - a)
    red = cell.copy()
    red[:, 1] -= red[:, 0]
    red[:, 2] += red[:, 1]
    assert abs(red[0, 1]) <= 0.5 * L and abs(red[1, 2]) <= 0.5 * L
    uv2, _ = O.energies_refs(sysm, p, red, pos, qs, np.zeros(3, np.int32), slt[None], nthreads=2)
    np.testing.assert_allclose(uv2[0], want, rtol=1e-12, atol=1e-12)


def test_triclinic_pme_pair_energy_converges_to_ewald():
    """PME in the skewed cell (engfac from the full inverse cell, splines of
    the fractional coordinates, recpcal.F90:121-163, 504-538) against a
    brute-force Ewald sum over reciprocal vectors of that lattice."""
    rng = np.random.default_rng(19)
    L = 30.0
    cell = _triclinic(L)
    sites = (rng.uniform(0, 1, (2, 3)) @ cell.T)
    q = np.array([0.5, -0.5]) * CH
    alpha = 0.35
    p = O.make_params(screen=alpha, splodr=8, ms=(64, 64, 64), elecut=12.0)
    sysm = O.OracleSystem(np.array([0, 1]), q[1:], np.zeros(1, np.int32), np.array([[0.0]]), np.array([[1.0]]))
    uv, _ = O.energies_refs(sysm, p, cell, sites[1:], q[:1], np.zeros(1, np.int32), sites[None, :1], nthreads=2)
    # real-space part over images (the oracle's pair term), plus the reciprocal sum
    vol = abs(np.linalg.det(cell))
    recip = 2 * np.pi * np.linalg.inv(cell).T
    m = np.array([[i, j, k] for i in range(-12, 13) for j in range(-12, 13) for k in range(-12, 13)
                  if (i, j, k) != (0, 0, 0)])
    kv = m @ recip.T
    k2 = np.einsum("ij,ij->i", kv, kv)
    d = sites[0] - sites[1]
    rec = q[0] * q[1] * 4 * np.pi / vol * np.sum(np.exp(-k2 / (4 * alpha ** 2)) / k2 * np.cos(kv @ d))
    shifts = np.array([[i, j, k] for i in (-2, -1, 0, 1, 2) for j in (-2, -1, 0, 1, 2) for k in (-2, -1, 0, 1, 2)])
    r = np.linalg.norm(d + shifts @ cell.T, axis=1)
    real = np.sum(q[0] * q[1] * erfc(alpha * r[r < 12.0]) / r[r < 12.0])
    resid = -np.pi * q[0] * q[1] / alpha ** 2 / vol
    np.testing.assert_allclose(uv[0, 0], real + rec + resid, rtol=0, atol=2e-6)

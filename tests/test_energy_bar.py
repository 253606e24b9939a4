"""The per-energy bar of the GPU parity tests (erutil) on the CPU: its two
terms, the solute self-energy term, the per-decade report and the solute
site counts it is sized by."""
import numpy as np

import erutil as U
from erdst_gpu import synth


def test_bar_terms():
    e = np.array([0.0, -0.5, 3.0, -250.0, 1.0e6])
    for n in (1, 5, 45, 150):
        tol = U.energy_tolerance(e, n)
        np.testing.assert_allclose(tol, U.ATOL_SITE * np.sqrt(n) + U.RTOL * np.abs(e), rtol=1e-15)
    # nslt below one site is held to the one-site bar
    np.testing.assert_array_equal(U.energy_tolerance(e, 0), U.energy_tolerance(e, 1))


def test_mismatch_counts_the_energies_beyond_the_bar():
    ref = np.array([[0.0, 10.0], [-1000.0, 0.25]])
    n = 5
    tol = U.energy_tolerance(ref, n)
    gpu = ref + np.array([[0.5, 0.9], [-0.99, 1.5]]) * tol   # two beyond (factors > 1)
    nbad, worst, at = U.energy_mismatch(gpu, ref, n)
    assert nbad == 1 and at == (1, 1)
    np.testing.assert_allclose(worst, 0.5 * tol[1, 1], rtol=1e-9)
    gpu[0, 0] = ref[0, 0] + 1.2 * tol[0, 0]
    assert U.energy_mismatch(gpu, ref, n)[0] == 2
    # a looser relative term for a comparison against other inputs
    assert U.energy_mismatch(ref * (1 + 5e-5), ref, n, rtol=1e-4)[0] == 0
    assert U.energy_mismatch(ref * (1 + 5e-5), ref, n)[0] == 2   # the 10 and -1000 entries


def test_self_bar_grows_with_the_sites():
    u0 = np.array([0.01, -0.2])
    for n in (5, 25, 150):
        edge = U.energy_tolerance(u0, n) + U.SELF_ATOL_SITE * n
        assert U.self_mismatch(u0 + 0.99 * edge, u0, n)[0] == 0
        assert U.self_mismatch(u0 + 1.01 * edge, u0, n)[0] == 2


def test_error_by_decade():
    ref = np.array([1e-5, 0.02, 0.5, 3.0, 40.0, 5e3, 2e7])
    err = np.array([1e-7, 2e-7, 3e-7, 4e-6, 5e-5, 6e-3, 7e1])
    rows = U.error_by_decade(ref + err, ref)
    got = {d: (cnt, ab, rel) for d, cnt, ab, rel in rows}
    assert got[-3][:2] == (1, 1e-7) and got[-3][2] is None          # below 1e-2: absolute only
    assert got[-2][0] == 1 and np.isclose(got[-2][1], 2e-7)
    assert got[0][0] == 1 and np.isclose(got[0][2], 4e-6 / 3.0, rtol=1e-3)
    assert got[3][0] == 1 and np.isclose(got[3][2], 6e-3 / 5e3, rtol=1e-3)
    assert got[6][0] == 1 and np.isclose(got[6][2], 7e1 / 2e7, rtol=1e-3)   # the top decade holds all above
    assert sum(r[1] for r in rows) == ref.size


def test_site_counts_of_the_workloads():
    assert U.nsites(synth.config0()) == 5       # the rigid tetrahedral test particle
    assert U.nsites(synth.config1()) == 20      # the solution run's chain molecule
    assert "sqrt(5 sites)" in U.bar_text(5)

"""erdst -> slvfe on the oracle: the solvation free energy of the 5-site
solute in 1000 waters from soln and refs runs, through engstore's files and
slvfe's merge.  Rounding every energy to float32 before binning (what the
FP32 GPU path computes at best) moves each block's free energy by far less
than the 0.01 kcal/mol bar."""
import numpy as np
import pytest

import erutil as U
from erdst_gpu import slvfe, synth


@pytest.fixture(scope="module")
def pair():
    return synth.dmu_pair(0)


def test_soln_energies_stay_in_the_linear_mesh(pair):
    soln, _ = pair
    p = U.workload_params(soln)
    top = soln.uvmax_uvsoft_uvcrd()[2][soln.uvmax_uvsoft_uvcrd()[1][0] - 1]
    for f in range(3):
        uv, _ = U.oracle_soln(soln, f, p)
        sp = U.species_of(soln.topo, 1, soln.topo.tagslt)[0]
        assert uv[0][sp > 0].max() < top and uv[0][sp > 0].min() > -20.0


def test_dmu_oracle_f64_vs_f32_energies(pair, tmp_path):
    soln, refs = pair
    kT = U.workload_params(refs)["temp"]
    mesh = refs.uvmax_uvsoft_uvcrd()
    s64, r64 = U.oracle_dmu_blocks(soln, refs, 5, 2, 100, nthreads=4)
    s32, r32 = U.oracle_dmu_blocks(soln, refs, 5, 2, 100, f32_energies=True, nthreads=4)
    a = slvfe.dmu_from_blocks(s64, r64, *mesh, kT, str(tmp_path / "f64"))
    b = slvfe.dmu_from_blocks(s32, r32, *mesh, kT, str(tmp_path / "f32"))
    assert a.numrun == 2 and np.all(np.isfinite(a.chmpt))
    print("dmu f64", a.chmpt[0, :, :].tolist(), "f32 diff", np.abs(a.chmpt - b.chmpt).max())
    assert np.abs(a.chmpt - b.chmpt).max() < 0.01
    # the self-energy term: voffset + kT log(avslf/engnorm) of the refs run
    assert a.slfeng == pytest.approx(r64[0]["voffset"] + kT * np.log(r64[0]["avslf"] / r64[0]["engnorm"]),
                                     rel=1e-6)

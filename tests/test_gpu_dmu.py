"""Solvation free energy from the GPU path's distributions against the same
from the FP64 oracle's, through engstore's files and slvfe's merge: the
5-site solute in 1000 waters, 2 solution blocks of 5 frames and 1000
reference insertions.

Bar: 0.01 kcal/mol on the reported free energy (the block average at the
picked grouping, pickgr = 3), the north star's figure.  Individual blocks
and groupings are held to 0.05: with 1000 insertions a single molecule that
changes bin moves a sparsely populated bin's density by a visible fraction,
and perturbing the FP64 energies by 3e-6 relative (below the FP32 path's
1e-5 bar) already moves single-block, single-grouping values by up to
0.01 kcal/mol on the oracle alone."""
import numpy as np
import pytest

import erutil as U
from erdst_gpu import capi, slvfe, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [capi.ER_BATCHED, capi.ER_PLAIN])
def test_dmu_gpu_vs_oracle(tmp_path, mode):
    soln, refs = synth.dmu_pair(0)
    kT = U.workload_params(refs)["temp"]
    mesh = refs.uvmax_uvsoft_uvcrd()
    sg, rg = U.gpu_dmu_blocks(soln, refs, 5, 2, 100, mode=mode)
    so, ro = U.oracle_dmu_blocks(soln, refs, 5, 2, 100)
    a = slvfe.dmu_from_blocks(sg, rg, *mesh, kT, str(tmp_path / "gpu"))
    b = slvfe.dmu_from_blocks(so, ro, *mesh, kT, str(tmp_path / "oracle"))
    err = np.abs(a.chmpt - b.chmpt)
    print(f"mode {mode}: dmu gpu {a.dmu:.6f} oracle {b.dmu:.6f} |err| {abs(a.dmu - b.dmu):.3e}, "
          f"per block and grouping max {err.max():.3e} kcal/mol")
    print(err[0])
    assert np.all(np.isfinite(a.chmpt))
    assert abs(a.dmu - b.dmu) < 0.01
    assert err.max() < 0.05
    assert abs(a.slfeng - b.slfeng) < 1e-4

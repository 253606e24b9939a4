"""er_run_noalloc (no device allocation; partial sums in the library's own
scratch array) against er_run through the C ABI: per-insertion energies and
every block accumulator equal, over two calls in a row, the second with
fewer insertions and another frame."""
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


def _same(ga, gb):
    for key in ("edens", "ecorr", "eself", "minuv", "maxuv"):
        np.testing.assert_array_equal(ga[key], gb[key], err_msg=key)
    for key in ("engnorm", "engsmpl", "avslf", "voffset"):
        assert ga[key] == gb[key], key


@pytest.fixture(scope="module")
def cfg0():
    return synth.refs_water(0, 1000, 31.04, 48, 2, 32)


@pytest.mark.parametrize("mode", [capi.ER_BATCHED, capi.ER_PLAIN])
def test_noalloc_equals_run_two_calls(cfg0, mode):
    w = cfg0
    p = U.workload_params(w)
    steps = [(0, 48), (1, 20)] if mode == capi.ER_BATCHED else [(0, 8), (1, 5)]
    ref, na = _ctx(w, p), _ctx(w, p)
    # size na's buffers with an er_run of the largest step, then clear it
    f0, n0 = steps[0]
    na.set_frame(w.frame(f0), w.cell)
    na.set_insertions(w.insertions(f0)[:n0])
    na.run(mode)
    na.sync()
    na.clear()
    for f, nins in steps:
        print(f"noalloc mode {mode}: frame {f}, {nins} insertions", flush=True)
        for ctx, run in ((ref, ref.run), (na, na.run_noalloc)):
            ctx.set_frame(w.frame(f), w.cell)
            ctx.set_insertions(w.insertions(f)[:nins])
            run(mode)
            ctx.sync()
        ua, u0a = ref.uvengy(nins)
        ub, u0b = na.uvengy(nins)
        np.testing.assert_array_equal(ua, ub)
        np.testing.assert_array_equal(u0a, u0b)
        _same(ref.fetch(), na.fetch())
        # and both against the FP64 oracle at the energy bar
        uo, _ = U.oracle_refs(w, f, p, nins=nins)
        assert U.energy_mismatch(ub, uo, U.nsites(w))[0] == 0
    ref.close()
    na.close()


def test_noalloc_needs_sized_buffers(cfg0):
    w = cfg0
    p = U.workload_params(w)
    ctx = _ctx(w, p)
    ctx.set_frame(w.frame(0), w.cell)
    ctx.set_insertions(w.insertions(0)[:16])
    with pytest.raises(capi.ErError) as e:
        ctx.run_noalloc(capi.ER_BATCHED)
    assert e.value.status == capi.ER_E_STATE
    ctx.run(capi.ER_BATCHED)
    ctx.sync()
    ctx.set_insertions(w.insertions(0)[:32])  # more insertions than sized for
    with pytest.raises(capi.ErError) as e:
        ctx.run_noalloc(capi.ER_BATCHED)
    assert e.value.status == capi.ER_E_STATE
    ctx.close()


def test_noalloc_rejects_partials_beyond_the_scratch():
    """1600 bins with ecorr, all 16 solutes in one context: every solute takes
    at least one slice of 1600^2 x 8 bytes of partial sums, so the run needs
    at least 16 x 20.5 MB, above the 64 MiB scratch."""
    cfg3 = synth.config3()
    ws = [cfg3.workload(k) for k in range(16)]
    p = U.workload_params(ws[0])
    ctx = capi.Context(ws[0].topo, p)
    ctx.set_mesh(*cfg3.uvmax_uvsoft_uvcrd())
    ctx.set_solutes([(w.topo.slt_ljtype, w.topo.slt_charge32) for w in ws])
    ctx.set_frame(cfg3.frame(0), cfg3.cell)
    for k, w in enumerate(ws):
        ctx.select_solute(k)
        ctx.set_insertions(w.insertions(0)[:24])
    with pytest.raises(capi.ErError) as e:
        ctx.run_noalloc(capi.ER_BATCHED)
    assert e.value.status == capi.ER_E_UNSUPPORTED
    ctx.close()

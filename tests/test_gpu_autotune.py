"""Launch-shape tuning through the C ABI (er_autotune, er_set_tuning,
er_get_tuning): no shape changes a result bit for bit, and tuning leaves the
block accumulators, voffset and minuv/maxuv as they were."""
import numpy as np
import pytest

import erutil as U
from erdst_gpu import capi, synth

pytestmark = pytest.mark.gpu

KEYS = ("edens", "ecorr", "eself", "minuv", "maxuv", "slnuv")
SCALARS = ("engnorm", "engsmpl", "avslf", "voffset")


def _ctx(w, p):
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*w.uvmax_uvsoft_uvcrd())
    if p["slttype"] == 1:
        ctx.set_solute_molecules(w.topo.tagslt)
    else:
        ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    return ctx


def _frame(ctx, w, f, nins=None):
    ctx.set_frame(w.frame(f), w.cell)
    if nins is not None:
        ctx.set_insertions(w.insertions(f)[:nins])


def _same_block(ga, gb):
    for k in KEYS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
    for k in SCALARS:
        assert ga[k] == gb[k], k


@pytest.fixture(scope="module")
def cfg0():
    return synth.config0()


def test_autotune_keeps_accumulators_and_results(cfg0):
    w, nins = cfg0, 100
    p = U.workload_params(w)
    plain, tuned = _ctx(w, p), _ctx(w, p)
    for c in (plain, tuned):
        _frame(c, w, 0, nins)
        c.run(capi.ER_BATCHED)
        c.sync()
        _frame(c, w, 1, nins)
    before = tuned.fetch()
    t = tuned.autotune(reps=1)
    print("tuning:", t, flush=True)
    assert t["chunk"] in (256, 512, 1024, 2048, 4096)
    assert t["hist_slices"] in (0, 1, 2, 4, 8, 16, 32, 64)
    assert len(t["pair_split"]) == 1 and t["pair_split"][0] in (0, 1, 2, 4, 8, 16, 32, 64)
    assert t["cell_width"] in (0.0, 5.0, 7.0, 10.0)
    _same_block(tuned.fetch(), before)
    # the energies tuning leaves behind are the current frame's
    plain.run(capi.ER_BATCHED)
    plain.sync()
    np.testing.assert_array_equal(tuned.uvengy(nins)[0], plain.uvengy(nins)[0])
    tuned.run(capi.ER_BATCHED)
    tuned.sync()
    for a, b in zip(tuned.uvengy(nins), plain.uvengy(nins)):
        np.testing.assert_array_equal(a, b)
    _same_block(tuned.fetch(), plain.fetch())


@pytest.mark.parametrize("chunk,split,slices,cellw", [(16, 1, 1, 5.0), (48, 3, 7, 10.0), (4096, 64, 64, 1.0)])
def test_explicit_tunings_equal_builtin(cfg0, chunk, split, slices, cellw):
    w, nins = cfg0, 100
    p = U.workload_params(w)
    ref, alt = _ctx(w, p), _ctx(w, p)
    want = {"chunk": chunk, "hist_slices": slices, "pair_split": [split], "cell_width": cellw}
    alt.set_tuning(want)
    assert alt.tuning() == want
    for f in (0, 1):
        for c in (ref, alt):
            _frame(c, w, f, nins)
            c.run(capi.ER_BATCHED)
            c.sync()
        for a, b in zip(alt.uvengy(nins), ref.uvengy(nins)):
            np.testing.assert_array_equal(a, b)
    _same_block(alt.fetch(), ref.fetch())
    alt.set_tuning({"chunk": 0, "hist_slices": 0, "pair_split": [0]})
    assert alt.tuning() == {"chunk": chunk, "hist_slices": 0, "pair_split": [0], "cell_width": 0.0}


@pytest.mark.parametrize("over", [dict(cltype=0, boxshp=0, estype=1), dict(cltype=0), dict(cltype=3)],
                         ids=["open-bare", "bare", "pppm"])
def test_autotune_other_interactions(cfg0, over):
    """Open systems (no pair split to tune), bare Coulomb and PPPM: tuned
    runs give the untuned runs' results bit for bit."""
    w, nins = cfg0, 40
    p = U.workload_params(w, **over)
    ref, tuned = _ctx(w, p), _ctx(w, p)
    for f in (0, 1):
        for c in (ref, tuned):
            _frame(c, w, f, nins)
        if f == 0:
            t = tuned.autotune(reps=1)
            if over.get("boxshp") == 0:
                assert t["pair_split"] == [0]
        for c in (ref, tuned):
            c.run(capi.ER_BATCHED)
            c.sync()
        for a, b in zip(tuned.uvengy(nins), ref.uvengy(nins)):
            np.testing.assert_array_equal(a, b)
    _same_block(tuned.fetch(), ref.fetch())


def test_tuning_arguments_rejected(cfg0):
    ctx = _ctx(cfg0, U.workload_params(cfg0))
    for bad in ({"chunk": 17, "pair_split": [0]}, {"chunk": 8, "pair_split": [0]},
                {"chunk": 0, "pair_split": [65]}, {"chunk": 0, "pair_split": [-1]},
                {"chunk": 0, "pair_split": [0], "hist_slices": 65}):
        with pytest.raises(capi.ErError):
            ctx.set_tuning(bad)
    with pytest.raises(capi.ErError):
        ctx.autotune(reps=1)          # no frame yet
    _frame(ctx, cfg0, 0, 20)
    with pytest.raises(capi.ErError):
        ctx.autotune(reps=101)


def test_autotune_several_solutes():
    """The smallest, a middle and the largest solute of configuration 3 (1,
    7 and 45 sites) in one context: one split each, results equal to the
    untuned context's."""
    cfg3 = synth.config3()
    ks = sorted(range(16), key=lambda k: cfg3.workload(k).topo.slt_charge32.size)
    ks = [ks[0], ks[7], ks[-1]]
    ws = [cfg3.workload(k) for k in ks]
    p = U.workload_params(ws[0])
    nins = 24
    ctxs = []
    for _ in range(2):
        c = capi.Context(ws[0].topo, p)
        c.set_mesh(*cfg3.uvmax_uvsoft_uvcrd())
        c.set_solutes([(w.topo.slt_ljtype, w.topo.slt_charge32) for w in ws])
        ctxs.append(c)
    ref, tuned = ctxs
    for f in (0, 1):
        for c in ctxs:
            c.set_frame(cfg3.frame(f), cfg3.cell)
            for k, w in enumerate(ws):
                c.select_solute(k)
                c.set_insertions(w.insertions(f)[:nins])
        if f == 0:
            t = tuned.autotune(reps=1)
            print("tuning:", t, flush=True)
            assert len(t["pair_split"]) == 3 and all(s in (0, 1, 2, 4, 8, 16, 32, 64) for s in t["pair_split"])
        for c in ctxs:
            c.run(capi.ER_BATCHED)
            c.sync()
    for k in range(3):
        ref.select_solute(k)
        tuned.select_solute(k)
        for a, b in zip(tuned.uvengy(nins), ref.uvengy(nins)):
            np.testing.assert_array_equal(a, b)
        _same_block(tuned.fetch(), ref.fetch())


def test_autotune_soln():
    """A soln run: tuned shapes, and histogram slices set by hand (which soln
    runs do not follow), give the untuned run's results bit for bit."""
    w = synth.config1()
    p = U.workload_params(w, corrcal=1)
    ref, tuned = _ctx(w, p), _ctx(w, p)
    tuned.set_tuning({"chunk": 0, "hist_slices": 5, "pair_split": [0]})
    for f in (0, 1):
        for c in (ref, tuned):
            _frame(c, w, f)
        if f == 1:
            tuned.autotune(reps=1)
        for c in (ref, tuned):
            c.run(capi.ER_BATCHED)
            c.sync()
    n = len(w.topo.tagslt)
    for a, b in zip(tuned.uvengy(n), ref.uvengy(n)):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(tuned.flceng(n), ref.flceng(n))
    _same_block(tuned.fetch(), ref.fetch())

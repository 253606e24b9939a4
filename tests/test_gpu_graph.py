"""HIP graphs of whole runs (er_graph_begin / er_graph_end / er_graph_launch)
through the C ABI: a replayed graph of several frames adds to the block
accumulators exactly what the same calls add when executed one by one, over
several replays with eager work and synchronisation between them; and the
refusals that keep a replay consistent."""
import numpy as np
import pytest

import erutil as U
from erdst_gpu import capi, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def _same(ga, gb):
    for key in ("edens", "ecorr", "eself", "minuv", "maxuv", "slnuv"):
        np.testing.assert_array_equal(ga[key], gb[key], err_msg=key)
    for key in ("engnorm", "engsmpl", "avslf", "voffset"):
        assert ga[key] == gb[key], key


def _refs_ctx(w, p, solutes=None, mesh=None):
    ctx = capi.Context(w.topo, p)
    ctx.set_mesh(*(mesh or w.uvmax_uvsoft_uvcrd()))
    if solutes is None:
        ctx.set_solute(w.topo.slt_ljtype, w.topo.slt_charge32)
    else:
        ctx.set_solutes(solutes)
    return ctx


def _frames(ctx, cell, pos, ins, nins):
    """The calls of a batch: per frame, the frame and every solute's insertions, one batched run."""
    for f in range(len(pos)):
        ctx.set_frame_device(pos[f].data_ptr(), cell)
        for k, x in enumerate(ins[f]):
            if len(ins[f]) > 1:
                ctx.select_solute(k)
            ctx.set_insertions_device(nins, x.data_ptr())
        ctx.run(capi.ER_BATCHED)


@pytest.fixture(scope="module")
def cfg0():
    return synth.refs_water(0, 1000, 31.04, 48, 4, 32)


def test_graph_replays_equal_eager_runs(cfg0):
    """Four frames of 48 insertions recorded once and replayed three times;
    after each replay the accumulators and the last frame's energies equal
    those of an eager context running the same frames, with er_clear (an
    eager memset) and host reads between the replays."""
    w, nins = cfg0, 48
    p = U.workload_params(w)
    pos = [_dev(w.frame(f)) for f in range(4)]
    ins = [[_dev(w.insertions(f)[:nins])] for f in range(4)]
    torch.cuda.synchronize()
    eager, gr = _refs_ctx(w, p), _refs_ctx(w, p)
    # one eager run sizes the graph context's buffers; its accumulators are cleared
    _frames(gr, w.cell, pos[:1], ins[:1], nins)
    gr.sync()
    gr.clear()
    gr.graph_begin()
    _frames(gr, w.cell, pos, ins, nins)
    h = gr.graph_end()
    assert h >= 1
    for rep in range(3):
        print(f"graph replay {rep}", flush=True)
        _frames(eager, w.cell, pos, ins, nins)
        eager.sync()
        gr.graph_launch(h)
        gr.sync()
        _same(eager.fetch(), gr.fetch())
        ue, u0e = eager.uvengy(nins)
        ug, u0g = gr.uvengy(nins)
        np.testing.assert_array_equal(ue, ug)
        np.testing.assert_array_equal(u0e, u0g)
        if rep == 1:
            eager.clear()
            gr.clear()
    # the replayed energies are the FP64 oracle's to the energy bar
    uo, _ = U.oracle_refs(w, 3, p, nins=nins)
    assert U.energy_mismatch(ug, uo, U.nsites(w))[0] == 0
    gr.graph_destroy(h)
    eager.close()
    gr.close()


def test_graph_of_all_solutes_equals_eager(request):
    """Configuration 3's 16 solutes (1 to 45 sites) in one context, two
    frames of 24 insertions each per solute: one launch per kernel class per
    frame, recorded and replayed, equal to the eager runs for every solute."""
    cfg3 = synth.config3()
    nins = 24
    ws = [cfg3.workload(k) for k in range(16)]
    p = U.workload_params(ws[0])
    sol = [(w.topo.slt_ljtype, w.topo.slt_charge32) for w in ws]
    pos = [_dev(cfg3.frame(f)) for f in range(2)]
    ins = [[_dev(w.insertions(f)[:nins]) for w in ws] for f in range(2)]
    torch.cuda.synchronize()
    mesh = cfg3.uvmax_uvsoft_uvcrd()
    eager, gr = _refs_ctx(ws[0], p, sol, mesh), _refs_ctx(ws[0], p, sol, mesh)
    _frames(gr, cfg3.cell, pos[:1], ins[:1], nins)
    gr.sync()
    gr.clear()
    gr.graph_begin()
    _frames(gr, cfg3.cell, pos, ins, nins)
    h = gr.graph_end()
    for rep in range(2):
        print(f"all-solute graph replay {rep}", flush=True)
        _frames(eager, cfg3.cell, pos, ins, nins)
        eager.sync()
        gr.graph_launch(h)
        gr.sync()
        for k in range(16):
            eager.select_solute(k)
            gr.select_solute(k)
            _same(eager.fetch(), gr.fetch())
            np.testing.assert_array_equal(eager.uvengy(nins)[0], gr.uvengy(nins)[0])
    eager.close()
    gr.close()


def test_graph_soln_equals_eager():
    """A solution run (one flexible solute in 8000 waters): three frames
    recorded and replayed twice."""
    w = synth.config1()
    p = U.workload_params(w)
    pos = [_dev(w.frame(f)) for f in range(3)]
    torch.cuda.synchronize()

    def ctx():
        c = capi.Context(w.topo, p)
        c.set_mesh(*w.uvmax_uvsoft_uvcrd())
        c.set_solute_molecules(w.topo.tagslt)
        return c

    def frames(c):
        for x in pos:
            c.set_frame_device(x.data_ptr(), w.cell)
            c.run(capi.ER_BATCHED)

    eager, gr = ctx(), ctx()
    gr.set_frame_device(pos[0].data_ptr(), w.cell)
    gr.run(capi.ER_BATCHED)
    gr.sync()
    gr.clear()
    gr.graph_begin()
    frames(gr)
    h = gr.graph_end()
    for rep in range(2):
        frames(eager)
        eager.sync()
        gr.graph_launch(h)
        gr.sync()
        _same(eager.fetch(), gr.fetch())
        nslt = len(w.topo.tagslt)
        np.testing.assert_array_equal(eager.flceng(nslt), gr.flceng(nslt))
    eager.close()
    gr.close()


def test_graph_refusals(cfg0):
    w, nins = cfg0, 16
    p = U.workload_params(w)
    pos = [_dev(w.frame(f)) for f in range(2)]
    ins = [[_dev(w.insertions(f)[:nins])] for f in range(2)]
    torch.cuda.synchronize()
    ctx = _refs_ctx(w, p)
    # recording before any run: nothing is sized yet
    ctx.graph_begin()
    with pytest.raises(capi.ErError) as e:
        _frames(ctx, w.cell, pos[:1], ins[:1], nins)
    assert e.value.status == capi.ER_E_STATE
    with pytest.raises(capi.ErError):
        ctx.graph_end()
    _frames(ctx, w.cell, pos[:1], ins[:1], nins)
    ctx.sync()
    # host-buffer setters, ER_PLAIN runs and reads are not recorded
    ctx.graph_begin()
    for call in (lambda: ctx.set_frame(w.frame(0), w.cell), lambda: ctx.set_insertions(w.insertions(0)[:nins]),
                 lambda: ctx.run(capi.ER_PLAIN), lambda: ctx.sync(), lambda: ctx.fetch(), lambda: ctx.clear()):
        with pytest.raises(capi.ErError) as e:
            call()
        assert e.value.status == capi.ER_E_STATE
    with pytest.raises(capi.ErError):
        ctx.graph_end()  # spoiled by the refused calls
    ctx.graph_begin()
    _frames(ctx, w.cell, pos, ins, nins)
    h = ctx.graph_end()
    ctx.graph_launch(h)
    ctx.sync()
    # another insertion count changes the run's shapes: the graph is refused
    ctx.set_insertions_device(nins // 2, ins[0][0].data_ptr())
    ctx.run(capi.ER_BATCHED)
    ctx.sync()
    with pytest.raises(capi.ErError) as e:
        ctx.graph_launch(h)
    assert e.value.status == capi.ER_E_STATE
    # back to the recorded shapes: it replays again
    ctx.set_insertions_device(nins, ins[0][0].data_ptr())
    ctx.run(capi.ER_BATCHED)
    ctx.sync()
    ctx.graph_launch(h)
    ctx.sync()
    ctx.close()


def test_graph_replay_reports_a_halt(cfg0):
    """A replayed run whose energies fall below the mesh: er_sync reports
    eng_min as after an eager run."""
    w, nins = cfg0, 4
    p = U.workload_params(w)
    uvmax, uvsoft, uvcrd = w.uvmax_uvsoft_uvcrd()
    pos = [_dev(w.frame(0))]
    ins = [[_dev(w.insertions(0)[:nins])]]
    torch.cuda.synchronize()
    ctx = _refs_ctx(w, p, mesh=(uvmax, uvsoft, uvcrd + 1.0e3))
    _frames(ctx, w.cell, pos, ins, nins)
    with pytest.raises(capi.ErError):
        ctx.sync()  # the eager run halts already
    ctx.graph_begin()
    _frames(ctx, w.cell, pos, ins, nins)
    h = ctx.graph_end()
    ctx.graph_launch(h)
    with pytest.raises(capi.ErError) as e:
        ctx.sync()
    assert e.value.status == capi.ER_E_HALT
    assert ctx.last_error()[0] == "eng_min"
    ctx.close()

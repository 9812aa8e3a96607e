"""Batched persistent sweeps (many problems in one launch), their HIP-graph
replay, and the Tgen_QDf step (Temp prescribed, one sweep per step):

* a problem swept in a batch gives bit-for-bit the intensities and moments it
  gives swept alone (the tile arithmetic does not depend on the batch);
* graph replays equal eager batched sweeps, with host work between replays;
* the Tgen_QDf step matches the FP64 oracle (MATERIAL_UPDATE, TRANSPORT_SCB,
  COLLAPSE_INTENSITIES, fg_Calc, Cg_Calc; ALGORITHMS.f08:1031-1095).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

# uneven grids, group counts and quadratures
MIX = [dict(nx=13, ny=11, order=4, ng=5), dict(nx=20, ny=9, order=8, ng=3), dict(nx=7, ny=16, order=12, ng=9),
       dict(nx=10, ny=10, order=4, ng=40)]


def _probs(trt):
    out = []
    for k, kw in enumerate(MIX):
        p = trt.problem.small(seed=11 + k, **kw)
        p.bcT = (1000.0, 0.0, 300.0, 0.0)
        out.append(p)
    return out


def _randomize(ctx, prob, seed):
    """A non-uniform temperature, hence non-uniform opacities and sources."""
    rng = np.random.default_rng(seed)
    ctx.push("Temp", rng.uniform(0.2, 1.2, (prob.nx, prob.ny)) * 1000.0)
    ctx.call("material_update")


@pytest.mark.parametrize("precision,tile,seg,arg", [(1, (4, 2), None, 0), (0, (2, 4), None, 0),
                                                   (1, (2, 2), None, 0), (1, (4, 2), "2", 0), (0, (2, 4), "3", 0),
                                                   (2, (2, 2), "4", 0), (1, (2, 8), None, 8), (0, (4, 2), None, 1),
                                                   (1, (4, 4), "1", 3)])
def test_batch_sweep_equals_single(trt, monkeypatch, precision, tile, seg, arg):
    """seg forces the batch's column-segment length through TRT_SEG_LEN, arg through
    trt_batch_create_seg (which wins over the environment); by default the batch
    picks tiles or segments from the available parallelism."""
    if seg:
        monkeypatch.setenv("TRT_SEG_LEN", seg)
    probs = _probs(trt)
    alone = [trt.native.Context(p, precision=precision, tile=tile) for p in probs]
    mixed = [trt.native.Context(p, precision=precision, tile=tile) for p in probs]
    for k, p in enumerate(probs):
        _randomize(alone[k], p, k)
        _randomize(mixed[k], p, k)
    for c in alone:
        c.sweep_timed(2)
    with trt.native.Batch(mixed, seg=arg) as b:
        upd, grid, nitems = b.info()
        assert upd == sum(p.updates_per_sweep() for p in probs)
        assert grid >= 1 and nitems > 0
        if arg:
            assert b.segment_length == arg
        elif seg:
            assert b.segment_length == int(seg)
        b.sweep_timed(2)
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.pull("I_crn"), m.pull("I_crn"))
        for c in (a, m):
            c.call("collapse_intensities")
        for n in ("HO_Eg_avg", "HO_Eg_edgV", "HO_Eg_edgH", "HO_Fxg_edgV", "HO_Fyg_edgH"):
            assert np.array_equal(a.pull(n), m.pull(n)), n
    for c in alone + mixed:
        c.close()


def test_batch_graph_replays_equal_eager(trt):
    probs = _probs(trt)
    eager = [trt.native.Context(p, precision=1, tile=(4, 2)) for p in probs]
    graph = [trt.native.Context(p, precision=1, tile=(4, 2)) for p in probs]
    for k, p in enumerate(probs):
        _randomize(eager[k], p, 10 + k)
        _randomize(graph[k], p, 10 + k)
    be = trt.native.Batch(eager)
    bg = trt.native.Batch(graph, graph=True)
    for rep in range(4):
        be.sweep_timed(1)
        bg.sweep_timed(1)
        for e, g in zip(eager, graph):
            assert np.array_equal(e.pull("I_crn"), g.pull("I_crn")), rep
        # host-driven work between replays, as a time step does
        for k, (e, g) in enumerate(zip(eager, graph)):
            t = e.pull("Temp") * (1.0 + 0.01 * (rep + 1))
            for c in (e, g):
                c.push("Temp", t)
                c.call("material_update")
    be.close()
    bg.close()
    for c in eager + graph:
        c.close()


def _oracle_tgen_step(o):
    o.call("material_update")
    o["I_crn_old"][...] = o["I_crn"]
    o.call("transport_scb")
    o.call("collapse_intensities")
    o.call("fg_calc")
    o.call("cg_calc")


TGEN_OUT = ["HO_E_avg", "HO_Eg_avg", "HO_Eg_edgV", "HO_Eg_edgH", "fg_avg_xx", "fg_avg_yy", "fg_avg_xy", "fg_edgV_xx",
            "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy", "Cg_L", "Cg_B", "Cg_R", "Cg_T"]


def test_tgen_step_matches_oracle(trt, oracle_mod):
    prob = trt.problem.small(nx=12, ny=10, order=4, ng=3)
    o = oracle_mod.Oracle(prob, threads=3)
    ctx = trt.native.Context(prob)
    temp = np.random.default_rng(5).uniform(200.0, 1200.0, (prob.nx, prob.ny))
    o["Temp"][...] = temp
    ctx.push("Temp", temp)
    for s in range(3):
        _oracle_tgen_step(o)
        ctx.tgen_step()
        for n in TGEN_OUT:
            assert rel_err(ctx.pull(n), o[n].ravel(order="F")) < 1e-12, (s, n)
    ctx.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_tgen_equals_single(trt, graph):
    probs = _probs(trt)
    alone = [trt.native.Context(p, precision=1, tile=(4, 2)) for p in probs]
    mixed = [trt.native.Context(p, precision=1, tile=(4, 2)) for p in probs]
    for k, p in enumerate(probs):
        t = np.random.default_rng(20 + k).uniform(200.0, 1200.0, (p.nx, p.ny))
        alone[k].push("Temp", t)
        mixed[k].push("Temp", t)
    b = trt.native.Batch(mixed, graph=graph)
    for s in range(3):
        for c in alone:
            c.tgen_step()
        b.tgen_step()
    for a, m in zip(alone, mixed):
        assert a.time == m.time
        for n in TGEN_OUT:
            assert np.array_equal(a.pull(n), m.pull(n)), n
    b.close()
    for c in alone + mixed:
        c.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_whole_step_graphs_equal_eager_steps(trt, precision):
    """Batched Tgen_QDf steps replayed from whole-step graphs (every member's
    MATERIAL_UPDATE, the persistent sweep, every member's moments, QD factors and C_g in
    one graph launch; one graph per I_crn / I_crn_old parity) equal eager batched steps
    bitwise at every step, with host work (a new prescribed Temp) between the replays."""
    probs = _probs(trt)
    tile = (4, 2) if precision else (2, 4)
    eager = [trt.native.Context(p, precision=precision, tile=tile) for p in probs]
    graph = [trt.native.Context(p, precision=precision, tile=tile) for p in probs]
    temps = [np.random.default_rng(40 + k).uniform(200.0, 1200.0, (p.nx, p.ny)) for k, p in enumerate(probs)]
    for k in range(len(probs)):
        eager[k].push("Temp", temps[k])
        graph[k].push("Temp", temps[k])
    be = trt.native.Batch(eager)
    bg = trt.native.Batch(graph, graph=2)
    for s in range(5):
        be.tgen_step()
        bg.tgen_step()
        for e, g in zip(eager, graph):
            assert e.time == g.time
            for n in TGEN_OUT + ["I_crn", "Temp"]:
                assert np.array_equal(e.pull(n), g.pull(n)), (s, n)
        for k, (e, g) in enumerate(zip(eager, graph)):
            t = temps[k] * (1.0 + 0.05 * (s + 1))
            e.push("Temp", t)
            g.push("Temp", t)
    be.close()
    bg.close()
    for c in eager + graph:
        c.close()


def test_batch_rejects_mixed_tiles(trt):
    probs = _probs(trt)[:2]
    c1 = trt.native.Context(probs[0], precision=1, tile=(4, 2))
    c2 = trt.native.Context(probs[1], precision=1, tile=(2, 4))
    with pytest.raises(trt.native.TrtError, match="tile shape"):
        trt.native.Batch([c1, c2])
    c3 = trt.native.Context(probs[1], precision=0, tile=(4, 2))
    with pytest.raises(trt.native.TrtError, match="precision"):
        trt.native.Batch([c1, c3])
    for c in (c1, c2, c3):
        c.close()


def test_fused_collapse_equals_separate_routines(trt):
    """The step's collapse (COLLAPSE_INTENSITIES with fg_Calc fused) gives bit for bit
    what trt_collapse_intensities followed by trt_fg_calc gives."""
    prob = _probs(trt)[2]
    temp = np.random.default_rng(9).uniform(200.0, 1200.0, (prob.nx, prob.ny))
    a = trt.native.Context(prob, precision=1, tile=(4, 2))
    b = trt.native.Context(prob, precision=1, tile=(4, 2))
    for c in (a, b):
        c.push("Temp", temp)
    a.tgen_step()
    for r in ("material_update", "transport_scb", "collapse_intensities", "fg_calc", "cg_calc"):
        b.call(r)
    for n in TGEN_OUT + ["Hg_avg_xy", "Hg_edgV_xx", "HO_Fxg_edgV", "HO_Fyg_edgH", "HO_E_edgV"]:
        assert np.array_equal(a.pull(n), b.pull(n)), n
    a.close()
    b.close()


def test_batch_rejects_negative_segment_length(trt):
    probs = _probs(trt)[:2]
    ctxs = [trt.native.Context(p, precision=1, tile=(4, 2)) for p in probs]
    try:
        with pytest.raises(trt.native.TrtError, match="segment length"):
            trt.native.Batch(ctxs, seg=-1)
    finally:
        for c in ctxs:
            c.close()


def test_batch_spin_limit_exceeded_is_reported(trt, monkeypatch):
    """As for one problem: a batched sweep whose dependency waits exceed the spin
    limit (forced to 0 from the host) finishes and reports the error."""
    probs = [trt.problem.small(nx=48, ny=40, order=8, ng=3, seed=s) for s in (1, 2)]
    ctxs = [trt.native.Context(p, precision=1, tile=(4, 2)) for p in probs]
    try:
        monkeypatch.setenv("TRT_SPIN_LIMIT", "0")
        with trt.native.Batch(ctxs) as b:
            with pytest.raises(trt.native.TrtError, match="spin limit"):
                b.sweep_timed(1)
    finally:
        for c in ctxs:
            c.close()

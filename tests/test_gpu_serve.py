"""The request loop (continuous batching under an HBM budget): a request's
results are the same alone and in a mix that arrives around it; admission
honours the budget and the active cap; latency statistics are consistent."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUTS = ["Temp", "HO_E_avg", "HO_Eg_avg", "fg_avg_xx", "fg_avg_yy", "fg_avg_xy", "fg_edgV_xx", "fg_edgH_xy", "Cg_L",
        "Cg_T"]


def _mix(trt):
    specs = [dict(nx=24, ny=20, order=8, ng=8), dict(nx=13, ny=11, order=4, ng=5), dict(nx=30, ny=9, order=12, ng=3),
             dict(nx=16, ny=16, order=4, ng=40)]
    out = []
    for k, kw in enumerate(specs):
        p = trt.problem.small(seed=31 + k, **kw)
        p.bcT = (1000.0 + 100.0 * k, 0.0, 0.0, 0.0)
        out.append(p)
    return out


def test_request_same_alone_and_in_mix(trt):
    probs = _mix(trt)
    temp = np.random.default_rng(3).uniform(300.0, 1100.0, (probs[0].nx, probs[0].ny))
    with trt.serve.Server(8 << 30) as srv:
        a = srv.submit(probs[0], steps=5, temp=temp)
        srv.run()
        assert srv.request(a)["status"] == 0
        alone = {n: srv.output(a, n) for n in OUTS}
    with trt.serve.Server(8 << 30) as srv:
        # long-running requests around the one under test, which arrives while they run
        others = [srv.submit(probs[1], 60, 0.0), srv.submit(probs[2], 80, 0.002)]
        a = srv.submit(probs[0], 5, 0.01, temp=temp)
        others.append(srv.submit(probs[3], 40, 0.012))
        srv.run()
        st = srv.request(a)
        assert st["status"] == 0 and st["steps"] == 5
        assert st["updates"] == 5 * probs[0].updates_per_sweep()
        ost = [srv.request(i) for i in others]
        assert all(o["status"] == 0 for o in ost)
        # it shared ticks with at least one other request
        assert any(o["admit_s"] < st["done_s"] and st["admit_s"] < o["done_s"] for o in ost)
        for n in OUTS:
            assert np.array_equal(srv.output(a, n), alone[n]), n
        assert srv.summary()["peak_active"] >= 2


def test_server_runs_twice_on_one_arena(trt):
    """A second run serves only what was submitted after the first (a finished
    request that sorts after a new one by arrival time is not served again), on
    the arena the first run reserved; the same problem gives the same results."""
    probs = _mix(trt)
    with trt.serve.Server(8 << 30) as srv:
        a = srv.submit(probs[0], steps=3, arrival_s=0.05)
        srv.run()
        s1 = srv.summary()
        ra = srv.request(a)
        assert ra["status"] == 0 and s1["ticks"] == 3
        b = srv.submit(probs[0], steps=3, arrival_s=0.0)
        c = srv.submit(probs[1], steps=2, arrival_s=0.0)
        srv.run()
        s2 = srv.summary()
        assert srv.request(b)["status"] == 0 and srv.request(c)["status"] == 0
        assert srv.request(a) == ra                 # untouched by the second run
        assert s2["ticks"] == 3 and s2["updates"] == 3 * probs[0].updates_per_sweep() + 2 * probs[1].updates_per_sweep()
        assert s1["setup_s"] >= 0.0 and s2["setup_s"] == s1["setup_s"]
        for n in OUTS:
            assert np.array_equal(srv.output(a, n), srv.output(b, n)), n


def test_error_in_a_tick_fails_held_requests(trt, monkeypatch):
    """An error inside a tick (a batched sweep whose spin limit, forced to 0 from the
    host, is exceeded) ends the run with an error and fails the requests it held; the
    server then serves new requests normally, with the same results as before."""
    probs = _mix(trt)
    with trt.serve.Server(8 << 30) as srv:
        ref = srv.submit(probs[0], steps=2)
        srv.run()
        good = {n: srv.output(ref, n) for n in OUTS}
        monkeypatch.setenv("TRT_SPIN_LIMIT", "0")
        held = [srv.submit(p, steps=3) for p in probs[:2]]
        with pytest.raises(trt.native.TrtError, match="spin limit"):
            srv.run()
        for i in held:
            st = srv.request(i)
            assert st["status"] == -1 and "aborted" in st["error"], st
        monkeypatch.delenv("TRT_SPIN_LIMIT")
        again = srv.submit(probs[0], steps=2)
        srv.run()
        assert srv.request(again)["status"] == 0
        for n in OUTS:
            assert np.array_equal(srv.output(again, n), good[n]), n


def test_admission_honours_budget_and_cap(trt):
    p = trt.problem.small(nx=96, ny=96, order=8, ng=16, seed=41)
    need = trt.native.device_bytes(p, precision=1, tile=(4, 2))
    admit = need + need // 16 + (8 << 20)             # the loop's admission estimate
    budget = 2 * admit + need // 2                    # room for two requests, not three
    with trt.serve.Server(budget, max_active=8) as srv:
        ids = [srv.submit(p, steps=3, arrival_s=0.0) for _ in range(5)]
        srv.run()
        st = [srv.request(i) for i in ids]
        summ = srv.summary()
    assert all(s["status"] == 0 for s in st)
    assert summ["peak_active"] <= 2 and summ["peak_bytes"] <= budget
    first_done = min(s["done_s"] for s in st)
    late = sorted(st, key=lambda s: s["admit_s"])[2:]
    assert all(s["admit_s"] >= first_done for s in late)
    with trt.serve.Server(64 << 30, max_active=1) as srv:
        ids = [srv.submit(p, steps=2, arrival_s=0.0) for _ in range(3)]
        srv.run()
        assert srv.summary()["peak_active"] == 1
        assert all(srv.request(i)["status"] == 0 for i in ids)


def test_oversized_request_fails_alone(trt):
    small, big = _mix(trt)[1], _mix(trt)[0]
    need = trt.native.device_bytes(small, precision=1, tile=(4, 2))
    budget = 2 * need + need // 16 + (8 << 20)       # the small request's admission estimate, with room
    assert trt.native.device_bytes(big, precision=1, tile=(4, 2)) > 2 * need
    with trt.serve.Server(budget) as srv:
        b = srv.submit(big, 2)
        s = srv.submit(small, 2)
        srv.run()
        assert srv.request(b)["status"] == -1 and "budget" in srv.request(b)["error"]
        assert srv.request(s)["status"] == 0


def test_run_load_statistics(trt):
    probs = _mix(trt)
    res = trt.serve.run_load(rate=50.0, n=8, steps=3, budget_gib=16.0, problems=[probs[k % 4] for k in range(8)])
    assert res["completed"] == 8 and res["failed"] == 0
    assert 0.0 <= res["latency_p50_s"] <= res["latency_p99_s"]
    assert res["service_p50_s"] <= res["latency_p99_s"] + 1e-12
    assert res["ticks"] >= 3 and res["updates_per_s"] > 0


def test_autotuner_returns_a_listed_shape(trt):
    probs = _mix(trt)[:2]
    for prec in (1, 0):
        shapes = trt.native.autotune_shapes(prec)
        tile, best, table = trt.native.autotune_tile(probs, precision=prec, sweeps=2)
        assert tile in shapes and set(table) == set(shapes)
        assert best == min(table.values()) and best > 0


def test_batch_autotuner_covers_shapes_and_segments(trt):
    probs = _mix(trt)[:2]
    shapes, segs = trt.native.autotune_shapes(1), trt.native.autotune_segments()
    assert 1 in segs and len(segs) >= 3
    tile, seg, best, table = trt.native.autotune_batch(probs, precision=1, sweeps=2)
    assert tile in shapes and seg in segs
    assert set(table) == {(sh, sg) for sh in shapes for sg in segs}
    assert best == min(table.values()) == table[(tile, seg)] and best > 0

"""Batched and served full time steps (solves) on the GPU, through the C ABI: trt_batch_step
runs TRT_MLQD_ALGORITHM's step (MainSrc/ALGORITHMS.f08:396-837) for every member in lockstep
rounds of RT iterations, each round's sweeps in one persistent launch (members that need no
sweep in a round are skipped in it), each member's low-order work on its own stream from a
host thread of its own; the request loop's solve mode serves such steps with continuous
batching.  Each member's results must be bitwise those of trt_step on it alone, in any mix."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("Temp", "E_avg", "HO_E_avg", "MGQD_E_avg", "Eg_avg", "fg_avg_xx", "fg_edgV_xy", "Cg_L", "I_crn")


def _members(trt):
    P = trt.problem
    a = P.mapped(nx=16, order=4, ng=4, sigma=1.0)
    b = P.small(nx=12, ny=10, order=8, ng=3).with_(bc_type=(0, 0, 1, 0), bcT=(1000.0, 0.0, 400.0, 0.0))
    c = P.cfg0()
    d = P.small(nx=14, ny=9, order=4, ng=3, seed=3)
    return [a, b, c, d]


@pytest.mark.parametrize("graph", [0, 1])
def test_batched_steps_equal_steps_alone(trt, graph):
    probs = _members(trt)
    alone = [trt.native.Context(p, precision=1, tile=(4, 4)) for p in probs]
    mixed = [trt.native.Context(p, precision=1, tile=(4, 4)) for p in probs]
    its_alone = [[c.step() for _ in range(3)] for c in alone]
    batch = trt.native.Batch(mixed, graph=graph)
    its_mixed = [batch.step() for _ in range(3)]
    batch.close()
    rts = set()
    for i, c in enumerate(mixed):
        for s in range(3):
            assert tuple(its_mixed[s][i]) == tuple(its_alone[i][s]), (i, s)
            rts.add(its_mixed[s][i][0])
        for n in FIELDS:
            assert np.array_equal(c.pull(n), alone[i].pull(n)), (i, n)
        assert c.time == alone[i].time
    assert len(rts) > 1          # members left RT iterations at different rounds: some rounds skipped some
    for c in alone + mixed:
        c.close()


def test_served_solves_same_alone_and_in_mix(trt):
    serve = importlib.import_module("2d_trt_rom_system_amd.serve")
    probs = _members(trt)[:3]
    steps = [3, 2, 3]

    def outputs(srv, rid):
        return {n: srv.output(rid, n) for n in ("Temp", "E_avg", "HO_E_avg", "fg_avg_xx", "Cg_L")}

    alone = []
    for p, k in zip(probs, steps):
        with serve.Server(8 * 2**30, 8, precision=1, solves=True) as srv:
            rid = srv.submit(p, k, 0.0)
            srv.run()
            r = srv.request(rid)
            assert r["status"] == 0 and r["steps"] == k, r
            alone.append(outputs(srv, rid))
    with serve.Server(8 * 2**30, 8, precision=1, solves=True) as srv:
        ids = [srv.submit(p, k, a) for p, k, a in zip(probs, steps, (0.0, 0.0, 0.05))]
        srv.run()
        summ = srv.summary()
        for i, rid in enumerate(ids):
            r = srv.request(rid)
            assert r["status"] == 0 and r["steps"] == steps[i], (i, r)
            got = outputs(srv, rid)
            for n, v in alone[i].items():
                assert np.array_equal(got[n], v), (i, n)
    assert summ["peak_active"] >= 2 and summ["updates"] > 0


def test_solve_requests_reject_a_temperature_field(trt):
    serve = importlib.import_module("2d_trt_rom_system_amd.serve")
    p = trt.problem.small(nx=8, ny=6, order=4, ng=2)
    with serve.Server(2 * 2**30, 4, precision=1, solves=True) as srv:
        rid = srv.submit(p, 1, 0.0, temp=np.full(p.nx * p.ny, 2.0))
        srv.run()
        r = srv.request(rid)
        assert r["status"] == -1 and "initial state" in r["error"]

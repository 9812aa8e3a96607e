"""Request loop with continuous batching (the trt_server C ABI) and a Poisson
load generator that reports latency percentiles beside throughput.

A request is a problem, an optional prescribed temperature field and a number
of Tgen_QDf steps (MainSrc/ALGORITHMS.f08:1031-1095, one sweep per step), or, in
solve mode, a number of full time steps of TRT_MLQD_ALGORITHM from the initial
state.  The loop itself is native (csrc/serve.hip); this module submits requests,
runs the loop and summarises it:

    python -m 2d_trt_rom_system_amd.serve --rates 2 4 8 --requests 32
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys

import numpy as np

from . import native, problem


class Server:
    """A trt_server: submit requests, run them, read stats and outputs."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 64, precision: int = 1, tile=(4, 4),
                 solves: bool = False):
        """solves: requests are full time steps (trt_server_set_mode 1), not Tgen_QDf steps."""
        self.L = native.load()
        self.precision = precision
        self.tile = tile
        h = C.c_void_p()
        if self.L.trt_server_create(C.byref(h), int(hbm_budget_bytes), int(max_active)) != 0:
            raise native.TrtError("trt_server_create failed")
        self.h = h
        self.ids = []
        if solves:
            self._check(self.L.trt_server_set_mode(self.h, 1), "server_set_mode")

    def close(self):
        if getattr(self, "h", None):
            self.L.trt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise native.TrtError(f"{what} failed ({rc}): {self.L.trt_server_last_error(self.h).decode()}")

    def submit(self, prob, steps: int, arrival_s: float = 0.0, temp=None) -> int:
        s, keep = native.make_problem_struct(prob, self.precision, self.tile)
        tp = None
        if temp is not None:
            t = np.ascontiguousarray(np.asarray(temp, dtype=np.float64).ravel(order="F"))
            if t.size != prob.nx * prob.ny:
                raise ValueError("temp must have N_x*N_y values")
            keep.append(t)
            tp = t.ctypes.data_as(native._dp)
        rid = C.c_int()
        self._check(self.L.trt_server_submit(self.h, C.byref(s), tp, int(steps), float(arrival_s), C.byref(rid)),
                    "server_submit")
        self.ids.append(rid.value)
        return rid.value

    def run(self) -> None:
        self._check(self.L.trt_server_run(self.h), "server_run")

    def request(self, rid: int) -> dict:
        st = native.RequestStats()
        self._check(self.L.trt_server_request(self.h, int(rid), C.byref(st)), "server_request")
        d = {k: getattr(st, k) for k, _ in native.RequestStats._fields_}
        d["error"] = self.L.trt_server_request_error(self.h, int(rid)).decode()
        return d

    def summary(self) -> dict:
        st = native.ServerStats()
        self._check(self.L.trt_server_summary(self.h, C.byref(st)), "server_summary")
        return {k: getattr(st, k) for k, _ in native.ServerStats._fields_}

    def output(self, rid: int, name: str) -> np.ndarray:
        n = int(self.L.trt_server_output_size(self.h, int(rid), name.encode()))
        if n == 0:
            raise KeyError(f"request {rid} has no output {name}")
        out = np.empty(n, dtype=np.float64)
        self._check(self.L.trt_server_output(self.h, int(rid), name.encode(), out.ctypes.data_as(native._dp), n),
                    "server_output")
        return out


def percentile(xs, q: float) -> float:
    """q-th percentile (0..100), linear interpolation between order statistics."""
    a = np.sort(np.asarray(xs, dtype=np.float64))
    if a.size == 0:
        return math.nan
    return float(np.percentile(a, q))


def poisson_arrivals(rate: float, n: int, seed: int) -> np.ndarray:
    """n arrival times of a Poisson process of the given rate (requests/s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def run_load(rate: float, n: int, steps: int = 20, budget_gib: float = 200.0, precision: int = 1, tile=(4, 4),
             base_seed: int = 20240611, arrival_seed: int = 7, max_active: int = 64, problems=None,
             solves: bool = False) -> dict:
    """Serve n requests (problems of the batched distribution unless given)
    arriving as a Poisson process; returns latency and throughput figures."""
    arr = poisson_arrivals(rate, n, arrival_seed)
    probs = problems if problems is not None else [problem.cfg3_problem(base_seed, i, nsteps=steps) for i in range(n)]
    with Server(int(budget_gib * 2**30), max_active, precision, tile, solves=solves) as srv:
        ids = [srv.submit(p, steps, float(a)) for p, a in zip(probs, arr)]
        srv.run()
        reqs = [srv.request(i) for i in ids]
        summ = srv.summary()
    ok = [r for r in reqs if r["status"] == 0]
    lat = [r["done_s"] - r["arrival_s"] for r in ok]
    svc = [r["done_s"] - r["admit_s"] for r in ok]
    wait = [r["admit_s"] - r["arrival_s"] for r in ok]
    return {
        "rate_per_s": rate, "requests": n, "completed": len(ok), "failed": len(reqs) - len(ok),
        "steps_per_request": steps,
        "latency_p50_s": percentile(lat, 50), "latency_p99_s": percentile(lat, 99),
        "service_p50_s": percentile(svc, 50), "service_p99_s": percentile(svc, 99),
        "queue_wait_p50_s": percentile(wait, 50), "queue_wait_p99_s": percentile(wait, 99),
        "wall_s": summ["wall_s"], "ticks": summ["ticks"],
        "updates_per_s": summ["updates"] / summ["wall_s"] if summ["wall_s"] > 0 else math.nan,
        "sweep_updates_per_s": summ["updates"] / (summ["sweep_ms"] * 1e-3) if summ["sweep_ms"] > 0 else math.nan,
        "completed_per_s": len(ok) / summ["wall_s"] if summ["wall_s"] > 0 else math.nan,
        "peak_active": summ["peak_active"], "peak_device_gib": summ["peak_bytes"] / 2**30,
        "host_s": {k: summ[k] for k in ("admit_s", "rebuild_s", "step_s", "retire_s")},
        "budget_gib": budget_gib,
        "setup_s": summ["setup_s"],
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", type=float, nargs="+", default=[2.0, 4.0, 8.0])
    ap.add_argument("--requests", type=int, default=32)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--budget-gib", type=float, default=200.0)
    ap.add_argument("--precision", type=int, default=1)
    a = ap.parse_args(argv)
    tile = (4, 4) if a.precision else (2, 4)
    for r in a.rates:
        print(f"[serve] rate {r}/s, {a.requests} requests", file=sys.stderr, flush=True)
        print(json.dumps(run_load(r, a.requests, a.steps, a.budget_gib, a.precision, tile)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

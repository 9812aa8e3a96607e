#!/usr/bin/env python3
"""Sweep throughput benchmark (one GPU).

A step is one TRANSPORT_SCB sweep with the fused angular moments over the
256x256-cell, S16 (144 directions), 32-group problem, inputs resident in HBM.
The headline value is the FP32-sweep form (FP64 moment accumulation); the
plain FP64 form is measured in the same run and reported beside it.  Both run
the persistent tile engine (one kernel launch per sweep).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, AMD public specification
BASELINE_METRIC = "sweep cell\u00b7angle\u00b7group updates/s, FP32-batched vs plain FP64; p50/p99 solve latency"


def alg_bytes(prob, precision: int) -> int:
    """Algorithmic bytes of one sweep: read 4 old corners + write 4 new corners per
    (cell, direction, group), read KapE+1/(c dt) and the isotropic source per (cell, group),
    write 12 FP64 moments per (cell, group).  Precisions 1 and 2 store intensities in FP32."""
    s_i = 4 if precision in (1, 2) else 8
    nc = prob.nx * prob.ny
    return nc * prob.nm * prob.ng * 8 * s_i + nc * prob.ng * (2 * 8 + 12 * 8)


def measure(trt, prob, precision: int, steps: int, warmup: int, tune: bool = True):
    """The sweep of one problem in one precision form, with the tile shape the autotuner
    (trt_autotune_tile) finds fastest for it on this device."""
    table = {}
    tile = (0, 0)
    if tune:
        tile, _, table = trt.native.autotune_tile([prob], precision, sweeps=3)
    ctx = trt.native.Context(prob, precision=precision, tile=tile)
    try:
        if warmup > 0:
            ctx.sweep_timed(warmup)
        t0 = time.perf_counter()
        ms = ctx.sweep_timed(steps)                 # HIP events on the context stream
        host_s = time.perf_counter() - t0
        # per-launch duration of the dominant kernel: an event pair around each launch
        nsw = max(2, steps // 2)
        wall, kern, nl = ctx.sweep_kernel_timed(nsw)
        mode, tx, ty, grid = ctx.sweep_config()
    finally:
        ctx.close()
    upd = prob.updates_per_sweep()
    b = alg_bytes(prob, precision)
    return {
        "ms_per_step": ms / steps,
        "updates_per_s": upd * steps / (ms * 1e-3),
        "host_s": host_s,
        "kernel_ms_per_sweep": kern / nsw,
        "launch_avg_ms": kern / nl,
        "launches_per_sweep": nl // nsw,
        "alg_bytes_per_sweep": b,
        "achieved_gbs": b / (kern / nsw * 1e-3) / 1e9,
        "kernel": kernel_name(precision, mode, tx, ty),
        "sweep_mode": mode, "tile": [tx, ty], "grid": grid,
        "autotune_ms_per_sweep": {f"{t[0]}x{t[1]}": round(v, 3) for t, v in table.items()},
    }


def kernel_name(precision: int, mode: int, tx: int, ty: int) -> str:
    """Template name of the sweep kernel a context runs (as rocprofv3 reports it)."""
    ti = "double" if precision == 0 else "float"
    tc = "double" if precision != 1 else "float"
    form = {2: "sweep_pk", 3: "sweep_diag", 4: "sweep_seg"}.get(mode)
    return f"{form}<{ti}, {tc}, {tx}, {ty}>" if form else "sweep_kernel"


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of this sweep kernel (template name, tile shape included) from
    the committed rocprofv3 counter summary, if it holds that kernel."""
    path = os.path.join(ROOT, "profiles", "sweep_cfg1_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    for v in d.values():
        if isinstance(v, dict) and v.get("kernel") == kernel:
            return v.get("hbm_bytes_per_launch")
    return None


def batch_mix(trt, budget_gib: float, base_seed: int = 20240611):
    """The longest prefix of the batched distribution's draws (grid 64..512, 16..64 groups,
    S8..S16) whose FP64 corner intensities (I_crn and I_crn_old, 2 x 4 corners x cells x
    directions x groups x 8 bytes) fit budget_gib together.  The rule depends on the
    problems only, so the mix stays the same when the library's own buffers change."""
    probs, tot = [], 0
    for i in range(64):
        p = trt.problem.cfg3_problem(base_seed, i)
        b = 2 * 4 * p.nx * p.ny * p.nm * p.ng * 8
        if tot + b > budget_gib * 2**30:
            break
        probs.append(p)
        tot += b
    return probs


def measure_batch(trt, probs, precision: int, steps: int):
    """Batched persistent sweep of probs (tile shape and column-segment length chosen by
    the autotuner on these problems) vs the same problems swept one at a time."""
    tile, seg, _, table = trt.native.autotune_batch(probs, precision, sweeps=3)
    ctxs = [trt.native.Context(p, precision=precision, tile=tile) for p in probs]
    try:
        upd = sum(p.updates_per_sweep() for p in probs)
        alone_ms = 0.0
        balg = sum(alg_bytes(p, precision) for p in probs)
        for c in ctxs:
            c.sweep_timed(1)
            alone_ms += c.sweep_timed(2) / 2
        with trt.native.Batch(ctxs, graph=True, seg=seg) as b:
            b.sweep_timed(1)
            ms = b.sweep_timed(steps) / steps
            _, grid, nitems = b.info()
            seg = b.segment_length
    finally:
        for c in ctxs:
            c.close()
    return {"updates_per_sweep": upd, "ms_per_sweep": ms, "updates_per_s": upd / (ms * 1e-3),
            "one_at_a_time_updates_per_s": upd / (alone_ms * 1e-3), "tile": list(tile), "segment_length": seg,
            "grid": grid, "work_items": nitems, "achieved_alg_gbs": balg / (ms * 1e-3) / 1e9,
            "autotune_ms_per_sweep": {f"{t[0]}x{t[1]}/seg{sg}": round(v, 3) for (t, sg), v in table.items()}}


def measure_steps(trt, precision: int, steps: int = 10, qd_solver: int = 0):
    """Full MLQD time steps (the drop-in's whole iteration: sweeps, QD factors, MLOQD and
    GLOQD solves, temperature update) of the 256x256, S16, 32-group problem with uniform
    material: cfg1's lognormal(0, 1.5) map stops the reference algorithm in its first step.
    qd_solver 0 reuses the nested-dissection factors within a step, 3 factorizes every solve."""
    prob = trt.problem.mapped(nx=256, order=16, ng=32, sigma=0.0, nsteps=steps + 1)
    ctx = trt.native.Context(prob, precision=precision, qd_solver=qd_solver)
    try:
        ctx.step()                                  # warm-up (plans, handles, first factorisations)
        q0 = ctx.qd_stats()
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            its.append(list(ctx.step()))
        dt = (time.perf_counter() - t0) / steps
        q1 = ctx.qd_stats()
    finally:
        ctx.close()
    sweeps = sum(i[3] for i in its)
    return {"s_per_step": dt, "steps": steps, "iterations_per_step": its[-1], "qd_solver": qd_solver,
            "direct_solves_per_step": {"factorized": (q1[0] - q0[0]) / steps, "reused_factors": (q1[1] - q0[1]) / steps,
                                       "refinement_steps": (q1[2] - q0[2]) / steps},
            "end_to_end_updates_per_s": prob.updates_per_sweep() * sweeps / (dt * steps),
            "note": "iterations of the step: RT, MLOQD, grey Newton, sweeps; end-to-end updates/s = sweep updates "
                    "of the timed steps over their whole wall time (low-order solves included)"}


def draws(trt, n: int = 64, steps: int = 20, base_seed: int = 20240611):
    return [trt.problem.cfg3_problem(base_seed, i, nsteps=steps) for i in range(n)]


def measure_waves(trt, budget_gib: float = 200.0, steps: int = 20, n: int = 64):
    """The batched distribution in full: all 64 draws (64..512 grids, S8..S16, 16..64 groups,
    their own lognormal(0, 1.5) maps and boundary temperatures) submitted at once to the
    request loop, which admits them in waves as the HBM budget frees up and sweeps every
    admitted problem in one persistent launch per step (FP32 sweep, FP64 moments); 20
    Tgen_QDf steps each (sweep + moments + QD factors; the maps stop full MLQD steps)."""
    sv = trt.serve.run_load(1e9, n, steps=steps, budget_gib=budget_gib, precision=1, tile=(4, 4),
                            problems=draws(trt, n, steps))
    sv["note"] = (f"{n} draws of the batched distribution submitted at t=0 under a {budget_gib:g} GiB budget, "
                  f"{steps} Tgen_QDf steps each, FP32 sweep, 4x4 tiles; updates_per_s = all sweep updates over the "
                  "wall time (admission, context creation and output copies included)")
    return sv


def solve_draws(trt, n: int = 8, nmax: int = 64, base_seed: int = 20240611):
    """Draws of the batched distribution (in draw order) of at most nmax^2 cells, with
    uniform material: the distribution's lognormal(0, 1.5) maps stop full MLQD steps of the
    reference algorithm (negative group energies at the opacity jumps), and 128^2 S8 draws
    reach the cold-cell 0/0 of fg_Calc within three steps; the 64^2 draws run."""
    import numpy as np
    out, i = [], 0
    while len(out) < n and i < 64:
        p = trt.problem.cfg3_problem(base_seed, i)
        if p.nx <= nmax:
            out.append(p.with_(kmult=np.ones(len(p.kmult))))
        i += 1
    return out


def measure_batched_solves(trt, steps: int = 5, n: int = 8):
    """Full MLQD time steps of n draws (64^2, S8..S16, 16..64 groups, uniform material):
    batched (trt_batch_step: each round of RT iterations' sweeps in one launch, every
    member's low-order work on its own stream from its own host thread) against the same
    problems stepped one at a time, FP32 sweep; and one at a time in plain FP64."""
    probs = solve_draws(trt, n)
    upd = 0
    res = {"problems": [f"{p.nx}x{p.ny} S{p.order} G{p.ng}" for p in probs], "steps": steps}
    for prec, key in ((1, "fp32_one_at_a_time"), (0, "fp64_one_at_a_time")):
        t = 0.0
        upd = 0
        for p in probs:
            ctx = trt.native.Context(p, precision=prec, tile=(4, 4) if prec else (0, 0))
            try:
                ctx.step()                            # warm-up: plans and handles (not timed)
                t0 = time.perf_counter()
                for _ in range(steps):
                    upd += p.updates_per_sweep() * ctx.step()[3]
                t += time.perf_counter() - t0
            finally:
                ctx.close()
        res[key] = {"wall_s": t, "updates_per_s": upd / t, "solves_steps_per_s": len(probs) * steps / t}
    ctxs = [trt.native.Context(p, precision=1, tile=(4, 4)) for p in probs]
    try:
        with trt.native.Batch(ctxs) as b:
            b.step()                                  # warm-up
            t0 = time.perf_counter()
            upd = 0
            for _ in range(steps):
                its = b.step()
                upd += sum(p.updates_per_sweep() * i[3] for p, i in zip(probs, its))
            t = time.perf_counter() - t0
    finally:
        for c in ctxs:
            c.close()
    res["fp32_batched"] = {"wall_s": t, "updates_per_s": upd / t, "solves_steps_per_s": len(probs) * steps / t}
    res["batched_over_one_at_a_time_fp32"] = res["fp32_one_at_a_time"]["wall_s"] / t
    res["batched_fp32_over_plain_fp64"] = res["fp64_one_at_a_time"]["wall_s"] / t
    res["note"] = ("full time steps (sweeps, QD factors, MLOQD and grey Newton iterations); updates_per_s = sweep "
                   "updates over the wall time of the steps, low-order solves included; results of the batched "
                   "members are bitwise those of the same problems alone (tests/test_gpu_batch_steps.py)")
    return res


def measure_served_solves(trt, n: int = 8, steps: int = 5, budget_gib: float = 64.0):
    """The request loop in solve mode: n requests (the draws of measure_batched_solves),
    each `steps` full time steps, at Poisson rates that are fractions of the completion rate
    of the all-at-once run, FP32 sweep."""
    probs = solve_draws(trt, n)
    out = {"rates": []}
    allat = trt.serve.run_load(1e9, n, steps=steps, budget_gib=budget_gib, problems=probs, solves=True)
    out["all_at_once"] = allat
    sat = allat["completed_per_s"]
    for f in (0.25, 0.5, 1.0):
        r = round(f * sat, 3)
        print(f"[bench] served solves, {n} requests at {r:g}/s", file=sys.stderr, flush=True)
        out["rates"].append(trt.serve.run_load(r, n, steps=steps, budget_gib=budget_gib, problems=probs,
                                               solves=True))
    out["note"] = (f"requests = {steps} full MLQD time steps from the initial state (solve mode), the draws of "
                   "batched_solves, Poisson arrivals at 0.25, 0.5 and 1.0 of the all-at-once completion rate, "
                   f"HBM budget {budget_gib:g} GiB, FP32 sweep; latency = completion - arrival")
    return out


def measure_serving(trt, rates, n: int, steps: int = 20, budget_gib: float = 200.0):
    """Poisson arrivals at each rate (requests/s) of requests drawn from the batched
    distribution, FP32 continuous batching; then the same requests solved one at a time in
    plain FP64 (a context per request, its steps run back to back), the reference point."""
    out = {"rates": []}
    for r in rates:
        print(f"[bench] request loop, {n} requests at {r:g}/s", file=sys.stderr, flush=True)
        sv = trt.serve.run_load(r, n, steps=steps, budget_gib=budget_gib, precision=1, tile=(4, 4),
                                problems=draws(trt, n, steps))
        out["rates"].append(sv)
    print(f"[bench] {n} requests one at a time, plain FP64", file=sys.stderr, flush=True)
    lat, upd = [], 0
    t_all = time.perf_counter()
    for p in draws(trt, n, steps):
        t0 = time.perf_counter()
        ctx = trt.native.Context(p, precision=0)
        try:
            for _ in range(steps):
                ctx.tgen_step()
            ctx.pull("Temp")                        # the request's result on the host
        finally:
            ctx.close()
        lat.append(time.perf_counter() - t0)
        upd += p.updates_per_sweep() * steps
    wall = time.perf_counter() - t_all
    out["fp64_one_at_a_time"] = {"requests": n, "wall_s": wall, "updates_per_s": upd / wall,
                                 "service_p50_s": trt.serve.percentile(lat, 50),
                                 "service_p99_s": trt.serve.percentile(lat, 99),
                                 "note": "each request alone: context creation, 20 FP64 Tgen_QDf steps, result copy; "
                                         "latency under arrivals at rate r would add the queueing of a single "
                                         "server (this is its service time)"}
    return out


def measure_dropin_sweep(trt, prob, reps: int = 2):
    """TRANSPORT_SCB through the host-array boundary, as the Fortran drop-in shim calls it
    (fortran/trt_dropin.f90): push KapE, RT_Src and I_crn_old, sweep, pull I_crn and the
    five intensity views.  FP64; the host buffers are allocated once, as a caller's arrays
    are, and hold the context's own values."""
    import ctypes as C
    import numpy as np
    ins = ("KapE", "RT_Src", "I_crn_old")
    outs = ("I_crn", "I_avg", "I_edgV", "I_edgH", "Ic_edgV", "Ic_edgH")
    ctx = trt.native.Context(prob, precision=0)
    try:
        L, dp = ctx.L, C.POINTER(C.c_double)
        bufs = {n: ctx.pull(n) for n in ins}
        bufs.update({n: np.zeros(ctx.size(n)) for n in outs})

        def call():
            for n in ins:
                ctx._check(L.trt_push(ctx.h, n.encode(), bufs[n].ctypes.data_as(dp), bufs[n].size), "push " + n)
            ctx._check(L.trt_transport_scb(ctx.h), "transport_scb")
            for n in outs:
                ctx._check(L.trt_pull(ctx.h, n.encode(), bufs[n].ctypes.data_as(dp), bufs[n].size), "pull " + n)

        call()                                      # first touch of the host buffers
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        dt = (time.perf_counter() - t0) / reps
        moved = 8 * sum(bufs[n].size for n in ins + outs)
    finally:
        ctx.close()
    return {"s_per_sweep": dt, "updates_per_s": prob.updates_per_sweep() / dt, "host_bytes_per_sweep": moved,
            "host_gbs": moved / dt / 1e9,
            "note": "per call: push KapE, RT_Src (per cell and group), I_crn_old; one FP64 sweep; pull I_crn, "
                    "I_avg, I_edgV, I_edgH, Ic_edgV, Ic_edgH into pageable host arrays"}


def measure_pod(trt, prob, rank: int = 32, nt: int = 10, reps: int = 20):
    """POD reconstruction (POD_RECONSTRUCT) of one QD-factor field of the ROM runs'
    step at the problem's size: fg_avg_xx (N_x N_y N_g values), a per-group basis of
    the given rank; synthetic basis.  Bytes per reconstruction: U + C read, the field
    written."""
    import numpy as np
    ctx = trt.native.Context(prob, precision=0)
    pod = None
    try:
        n = ctx.size("fg_avg_xx_next")
        rng = np.random.default_rng(3)
        ng = prob.ng
        U = rng.standard_normal(n * rank)
        C = rng.uniform(0.2, 0.4, n)
        S = np.sort(rng.uniform(0.0, 1.0, ng * rank))[::-1].copy()
        V = rng.standard_normal(ng * rank * nt)
        pod = trt.native.PodBasis(ctx, "fg_avg_xx_next", C, S, U, V, np.full(ng, rank), nt, 0)
        pod.reconstruct(1)
        ms = pod.timed(reps)
        b = pod.basis_bytes + 2 * 8 * n
    finally:
        if pod is not None:
            pod.close()
        ctx.close()
    gbs = b / (ms * 1e-3) / 1e9
    return {"field": f"fg_avg_xx ({prob.nx}x{prob.ny}x{prob.ng})", "rank_per_group": rank, "ms": ms,
            "bytes": b, "gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}


def measure_restart(trt):
    """Checkpoint and resume (RESTART_OUT / RESTART_IN records) of the 128x128, S16,
    32-group FP64 state after one time step: 4.8 GB in 25 records, I_crn and RT_Src
    (2.4 GB each) split into subrecords.  The file goes to a temporary directory
    (TMPDIR) and is deleted afterwards; the resumed context's arrays are compared
    with the writer's."""
    import numpy as np
    prob = trt.problem.mapped(nx=128, order=16, ng=32, sigma=0.0, nsteps=2)
    d = tempfile.mkdtemp(prefix="trt_restart_")
    f = os.path.join(d, "state.rst")
    a = b = None
    try:
        a = trt.native.Context(prob)
        a.step()
        t0 = time.perf_counter()
        a.restart_out(f)
        tw = time.perf_counter() - t0
        size = os.path.getsize(f)
        b = trt.native.Context(prob)
        t0 = time.perf_counter()
        b.restart_in(f)
        tr = time.perf_counter() - t0
        same = bool(all(np.array_equal(a.pull(n), b.pull(n)) for n in ("I_crn", "Temp", "Eg_avg", "KapE", "Cg_T")))
    finally:
        for x in (a, b):
            if x is not None:
                x.close()
        shutil.rmtree(d, ignore_errors=True)
    return {"problem": "128x128 cells, S16, 32 groups, FP64 context after one time step", "file_bytes": size,
            "write_s": tw, "read_s": tr, "write_gbs": size / tw / 1e9, "read_gbs": size / tr / 1e9,
            "resumed_equal": same,
            "note": "write = device-to-host pulls + reference layout + buffered file writes (no fsync); read = "
                    "whole-file check, then host-to-device pushes; the read follows the write, so it is served "
                    "from the page cache"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(trt, budget_s: float = 12.0):
    """The FP64 CPU restatement (oracle) of TRANSPORT_SCB + COLLAPSE_INTENSITIES on the same
    256x256 S16 grid, OpenMP over groups as the reference, one group per thread."""
    import oracle
    oracle.build()
    cores = os.cpu_count() or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else cores
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = max(1, min(avail, int(omp) if omp else avail, 32))
    prob = trt.problem.cfg1()
    ng = threads
    prob = prob.with_(ng=ng, nu=prob.nu[:ng + 1])
    o = oracle.Oracle(prob, threads=threads)
    n = 0
    t_total = 0.0
    while n < 200 and (n == 0 or t_total < budget_s):
        t0 = time.perf_counter()
        o.call("transport_scb")
        o.call("collapse_intensities")
        t_total += time.perf_counter() - t0
        n += 1
    del o
    upd = prob.updates_per_sweep() * n
    return {
        "value": upd / t_total,
        "unit": "cell*angle*group updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{ng} of the 32 groups: {n} x (TRANSPORT_SCB + COLLAPSE_INTENSITIES) at 256x256 cells, S16, "
                   f"FP64, one group per OpenMP thread as in the reference, gcc -O2 -fopenmp (the reference's "
                   f"-fcheck=all omitted); {t_total:.1f} s"),
        "cpu_model": cpu_model(), "nproc": cores, "cpus_available": avail, "omp_num_threads": omp,
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64", "fp32s"], default=None,
                    help="measure one form only (profiling runs)")
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-problem batch")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-loop load")
    ap.add_argument("--no-steps", action="store_true", help="skip the full MLQD time steps")
    ap.add_argument("--no-restart", action="store_true", help="skip the checkpoint/resume measurement")
    ap.add_argument("--no-dropin", action="store_true", help="skip the host-array (PCIe-inclusive) sweep")
    ap.add_argument("--no-pod", action="store_true", help="skip the POD reconstruction measurement")
    ap.add_argument("--batch-gib", type=float, default=96.0,
                    help="budget of the batch mix's FP64 corner intensities (I_crn and I_crn_old)")
    ap.add_argument("--serve-requests", type=int, default=24)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path is single-GPU; "
              "it would shard by energy group, one all-reduce of the grey sums per MLOQD iteration)", file=sys.stderr)
        return 2
    trt = importlib.import_module("2d_trt_rom_system_amd")
    prob = trt.problem.cfg1()
    res = {}
    # fp32: FP32 intensities and arithmetic; fp64: plain FP64; fp32s: FP32 storage, FP64 arithmetic
    for prec, key in ((1, "fp32"), (0, "fp64"), (2, "fp32s")):
        if args.only and args.only != key:
            continue
        print(f"[bench] {key} sweep, {args.warmup} warmup + {args.steps} timed", file=sys.stderr, flush=True)
        res[key] = measure(trt, prob, prec, args.steps, args.warmup)
    head_key = "fp32" if "fp32" in res else ("fp64" if "fp64" in res else "fp32s")
    head = res[head_key]
    prec = {"fp32": 1, "fp64": 0, "fp32s": 2}[head_key]
    out = {
        "metric": BASELINE_METRIC,
        "value": head["updates_per_s"],
        "unit": "cell*angle*group updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "fp32 (FP64 moment accumulation)", "fp64": "fp64",
                  "fp32s": "fp32 storage, fp64 arithmetic"}[head_key],
        "data": "synthetic",
        "config": {
            "workload": "TRANSPORT_SCB sweep + fused angular moments; 256x256 cells, S16 (144 directions), "
                        "32 log-spaced groups 1e-3..1e2 keV, 8x8 lognormal(0,1.5) opacity map",
            "cells": prob.nx * prob.ny, "directions": prob.nm, "groups": prob.ng,
            "updates_per_step": prob.updates_per_sweep(),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": head["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(head["kernel"]),
            "kernel": head["kernel"],
            "alg_bytes_per_launch": head["alg_bytes_per_sweep"] / head["launches_per_sweep"],
            "avg_launch_ms": head["launch_avg_ms"],
        },
    }
    tr = out["roofline"]["traffic"]
    if tr:
        # HBM bytes actually moved per launch (calibrated counters) over the live launch time
        out["roofline"]["traffic_gbs"] = tr / (head["launch_avg_ms"] * 1e-3) / 1e9
        out["roofline"]["traffic_frac"] = out["roofline"]["traffic_gbs"] / HBM_PEAK_GBS
    for k, v in res.items():
        out[k] = {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
    if "fp32" in res and "fp64" in res:
        out["fp32_over_fp64"] = res["fp32"]["updates_per_s"] / res["fp64"]["updates_per_s"]
    for k in res:
        out[k]["traffic_bytes_per_launch"] = traffic_from_profiles(res[k]["kernel"])
    if not args.no_batch and not args.only:
        probs = batch_mix(trt, args.batch_gib)
        print(f"[bench] batch of {len(probs)} problems", file=sys.stderr, flush=True)
        bt = {"problems": len(probs),
              "mix": [f"{p.nx}x{p.ny} S{p.order} G{p.ng}" for p in probs],
              "note": "the longest prefix of the batched distribution's draws (grid 64..512, 16..64 groups, "
                      f"S8..S16) whose FP64 I_crn and I_crn_old fit {args.batch_gib:g} GiB; one persistent launch "
                      "per sweep, HIP-graph replay; tile shape and column-segment length from the autotuner "
                      "(trt_autotune_batch)"}
        for prec, key in ((1, "fp32"), (0, "fp64")):
            bt[key] = measure_batch(trt, probs, prec, max(2, args.steps // 2))
        bt["fp32_over_fp64"] = bt["fp32"]["updates_per_s"] / bt["fp64"]["updates_per_s"]
        out["batched"] = bt
        if "fp64" in res:
            # the batched FP32 sweep set against the plain FP64 form (one 256x256 problem per launch)
            out["fp32_batched_over_plain_fp64"] = bt["fp32"]["updates_per_s"] / res["fp64"]["updates_per_s"]
    sat = None
    if not args.no_batch and not args.only:
        print("[bench] 64 draws of the batched distribution in waves under the HBM budget", file=sys.stderr,
              flush=True)
        wv = measure_waves(trt)
        out["batched_64_waves"] = wv
        sat = wv["completed_per_s"]
    if not args.no_serve and not args.only:
        # five arrival rates up to saturation: fractions of the completion rate of the
        # all-at-once run above (or fixed rates without it)
        rates = [round(f * sat, 3) for f in (0.25, 0.5, 0.75, 1.0, 1.5)] if sat else [2.0, 4.0, 8.0, 12.0, 16.0]
        sv = measure_serving(trt, rates, args.serve_requests)
        sv["note"] = ("requests drawn from the batched distribution (the first draws, in order), 20 Tgen_QDf steps "
                      "each (one sweep per step), Poisson arrivals, HBM budget 200 GiB, FP32 sweep, continuous "
                      "batching; latency = completion - arrival; rates are fractions 0.25..1.5 of the completion "
                      "rate of the all-at-once run")
        out["serving"] = sv
    if not args.no_steps and not args.no_batch and not args.only:
        print("[bench] batched full time steps (solves)", file=sys.stderr, flush=True)
        out["batched_solves"] = measure_batched_solves(trt)
    if not args.no_steps and not args.no_serve and not args.only:
        print("[bench] served solves", file=sys.stderr, flush=True)
        out["served_solves"] = measure_served_solves(trt)
    if not args.no_steps and not args.only:
        print("[bench] full MLQD time steps (256x256, uniform material)", file=sys.stderr, flush=True)
        out["full_step"] = {"problem": "256x256 cells, S16, 32 groups 1e-3..1e2 keV, uniform material, "
                                       "dt 2e-3 ns, QD systems by nested dissection",
                            "fp64": measure_steps(trt, 0), "fp32": measure_steps(trt, 1),
                            "fp64_factorizing_every_solve": measure_steps(trt, 0, qd_solver=3)}
    if not args.no_dropin and not args.only:
        print("[bench] drop-in TRANSPORT_SCB with host arrays (PCIe-inclusive)", file=sys.stderr, flush=True)
        out["dropin_sweep"] = measure_dropin_sweep(trt, prob)
    if not args.no_pod and not args.only:
        print("[bench] POD reconstruction (ROM runs)", file=sys.stderr, flush=True)
        out["pod_reconstruct"] = measure_pod(trt, prob)
    if not args.no_restart and not args.only:
        print("[bench] checkpoint and resume (128x128, S16, 32 groups)", file=sys.stderr, flush=True)
        try:
            out["restart"] = measure_restart(trt)
        except (trt.native.TrtError, OSError) as e:     # e.g. a temporary directory too small for 4.8 GB
            out["restart"] = {"error": str(e)}
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(trt)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

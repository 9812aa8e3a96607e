"""FP32-sweep accuracy against the FP64 oracle over full time steps.

Runs the same problem four ways -- the FP64 oracle (CPU), and on the GPU the
FP64 sweep (fp64), the FP32 sweep (fp32: FP32 intensities and corner-solve
arithmetic) and FP32 storage with FP64 arithmetic (fp32s); all with FP64
moments and FP64 low-order solves -- for a number of MLQD time steps, and
reports after every step the relative L2 error of E (grey, cell averages), T
and the Eddington tensor components f_xx, f_yy, f_xy of each GPU form against
the oracle, plus the f errors weighted by the group energy density:

    python tests/precision_report.py --problem map48 --steps 10

Problems too large for the oracle compare the FP32 forms with the FP64 GPU run
(--vs-fp64 EVERY).  Long runs split the work: the oracle's trajectory (T and E
after every step) on the CPU, the GPU forms' trajectories on the GPU, then the
comparison:

    python tests/precision_report.py --problem s16g32n64u --steps 40 --traj oracle --out oracle.npz
    python tests/precision_report.py --problem s16g32n64u --steps 40 --traj gpu --out gpu.npz
    python tests/precision_report.py --compare oracle.npz gpu.npz

This is test infrastructure: the oracle is the checker here, never the thing
measured.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
native = importlib.import_module("2d_trt_rom_system_amd.native")
problem = importlib.import_module("2d_trt_rom_system_amd.problem")

QUANTITIES = {"E": "E_avg", "T": "Temp", "f_xx": "fg_avg_xx", "f_yy": "fg_avg_yy", "f_xy": "fg_avg_xy"}


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (d if d > 0 else 1.0)


def build_problem(name: str, tight: bool = False):
    if name == "cfg0":
        p = problem.cfg0()
    elif name == "cfg1":
        p = problem.cfg1()
    elif name in ("cfg1map1", "cfg1map0"):
        # the 256x256, S16, 32-group problem with a lognormal(0, 1) map, or uniform material
        p = problem.mapped(nx=256, order=16, ng=32, sigma=1.0 if name == "cfg1map1" else 0.0, nsteps=50)
    elif name.startswith("s16g32n"):
        # the 256x256 problem's physics (S16, 32 groups, lognormal(0, sigma) map) on an
        # n x n grid: "s16g32n128" (sigma 1) or "s16g32n128u" (uniform material)
        n = int(name[7:].rstrip("u"))
        p = problem.mapped(nx=n, order=16, ng=32, sigma=0.0 if name.endswith("u") else 1.0, nsteps=50)
    elif name.startswith("map"):
        # S8 and 16 groups keep the oracle's band solves to seconds per step
        p = problem.mapped(nx=int(name[3:]), order=8, ng=16, sigma=1.0)
    else:
        raise ValueError(name)
    if tight:
        # iterations converged far below the sweep's FP32 rounding, so that the
        # FP32 and FP64 runs are compared at (nearly) the same fixed point rather
        # than at iterates that stopped at different distances from it
        p.conv_ho, p.conv_lo, p.conv_gr1, p.conv_gr2 = 1e-11, 1e-12, 1e-13, 1e-13
    return p


def weighted_rel(a, b, w) -> float:
    """sum w |a - b| / sum w |b|: errors of a ratio weighted by the energy it carries."""
    a, b, w = (np.asarray(x, dtype=np.float64).ravel() for x in (a, b, w))
    den = float(np.sum(np.abs(w * b)))
    return float(np.sum(np.abs(w * (a - b)))) / (den if den > 0 else 1.0)


def run(name: str, steps: int, threads: int = 8, tight: bool = False) -> dict:
    import oracle
    prob = build_problem(name, tight)
    o = oracle.Oracle(prob, threads=threads)
    forms = {"fp64": native.Context(prob, precision=0), "fp32": native.Context(prob, precision=1),
             "fp32s": native.Context(prob, precision=2)}
    rows = []
    try:
        for s in range(1, steps + 1):
            t0 = time.perf_counter()
            o.step()
            t1 = time.perf_counter()
            row = {"step": s, "oracle_s": t1 - t0}
            for k, g in forms.items():
                t2 = time.perf_counter()
                g.step()
                row[f"{k}_gpu_s"] = time.perf_counter() - t2
                print(f"[precision] step {s} {k} done", file=sys.stderr, flush=True)
            eg = np.abs(o["HO_Eg_avg"].ravel(order="F"))
            for k, g in forms.items():
                for q, arr in QUANTITIES.items():
                    row[f"{k}_{q}"] = rel_l2(g.pull(arr), o[arr].ravel(order="F"))
                # f weighted by the group energy density: cells and groups where E_g
                # is near zero (or changes sign) make f = H/E meaningless there
                for q in ("f_xx", "f_yy", "f_xy"):
                    arr = QUANTITIES[q]
                    row[f"{k}_{q}_Ew"] = weighted_rel(g.pull(arr), o[arr].ravel(order="F"), eg)
            rows.append(row)
            print(json.dumps(row), file=sys.stderr, flush=True)
    finally:
        for g in forms.values():
            g.close()
    keys = [k for k in rows[0] if k.startswith("fp") and not k.endswith("gpu_s")]
    worst = {k: max(r[k] for r in rows) for k in keys}
    return {"problem": name, "nx": prob.nx, "ny": prob.ny, "order": prob.order, "ng": prob.ng, "steps": steps,
            "tight": tight, "conv": [prob.conv_ho, prob.conv_lo, prob.conv_gr1, prob.conv_gr2], "rows": rows,
            "worst": worst}


def run_vs_fp64(name: str, steps: int, every: int) -> dict:
    """Problems too large for the oracle: the FP32 forms against the FP64 GPU run
    (which is itself pinned to the oracle on smaller problems), every `every` steps."""
    prob = build_problem(name)
    forms = {"fp64": native.Context(prob, precision=0), "fp32": native.Context(prob, precision=1),
             "fp32s": native.Context(prob, precision=2)}
    rows = []
    failed = None
    try:
        for s in range(1, steps + 1):
            row = {"step": s}
            for k, g in forms.items():
                t0 = time.perf_counter()
                try:
                    g.step()
                except native.TrtError as e:
                    failed = {"step": s, "form": k, "error": str(e)}
                    break
                row[f"{k}_gpu_s"] = time.perf_counter() - t0
            if failed:
                print(f"[precision] step {s}: {failed['form']} failed: {failed['error']}", file=sys.stderr, flush=True)
                break
            print(f"[precision] step {s} " + " ".join(f"{k} {row[k + '_gpu_s']:.2f}s" for k in forms),
                  file=sys.stderr, flush=True)
            for k, g in forms.items():
                # cells and groups whose radiation energy density is not positive (negative
                # intensities of the SCB sweep) make the Eddington factors H/E unbounded
                row[f"{k}_Eg_nonpos"] = int(np.count_nonzero(g.pull("HO_Eg_avg") <= 0.0))
            if s % every == 0 or s == steps:
                ref = {q: forms["fp64"].pull(arr) for q, arr in QUANTITIES.items()}
                eg = np.abs(forms["fp64"].pull("HO_Eg_avg"))
                for k in ("fp32", "fp32s"):
                    for q, arr in QUANTITIES.items():
                        a = forms[k].pull(arr)
                        row[f"{k}_{q}"] = rel_l2(a, ref[q])
                        row[f"{k}_{q}_max"] = float(np.max(np.abs(a - ref[q])) / max(np.max(np.abs(ref[q])), 1e-300))
                        if q.startswith("f_"):
                            # weighted by the group energy density, as against the oracle
                            row[f"{k}_{q}_Ew"] = weighted_rel(a, ref[q], eg)
                print(json.dumps(row), file=sys.stderr, flush=True)
            rows.append(row)
    finally:
        for g in forms.values():
            g.close()
    return {"problem": name, "nx": prob.nx, "ny": prob.ny, "order": prob.order, "ng": prob.ng, "steps": steps,
            "every": every, "reference": "fp64 GPU run", "failed": failed, "rows": rows}


FORMS = (("fp64", 0), ("fp32", 1), ("fp32s", 2))


def trajectory(name: str, steps: int, which: str, out: str, threads: int = 8) -> None:
    """T and E (grey cell averages) after every step: of the oracle (which =
    "oracle", CPU) or of the three GPU forms (which = "gpu"), saved to `out`."""
    prob = build_problem(name)
    save = {}
    if which == "oracle":
        import oracle
        o = oracle.Oracle(prob, threads=threads)
        for s in range(1, steps + 1):
            t0 = time.perf_counter()
            o.step()
            save[f"oracle_T{s}"] = np.array(o["Temp"]).ravel(order="F")
            save[f"oracle_E{s}"] = np.array(o["E_avg"]).ravel(order="F")
            print(f"[precision] oracle step {s} {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
            np.savez(out, **save)
        return
    forms = {k: native.Context(prob, precision=pr) for k, pr in FORMS}
    try:
        for s in range(1, steps + 1):
            for k, g in forms.items():
                g.step()
                save[f"{k}_T{s}"] = g.pull("Temp")
                save[f"{k}_E{s}"] = g.pull("E_avg")
            print(f"[precision] gpu step {s}", file=sys.stderr, flush=True)
    finally:
        for g in forms.values():
            g.close()
    np.savez(out, **save)


def compare(oracle_npz: str, gpu_npz: str) -> dict:
    """Relative L2 and max errors of E and T of each GPU form against the oracle, per step."""
    o, g = np.load(oracle_npz), np.load(gpu_npz)
    rows = []
    s = 1
    while f"oracle_T{s}" in o and f"fp64_T{s}" in g:
        row = {"step": s}
        for k, _ in FORMS:
            for q in ("E", "T"):
                a, b = g[f"{k}_{q}{s}"], o[f"oracle_{q}{s}"]
                row[f"{k}_{q}"] = rel_l2(a, b)
                row[f"{k}_{q}_max"] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        rows.append(row)
        s += 1
    worst = {k: max(r[k] for r in rows) for k in rows[0] if k != "step"} if rows else {}
    return {"oracle": oracle_npz, "gpu": gpu_npz, "steps": len(rows), "rows": rows, "worst": worst}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problem", default="map48")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--tight", action="store_true", help="converge the iterations to 1e-11..1e-13")
    ap.add_argument("--vs-fp64", type=int, default=0, metavar="EVERY",
                    help="compare with the FP64 GPU run every EVERY steps instead of the oracle (large problems)")
    ap.add_argument("--traj", choices=("oracle", "gpu"), help="save the trajectory of the oracle or the GPU forms")
    ap.add_argument("--out", default="trajectory.npz")
    ap.add_argument("--compare", nargs=2, metavar=("ORACLE_NPZ", "GPU_NPZ"))
    a = ap.parse_args(argv)
    if a.compare:
        print(json.dumps(compare(*a.compare)))
    elif a.traj:
        trajectory(a.problem, a.steps, a.traj, a.out, a.threads)
    elif a.vs_fp64:
        print(json.dumps(run_vs_fp64(a.problem, a.steps, a.vs_fp64)))
    else:
        print(json.dumps(run(a.problem, a.steps, a.threads, a.tight)))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Driver for counter profiles of the sweep (run under rocprofv3).

    python3 2d_trt_rom_system_amd/prof_sweep.py --precision 1 --tile 4 4 --sweeps 2

runs, on the 256x256-cell, S16, 32-group problem: one plain device copy of
--copy-mib MiB (calib_copy_kernel, the reference for byte counters: it reads
and writes exactly that many bytes), then one warm-up sweep and --sweeps timed
sweeps.  summarize() turns the rocprofv3 counter CSVs of a FETCH_SIZE pass and a
WRITE_SIZE pass into HBM bytes per sweep-kernel launch, each counter scaled by
what it reads for the copy kernel.
"""
from __future__ import annotations

import argparse
import csv
import importlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def run(precision: int, sweeps: int, copy_mib: int, tile=(0, 0), mode: int = 0) -> None:
    trt = importlib.import_module("2d_trt_rom_system_amd")
    trt.native.calib_copy(copy_mib << 20, 1)
    prob = trt.problem.cfg1()
    ctx = trt.native.Context(prob, precision=precision, tile=tile, sweep_mode=mode)
    ctx.sweep_timed(1)
    ms = ctx.sweep_timed(sweeps)
    print(f"precision={precision} config={ctx.sweep_config()} {ms / sweeps:.3f} ms/sweep", flush=True)
    ctx.close()


def run_steps(precision: int, steps: int) -> None:
    """Full MLQD time steps of the 256x256, S16, 32-group problem with uniform
    material (cfg1's multiplier map drives the grey Newton iteration to NaN):
    one warm-up step, then `steps` timed ones."""
    import time
    trt = importlib.import_module("2d_trt_rom_system_amd")
    prob = trt.problem.mapped(nx=256, order=16, ng=32, sigma=0.0, nsteps=steps + 1)
    ctx = trt.native.Context(prob, precision=precision)
    ctx.step()
    for k in range(steps):
        t0 = time.perf_counter()
        its = ctx.step()
        print(f"step {k + 2}: {time.perf_counter() - t0:.3f} s, iterations {its}", flush=True)
    ctx.close()


def _per_kernel(path: str, counter: str):
    vals = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return vals


def summarize(fetch_csv: str, write_csv: str, copy_mib: int, kernel_prefix: str) -> dict:
    """HBM bytes per launch of the kernel whose name starts with kernel_prefix."""
    nbytes = float(copy_mib << 20)
    out = {}
    for name, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        v = _per_kernel(path, counter)
        cal = [x for k, xs in v.items() if "calib_copy" in k for x in xs]
        sw = [x for k, xs in v.items() if k.startswith(kernel_prefix) or kernel_prefix in k for x in xs]
        if not cal or not sw:
            raise RuntimeError(f"{counter}: no calibration or sweep dispatches in {path}")
        scale = nbytes / (cal[0] * 1024.0)          # true bytes per counted byte
        # the first sweep dispatch is the warm-up; average the rest
        per = sw[1:] if len(sw) > 1 else sw
        out[name] = {"counter_kib_per_launch": sum(per) / len(per), "copy_counter_kib": cal[0],
                     "copy_true_bytes": nbytes, "scale": scale,
                     "bytes_per_launch": sum(per) / len(per) * 1024.0 * scale}
    out["hbm_bytes_per_launch"] = out["fetch"]["bytes_per_launch"] + out["write"]["bytes_per_launch"]
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--tile", type=int, nargs=2, default=(0, 0), help="tile shape (0 0: the default)")
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--steps", type=int, default=0, help="full time steps instead of sweeps (uniform material)")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--summarize", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "KERNEL"))
    a = ap.parse_args()
    if a.summarize:
        print(json.dumps(summarize(a.summarize[0], a.summarize[1], a.copy_mib, a.summarize[2]), indent=1))
    elif a.steps:
        run_steps(a.precision, a.steps)
    else:
        run(a.precision, a.sweeps, a.copy_mib, tuple(a.tile))

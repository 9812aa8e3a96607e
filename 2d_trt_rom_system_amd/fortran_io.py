"""Input/output files of the Fortran host program trt_mlqd (fortran/trt_mlqd_main.f90)."""
from __future__ import annotations

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "trt_mlqd")

DROPIN = 1      # reference call sequence over host arrays through the drop-in modules
RESIDENT = 2    # device-resident trt_step


def write_problem(path: str, prob, mode: int = DROPIN, precision: int = 0) -> None:
    ints = [prob.nx, prob.ny, prob.nm, prob.ng, prob.nmat, prob.nsteps, prob.opacity_model, prob.maxit_rte,
            prob.maxit_mloqd, prob.maxit_gloqd, prob.conv_type, prob.use_line_search, prob.kapE_dT_flag,
            *[int(b) for b in prob.bc_type], precision, mode]
    dbl = [prob.c, prob.dt, prob.theta, prob.comp_unit, prob.cv, prob.kap0, prob.tini, *[float(b) for b in prob.bcT],
           prob.conv_ho, prob.conv_lo, prob.conv_gr1, prob.conv_gr2, prob.chi, prob.line_src]
    with open(path, "wb") as f:
        f.write(np.asarray(ints, dtype=np.int32).tobytes())
        f.write(np.asarray(dbl, dtype=np.float64).tobytes())
        for a in (prob.dx, prob.dy, prob.ox, prob.oy, prob.w, prob.nu, prob.kmult):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write((np.asfortranarray(prob.mat).ravel(order="F") + 1).astype(np.int32).tobytes())


def read_results(path: str, prob):
    """List of per-step dicts: RT_Its, MGQD_Its and the arrays written by trt_mlqd."""
    nx, ny, ng = prob.nx, prob.ny, prob.ng
    cell, cellg = nx * ny, nx * ny * ng
    raw = open(path, "rb").read()
    rec = 8 + 8 * (4 * cell + 3 * cellg)
    out = []
    for k in range(len(raw) // rec):
        b = raw[k * rec:(k + 1) * rec]
        its = np.frombuffer(b[:8], dtype=np.int32)
        v = np.frombuffer(b[8:], dtype=np.float64)
        d = {"RT_Its": int(its[0]), "MGQD_Its": int(its[1])}
        off = 0
        for name, n, shape in (("Temp", cell, (nx, ny)), ("E_avg", cell, (nx, ny)), ("HO_E_avg", cell, (nx, ny)),
                               ("MGQD_E_avg", cell, (nx, ny)), ("fg_avg_xx", cellg, (nx, ny, ng)),
                               ("fg_avg_yy", cellg, (nx, ny, ng)), ("fg_avg_xy", cellg, (nx, ny, ng))):
            d[name] = v[off:off + n].reshape(shape, order="F")
            off += n
        out.append(d)
    return out


def run(prob, workdir: str, mode: int = DROPIN, precision: int = 0, timeout: float = 600.0,
        restart_in: str | None = None, restart_out: str | None = None, restart_freq: int = 0,
        outfile: str | None = None, out_flags: str | None = None):
    """Run the Fortran host program on prob; returns (per-step results, stdout).  In mode
    RESIDENT the state can be read from a restart file before the time loop (restart_in)
    and written to one every restart_freq-th step (restart_out), and the reference's NetCDF
    output file written to outfile (out_flags: 18 digits 0/1, the OUTPUT_OPTS flags)."""
    os.makedirs(workdir, exist_ok=True)
    pin = os.path.join(workdir, "problem.bin")
    pout = os.path.join(workdir, "results.bin")
    write_problem(pin, prob, mode, precision)
    args = [EXE, pin, pout]
    if restart_in or restart_freq > 0 or outfile:
        args.append(restart_in or "-")
    if restart_freq > 0 or outfile:
        args += [restart_out or "-", str(int(restart_freq))]
    if outfile:
        args.append(outfile)
        if out_flags:
            args.append(out_flags)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"trt_mlqd failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return read_results(pout, prob), r.stdout

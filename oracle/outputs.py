"""The output file's derived quantities restated in Python (test infrastructure only).

XWAVE_SPEED_edg and XWAVE_SPEED_cnt (MainSrc/WAVEPROP_TOOLS.f08:44-125) with YMEAN
(:14-27): the speed of a front moving in x, sum_i |d(ymean x)/dt| / sum_i |d(ymean x)/dx|,
written by TRT_MLQD_ALGORITHM after every time step (MainSrc/ALGORITHMS.f08:859-864) as
E_xWave_Speed (from E_avg, E_edgV, E_avg_old) and Temp_xWave_Speed (from Temp,
Temp_old).  YMEAN reads its argument as a flat array with row stride N_x, also the
N_x+1-long rows of E_edgV, as the reference does.

The mlqd output time axis (MainSrc/INPUTS.f08:120-151): N_t and tpts.

Parity pinning (tests/test_outputs_format.py): a field that is linear in x and moves at
speed v gives v exactly (both functions, N_y = 1 for the edge form, where the stride
does not matter, any N_y for the cell form); a field that does not move gives 0; and the
edge form's reads of E_edgV follow the flat stride-N_x walk (a field that is nonzero in
one flat position only).
"""
from __future__ import annotations

import numpy as np


def ymean(x, dy, ylen, ny, nx, xp):
    """YMEAN: x flat (Fortran order), xp 1-based."""
    x = np.asarray(x, dtype=np.float64).ravel(order="F")
    s = 0.0
    p = xp - 1
    for j in range(ny):
        s = s + x[p] * dy[j]
        p += nx
    return s / ylen


def xwave_speed_edg(xc, xe, xold, dy, dx, dt, ylen, ny, nx):
    tsum = 0.0
    xsum = 0.0
    xr = ymean(xe, dy, ylen, ny, nx, 1)
    for i in range(1, nx + 1):
        xl = xr
        xr = ymean(xe, dy, ylen, ny, nx, i + 1)
        xsum = xsum + abs((xr - xl) / dx[i - 1])
        tl = ymean(xold, dy, ylen, ny, nx, i)
        tr = ymean(xc, dy, ylen, ny, nx, i)
        tsum = tsum + abs((tr - tl) / dt)
    return tsum / xsum


def xwave_speed_cnt(xc, xold, dy, dx, dt, ylen, ny, nx):
    tsum = 0.0
    xsum = 0.0
    xl = ymean(xc, dy, ylen, ny, nx, 1)
    xr = ymean(xc, dy, ylen, ny, nx, 2)
    xsum = xsum + abs((xr - xl) / dx[0])
    for i in range(2, nx):
        xl = ymean(xc, dy, ylen, ny, nx, i - 1)
        xr = ymean(xc, dy, ylen, ny, nx, i + 1)
        xsum = xsum + abs((xr - xl) / dx[i - 1] / 2.0)
    xl = ymean(xc, dy, ylen, ny, nx, nx - 1)
    xr = ymean(xc, dy, ylen, ny, nx, nx)
    xsum = xsum + abs((xr - xl) / dx[nx - 1])
    for i in range(1, nx + 1):
        tl = ymean(xold, dy, ylen, ny, nx, i)
        tr = ymean(xc, dy, ylen, ny, nx, i)
        tsum = tsum + abs((tr - tl) / dt)
    return tsum / xsum


def time_axis(tlen, start_time, dt, init_out):
    """N_t and tpts of INPUT (MainSrc/INPUTS.f08:120-151)."""
    q = (tlen - start_time) / dt
    nt = int(np.sign(q) * np.floor(abs(q) + 0.5)) + (1 if init_out else 0)   # NINT
    tpts = np.empty(nt)
    tpts[0] = start_time if init_out else start_time + dt
    for i in range(1, nt):
        tpts[i] = tpts[i - 1] + dt
    return nt, tpts

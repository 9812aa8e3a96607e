"""The ROM runs' POD database input: INPUT_fg_POD, READ_POD_FULL, READ_POD_GPART and
RRANK_CALC (MainSrc/POD_ROUTINES.f08:437-822), through the library's classic-format reader.

A database holds, for BCg and the six QD-factor fields, the centering vector C_<f>, the
singular values S_<f>, the left vectors U_<f> and the right vectors Vt_<f> of a
decomposition of snapshots on the data grid (N_x, N_y, N_g, N_t dimensions; Delx, Dely,
xlen, ylen, tlen, Delt, bcT_*, Tini; the tpts bnds attribute gives the start time), per
group (global attribute dcmp_type "PODg", PODgsum 0) or over all groups ("POD", PODgsum 1).
The reader truncates each decomposition to the rank RRANK_CALC picks for eps and returns it
in the storage trt_pod_create takes: C; S; U as rank columns of clen values per group; V
t-major per group.  The file must be a classic (CDF-1/2/5) netCDF file: this image has no
HDF5 library for NETCDF4 ones.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict

import numpy as np

from . import native

FIELDS = ("BCg", "fg_avg_xx", "fg_edgV_xx", "fg_avg_yy", "fg_edgH_yy", "fg_edgV_xy", "fg_edgH_xy")


def rrank_calc(sig, eps: float) -> int:
    """RRANK_CALC (POD_ROUTINES.f08:791-822): the first rank i whose discarded singular values
    have a 2-norm at most eps times the 2-norm of all of them (the whole length if none)."""
    sig = np.asarray(sig, dtype=np.float64)
    n = len(sig)
    sum1 = 0.0
    for s in sig:
        sum1 = sum1 + s ** 2
    eps2 = eps * np.sqrt(sum1)
    for i in range(1, n + 1):
        sum2 = 0.0
        for s in sig[i:]:
            sum2 = sum2 + s ** 2
        if np.sqrt(sum2) <= eps2:
            return i
    return n


def field_length(name: str, nx: int, ny: int) -> int:
    """Values per group of a field on an nx x ny data grid (INPUT_fg_POD's clen)."""
    return {"BCg": 2 * (nx + ny), "fg_avg_xx": nx * ny, "fg_avg_yy": nx * ny, "fg_edgV_xx": (nx + 1) * ny,
            "fg_edgV_xy": (nx + 1) * ny, "fg_edgH_yy": nx * (ny + 1), "fg_edgH_xy": nx * (ny + 1)}[name]


@dataclass
class PodDatabase:
    nx: int
    ny: int
    ng: int
    nt: int
    gsum: int
    tlen: float
    xlen: float
    ylen: float
    delt: float
    start_time: float
    tini: float
    delx: np.ndarray
    dely: np.ndarray
    bcT: np.ndarray
    bases: Dict[str, tuple] = field(default_factory=dict)   # name -> (C, S, U, V, rank)

    def create_bases(self, ctx, prefix: str = "dat_"):
        """One native.PodBasis per field, reconstructing into the context arrays
        prefix + name (the dat_* arrays of a native.RomGrid)."""
        out = {}
        for name, (C, S, U, V, rank) in self.bases.items():
            out[name] = native.PodBasis(ctx, prefix + name, C, S, U, V, rank, self.nt, self.gsum)
        return out


def read_pod_database(path: str, eps: float) -> PodDatabase:
    """INPUT_fg_POD (POD_ROUTINES.f08:437-562)."""
    r = native.NcReader(path)
    try:
        nx, ny, ng, nt = (r.dim(d) for d in ("N_x", "N_y", "N_g", "N_t"))
        scal = lambda n: float(r.get(n, [], []))
        delt = scal("Delt")
        bnds = r.att("bnds", "tpts")
        dcmp = r.att("dcmp_type")
        if dcmp[:4] == "PODg":
            gsum = 0
        elif dcmp[:3] == "POD":
            gsum = 1
        else:
            raise native.TrtError(f"INPUT_fg_POD: unrecognized dcmp_type ({dcmp}) in {path}")
        db = PodDatabase(nx=nx, ny=ny, ng=ng, nt=nt, gsum=gsum, tlen=scal("tlen"), xlen=scal("xlen"),
                         ylen=scal("ylen"), delt=delt, start_time=float(bnds[0]) - delt, tini=scal("Tini"),
                         delx=r.get("Delx", [1], [nx]), dely=r.get("Dely", [1], [ny]),
                         bcT=np.array([scal(f"bcT_{s}") for s in ("left", "bottom", "right", "top")]))
        for name in FIELDS:
            clen = field_length(name, nx, ny)
            if gsum == 1:
                db.bases[name] = _read_full(r, name, clen * ng, nt, eps)
            else:
                db.bases[name] = _read_gpart(r, name, clen, nt, ng, eps)
        return db
    finally:
        r.close()


def _read_full(r, name, clen, nt, eps):
    """READ_POD_FULL (:654-698): one decomposition of the whole vector."""
    rank = min(clen, nt)
    C = r.get("C_" + name, [1], [clen])
    S = r.get("S_" + name, [1], [rank])
    rr = rrank_calc(S, eps)
    U = r.get("U_" + name, [1, 1], [clen, rr]).ravel(order="F")
    V = r.get("Vt_" + name, [1, 1], [rr, nt]).ravel(order="F")
    return C, S[:rr].copy(), U, V, np.array([rr], dtype=np.int32)


def _read_gpart(r, name, clen, nt, ng, eps):
    """READ_POD_GPART (:700-789): a decomposition per group, concatenated."""
    rank = min(clen, nt)
    Cs, Ss, Us, Vs, rr = [], [], [], [], []
    for g in range(1, ng + 1):
        Cs.append(r.get("C_" + name, [1, g], [clen, 1]).ravel(order="F"))
        s = r.get("S_" + name, [1, g], [rank, 1]).ravel(order="F")
        k = rrank_calc(s, eps)
        rr.append(k)
        Ss.append(s[:k])
        Us.append(r.get("U_" + name, [1, 1, g], [clen, k, 1]).ravel(order="F"))
        Vs.append(r.get("Vt_" + name, [1, 1, g], [k, nt, 1]).ravel(order="F"))
    return (np.concatenate(Cs), np.concatenate(Ss), np.concatenate(Us), np.concatenate(Vs),
            np.array(rr, dtype=np.int32))

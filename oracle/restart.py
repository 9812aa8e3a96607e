"""Restatement of the reference's restart file in numpy (test infrastructure only).

RESTART_OUT (MainSrc/OUTPUTS.f08:1624-1674) opens the file FORM='UNFORMATTED'
(sequential access) and WRITEs 25 FP64 records in the order of RECORDS;
RESTART_IN (MainSrc/INPUTS.f08:1734-1783) READs them back in the same order and
RESTART_TIME_IN (:1785-1803) only the first.  Every array is written whole, in
its Fortran shape: RT_Src and I_crn as (2N_x, 2N_y, N_m, N_g)
(MainSrc/INITIALIZERS.f08:46,124), the others as the low-order arrays they are.

Record framing is the Fortran runtime's, not the reference's code: each record
is one or more subrecords of <int32 length><data><int32 length>; a subrecord
carries at most 2147483639 data bytes, and in a split record the leading marker
of every subrecord but the last and the trailing marker of every subrecord but
the first are negated (gfortran's documented layout of unformatted sequential
files).

Parity pinning: tests/test_restart_format.py checks frame()/unframe() against
files that a Fortran program compiled with amdflang writes and reads (single
subrecords: the real runtime), the split rule against hand-built byte strings
that follow the documented rule above (no compiler here splits records, so the
split layout is pinned by the documentation only), and restart_arrays() against
the record list of RESTART_OUT above.
"""
from __future__ import annotations

import struct

import numpy as np

RECORDS = ("Time", "Temp", "I_crn", "fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy",
           "RT_Src", "MGQD_Src", "GQD_Src", "E_avg", "Eg_avg", "Eg_edgV", "Eg_edgH", "Fxg_edgV", "Fyg_edgH",
           "KapE_Bar", "KapE", "KapR", "Cg_L", "Cg_B", "Cg_R", "Cg_T")
MAX_SUB = 2147483639


def frame(data: bytes, max_sub: int = MAX_SUB) -> bytes:
    """One record holding `data`, split into subrecords of at most max_sub bytes."""
    out = []
    n = len(data)
    if n == 0:
        return struct.pack("<ii", 0, 0)
    pos, first = 0, True
    while pos < n:
        k = min(max_sub, n - pos)
        more = pos + k < n
        out.append(struct.pack("<i", -k if more else k))
        out.append(data[pos:pos + k])
        out.append(struct.pack("<i", k if first else -k))
        pos += k
        first = False
    return b"".join(out)


def unframe(buf: bytes, pos: int = 0):
    """(data of the record starting at pos, position after it)."""
    parts, first = [], True
    while True:
        (m,) = struct.unpack_from("<i", buf, pos)
        k = abs(m)
        parts.append(buf[pos + 4:pos + 4 + k])
        (t,) = struct.unpack_from("<i", buf, pos + 4 + k)
        if abs(t) != k or (k and (t < 0) == first):
            raise ValueError(f"record markers disagree at byte {pos}")
        pos += 8 + k
        first = False
        if m >= 0:
            return b"".join(parts), pos


def restart_arrays(o, nm: int):
    """The 25 arrays of RESTART_OUT from an oracle state, in record order, each in its
    Fortran shape.  The oracle keeps RT_Src per (cell, group); the file holds it per
    corner and direction (MATERIAL_UPDATE writes the same value to all of them,
    MainSrc/UPDATES.f08:141-150)."""
    out = []
    for name in RECORDS:
        if name == "Time":
            out.append(np.array([o.time], dtype=np.float64))
        elif name == "RT_Src":
            src = np.asarray(o["RT_Src"])
            if src.ndim == 2:
                src = src[:, :, None]
            nx, ny, ng = src.shape
            full = np.repeat(np.repeat(src, 2, axis=0), 2, axis=1)          # (2nx, 2ny, ng)
            out.append(np.broadcast_to(full[:, :, None, :], (2 * nx, 2 * ny, nm, ng)).copy(order="F"))
        else:
            out.append(np.array(o[name], dtype=np.float64, order="F"))
    return out


def write_restart(path: str, arrays, max_sub: int = MAX_SUB) -> None:
    with open(path, "wb") as f:
        for a in arrays:
            f.write(frame(np.asarray(a, dtype=np.float64).tobytes(order="F"), max_sub))


def read_restart(path: str, shapes):
    """Arrays of the first len(shapes) records, each reshaped (Fortran order)."""
    buf = open(path, "rb").read()
    pos, out = 0, []
    for shp in shapes:
        data, pos = unframe(buf, pos)
        a = np.frombuffer(data, dtype=np.float64)
        if a.size != int(np.prod(shp)):
            raise ValueError(f"record holds {a.size} values, expected {int(np.prod(shp))}")
        out.append(a.reshape(shp, order="F"))
    return out

"""Writes tests/golden/cfg0_steps.npz: the FP64 oracle's cfg0 run (20x20 cells, S4, 4 groups
1e-2..1e1 keV, 1 keV left boundary, 10 BDF-1 steps of 2e-4 sh) with the iterations converged
tightly (conv_ho 1e-13, conv_lo 1e-14, conv_gr1 1e-15, line search off) -- per step T, the grey
E, HO_E and the cell-average Eddington factors f_xx, f_yy, f_xy (N_x, N_y, N_g), and the
iteration counts (RT, MLOQD, Newton, sweeps).  Run from the repository root:

    python tests/golden/make_cfg0_steps.py
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

P = importlib.import_module("2d_trt_rom_system_amd.problem")

FIELDS = ("Temp", "E_avg", "HO_E_avg", "fg_avg_xx", "fg_avg_yy", "fg_avg_xy")


def cfg0_tight():
    p = P.cfg0()
    p.use_line_search = 0
    p.conv_ho, p.conv_lo, p.conv_gr1 = 1e-13, 1e-14, 1e-15
    return p


def run(steps=10):
    oracle.build()
    o = oracle.Oracle(cfg0_tight(), threads=4)
    out = {n: [] for n in FIELDS}
    its = []
    for _ in range(steps):
        its.append(o.step())
        for n in FIELDS:
            out[n].append(np.array(o[n], dtype=np.float64))
    res = {n: np.stack(v) for n, v in out.items()}
    res["iterations"] = np.array(its, dtype=np.int32)
    return res


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "cfg0_steps.npz"), **run())
    print("wrote", os.path.join(HERE, "cfg0_steps.npz"))

"""The committed cfg0 fixture (tests/golden/cfg0_steps.npz, written by
tests/golden/make_cfg0_steps.py from the FP64 oracle, iterations converged tightly) is what the
oracle computes now: a guard against changes to the restatement, CPU only.  The GPU runs are
compared with the same fixture in tests/test_gpu_golden.py."""
import os

import numpy as np

from conftest import ROOT

FIX = os.path.join(ROOT, "tests", "golden", "cfg0_steps.npz")


def _l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_oracle_reproduces_the_cfg0_fixture(oracle_mod):
    import runpy
    mk = runpy.run_path(os.path.join(ROOT, "tests", "golden", "make_cfg0_steps.py"))
    got = mk["run"](steps=10)
    want = np.load(FIX)
    assert np.array_equal(got["iterations"], want["iterations"])
    for n in mk["FIELDS"]:
        for s in range(10):
            assert _l2(got[n][s], want[n][s]) < 1e-13, (n, s)

"""The measurement rules of bench.py that its figures depend on (CPU only): the batched mix
is the longest prefix of the batched distribution's draws whose FP64 corner intensities fit
the budget, the algorithmic bytes of a sweep follow the stated count, and the HBM traffic of
a sweep kernel comes from the committed counter summary for that kernel alone."""
import importlib
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
bench = importlib.import_module("bench")
trt = importlib.import_module("2d_trt_rom_system_amd")


def _icrn_bytes(p):
    return 2 * 4 * p.nx * p.ny * p.nm * p.ng * 8


def test_batch_mix_is_the_longest_prefix_within_budget():
    budget = 96.0
    mix = bench.batch_mix(trt, budget)
    draws = [trt.problem.cfg3_problem(20240611, i) for i in range(len(mix) + 1)]
    for a, b in zip(mix, draws):
        assert (a.nx, a.ny, a.nm, a.ng) == (b.nx, b.ny, b.nm, b.ng)
    used = sum(_icrn_bytes(p) for p in mix)
    assert used <= budget * 2**30
    assert used + _icrn_bytes(draws[len(mix)]) > budget * 2**30
    # the 96 GiB mix: draws 0-6 (64^2 .. 512^2, S8 .. S16, 16 .. 64 groups), 1.49e9 updates
    assert len(mix) == 7
    assert sum(p.updates_per_sweep() for p in mix) == 1490026496


@pytest.mark.parametrize("budget", [0.5, 20.0, 200.0])
def test_batch_mix_prefix_rule_at_other_budgets(budget):
    mix = bench.batch_mix(trt, budget)
    used = sum(_icrn_bytes(p) for p in mix)
    assert used <= budget * 2**30
    if len(mix) < 64:
        nxt = trt.problem.cfg3_problem(20240611, len(mix))
        assert used + _icrn_bytes(nxt) > budget * 2**30


def test_algorithmic_bytes_of_the_headline_sweep():
    p = trt.problem.cfg1()
    n = p.nx * p.ny * p.nm * p.ng
    assert bench.alg_bytes(p, 1) == n * 32 + p.nx * p.ny * p.ng * 112 == 9898557440
    assert bench.alg_bytes(p, 0) == n * 64 + p.nx * p.ny * p.ng * 112 == 19562233856


def test_traffic_comes_from_the_kernels_own_counter_entry():
    with open(os.path.join(ROOT, "profiles", "sweep_cfg1_hbm_traffic.json")) as f:
        d = json.load(f)
    entries = {v["kernel"]: v["hbm_bytes_per_launch"] for v in d.values() if isinstance(v, dict) and "kernel" in v}
    assert "sweep_pk<float, float, 4, 4>" in entries
    for kern, b in entries.items():
        assert bench.traffic_from_profiles(kern) == b
    assert bench.traffic_from_profiles("sweep_pk<float, float, 3, 3>") is None

"""Host logic of the accuracy report: the error measures and the trajectory
comparison, on synthetic data with known answers (no GPU, no oracle run)."""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def pr():
    spec = importlib.util.spec_from_file_location("precision_report", os.path.join(HERE, "precision_report.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_rel_l2_known_values(pr):
    b = np.array([3.0, 4.0])
    assert pr.rel_l2(b, b) == 0.0
    assert pr.rel_l2(b * (1 + 1e-3), b) == pytest.approx(1e-3, rel=1e-12)
    # a zero reference falls back to the absolute norm
    assert pr.rel_l2(np.array([0.0, 2.0]), np.zeros(2)) == pytest.approx(2.0)


def test_weighted_rel_ignores_weightless_entries(pr):
    """An entry with zero weight (a group with no energy) cannot move the measure;
    one with weight moves it by its share."""
    b = np.array([1.0, 1.0, 1.0])
    w = np.array([1.0, 1.0, 0.0])
    a = np.array([1.0, 1.0, 1e6])
    assert pr.weighted_rel(a, b, w) == 0.0
    a2 = np.array([1.1, 1.0, 1.0])
    assert pr.weighted_rel(a2, b, w) == pytest.approx(0.1 / 2.0, rel=1e-12)


def test_compare_trajectories(pr, tmp_path):
    """compare() pairs steps present in both files and reports the planted errors."""
    rng = np.random.default_rng(1)
    T = [rng.uniform(1.0, 2.0, 20) for _ in range(3)]
    E = [rng.uniform(1.0, 2.0, 20) for _ in range(3)]
    o = {f"oracle_T{s + 1}": T[s] for s in range(3)} | {f"oracle_E{s + 1}": E[s] for s in range(3)}
    g = {}
    for s in range(2):          # the GPU file holds only two steps
        for k, eps in (("fp64", 0.0), ("fp32", 1e-6), ("fp32s", 1e-8)):
            g[f"{k}_T{s + 1}"] = T[s] * (1 + eps)
            g[f"{k}_E{s + 1}"] = E[s] * (1 - eps)
    np.savez(tmp_path / "o.npz", **o)
    np.savez(tmp_path / "g.npz", **g)
    d = pr.compare(str(tmp_path / "o.npz"), str(tmp_path / "g.npz"))
    assert d["steps"] == 2
    assert d["worst"]["fp64_T"] == 0.0
    assert d["worst"]["fp32_T"] == pytest.approx(1e-6, rel=1e-6)
    assert d["worst"]["fp32s_E"] == pytest.approx(1e-8, rel=1e-6)
    assert d["worst"]["fp32_E_max"] == pytest.approx(1e-6, rel=1e-2)

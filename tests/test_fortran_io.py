"""The Fortran host program is built and fails loudly (non-zero exit, message) without a GPU (CPU only)."""
import os
import subprocess

import pytest


def test_problem_file_layout(trt, tmp_path):
    fio = __import__("importlib").import_module("2d_trt_rom_system_amd.fortran_io")
    prob = trt.problem.cfg0()
    p = tmp_path / "p.bin"
    fio.write_problem(str(p), prob)
    n_int, n_dbl = 19, 17
    expected = 4 * n_int + 8 * n_dbl + 8 * (prob.nx + prob.ny + 3 * prob.nm + prob.ng + 1 + prob.nmat) + 4 * prob.nx * prob.ny
    assert p.stat().st_size == expected


def test_driver_stops_with_error_without_gpu(trt, tmp_path):
    fio = __import__("importlib").import_module("2d_trt_rom_system_amd.fortran_io")
    if not os.path.exists(fio.EXE):
        pytest.skip("trt_mlqd not built")
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    prob = trt.problem.cfg0()
    p = tmp_path / "p.bin"
    fio.write_problem(str(p), prob)
    r = subprocess.run([fio.EXE, str(p), str(tmp_path / "r.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "trt_ctx_create" in r.stdout + r.stderr

"""The Fortran host program drives the GPU through the ISO_C_BINDING drop-in modules with the
reference call sequence (and through the device-resident step); results match the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("mode", [1, 2])
def test_fortran_driver_matches_oracle(trt, oracle_mod, tmp_path, mode):
    prob = trt.problem.cfg0()
    prob.nsteps = 3
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    fio = __import__("importlib").import_module("2d_trt_rom_system_amd.fortran_io")
    res, out = fio.run(prob, str(tmp_path), mode=mode)
    assert len(res) == prob.nsteps
    o = oracle_mod.Oracle(prob, threads=4)
    for s in range(prob.nsteps):
        o.step()
        for n in ("Temp", "E_avg", "HO_E_avg"):
            assert _l2(res[s][n], o[n]) < 1e-8, (mode, s, n)
        assert np.max(np.abs(res[s]["fg_avg_xx"] - o["fg_avg_xx"])) < 1e-8


def test_fortran_driver_restart_resumes_bitwise(trt, tmp_path):
    """The Fortran host program checkpoints through TRT_GPU_RESTART_OUT where the reference
    calls RESTART_OUT and resumes through TRT_GPU_RESTART_IN where it calls RESTART_IN:
    4 steps == 2 steps + checkpoint, then 2 steps from the checkpoint, bitwise."""
    import oracle.restart as orst
    fio = __import__("importlib").import_module("2d_trt_rom_system_amd.fortran_io")
    prob = trt.problem.cfg0()
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    prob.nsteps = 4
    full, _ = fio.run(prob, str(tmp_path / "full"), mode=fio.RESIDENT)
    rst = str(tmp_path / "state.rst")
    prob.nsteps = 2
    first, _ = fio.run(prob, str(tmp_path / "a"), mode=fio.RESIDENT, restart_out=rst, restart_freq=2)
    t, temp = orst.read_restart(rst, [(1,), (prob.nx, prob.ny)])
    assert t[0] == pytest.approx(2 * prob.dt)
    assert trt.native.restart_time_in(rst) == t[0]
    second, _ = fio.run(prob, str(tmp_path / "b"), mode=fio.RESIDENT, restart_in=rst)
    assert np.array_equal(temp, first[1]["Temp"])
    for s, res in enumerate(first + second):
        for k, v in full[s].items():
            assert np.array_equal(np.asarray(res[k]), np.asarray(v)), (s, k)


def test_fortran_driver_writes_the_output_file(trt, tmp_path):
    """The Fortran host program writes the reference's NetCDF output file through
    TRT_GPU_OUTFILE_INIT / TRT_GPU_TIMESTEP_OUTS where the reference calls OUTFILE_INIT and
    TIMESTEP_OUTS: the initial state and every step, equal to the step results it pulls."""
    from scipy.io import netcdf_file
    fio = __import__("importlib").import_module("2d_trt_rom_system_amd.fortran_io")
    prob = trt.problem.cfg0()
    prob.nsteps = 2
    nc = str(tmp_path / "out.nc")
    flags = "0" + "1" * 17                      # I_out 0, the reference deck's choice
    res, _ = fio.run(prob, str(tmp_path), mode=fio.RESIDENT, outfile=nc, out_flags=flags)
    with netcdf_file(nc, "r", mmap=False) as f:
        assert f.dimensions["N_t"] == prob.nsteps + 1
        assert "I_avg" not in f.variables and "fg_avg_xx" in f.variables
        for s in range(prob.nsteps):
            for k, name in (("Temp", "Temperature"), ("E_avg", "E_avg_Grey"), ("HO_E_avg", "E_avg_HO")):
                got = np.array(f.variables[name][s + 1]).T
                assert np.array_equal(got, res[s][k]), (s, k)
        assert list(f.variables["RT_Its-count"][1:]) == [r["RT_Its"] for r in res]

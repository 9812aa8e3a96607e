"""The output file's host parts, without a GPU.

* The library's classic-format reader (trt_nc_*, the reads of MainSrc/NCDF_IO.f08:339-460)
  against files written by scipy's independent implementation of the format (CDF-1 and
  CDF-2): every type, scalars, hyperslabs in Fortran order with 1-based starts.
* The wave-speed restatement (oracle/outputs.py, MainSrc/WAVEPROP_TOOLS.f08) pinned by
  closed forms: a front linear in x moving at speed v gives v.
* The golden table of the reference's output definitions (tests/golden/outputs_table.json,
  made from MainSrc/OUTPUTS.f08 by tests/golden/make_outputs_table.py) holds what the GPU
  tests compare files against.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "outputs_table.json")


def _native(trt):
    from importlib import import_module
    return import_module("2d_trt_rom_system_amd.native")


@pytest.fixture(scope="module")
def native(trt):
    n = _native(trt)
    if not os.path.exists(n.LIBPATH):
        pytest.skip("libtrt_hip.so is not built")
    return n


@pytest.mark.parametrize("version", [1, 2])
def test_reader_reads_scipy_files(native, tmp_path, version):
    from scipy.io import netcdf_file
    path = str(tmp_path / f"v{version}.nc")
    rng = np.random.default_rng(version)
    T = rng.standard_normal((4, 3, 5))            # netCDF (N_t, N_y, N_x) = Fortran (N_x, N_y, N_t)
    I = rng.integers(-1000, 1000, size=(2, 7)).astype(np.int32)
    F = rng.standard_normal(6).astype(np.float32)
    with netcdf_file(path, "w", version=version) as f:
        f.createDimension("N_t", 4)
        f.createDimension("N_y", 3)
        f.createDimension("N_x", 5)
        f.createDimension("a", 2)
        f.createDimension("b", 7)
        f.createDimension("c", 6)
        v = f.createVariable("Temperature", "d", ("N_t", "N_y", "N_x"))
        v[:] = T
        v.Units = "ev/cm^3"
        f.createVariable("counts", "i", ("a", "b"))[:] = I
        f.createVariable("norms", "f", ("c",))[:] = F
        s = f.createVariable("pi", "d", ())
        s.data[()] = np.pi
        f.run_type = "mlqd"
    r = native.NcReader(path)
    try:
        assert r.dim("N_t") == 4 and r.dim("N_x") == 5
        # the whole variable, Fortran order
        got = r.get("Temperature", [1, 1, 1], [5, 3, 4])
        np.testing.assert_array_equal(got, T.transpose(2, 1, 0))
        # one time slice, as Tgen_QDf reads Temperature (ALGORITHMS.f08:1038)
        np.testing.assert_array_equal(r.get("Temperature", [1, 1, 3], [5, 3, 1])[:, :, 0], T[2].T)
        # an interior hyperslab
        np.testing.assert_array_equal(r.get("Temperature", [2, 2, 2], [3, 2, 2]), T[1:3, 1:3, 1:4].transpose(2, 1, 0))
        np.testing.assert_array_equal(r.get("counts", [1, 1], [7, 2]), I.T.astype(np.float64))
        np.testing.assert_array_equal(r.get("norms", [2], [4]), F[1:5].astype(np.float64))
        assert r.get("pi", [], []) == np.pi
        with pytest.raises(native.TrtError):
            r.get("Temperature", [1, 1, 4], [5, 3, 2])      # beyond N_t
        with pytest.raises(native.TrtError):
            r.get("nope", [1], [1])
    finally:
        r.close()


def test_reader_rejects_other_files(native, tmp_path):
    p = tmp_path / "x.nc"
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))        # an HDF5 (NETCDF4) signature
    with pytest.raises(native.TrtError, match="classic"):
        native.NcReader(str(p))


def test_wave_speed_closed_forms():
    from oracle import outputs as ow
    rng = np.random.default_rng(5)
    for nx, ny in ((6, 1), (9, 1), (5, 4), (12, 7)):
        dx = np.full(nx, 0.3)
        dy = rng.uniform(0.1, 0.5, ny)
        ylen = dy.sum()
        b, v, dt, x0 = -2.5, 0.7, 0.01, 0.4
        xc_c = (np.arange(nx) + 0.5) * 0.3
        xe_c = np.arange(nx + 1) * 0.3
        new_c = np.repeat((b * (xc_c - v * dt - x0))[:, None], ny, axis=1)   # f(x - v t), t = dt
        old_c = np.repeat((b * (xc_c - x0))[:, None], ny, axis=1)           # f(x), t = 0
        # cell form: any N_y (rows of length N_x)
        assert ow.xwave_speed_cnt(new_c, old_c, dy, dx, dt, ylen, ny, nx) == pytest.approx(v, rel=1e-12)
        assert ow.xwave_speed_cnt(old_c, old_c, dy, dx, dt, ylen, ny, nx) == 0.0
        if ny == 1:
            # edge form: N_y = 1, where the stride-N_x walk over E_edgV is exact
            new_e = (b * (xe_c - v * dt - x0))[:, None]
            assert ow.xwave_speed_edg(new_c, new_e, old_c, dy, dx, dt, ylen, ny, nx) == pytest.approx(v, rel=1e-12)


def test_wave_speed_edge_reads_follow_the_flat_walk():
    """YMEAN walks E_edgV (N_x+1, N_y) with stride N_x (WAVEPROP_TOOLS.f08:21-24): column xp
    (1-based) reads the flat positions xp-1 + j N_x, j = 0..N_y-1, with weight dy(j), so a
    field that is 1 at flat position p only is seen by every xp with p = xp-1 + j N_x --
    also by xp = N_x+1 and by columns of the following rows of E_edgV."""
    from oracle import outputs as ow
    nx, ny = 4, 3
    dy = np.array([0.2, 0.3, 0.5])
    for p in range(nx * ny):
        e = np.zeros((nx + 1) * ny)
        e[p] = 1.0
        for xp in range(1, nx + 2):
            want = sum(dy[j] for j in range(ny) if p - j * nx == xp - 1)
            assert ow.ymean(e, dy, 1.0, ny, nx, xp) == want


def test_time_axis():
    from oracle import outputs as ow
    nt, t = ow.time_axis(2e-2, 0.0, 2e-3, 1)
    assert nt == 11 and t[0] == 0.0 and t[-1] == pytest.approx(2e-2)
    nt, t = ow.time_axis(2e-2, 0.0, 2e-3, 0)
    assert nt == 10 and t[0] == 2e-3


def test_golden_output_table():
    t = json.load(open(GOLDEN))
    vd = {v["name"]: v for v in t["OUTFILE_VARDEFS"]["vars"]}
    # a few definitions the reference's post-processing reads (Processing/processing.py:104-176)
    assert vd["Temperature"]["dims"] == ["N_x_ID", "N_y_ID", "N_t_ID"]
    assert vd["E_avg_Grey"]["cond"] == ["E_out .EQ. 1"]
    assert vd["fg_avg_xx"]["atts"]["grid2"] == "xpts_avg"
    assert vd["GQD_MatVec_prds"]["type"] == "Int" and vd["RT_Tnorm"]["type"] == "Real"
    # the reference's own quirk: KapE and KapR carry no Units (OUTPUTS.f08:1277-1287)
    assert "Units" in vd["KapB"]["atts"] and "Units" not in vd["KapE"]["atts"]
    assert len(t["OUTFILE_VARDEFS"]["vars"]) == 133
    assert {d["name"] for d in t["OUTFILE_INIT"]["dims"]} >= {"N_x", "N_t", "RT_Its-dim", "Boundaries"}

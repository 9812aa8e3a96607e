"""The reference's NetCDF output file written from device state (trt_outfile_*).

Checks, through the C ABI, that a run file (OUTFILE_INIT + OUTFILE_VARDEFS +
TIMESTEP_OUTS, MainSrc/OUTPUTS.f08) and a Tgen_QDf file (OUTFILE_INIT_Tgen):
* define every dimension, global attribute and variable the reference defines for the
  same output flags, with its type, shape and attributes (tests/golden/outputs_table.json,
  made from OUTPUTS.f08), and nothing else;
* hold, read back by scipy's independent netCDF reader, the context's arrays bitwise at
  every written time index, the wave speeds of the oracle's restatement and the step's
  iteration counts;
* can be read back by the library's own reader (Tgen_QDf's temperature input), also in
  the CDF-5 layout.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "outputs_table.json")))
FILL_D = 9.9692099683868690e+36


def ctx_array(nc_name):
    """The array TIMESTEP_OUTS writes under nc_name (its call at ALGORITHMS.f08:872-880)."""
    special = {"Temperature": "Temp", "Gold_L": "G_old", "Gold_hat": "Gold_Hat"}
    if nc_name in special:
        return special[nc_name]
    for suf in ("_Grey", "_MGQD", "_HO"):
        if nc_name.endswith(suf):
            base = nc_name[: -len(suf)]
            group = base.startswith(("Eg_", "Fxg_", "Fyg_"))
            if suf == "_Grey" or (suf == "_MGQD" and group):
                return base
            return ("MGQD_" if suf == "_MGQD" else "HO_") + base
    return nc_name


def cond_holds(cond, flags, run_type="mlqd"):
    for c in cond:
        c = c.strip()
        if c.startswith("run_type"):
            if c.split("'")[1] != run_type:
                return False
        elif c in ("Res_Calc", "fg_pod_out .EQ. 1", "fg_dmd_out .EQ. 1"):
            return False
        elif c.endswith(".EQ. 1"):
            if not flags.get(c.split()[0], 0):
                return False
        elif c.startswith("NOT "):
            continue
        elif c in ("kapE_dT_flag", "Use_Line_Search", "Use_Safety_Search"):
            if not flags.get(c, 0):
                return False
        else:
            raise AssertionError(f"unhandled condition {c!r}")
    return True


def check_structure(f, sections, flags, nt):
    dim_ids = GOLDEN["dim_ids"]
    want = {}
    for s in sections:
        for v in GOLDEN[s]["vars"]:
            if cond_holds(v["cond"], flags):
                want[v["name"]] = v
    got = set(f.variables)
    assert got == set(want), (sorted(got - set(want)), sorted(set(want) - got))
    tcode = {"Double": "d", "Real": "f", "Int": "i"}
    for name, v in want.items():
        var = f.variables[name]
        assert var.typecode() == tcode[v["type"]], name
        assert tuple(var.dimensions) == tuple(dim_ids[d] for d in reversed(v["dims"])), name
        atts = {k: val for k, val in var._attributes.items()}
        assert set(atts) == set(v["atts"]), (name, atts, v["atts"])
        for k, val in v["atts"].items():
            if isinstance(val, dict):
                continue
            got_v = atts[k].decode() if isinstance(atts[k], bytes) else atts[k]
            assert np.all(got_v == val), (name, k, got_v, val)
    assert f.dimensions["N_t"] == nt


def pull_all(ctx, f, names):
    return {n: ctx.pull(ctx_array(n)) for n in names if ctx_array(n) and ctx.size(ctx_array(n))}


def fslice(f, name, t):
    """Time slice t (1-based) of a variable as a flat Fortran-order vector."""
    a = np.array(f.variables[name][t - 1])
    return a.T.ravel(order="F") if a.ndim else a


def test_run_file(trt, tmp_path, oracle_mod):
    from scipy.io import netcdf_file
    from oracle import outputs as ow
    native = trt.native
    prob = trt.problem.small(nx=10, ny=8, order=4, ng=3)
    flags = {k: 1 for k in native.OUT_FLAGS}
    flags.update(kapE_dT_flag=prob.kapE_dT_flag, Use_Line_Search=prob.use_line_search,
                 Use_Safety_Search=prob.use_safety_search)
    path = str(tmp_path / "run.nc")
    nsteps = 2
    tlen = nsteps * prob.dt
    with native.Context(prob) as ctx:
        print(f"run file: {prob.nx}x{prob.ny} S{prob.order} {prob.ng} groups, {nsteps} steps", flush=True)
        of = native.OutFile(ctx, path, kind=0, run_type="mlqd", quadrature="LS4   ", enrgy_strc="custom", tlen=tlen,
                            init_out=1, **{k: 1 for k in native.OUT_FLAGS})
        names = [v["name"] for v in GOLDEN["OUTFILE_VARDEFS"]["vars"]]
        snaps, its, speeds = {}, {}, {}
        of.timestep_outs(1, after_step=False)
        snaps[1] = pull_all(ctx, None, names)
        for t in range(2, nsteps + 2):
            its[t] = ctx.step()
            of.timestep_outs(t)
            snaps[t] = pull_all(ctx, None, names)
            g = {n: ctx.pull(n) for n in ("E_avg", "E_edgV", "E_avg_old", "Temp", "Temp_Old", "Delx", "Dely")}
            args = (g["Dely"], g["Delx"], prob.dt, prob.ylen, prob.ny, prob.nx)
            speeds[t] = (ow.xwave_speed_edg(g["E_avg"], g["E_edgV"], g["E_avg_old"], *args),
                         ow.xwave_speed_cnt(g["Temp"], g["Temp_Old"], *args))
        of.close()
    with netcdf_file(path, "r", mmap=False) as f:
        assert open(path, "rb").read(4) == b"CDF\x02"
        check_structure(f, ("OUTFILE_INIT", "OUTFILE_VARDEFS"), flags, nsteps + 1)
        ga = f._attributes
        want_g = [g["name"] for g in GOLDEN["OUTFILE_INIT"]["gatts"] if isinstance(g["name"], str)
                  and cond_holds(g["cond"], flags)]
        assert set(ga) == set(want_g), (set(ga) ^ set(want_g))
        assert ga["run_type"] == b"mlqd" and ga["quadrature"] == b"LS4" and ga["use_grey"] == 1
        assert list(ga["BC_type"]) == list(prob.bc_type)
        assert list(ga["conv_gr"]) == [prob.conv_gr1, prob.conv_gr2]
        assert ga["kapE_dT_flag"] == (b"on" if prob.kapE_dT_flag else b"off")
        nt, tpts = ow.time_axis(tlen, 0.0, prob.dt, 1)
        np.testing.assert_array_equal(f.variables["tpts"][:], tpts)
        np.testing.assert_array_equal(f.variables["Delx"][:], np.full(prob.nx, prob.xlen / prob.nx))
        assert f.variables["c--lightspeed"].getValue() == prob.c
        assert f.variables["Computational_Unit"].getValue() == prob.comp_unit
        # every time slice equals the context's arrays at that time, bitwise
        compared = set()
        for t, snap in snaps.items():
            for name, want in snap.items():
                if name in f.variables:
                    np.testing.assert_array_equal(fslice(f, name, t), want, err_msg=f"{name} t={t}")
                    compared.add(name)
        assert len(compared) >= 80, sorted(compared)
        # wave speeds: fill at the initial state, the restatement's values after each step
        assert f.variables["E_xWave_Speed"][0] == FILL_D
        for t, (es, ts) in speeds.items():
            assert f.variables["E_xWave_Speed"][t - 1] == pytest.approx(es, rel=1e-14)
            assert f.variables["Temp_xWave_Speed"][t - 1] == pytest.approx(ts, rel=1e-14)
        # iteration history against the step's totals {RT, MLOQD, Newton, sweeps}
        rtc = f.variables["RT_Its-count"][:]
        mgc = np.array(f.variables["MGQD_Its-count"][:])
        gqc = np.array(f.variables["GQD_Its-count"][:])
        mgk = np.array(f.variables["MGQD_MatVec_prds"][:])
        for t, (rt, mg, nw, _) in its.items():
            assert rtc[t - 1] == rt
            assert mgc[t - 1, :rt].sum() == mg
            assert np.all(mgc[t - 1, rt:] == -2147483647)
            g = gqc[t - 1, :rt]
            assert g[g != -2147483647].sum() == nw
            assert np.all(mgk[t - 1, :rt] == -1)          # direct solves, the reference's -1 (QS:499)
            tn = np.array(f.variables["RT_Tnorm"][t - 1, :rt])
            assert np.all(np.isfinite(tn))
        assert rtc[0] == -2147483647                     # the initial state has no iterations
        assert f.dimensions["RT_Its-dim"] == max(i[0] for i in its.values())


def test_tgen_file_and_reader(trt, tmp_path):
    from scipy.io import netcdf_file
    native = trt.native
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=2)
    flags = {k: 1 for k in ("I_out", "HO_Eg_out", "HO_Fg_out", "HO_E_out", "HO_F_out", "QDfg_out")}
    path = str(tmp_path / "tgen.nc")
    rng = np.random.default_rng(11)
    tpts = np.array([0.5, 1.0, 1.5]) * prob.dt
    temps = {}
    with native.Context(prob) as ctx:
        of = native.OutFile(ctx, path, kind=1, run_type="Tgen_qdf", quadrature="LS4", enrgy_strc="custom",
                            tlen=tpts[-1], n_t=3, tpts=tpts, **flags)
        snaps = {}
        for t in (1, 2, 3):
            T = rng.uniform(1.0, 500.0, (prob.nx, prob.ny))
            ctx.push("Temp", T)
            ctx.tgen_step()
            of.tgen_outs(t)
            temps[t] = T
            snaps[t] = pull_all(ctx, None, [v["name"] for v in GOLDEN["OUTFILE_INIT_Tgen"]["vars"]])
        of.close()
    with netcdf_file(path, "r", mmap=False) as f:
        flags2 = dict(flags)
        check_structure(f, ("OUTFILE_INIT_Tgen",), flags2, 3)
        assert f._attributes["T_run_type"] == b"Tgen_qdf"
        np.testing.assert_array_equal(f.variables["tpts"][:], tpts)
        for t, snap in snaps.items():
            for name, want in snap.items():
                if name in f.variables:
                    np.testing.assert_array_equal(fslice(f, name, t), want, err_msg=f"{name} t={t}")
    # the library's reader: Temperature at time t, as Tgen_QDf reads it (ALGORITHMS.f08:1038)
    r = native.NcReader(path)
    for t, T in temps.items():
        np.testing.assert_array_equal(r.get("Temperature", [1, 1, t], [prob.nx, prob.ny, 1])[:, :, 0], T)
    r.close()


def test_cdf5_layout_and_rom_attributes(trt, tmp_path):
    native = trt.native
    prob = trt.problem.small(nx=6, ny=5, order=4, ng=2)
    path = str(tmp_path / "rom.nc")
    with native.Context(prob) as ctx:
        of = native.OutFile(ctx, path, kind=0, run_type="mg_pod", pod_type="fg", direc_diff=1, n_dsets=2,
                            dset_times=[0.1, 0.2], dsets=["a.h5", "b.h5"], tlen=prob.dt, format=5,
                            its_out=1)
        of.timestep_outs(1, after_step=False)
        with pytest.raises(native.TrtError, match="time index"):
            of.timestep_outs(2, after_step=False)                 # N_t = 1
        T = ctx.pull("Temp")
        of.close()
        with pytest.raises(native.TrtError, match="kind"):
            native.OutFile(ctx, str(tmp_path / "bad.nc"), kind=3)
    assert open(path, "rb").read(4) == b"CDF\x05"
    r = native.NcReader(path)
    assert r.dim("N_dsets") == 2 and r.dim("N_t") == 1
    np.testing.assert_array_equal(r.get("Temperature", [1, 1, 1], [prob.nx, prob.ny, 1]).ravel(order="F"), T)
    assert r.get("RT_Its-count", [1], [1])[0] == -2147483647
    r.close()

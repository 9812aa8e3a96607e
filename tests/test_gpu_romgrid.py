"""The ROM runs' data grid on the GPU (trt_romgrid_*) against the oracle's restatement
(oracle/romgrid.py): MAP_GRIDS maps and TMap equal, the boundary appends + FMAP of every field
equal, and a ROM step fed by POD bases on a data grid other than the simulation grid
(POD_RECONSTRUCT_fg + POD_RECONSTRUCT_BCg, MainSrc/POD_ROUTINES.f08:28-353) equal to the
oracle's same step."""
import numpy as np
import pytest

from oracle import rom
from oracle import romgrid as R

pytestmark = pytest.mark.gpu

FIELDS = ("fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy")


def _rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _data_grid(n, seed, length=1.0):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.6, 1.4, n)
    return d * (length / d.sum())


def _dat_shapes(dnx, dny, ng):
    return {"fg_avg_xx": (dnx, dny, ng), "fg_avg_yy": (dnx, dny, ng), "fg_edgV_xx": (dnx + 1, dny, ng),
            "fg_edgV_xy": (dnx + 1, dny, ng), "fg_edgH_yy": (dnx, dny + 1, ng), "fg_edgH_xy": (dnx, dny + 1, ng),
            "BCg": (2 * (dnx + dny), ng)}


@pytest.mark.parametrize("dn,n", [((9, 9), (12, 12)), ((7, 10), (11, 8))])
def test_maps_and_fields_match_the_oracle(trt, dn, n):
    native = trt.native
    dnx, dny = dn
    prob = trt.problem.small(nx=n[0], ny=n[1], order=4, ng=3)
    dxd, dyd = _data_grid(dnx, 1), _data_grid(dny, 2)
    print(f"romgrid: data {dnx}x{dny} -> simulation {prob.nx}x{prob.ny}", flush=True)
    with native.Context(prob) as ctx:
        rg = native.RomGrid(ctx, dxd, dyd, 1.0, 1.0, 4, 1.5 * prob.dt, 0.0, 6, 0.0, prob.xlen, prob.ylen)
        G = R.Grids(dxd, dyd, prob.dx, prob.dy, 1.0, 1.0, prob.xlen, prob.ylen)
        square = dnx == dny and prob.nx == prob.ny
        layouts = [(G.sim_avg, G.dat_avg, G.map_avg), (G.sim_edgv, G.dat_edgv, G.map_edgv),
                   (G.sim_edgh, G.dat_edgh, G.map_edgh)]
        if square:
            layouts += [(G.sim_bnds[s], G.dat_bnds[s], G.map_bnds[s]) for s in range(4)]
        for k, (sg, dg, (gm, vm)) in enumerate(layouts):
            ggm, gvm, gsg, gdg = rg.maps(k, len(sg) // 2, len(dg) // 2)
            np.testing.assert_array_equal(gsg, sg, err_msg=f"layout {k} simulation points")
            np.testing.assert_array_equal(gdg, dg, err_msg=f"layout {k} data points")
            np.testing.assert_array_equal(ggm[:len(gm)], gm, err_msg=f"layout {k} GMap")
            np.testing.assert_array_equal(gvm[:len(vm)], vm, err_msg=f"layout {k} VMap")
        tm, simt, datt = R.generate_grids_t(1.5 * prob.dt, prob.dt, 0.0, 0.0, 4, 6)
        for t in range(1, 7):
            t1, t2, x, x1, x2 = rg.tmap(t)
            assert (t1, t2) == (tm[2 * t - 2], tm[2 * t - 1])
            assert (x, x1, x2) == (simt[t - 1], datt[t1 - 1], datt[t2 - 1])
        rng = np.random.default_rng(dnx * 100 + dny)
        dat = {k: rng.uniform(0.1, 0.5, s) for k, s in _dat_shapes(dnx, dny, prob.ng).items()}
        for k, v in dat.items():
            ctx.push("dat_" + k, v)
        rg.map(with_bcg=square)
        flat = {k: v.ravel(order="F") for k, v in dat.items()}
        ref = R.map_fields(G, flat, prob.ng) if square else _fields_only(G, dat, prob.ng)
        for k in FIELDS:
            got = ctx.pull(k + "_next")
            want = ref[k]
            assert _rel(got, want) < 1e-14, (k, _rel(got, want))
        if square:
            for k in ("Cg_L", "Cg_B", "Cg_R", "Cg_T"):
                assert _rel(ctx.pull(k + "_next"), ref[k]) < 1e-14, k
        else:
            with pytest.raises(native.TrtError, match="N_x = N_y"):
                rg.map(with_bcg=True)
        rg.close()


def _fields_only(G, dat, ng):
    """map_fields without C_g (any grid shape)."""
    d = {k: v.ravel(order="F") for k, v in dat.items()}
    dnx, dny, nx, ny = G.dnx, G.dny, G.nx, G.ny
    out = {}
    xxbnd = R.collect_lr(d["fg_edgV_xx"], dnx + 1, dny, ng)
    for n in ("fg_edgV_xx", "fg_edgV_xy"):
        out[n] = R.fmap(R.extnd_bt(d[n], dnx + 1, dny, ng), G.dat_edgv, G.sim_edgv, *G.map_edgv,
                        (dnx + 1) * (dny + 2), (nx + 1) * ny, ng)
    yybnd = R.collect_bt(d["fg_edgH_yy"], dnx, dny + 1, ng)
    for n in ("fg_edgH_yy", "fg_edgH_xy"):
        out[n] = R.fmap(R.extnd_lr(d[n], dnx, dny + 1, ng), G.dat_edgh, G.sim_edgh, *G.map_edgh,
                        (dnx + 2) * (dny + 1), nx * (ny + 1), ng)
    a = d["fg_avg_xx"]
    out["fg_avg_xx"] = R.fmap(R.avg_apnd(a, xxbnd, R.collect_bt(a, dnx, dny, ng), R.collect_crn(a, dnx, dny, ng),
                                         dnx, dny, ng), G.dat_avg, G.sim_avg, *G.map_avg, (dnx + 2) * (dny + 2),
                              nx * ny, ng)
    a = d["fg_avg_yy"]
    out["fg_avg_yy"] = R.fmap(R.avg_apnd(a, R.collect_lr(a, dnx, dny, ng), yybnd, R.collect_crn(a, dnx, dny, ng),
                                         dnx, dny, ng), G.dat_avg, G.sim_avg, *G.map_avg, (dnx + 2) * (dny + 2),
                              nx * ny, ng)
    return out


def test_rom_step_from_bases_on_a_data_grid(trt, oracle_mod):
    """POD bases of synthetic fields on a 7x7 non-uniform data grid with 3 data times ->
    reconstruction at the data times TMap picks for simulation time 1 (interpolated) ->
    boundary appends and FMAP onto the 10x10 simulation grid -> one ROM step; the same chain
    through the oracle (numpy reconstruction and map, C oracle step)."""
    native = trt.native
    prob = trt.problem.small(nx=10, ny=10, order=4, ng=2)
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    dnx = dny = 7
    dxd, dyd = _data_grid(dnx, 5), _data_grid(dny, 6)
    G = R.Grids(dxd, dyd, prob.dx, prob.dy, 1.0, 1.0, prob.xlen, prob.ylen)
    shapes = _dat_shapes(dnx, dny, prob.ng)
    rng = np.random.default_rng(9)
    nt = 3
    # snapshots near the equilibrium factors (1/3 on the diagonal, 0 off it, -/+1/2 for C_g)
    base = {"fg_avg_xx": 1 / 3, "fg_avg_yy": 1 / 3, "fg_edgV_xx": 1 / 3, "fg_edgV_xy": 0.0, "fg_edgH_yy": 1 / 3,
            "fg_edgH_xy": 0.0}
    snaps = {}
    for k, s in shapes.items():
        if k == "BCg":
            side = np.concatenate([-0.5 * np.ones(dny), -0.5 * np.ones(dnx), 0.5 * np.ones(dny), 0.5 * np.ones(dnx)])
            b = np.repeat(side[:, None], prob.ng, axis=1).ravel(order="F")
        else:
            b = np.full(int(np.prod(s)), base[k])
        snaps[k] = np.stack([b + 1e-3 * rng.standard_normal(b.size) for _ in range(nt)], axis=1)
    o = oracle_mod.Oracle(prob, threads=4)
    with native.Context(prob) as ctx:
        o.step()
        ctx.step()
        rg = native.RomGrid(ctx, dxd, dyd, 1.0, 1.0, nt, 0.7 * prob.dt, prob.dt, 1, ctx.time, prob.xlen, prob.ylen)
        t1, t2, x, x1, x2 = rg.tmap(1)
        assert t1 != t2                           # the simulation time lies between two data times
        dat_ref, pods = {}, []
        for k in shapes:
            C, S, U, V, rk = rom.basis_from_snapshots(snaps[k], prob.ng, 0)
            pod = native.PodBasis(ctx, "dat_" + k, C, S, U, V, rk, nt, 0)
            pod.reconstruct(t1, t2, x, x1, x2)
            pods.append(pod)
            y1 = rom.pod_reconstruct(C, S, U, V, rk, snaps[k].shape[0], prob.ng, nt, t1, 0)
            y2 = rom.pod_reconstruct(C, S, U, V, rk, snaps[k].shape[0], prob.ng, nt, t2, 0)
            dat_ref[k] = y1 + (x - x1) * (y2 - y1) / (x2 - x1)
            assert _rel(ctx.pull("dat_" + k), dat_ref[k]) < 1e-13, k
        rg.map(with_bcg=True)
        mapped = R.map_fields(G, dat_ref, prob.ng)
        for k in FIELDS:
            assert _rel(ctx.pull(k + "_next"), mapped[k]) < 1e-13, k
            o[k + "_next"][...] = mapped[k].reshape(np.asarray(o[k]).shape, order="F")
        for k in ("Cg_L", "Cg_B", "Cg_R", "Cg_T"):
            assert _rel(ctx.pull(k + "_next"), mapped[k]) < 1e-13, k
            o[k + "_next"][...] = mapped[k].reshape(np.asarray(o[k]).shape, order="F")
        it_o = o.step_given_qdf()
        it_g = ctx.step_given_qdf()
        for n in ("Temp", "E_avg", "MGQD_E_avg", "Eg_avg"):
            ref = np.asarray(o[n]).ravel(order="F")
            err = float(np.sqrt(np.sum((ctx.pull(n) - ref) ** 2) / np.sum(ref ** 2)))
            assert err < 1e-8, (n, err, it_o, it_g)
        for p in pods:
            p.close()
        rg.close()


def test_simulation_grid_outside_the_data_grid_is_an_error(trt):
    native = trt.native
    prob = trt.problem.small(nx=8, ny=8, order=4, ng=2)
    with native.Context(prob) as ctx:
        with pytest.raises(native.TrtError, match="outside the data grid"):
            native.RomGrid(ctx, np.full(6, 0.1), np.full(6, 0.1), 0.6, 0.6, 2, prob.dt, 0.0, 2, 0.0, prob.xlen,
                           prob.ylen)


def test_database_to_simulation_grid(trt, tmp_path):
    """A POD database file (classic netCDF, the decomposer's layout) read by rom_db, its bases
    reconstructed on the GPU into the data grid at the data times TMap picks, and mapped onto
    the simulation grid: equal to the same chain in numpy (oracle.rom + oracle.romgrid)."""
    import importlib
    from test_rom_db import _write_database
    native = trt.native
    rom_db = importlib.import_module("2d_trt_rom_system_amd.rom_db")
    dn, nt, ng = 6, 4, 2
    prob = trt.problem.small(nx=9, ny=9, order=4, ng=ng)
    rng = np.random.default_rng(21)
    snaps = {n: 0.3 + 0.01 * rng.standard_normal((rom_db.field_length(n, dn, dn) * ng, nt)) for n in rom_db.FIELDS}
    path = str(tmp_path / "db.nc")
    _write_database(path, snaps, dn, dn, ng, nt, 0, delt=0.7 * prob.dt, t0=0.0)
    db = rom_db.read_pod_database(path, eps=1e-10)
    with native.Context(prob) as ctx:
        rg = native.RomGrid(ctx, db.delx, db.dely, db.xlen, db.ylen, db.nt, db.delt, db.start_time, 3, 0.0,
                            prob.xlen, prob.ylen)
        bases = db.create_bases(ctx)
        G = R.Grids(db.delx, db.dely, prob.dx, prob.dy, db.xlen, db.ylen, prob.xlen, prob.ylen)
        for t in (1, 2, 3):
            t1, t2, x, x1, x2 = rg.tmap(t)
            dat_ref = {}
            for name, b in bases.items():
                b.reconstruct(t1, t2, x, x1, x2)
                C, S, U, V, rank = db.bases[name]
                y1 = rom.pod_reconstruct(C, S, U, V, rank, len(C), ng, nt, t1, 0)
                y2 = rom.pod_reconstruct(C, S, U, V, rank, len(C), ng, nt, t2, 0)
                dat_ref[name] = y1 if t1 == t2 else y1 + (x - x1) * (y2 - y1) / (x2 - x1)
            rg.map(with_bcg=True)
            ref = R.map_fields(G, dat_ref, ng)
            for k in FIELDS + ("Cg_L", "Cg_B", "Cg_R", "Cg_T"):
                assert _rel(ctx.pull(k + "_next"), ref[k]) < 1e-13, (t, k)
        for b in bases.values():
            b.close()
        rg.close()

"""GPU parity of the low-order routines (OLD_*_COEFS, MLOQD_FV, COLLAPSE_MG_EF,
Cbar_Calc, GREY_COEFS, EGP_FV_NEWT, CONVERGENCE) and of whole time steps
against the FP64 oracle."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

SKIP_SYNC = {"Delx", "Dely", "Omega_x", "Omega_y", "quad_weight", "nu_g", "Kappa_Mult", "A"}


def sync_to_gpu(o, ctx, names):
    for n in names:
        if n in SKIP_SYNC:
            continue
        ctx.push(n, o[n])


def _pair(trt, oracle_mod, steps=1, lin_maxit=0, qd_solver=0, **kw):
    prob = trt.problem.small(nx=9, ny=7, order=8, ng=3)
    for k, v in kw.items():
        setattr(prob, k, v)
    o = oracle_mod.Oracle(prob, threads=3)
    for _ in range(steps):
        o.step()
    ctx = trt.native.Context(prob, lin_maxit=lin_maxit, qd_solver=qd_solver)
    return prob, o, ctx


ML_INPUTS = ["fg_avg_xx", "fg_avg_yy", "fg_avg_xy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy",
             "Cg_L", "Cg_B", "Cg_R", "Cg_T", "Eg_in_L", "Eg_in_B", "Eg_in_R", "Eg_in_T", "Fg_in_L", "Fg_in_B",
             "Fg_in_R", "Fg_in_T", "MGQD_Src", "KapE", "KapR", "G_old", "Pold_L", "Pold_B", "Pold_R", "Pold_T",
             "Eg_avg_old", "Fxg_edgV_old", "Fyg_edgH_old", "Eg_avg", "Eg_edgV", "Eg_edgH"]
OLD_INPUTS = ["Eg_avg_old", "Eg_edgV_old", "Eg_edgH_old", "Fxg_edgV_old", "Fyg_edgH_old", "fg_avg_xx_old",
              "fg_avg_yy_old", "fg_edgV_xx_old", "fg_edgV_xy_old", "fg_edgH_yy_old", "fg_edgH_xy_old", "KapE_old",
              "KapR_old", "MGQD_Src_old", "Temp_Old", "KapE_Bar_old", "E_avg_old", "GQD_Src_old"]


def test_old_coefs_match(trt, oracle_mod):
    prob, o, ctx = _pair(trt, oracle_mod, steps=2)
    sync_to_gpu(o, ctx, OLD_INPUTS)
    for x in (o, ctx):
        x.call("old_mgqd_coefs")
    sync_to_gpu(o, ctx, ["G_old"])
    for x in (o, ctx):
        x.call("old_grey_coefs")
    for n in ("G_old", "Pold_L", "Pold_B", "Pold_R", "Pold_T", "Gold_Hat", "Rhat_old"):
        ref = o[n].ravel(order="F")
        got = ctx.pull(n)
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(np.max(np.abs(ref)), 1e-300), n
    ctx.close()


@pytest.mark.parametrize("qd_solver,lin_maxit", [(0, 0), (1, 0), (1, 1), (2, 0)])
def test_mloqd_fv_matches(trt, oracle_mod, qd_solver, lin_maxit):
    """qd_solver 0 solves every group directly (nested dissection, reported as -1
    iterations), qd_solver 2 by block LU; with BiCGSTAB (qd_solver 1) lin_maxit = 1
    starves it (one iteration per restart), so groups go through the direct fallback."""
    prob, o, ctx = _pair(trt, oracle_mod, steps=1, lin_maxit=lin_maxit, qd_solver=qd_solver)
    sync_to_gpu(o, ctx, ML_INPUTS)
    if lin_maxit == 1:
        # an initial guess that is not already the solution
        for n in ("Eg_avg", "Eg_edgV", "Eg_edgH"):
            ctx.push(n, 1.5 * o[n])
    o.call("mloqd_fv")
    kits = ctx.mloqd_fv()
    assert len(kits) == prob.ng, kits
    if qd_solver != 1:
        assert all(k == -1 for k in kits), kits
    else:
        assert (-1 in kits) if lin_maxit == 1 else all(k >= 0 for k in kits), kits
    for n in ("Eg_avg", "Eg_edgV", "Eg_edgH", "Fxg_edgV", "Fyg_edgH"):
        assert rel_err(ctx.pull(n), o[n].ravel(order="F")) < 1e-10, n
    ctx.close()


@pytest.mark.parametrize("nx,ny", [(2, 2), (3, 17), (17, 3), (5, 5), (8, 8), (33, 6)])
def test_mloqd_fv_grid_shapes(trt, oracle_mod, nx, ny):
    """Nested dissection on grids whose trees differ: one leaf holding everything
    (2x2: no border), thin strips, ragged leaves, a power-of-two square."""
    prob = trt.problem.small(nx=nx, ny=ny, order=4, ng=2, seed=nx * 31 + ny)
    o = oracle_mod.Oracle(prob, threads=2)
    o.step()
    ctx = trt.native.Context(prob, qd_solver=0)
    sync_to_gpu(o, ctx, ML_INPUTS)
    o.call("mloqd_fv")
    kits = ctx.mloqd_fv()
    assert all(k == -1 for k in kits), kits
    for n in ("Eg_avg", "Eg_edgV", "Eg_edgH", "Fxg_edgV", "Fyg_edgH"):
        assert rel_err(ctx.pull(n), o[n].ravel(order="F")) < 1e-10, n
    ctx.close()


def test_grey_chain_matches(trt, oracle_mod):
    prob, o, ctx = _pair(trt, oracle_mod, steps=1)
    names = ML_INPUTS + ["Fxg_edgV", "Fyg_edgH", "KapE_Bar"]
    sync_to_gpu(o, ctx, names)
    for x in (o, ctx):
        x.call("collapse_mg_ef")
        x.call("cbar_calc")
        x.call("grey_coefs")
    for n in ("MGQD_E_avg", "MGQD_E_edgV", "MGQD_E_edgH", "MGQD_Fx_edgV", "MGQD_Fy_edgH", "Cb_L", "Cb_B", "Cb_R",
              "Cb_T", "KapE_Bar", "DC_xx", "DL_xx", "DR_xx", "DC_yy", "DB_yy", "DT_yy", "DL_xy", "DR_xy", "DB_xy",
              "DT_xy", "PL", "PR", "PB", "PT"):
        ref = o[n].ravel(order="F")
        assert np.max(np.abs(ctx.pull(n) - ref)) <= 1e-13 * max(np.max(np.abs(ref)), 1e-300), n
    ctx.close()


GREY_STATE = ["E_avg", "E_edgV", "E_edgH", "Fx_edgV", "Fy_edgH", "Temp", "KapE_Bar", "GQD_Src", "Rhat_old",
              "Gold_Hat", "DC_xx", "DL_xx", "DR_xx", "DC_yy", "DB_yy", "DT_yy", "DL_xy", "DR_xy", "DB_xy", "DT_xy",
              "PL", "PR", "PB", "PT", "Cb_L", "Cb_B", "Cb_R", "Cb_T", "E_in_L", "E_in_B", "E_in_R", "E_in_T",
              "F_in_L", "F_in_B", "F_in_R", "F_in_T", "Temp_MGQDold2", "KapE_bar_MGQDold"]


@pytest.mark.parametrize("mgqd_it,qd_solver,lin_maxit", [(1, 0, 0), (2, 0, 0), (1, 1, 0), (2, 1, 0), (1, 1, 1),
                                                        (2, 2, 0)])
def test_egp_newton_matches(trt, oracle_mod, mgqd_it, qd_solver, lin_maxit):
    """Line search off: both Newton iterations reach the same grey fixed point (the line
    search's accept/reject decisions follow rounding noise once a delta is tiny).
    With BiCGSTAB, lin_maxit = 1 sends the grey linear solves through the direct fallback."""
    prob, o, ctx = _pair(trt, oracle_mod, steps=1, lin_maxit=lin_maxit, qd_solver=qd_solver, use_line_search=0)
    o.call("mloqd_fv")
    o.call("collapse_mg_ef")
    o.call("cbar_calc")
    o.call("grey_coefs")
    for a, b in (("E_avg", "MGQD_E_avg"), ("E_edgV", "MGQD_E_edgV"), ("E_edgH", "MGQD_E_edgH"),
                 ("Fx_edgV", "MGQD_Fx_edgV"), ("Fy_edgH", "MGQD_Fy_edgH")):
        o[a][...] = o[b]
    if mgqd_it > 1:
        o["Temp_MGQDold2"][...] = o["Temp"] * 0.97
        o["KapE_bar_MGQDold"][...] = o["KapE_Bar"] * 1.05
    sync_to_gpu(o, ctx, GREY_STATE)
    n_ref = o.L.orc_egp_fv_newt(o.h, mgqd_it)
    n_gpu = ctx.egp_fv_newt(mgqd_it)
    assert abs(n_ref - n_gpu) <= 1
    for n in ("Temp", "E_avg", "E_edgV", "E_edgH", "Fx_edgV", "Fy_edgH", "KapE_Bar", "GQD_Src"):
        assert rel_err(ctx.pull(n), o[n].ravel(order="F")) < 1e-9, n
    ctx.close()


def test_convergence_matches(trt, oracle_mod):
    prob, o, ctx = _pair(trt, oracle_mod, steps=0)
    rng = np.random.default_rng(3)
    shp = (prob.nx, prob.ny)
    x, xo, xo2 = (rng.uniform(1, 2, shp) for _ in range(3))
    ctx.push("Temp", x)
    ctx.push("Temp_RTold", xo)
    ctx.push("Temp_RTold2", xo2)
    for ct in (1, 2, 3):
        for its in (1, 3):
            ref = oracle_mod.convergence(ct, 1e-2, x, xo, xo2, its)
            got = ctx.convergence("Temp", "Temp_RTold", "Temp_RTold2", ct, 1e-2, its)
            assert got[0] == ref[0] and got[1] == pytest.approx(ref[1], rel=1e-15) and got[2] == pytest.approx(ref[2], rel=1e-15)
    ctx.close()


def _l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("name,kw,tol,qd_solver", [
    ("small", dict(use_line_search=0, conv_ho=1e-11, conv_lo=1e-12, conv_gr1=1e-13), 1e-8, 0),
    ("small", dict(), 1e-5, 0),
    ("cfg0", dict(use_line_search=0, conv_ho=1e-11, conv_lo=1e-12, conv_gr1=1e-13), 1e-8, 0),
    ("cfg0", dict(use_line_search=0, conv_ho=1e-11, conv_lo=1e-12, conv_gr1=1e-13), 1e-8, 1),
    ("map20", dict(use_line_search=0, conv_ho=1e-11, conv_lo=1e-12, conv_gr1=1e-13), 1e-8, 0),
    ("map20", dict(use_line_search=0, conv_ho=1e-11, conv_lo=1e-12, conv_gr1=1e-13), 1e-8, 3),
])
def test_time_steps_match(trt, oracle_mod, name, kw, tol, qd_solver):
    prob = {"small": lambda: trt.problem.small(nx=12, ny=10, order=4, ng=3), "cfg0": trt.problem.cfg0,
            "map20": lambda: trt.problem.mapped(nx=20, order=4, ng=6)}[name]()
    for k, v in kw.items():
        setattr(prob, k, v)
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob, qd_solver=qd_solver)
    for s in range(3):
        it_o = o.step()
        it_g = ctx.step()
        for n in ("Temp", "E_avg", "HO_E_avg", "MGQD_E_avg"):
            err = _l2(ctx.pull(n), o[n].ravel(order="F"))
            assert err < tol, (s, n, err, it_o, it_g)
        for n in ("fg_avg_xx", "fg_avg_yy", "fg_avg_xy"):
            err = rel_err(ctx.pull(n), o[n].ravel(order="F"))
            assert err < 10 * tol, (s, n, err)
    assert ctx.time == pytest.approx(o.time)
    ctx.close()


def test_reused_factors_equal_fresh_factorizations(trt):
    """qd_solver 0 reuses the nested-dissection factors within a time step (iterative
    refinement to the direct solve's backward error 1e-14), qd_solver 3 factorizes every
    solve: three steps of a 24x24 map problem agree to within the iteration tolerances (the
    two forms differ by rounding, which can move a convergence test by an iteration), and
    most solves reused."""
    prob = trt.problem.mapped(nx=24, order=4, ng=6, sigma=1.0)
    prob.use_line_search, prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 0, 1e-11, 1e-12, 1e-13
    a = trt.native.Context(prob, qd_solver=0)
    b = trt.native.Context(prob, qd_solver=3)
    for s in range(3):
        a.step()
        b.step()
        for n in ("Temp", "E_avg", "HO_E_avg", "MGQD_E_avg", "fg_avg_xx", "Eg_avg"):
            assert rel_err(a.pull(n), b.pull(n)) < 1e-9, (s, n)
    fact, reused, steps = a.qd_stats()
    assert reused > fact > 0 and steps >= reused
    assert b.qd_stats()[1] == 0
    a.close()
    b.close()


def test_small_front_forms_bitwise(trt, monkeypatch):
    """The small-front factorization compiled for a level's pivot count (48 at the 4x4-cell
    leaves of a 32x32 grid, 4 and 8 at the separator levels above them) and the form with
    runtime sizes (TRT_ND_SMALL_GENERIC=1) do the same operations on every stored entry:
    two steps that factorize every solve end in bitwise equal states."""
    prob = trt.problem.mapped(nx=32, order=4, ng=6, sigma=1.0)
    prob.use_line_search = 0
    names = ("Temp", "E_avg", "HO_E_avg", "MGQD_E_avg", "Eg_avg", "Eg_edgV", "Eg_edgH")
    out = []
    for generic in ("0", "1"):
        monkeypatch.setenv("TRT_ND_SMALL_GENERIC", generic)
        c = trt.native.Context(prob, qd_solver=3)
        for _ in range(2):
            c.step()
        out.append({n: c.pull(n).copy() for n in names})
        c.close()
    for n in names:
        assert np.all(np.isfinite(out[0][n])), n
        assert np.array_equal(out[0][n], out[1][n]), n

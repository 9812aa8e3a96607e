"""Pins the oracle's reduced QD system (QD_FV), MLOQD_FV, EGP_FV_NEWT,
CONVERGENCE and the time step, through conservation residuals of the
un-reduced equations, direct sparse solves and equilibrium (CPU only)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _rand_coefs(nx, ny, seed):
    rng = np.random.default_rng(seed)
    names = [
        "EB_L", "EB_B", "EB_C", "EB_R", "EB_T", "MBx_C", "MBx_R", "MBx_B", "MBx_T",
        "MBy_C", "MBy_T", "MBy_L", "MBy_R", "EB_RHS", "MBx_RHS", "MBy_RHS"]
    coef = {n: rng.uniform(-1, 1, (nx, ny)) for n in names}
    coef["EB_C"] = rng.uniform(6, 9, (nx, ny))
    coef["MBx_R"] = -rng.uniform(6, 9, (nx, ny))
    coef["MBy_T"] = -rng.uniform(6, 9, (nx, ny))
    bnd = {n: rng.uniform(-1, 1, nx if n[-1] in "BT" else ny) for n in
           ("Cp_L", "Cp_B", "Cp_R", "Cp_T", "BC_L", "BC_B", "BC_R", "BC_T")}
    for n in ("Cp_L", "Cp_B", "Cp_R", "Cp_T"):
        bnd[n] = -5 - np.abs(bnd[n])
    return coef, bnd


@pytest.mark.parametrize("nx,ny", [(2, 2), (3, 5), (7, 4)])
def test_qd_csr_shape_and_direct_solve(oracle_mod, nx, ny):
    coef, bnd = _rand_coefs(nx, ny, nx * 10 + ny)
    ia, ja, a, rhs = oracle_mod.qd_build_csr(nx, ny, coef, bnd)
    nrows = nx + ny + 3 * nx * ny                  # QD_SOLVERS.f08:62
    assert ia.shape == (nrows + 1,) and ia[-1] == len(a) <= 19 * nx * ny + nx + ny   # :63
    A = sp.csr_matrix((a, ja, ia), shape=(nrows, nrows))
    # every unknown owns exactly one equation; rows are never empty
    assert np.all(np.diff(ia) > 0)
    x = spla.spsolve(A.tocsc(), rhs)
    Ea, Ev, Eh = oracle_mod.qd_fv(nx, ny, coef, bnd)
    # scatter in the reference ordering (QS:512-534)
    sol = np.empty(nrows)
    sol[0] = Ev[0, 0]
    for i in range(1, nx + 1):
        k = 4 * i + 1
        sol[k - 4], sol[k - 3], sol[k - 2], sol[k - 1] = Eh[i - 1, 0], Ea[i - 1, 0], Ev[i, 0], Eh[i - 1, 1]
    row_n = 4 * nx + 1
    for j in range(2, ny + 1):
        sol[row_n + 4 - 4] = Ev[0, j - 1]
        for i in range(1, nx + 1):
            k = row_n + 3 * i + 1
            sol[k - 3], sol[k - 2], sol[k - 1] = Ea[i - 1, j - 1], Ev[i, j - 1], Eh[i - 1, j]
        row_n += 3 * nx + 1
    assert np.max(np.abs(sol - x)) <= 1e-12 * np.max(np.abs(x))


def _stepped(trt, oracle_mod, steps=1, **kw):
    prob = trt.problem.small(nx=7, ny=6, order=8, ng=3)
    for k, v in kw.items():
        setattr(prob, k, v)
    o = oracle_mod.Oracle(prob, threads=3)
    for _ in range(steps):
        o.step()
    return prob, o


def test_mloqd_satisfies_unreduced_balances(trt, oracle_mod):
    """After MLOQD_FV the recovered E_g, F_g satisfy all five balance equations and the
    boundary conditions of MLOQD_SOLVES.f08:280-339 to rounding."""
    prob, o = _stepped(trt, oracle_mod)
    o.call("mloqd_fv")
    c, dt, th = prob.c, prob.dt, prob.theta
    dx, dy, A = o["Delx"], o["Dely"], o["A"]
    for g in range(prob.ng):
        E, Ev, Eh = o["Eg_avg"][:, :, g], o["Eg_edgV"][:, :, g], o["Eg_edgH"][:, :, g]
        Fx, Fy = o["Fxg_edgV"][:, :, g], o["Fyg_edgH"][:, :, g]
        fxx, fyy = o["fg_avg_xx"][:, :, g], o["fg_avg_yy"][:, :, g]
        fvxx, fvxy = o["fg_edgV_xx"][:, :, g], o["fg_edgV_xy"][:, :, g]
        fhyy, fhxy = o["fg_edgH_yy"][:, :, g], o["fg_edgH_xy"][:, :, g]
        KE, KR, S = o["KapE"][:, :, g], o["KapR"][:, :, g], o["MGQD_Src"][:, :, g]
        Eo, Fxo, Fyo = o["Eg_avg_old"][:, :, g], o["Fxg_edgV_old"][:, :, g], o["Fyg_edgH_old"][:, :, g]
        G, PL, PB, PR, PT = (o[n][:, :, g] for n in ("G_old", "Pold_L", "Pold_B", "Pold_R", "Pold_T"))
        DY, DX = dy[None, :], dx[:, None]
        t_eb = [A * (1 / (th * dt) + c * KE) * E, DY * (Fx[1:] - Fx[:-1]), DX * (Fy[:, 1:] - Fy[:, :-1]),
                A * S + A * Eo / (th * dt) + G]
        res1 = t_eb[0] + t_eb[1] + t_eb[2] - t_eb[3]
        xy_h = c * DX * (fhxy[:, 1:] * Eh[:, 1:] - fhxy[:, :-1] * Eh[:, :-1]) / 2
        xy_v = c * DY * (fvxy[1:] * Ev[1:] - fvxy[:-1] * Ev[:-1]) / 2
        k2 = A * (1 / (th * c * dt) + KR) / 2
        res2 = k2 * Fx[:-1] + c * DY * (fxx * E - fvxx[:-1] * Ev[:-1]) + xy_h - (A * Fxo[:-1] / (2 * th * c * dt) + PL)
        res3 = k2 * Fy[:, :-1] + c * DX * (fyy * E - fhyy[:, :-1] * Eh[:, :-1]) + xy_v - (A * Fyo[:, :-1] / (2 * th * c * dt) + PB)
        res4 = k2 * Fx[1:] + c * DY * (fvxx[1:] * Ev[1:] - fxx * E) + xy_h - (A * Fxo[1:] / (2 * th * c * dt) + PR)
        res5 = k2 * Fy[:, 1:] + c * DX * (fhyy[:, 1:] * Eh[:, 1:] - fyy * E) + xy_v - (A * Fyo[:, 1:] / (2 * th * c * dt) + PT)
        scale = max(np.max(np.abs(t)) for t in t_eb)
        assert np.max(np.abs(res1)) < 1e-10 * scale
        mscale = np.max(np.abs(k2 * Fx[:-1])) + np.max(np.abs(c * DY * fxx * E))
        for r in (res2, res3, res4, res5):
            assert np.max(np.abs(r)) < 1e-10 * mscale
        bl = Fx[0] - c * o["Cg_L"][:, g] * (Ev[0] - o["Eg_in_L"][:, g]) - o["Fg_in_L"][:, g]
        br = Fx[-1] - c * o["Cg_R"][:, g] * (Ev[-1] - o["Eg_in_R"][:, g]) - o["Fg_in_R"][:, g]
        bb = Fy[:, 0] - c * o["Cg_B"][:, g] * (Eh[:, 0] - o["Eg_in_B"][:, g]) - o["Fg_in_B"][:, g]
        bt = Fy[:, -1] - c * o["Cg_T"][:, g] * (Eh[:, -1] - o["Eg_in_T"][:, g]) - o["Fg_in_T"][:, g]
        fs = np.max(np.abs(Fx)) + np.max(np.abs(Fy)) + c * np.max(np.abs(Ev))
        for r in (bl, br, bb, bt):
            assert np.max(np.abs(r)) < 1e-10 * fs


def test_egp_newton_reaches_the_grey_fixed_point(trt, oracle_mod):
    """With dKapE_Bar/dT = 0 (first MLOQD iterate) the grey system is linear in (E, F, T);
    the Newton result satisfies every residual of GLOQD_SOLVES.f08:920-994 and the grey
    boundary conditions.  The line search is off here: it rescales every delta whenever
    one delta norm grows (GL:604-636), and once a delta is at rounding level that decision
    follows rounding noise, which can stop the iteration short of the fixed point."""
    prob, o = _stepped(trt, oracle_mod, use_line_search=0)
    o.call("mloqd_fv")
    o.call("collapse_mg_ef")
    o.call("cbar_calc")
    o.call("grey_coefs")
    for a, b in (("E_avg", "MGQD_E_avg"), ("E_edgV", "MGQD_E_edgV"), ("E_edgH", "MGQD_E_edgH"),
                 ("Fx_edgV", "MGQD_Fx_edgV"), ("Fy_edgH", "MGQD_Fy_edgH")):
        o[a][...] = o[b]
    its = o.L.orc_egp_fv_newt(o.h, 1)
    assert 1 <= its < prob.maxit_gloqd
    c, dt, th, cv = prob.c, prob.dt, prob.theta, prob.cv
    dx, dy, A = o["Delx"][:, None], o["Dely"][None, :], o["A"]
    T, E, Ev, Eh, Fx, Fy = (o[n] for n in ("Temp", "E_avg", "E_edgV", "E_edgH", "Fx_edgV", "Fy_edgH"))
    KB, Q = o["KapE_Bar"], o["GQD_Src"]
    D = {n: o[n] for n in ("DC_xx", "DL_xx", "DR_xx", "DC_yy", "DB_yy", "DT_yy", "DL_xy", "DR_xy", "DB_xy", "DT_xy",
                           "PL", "PR", "PB", "PT")}
    drT = cv / (th * dt) * T + Q - c * KB * E - o["Rhat_old"]
    drB = A * (1 / (th * dt) + c * KB) * E + dy * (Fx[1:] - Fx[:-1]) + dx * (Fy[:, 1:] - Fy[:, :-1]) - A * Q - o["Gold_Hat"]
    drML = A * Fx[:-1] / 2 + c * dy * (D["DC_xx"] * E - D["DL_xx"] * Ev[:-1]) + c * dx * (D["DT_xy"] * Eh[:, 1:] - D["DB_xy"] * Eh[:, :-1]) / 2 - D["PL"]
    drMR = A * Fx[1:] / 2 + c * dy * (D["DR_xx"] * Ev[1:] - D["DC_xx"] * E) + c * dx * (D["DT_xy"] * Eh[:, 1:] - D["DB_xy"] * Eh[:, :-1]) / 2 - D["PR"]
    drMB = A * Fy[:, :-1] / 2 + c * dx * (D["DC_yy"] * E - D["DB_yy"] * Eh[:, :-1]) + c * dy * (D["DR_xy"] * Ev[1:] - D["DL_xy"] * Ev[:-1]) / 2 - D["PB"]
    drMT = A * Fy[:, 1:] / 2 + c * dx * (D["DT_yy"] * Eh[:, 1:] - D["DC_yy"] * E) + c * dy * (D["DR_xy"] * Ev[1:] - D["DL_xy"] * Ev[:-1]) / 2 - D["PT"]
    assert np.max(np.abs(drT)) < 1e-9 * np.max(np.abs(cv / (th * dt) * T))
    assert np.max(np.abs(drB)) < 1e-9 * np.max(np.abs(A * (1 / (th * dt) + c * KB) * E))
    ms = np.max(np.abs(A * Fx[:-1])) + np.max(np.abs(c * dy * D["DC_xx"] * E))
    for r in (drML, drMR, drMB, drMT):
        assert np.max(np.abs(r)) < 1e-9 * ms
    # the exact-Newton converged state satisfies the grey boundary conditions F = c Cb (E - E_in) + F_in
    bl = Fx[0] - c * o["Cb_L"] * (Ev[0] - o["E_in_L"]) - o["F_in_L"]
    assert np.max(np.abs(bl)) < 1e-9 * (np.max(np.abs(Fx)) + c * np.max(np.abs(Ev)))


def test_convergence_hand_cases(oracle_mod):
    x = np.array([1.0, 2.0, 4.0])
    xo = np.array([1.0, 2.1, 4.0])
    xo2 = np.array([1.0, 2.5, 4.0])
    ok, nrm, rho = oracle_mod.convergence(1, 1e-3, x, xo, xo2, 1)
    assert rho == 2.0 and nrm == pytest.approx(0.1 / 4.0) and not ok
    ok, nrm, rho = oracle_mod.convergence(1, 1.0, x, xo, xo2, 2)
    assert rho == pytest.approx(0.1 / 0.4) and ok == (0.025 <= abs(1 / rho - 1) * 1.0)
    ok, nrm, rho = oracle_mod.convergence(2, 1.0, x, xo, xo2, 2)
    assert nrm == pytest.approx(abs(1 - 2.1 / 2.0))
    ok, nrm, rho = oracle_mod.convergence(3, 1.0, x, xo, xo2, 2)
    assert nrm == pytest.approx(0.1 / 1.0)


def test_whole_step_equilibrium_with_telescoping_groups(trt, oracle_mod):
    """Groups spanning [0, inf), unit multipliers and the Planck-mean opacity: sum_g 4 pi kap_g B_g
    equals the analytic grey emission (GLOQD_SOLVES.f08:439), so a problem in equilibrium stays there."""
    nu = np.array([0.0, 0.4, 1.3, 3.0, 8.0, 1e10])
    prob = trt.problem.Problem(nx=5, ny=4, order=4, ng=5, nu=nu, nsteps=2, bcT=(1.0,) * 4, tini=1.0,
                               opacity_model=0, name="eq")
    o = oracle_mod.Oracle(prob, threads=2)
    T0 = o["Temp"].copy()
    E0 = o["E_avg"].copy()
    for _ in range(2):
        o.step()
    assert np.max(np.abs(o["Temp"] / T0 - 1)) < 1e-9
    assert np.max(np.abs(o["E_avg"] / E0 - 1)) < 1e-9
    assert np.max(np.abs(o["fg_avg_xy"])) < 1e-12


def test_cfg0_runs_and_heats_from_the_left(trt, oracle_mod):
    prob = trt.problem.cfg0()
    o = oracle_mod.Oracle(prob, threads=4)
    for _ in range(3):
        its = o.step()
        assert its[0] >= 2 and its[3] >= 1
    T = o["Temp"]
    assert T[0].mean() > T[-1].mean()
    assert np.all(np.isfinite(T)) and np.all(T > 0)

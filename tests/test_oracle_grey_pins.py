"""Pins the oracle's grey-level coefficient routines -- GREY_COEFS (GLOQD_SOLVES.f08:47-146),
Cbar_Calc (:177-231), OLD_GREY_COEFS (:12-42), COLLAPSE_MG_EF (MLOQD_SOLVES.f08:687-730)
and GQD_In_Calc (GLOQD_SOLVES.f08:263-290) -- by what the mathematics fixes:

* The grey QD equations are the group sums of the multigroup ones.  Each multigroup
  momentum equation divided by its kappa~_R,g, summed over g, is the grey momentum
  equation with the E_g-weighted D and P coefficients; the summed balance equations are
  the grey balance with kappa_bar_E E = sum_g kappa_E,g E_g and Gold_Hat; the summed
  boundary conditions are the grey ones with C_bar.  So once MLOQD_FV has solved the
  multigroup systems, (E, F) = (sum_g E_g, sum_g F_g) satisfies every grey residual
  (with A sum_g MGQD_Src in place of A Q_bar) to rounding, on any state.
* With theta = 1 the grey balance plus A times the material energy equation is a
  conservation law: sum over cells of A [(E - E_old) + c_v (T - T_old)] / dt plus the
  net outflow through the boundary vanishes at the Newton solution.

Each pin is shown to catch a planted mistake: the oracle compiled with ORC_MUTANT = k
(an index slip, a wrong weight, a wrong time level) fails it (CPU only)."""
import numpy as np
import pytest


def _prob(trt, theta=1.0):
    prob = trt.problem.cfg0()
    prob.theta = theta
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    return prob


def grey_sum_residuals(o, prob):
    """Relative max residuals of the grey equations evaluated on the collapsed multigroup
    solution, with the oracle's own grey coefficients."""
    o.call("mloqd_fv")                          # multigroup systems solved on the current factors
    o.call("collapse_mg_ef")
    o.call("cbar_calc")
    o.call("grey_coefs")
    ng = prob.ng
    # the old grey energy is the sum of the old group energies (true of the reference's
    # state only to the step's tolerance)
    o["E_avg_old"][...] = o["Eg_avg_old"].sum(axis=2)
    o.call("old_grey_coefs")
    c, dt, th = prob.c, prob.dt, prob.theta
    dx, dy, A = o["Delx"][:, None], o["Dely"][None, :], o["A"]
    E, Ev, Eh = o["MGQD_E_avg"], o["MGQD_E_edgV"], o["MGQD_E_edgH"]
    Fx, Fy = o["MGQD_Fx_edgV"], o["MGQD_Fy_edgH"]
    KB = o["KapE_Bar"]
    S = o["MGQD_Src"].sum(axis=2) if ng > 1 else o["MGQD_Src"]
    D = {n: o[n] for n in ("DC_xx", "DL_xx", "DR_xx", "DC_yy", "DB_yy", "DT_yy", "DL_xy", "DR_xy", "DB_xy", "DT_xy",
                           "PL", "PR", "PB", "PT")}
    bal = [A * (1 / (th * dt) + c * KB) * E, dy * (Fx[1:] - Fx[:-1]), dx * (Fy[:, 1:] - Fy[:, :-1]), A * S,
           o["Gold_Hat"]]
    drB = bal[0] + bal[1] + bal[2] - bal[3] - bal[4]
    xy_h = c * dx * (D["DT_xy"] * Eh[:, 1:] - D["DB_xy"] * Eh[:, :-1]) / 2
    xy_v = c * dy * (D["DR_xy"] * Ev[1:] - D["DL_xy"] * Ev[:-1]) / 2
    drML = A * Fx[:-1] / 2 + c * dy * (D["DC_xx"] * E - D["DL_xx"] * Ev[:-1]) + xy_h - D["PL"]
    drMR = A * Fx[1:] / 2 + c * dy * (D["DR_xx"] * Ev[1:] - D["DC_xx"] * E) + xy_h - D["PR"]
    drMB = A * Fy[:, :-1] / 2 + c * dx * (D["DC_yy"] * E - D["DB_yy"] * Eh[:, :-1]) + xy_v - D["PB"]
    drMT = A * Fy[:, 1:] / 2 + c * dx * (D["DT_yy"] * Eh[:, 1:] - D["DC_yy"] * E) + xy_v - D["PT"]
    ms = np.max(np.abs(A * Fx[:-1])) + np.max(np.abs(c * dy * D["DC_xx"] * E)) + np.max(np.abs(c * dx * D["DC_yy"] * E))
    out = {"drB": np.max(np.abs(drB)) / max(np.max(np.abs(b)) for b in bal)}
    for k, r in (("drML", drML), ("drMR", drMR), ("drMB", drMB), ("drMT", drMT)):
        out[k] = np.max(np.abs(r)) / ms
    fs = np.max(np.abs(Fx)) + np.max(np.abs(Fy)) + c * max(np.max(np.abs(Ev)), np.max(np.abs(Eh)))
    for side, F, Eb in (("L", Fx[0], Ev[0]), ("R", Fx[-1], Ev[-1]), ("B", Fy[:, 0], Eh[:, 0]), ("T", Fy[:, -1], Eh[:, -1])):
        r = F - c * o["Cb_" + side] * (Eb - o["E_in_" + side]) - o["F_in_" + side]
        out["bc" + side] = np.max(np.abs(r)) / fs
    return out


def energy_residual(o, prob):
    """One more step (theta = 1); relative defect of the discrete energy conservation law."""
    T_old = o["Temp"].copy()
    o.step()
    E_old, c, dt = o["E_avg_old"], prob.c, prob.dt
    dx, dy, A = o["Delx"], o["Dely"], o["A"]
    E, T, Fx, Fy = o["E_avg"], o["Temp"], o["Fx_edgV"], o["Fy_edgH"]
    store = A * (E - E_old) / dt
    heat = A * prob.cv * (T - T_old) / dt
    out_x = dy * (Fx[-1] - Fx[0])
    out_y = dx * (Fy[:, -1] - Fy[:, 0])
    total = store.sum() + heat.sum() + out_x.sum() + out_y.sum()
    scale = np.abs(store).sum() + np.abs(heat).sum() + np.abs(dy * Fx[-1]).sum() + np.abs(dy * Fx[0]).sum() \
        + np.abs(dx * Fy[:, -1]).sum() + np.abs(dx * Fy[:, 0]).sum()
    return abs(total) / scale


@pytest.mark.parametrize("theta", [1.0, 0.75])
def test_grey_equations_are_the_group_sums(trt, oracle_mod, theta):
    prob = _prob(trt, theta)
    o = oracle_mod.Oracle(prob, threads=4)
    for _ in range(2):
        o.step()
    if theta != 1.0:
        assert np.max(np.abs(o["Pold_L"])) > 0 and np.max(np.abs(o["G_old"])) > 0
    res = grey_sum_residuals(o, prob)
    assert max(res.values()) < 1e-10, res


def test_step_conserves_energy(trt, oracle_mod):
    prob = _prob(trt)
    o = oracle_mod.Oracle(prob, threads=4)
    o.step()
    assert energy_residual(o, prob) < 1e-9


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("mutant,what", [
    (1, "DR_xx reads face i instead of i+1"), (2, "DT_xy reads face j instead of j+1"),
    (3, "PL takes the old flux of face i+1"), (4, "DC_yy normalised by the V-face weight"),
    (5, "C_bar weights without E_in"), (6, "the collapsed F_x sums the old fluxes"),
    (7, "Gold_Hat takes the new grey energy"),
])
def test_grey_pin_catches_planted_mistake(trt, oracle_mod, mutant_dir, mutant, what):
    lib = oracle_mod.build_mutant(mutant, mutant_dir)
    prob = _prob(trt, 0.75 if mutant == 3 else 1.0)
    o = oracle_mod.Oracle(prob, threads=4, lib_path=lib)
    for _ in range(2):
        o.step()
    res = grey_sum_residuals(o, prob)
    assert max(res.values()) > 100 * 1e-10, (what, res)      # a hundred times the pin's tolerance


def test_energy_pin_catches_planted_mistake(trt, oracle_mod, mutant_dir):
    lib = oracle_mod.build_mutant(8, mutant_dir)          # Rhat_old with twice the heat capacity
    prob = _prob(trt)
    o = oracle_mod.Oracle(prob, threads=4, lib_path=lib)
    o.step()
    assert energy_residual(o, prob) > 100 * 1e-9

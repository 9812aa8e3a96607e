"""Residuals of the un-reduced low-order equations, evaluated on arrays pulled from either
implementation (a `get(name)` returning the reference's Fortran shape): the checks that hold
at sizes where no direct solve by the oracle is possible."""
import numpy as np


def _rel(num, den):
    """num / den, with 0 / 0 = 0 (a group whose every term is zero is satisfied exactly);
    a nonzero residual over a zero scale is inf, i.e. a failure."""
    num, den = float(num), float(den)
    if num == 0.0:
        return 0.0
    return num / den if den > 0.0 else float("inf")


def mloqd_residuals(get, prob, g):
    """Relative max residuals of the five balances and four boundary conditions of group g
    (MLOQD_SOLVES.f08:280-339) after MLOQD_FV."""
    c, dt, th = prob.c, prob.dt, prob.theta
    dx, dy, A = get("Delx"), get("Dely"), get("A")
    E, Ev, Eh = get("Eg_avg")[:, :, g], get("Eg_edgV")[:, :, g], get("Eg_edgH")[:, :, g]
    Fx, Fy = get("Fxg_edgV")[:, :, g], get("Fyg_edgH")[:, :, g]
    fxx, fyy = get("fg_avg_xx")[:, :, g], get("fg_avg_yy")[:, :, g]
    fvxx, fvxy = get("fg_edgV_xx")[:, :, g], get("fg_edgV_xy")[:, :, g]
    fhyy, fhxy = get("fg_edgH_yy")[:, :, g], get("fg_edgH_xy")[:, :, g]
    KE, KR, S = get("KapE")[:, :, g], get("KapR")[:, :, g], get("MGQD_Src")[:, :, g]
    Eo, Fxo, Fyo = get("Eg_avg_old")[:, :, g], get("Fxg_edgV_old")[:, :, g], get("Fyg_edgH_old")[:, :, g]
    G, PL, PB, PR, PT = (get(n)[:, :, g] for n in ("G_old", "Pold_L", "Pold_B", "Pold_R", "Pold_T"))
    DY, DX = dy[None, :], dx[:, None]
    t_eb = [A * (1 / (th * dt) + c * KE) * E, DY * (Fx[1:] - Fx[:-1]), DX * (Fy[:, 1:] - Fy[:, :-1]),
            A * S + A * Eo / (th * dt) + G]
    out = {"balance": _rel(np.max(np.abs(t_eb[0] + t_eb[1] + t_eb[2] - t_eb[3])), max(np.max(np.abs(t)) for t in t_eb))}
    xy_h = c * DX * (fhxy[:, 1:] * Eh[:, 1:] - fhxy[:, :-1] * Eh[:, :-1]) / 2
    xy_v = c * DY * (fvxy[1:] * Ev[1:] - fvxy[:-1] * Ev[:-1]) / 2
    k2 = A * (1 / (th * c * dt) + KR) / 2
    r = {"ML": k2 * Fx[:-1] + c * DY * (fxx * E - fvxx[:-1] * Ev[:-1]) + xy_h - (A * Fxo[:-1] / (2 * th * c * dt) + PL),
         "MB": k2 * Fy[:, :-1] + c * DX * (fyy * E - fhyy[:, :-1] * Eh[:, :-1]) + xy_v - (A * Fyo[:, :-1] / (2 * th * c * dt) + PB),
         "MR": k2 * Fx[1:] + c * DY * (fvxx[1:] * Ev[1:] - fxx * E) + xy_h - (A * Fxo[1:] / (2 * th * c * dt) + PR),
         "MT": k2 * Fy[:, 1:] + c * DX * (fhyy[:, 1:] * Eh[:, 1:] - fyy * E) + xy_v - (A * Fyo[:, 1:] / (2 * th * c * dt) + PT)}
    ms = np.max(np.abs(k2 * Fx[:-1])) + np.max(np.abs(c * DY * fxx * E)) + np.max(np.abs(c * DX * fyy * E))
    for k, v in r.items():
        out[k] = _rel(np.max(np.abs(v)), ms)
    fs = np.max(np.abs(Fx)) + np.max(np.abs(Fy)) + c * max(np.max(np.abs(Ev)), np.max(np.abs(Eh)))
    for side, F, Eb in (("L", Fx[0], Ev[0]), ("R", Fx[-1], Ev[-1]), ("B", Fy[:, 0], Eh[:, 0]), ("T", Fy[:, -1], Eh[:, -1])):
        res = F - c * get("Cg_" + side)[:, g] * (Eb - get("Eg_in_" + side)[:, g]) - get("Fg_in_" + side)[:, g]
        out["bc" + side] = _rel(np.max(np.abs(res)), fs)
    return out


def grey_residuals(get, prob):
    """Relative max residuals of the grey equations (GLOQD_SOLVES.f08:920-994) and the grey
    boundary conditions at the state EGP_FV_NEWT left."""
    c, dt, th, cv = prob.c, prob.dt, prob.theta, prob.cv
    dx, dy, A = get("Delx")[:, None], get("Dely")[None, :], get("A")
    T, E, Ev, Eh, Fx, Fy = (get(n) for n in ("Temp", "E_avg", "E_edgV", "E_edgH", "Fx_edgV", "Fy_edgH"))
    KB, Q = get("KapE_Bar"), get("GQD_Src")
    D = {n: get(n) for n in ("DC_xx", "DL_xx", "DR_xx", "DC_yy", "DB_yy", "DT_yy", "DL_xy", "DR_xy", "DB_xy", "DT_xy",
                             "PL", "PR", "PB", "PT")}
    out = {}
    out["drT"] = _rel(np.max(np.abs(cv / (th * dt) * T + Q - c * KB * E - get("Rhat_old"))), np.max(np.abs(cv / (th * dt) * T)))
    bal = A * (1 / (th * dt) + c * KB) * E
    drB = bal + dy * (Fx[1:] - Fx[:-1]) + dx * (Fy[:, 1:] - Fy[:, :-1]) - A * Q - get("Gold_Hat")
    out["drB"] = _rel(np.max(np.abs(drB)), np.max(np.abs(bal)))
    xy_h = c * dx * (D["DT_xy"] * Eh[:, 1:] - D["DB_xy"] * Eh[:, :-1]) / 2
    xy_v = c * dy * (D["DR_xy"] * Ev[1:] - D["DL_xy"] * Ev[:-1]) / 2
    ms = np.max(np.abs(A * Fx[:-1])) + np.max(np.abs(c * dy * D["DC_xx"] * E)) + np.max(np.abs(c * dx * D["DC_yy"] * E))
    for k, r in (("drML", A * Fx[:-1] / 2 + c * dy * (D["DC_xx"] * E - D["DL_xx"] * Ev[:-1]) + xy_h - D["PL"]),
                 ("drMR", A * Fx[1:] / 2 + c * dy * (D["DR_xx"] * Ev[1:] - D["DC_xx"] * E) + xy_h - D["PR"]),
                 ("drMB", A * Fy[:, :-1] / 2 + c * dx * (D["DC_yy"] * E - D["DB_yy"] * Eh[:, :-1]) + xy_v - D["PB"]),
                 ("drMT", A * Fy[:, 1:] / 2 + c * dx * (D["DT_yy"] * Eh[:, 1:] - D["DC_yy"] * E) + xy_v - D["PT"])):
        out[k] = _rel(np.max(np.abs(r)), ms)
    fs = np.max(np.abs(Fx)) + np.max(np.abs(Fy)) + c * max(np.max(np.abs(Ev)), np.max(np.abs(Eh)))
    for side, F, Eb in (("L", Fx[0], Ev[0]), ("R", Fx[-1], Ev[-1]), ("B", Fy[:, 0], Eh[:, 0]), ("T", Fy[:, -1], Eh[:, -1])):
        res = F - c * get("Cb_" + side) * (Eb - get("E_in_" + side)) - get("F_in_" + side)
        out["bc" + side] = _rel(np.max(np.abs(res)), fs)
    return out

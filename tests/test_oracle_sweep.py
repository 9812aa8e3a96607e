"""Pins the oracle's SCB sweep (TRANSPORT_SCB), COLLAPSE_INTENSITIES, fg/Cg and
MGQD_In_Calc by brute force and invariants (CPU only)."""
import math

import numpy as np
import pytest


def _prob(trt, nx=5, ny=4, order=4, ng=2, seed=3):
    return trt.problem.small(nx=nx, ny=ny, order=order, ng=ng, seed=seed)


def _randomize(o, prob, seed=0, geometry=True):
    rng = np.random.default_rng(seed)
    if geometry:   # non-uniform cells
        dx = rng.uniform(0.02, 0.3, prob.nx)
        dy = rng.uniform(0.02, 0.3, prob.ny)
        o["Delx"][...] = dx
        o["Dely"][...] = dy
        o["A"][...] = np.outer(dx, dy)
    o["Temp"][...] = rng.uniform(0.5, 900.0, (prob.nx, prob.ny))
    o.call("material_update")
    o["I_crn_old"][...] = rng.uniform(0.0, 3.0, o["I_crn_old"].shape) * o["Bg"].max()


def _brute_force_sweep(o, prob):
    """Solve the global SCB system of every (group, direction) directly.

    Unknowns: the 4 corner intensities of every cell.  The corner balances are
    the residual equations of TRANSPORT_SOLVES.f08:270-294 with the face values
    taken upwind: the neighbour's corner (or the boundary value) on incoming
    faces and the cell's own corner on outgoing faces."""
    nx, ny, nm, ng = prob.nx, prob.ny, prob.nm, prob.ng
    dx, dy, A = o["Delx"], o["Dely"], o["A"]
    c, dt = prob.c, prob.dt
    kap = o["KapE"] + 1.0 / (c * dt)
    src = o["RT_Src"]
    Iold = o["I_crn_old"]
    bc = np.array([[o["I_edgV"][0, 0, m, g] if prob.ox[m] > 0 else o["I_edgV"][nx, 0, m, g] for m in range(nm)]
                   for g in range(ng)])
    bcy = np.array([[o["I_edgH"][0, 0, m, g] if prob.oy[m] > 0 else o["I_edgH"][0, ny, m, g] for m in range(nm)]
                    for g in range(ng)])
    out = np.zeros((2 * nx, 2 * ny, nm, ng))
    LB, RB, RT, LT = 0, 1, 2, 3
    corner_xy = {LB: (0, 0), RB: (1, 0), RT: (1, 1), LT: (0, 1)}
    for g in range(ng):
        for m in range(nm):
            ox, oy = prob.ox[m], prob.oy[m]
            N = 4 * nx * ny
            M = np.zeros((N, N))
            b = np.zeros(N)
            idx = lambda i, j, k: 4 * (i + nx * j) + k
            for j in range(ny):
                for i in range(nx):
                    for k in range(4):
                        cx, cy = corner_xy[k]
                        row = idx(i, j, k)
                        S = src[i, j, g] + Iold[2 * i + cx, 2 * j + cy, m, g] / (c * dt)
                        # volume term and the two half-face terms inside the cell:
                        # (sx*Ox*dy + sy*Oy*dx + kap*A)/4 * I_k + (sx*Ox*dy)/4 I_xneighbour + ...
                        sx = 1 if cx == 0 else -1       # outward normal of the x-face of this corner is -sx
                        sy = 1 if cy == 0 else -1
                        M[row, row] += (sx * ox * dy[j] + sy * oy * dx[i] + kap[i, j, g] * A[i, j]) / 4.0
                        kxn = {LB: RB, RB: LB, RT: LT, LT: RT}[k]
                        kyn = {LB: LT, LT: LB, RB: RT, RT: RB}[k]
                        M[row, idx(i, j, kxn)] += sx * ox * dy[j] / 4.0
                        M[row, idx(i, j, kyn)] += sy * oy * dx[i] / 4.0
                        b[row] += A[i, j] * S / 4.0
                        # face term -sx*Ox*dy*Ic/2 on the corner's x-face, upwinded
                        xface_in = (cx == 0 and ox > 0) or (cx == 1 and ox < 0)
                        coef = -sx * ox * dy[j] / 2.0
                        if xface_in:
                            inb = i - 1 if cx == 0 else i + 1
                            if 0 <= inb < nx:
                                M[row, idx(inb, j, kxn)] += coef
                            else:
                                b[row] -= coef * bc[g, m]
                        else:
                            M[row, row] += coef
                        yface_in = (cy == 0 and oy > 0) or (cy == 1 and oy < 0)
                        coef = -sy * oy * dx[i] / 2.0
                        if yface_in:
                            jnb = j - 1 if cy == 0 else j + 1
                            if 0 <= jnb < ny:
                                M[row, idx(i, jnb, kyn)] += coef
                            else:
                                b[row] -= coef * bcy[g, m]
                        else:
                            M[row, row] += coef
            x = np.linalg.solve(M, b)
            for j in range(ny):
                for i in range(nx):
                    for k in range(4):
                        cx, cy = corner_xy[k]
                        out[2 * i + cx, 2 * j + cy, m, g] = x[idx(i, j, k)]
    return out


@pytest.mark.parametrize("order", [4, 8])
def test_sweep_matches_global_brute_force(trt, oracle_mod, order):
    prob = _prob(trt, order=order)
    o = oracle_mod.Oracle(prob)
    _randomize(o, prob)
    o.call("transport_scb")
    ref = _brute_force_sweep(o, prob)
    got = o["I_crn"]
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


def test_sweep_face_and_average_bookkeeping(trt, oracle_mod):
    """I_avg = corner mean; interior I_edgV/I_edgH = mean of the upwind corner-edge values (TS:86-96)."""
    prob = _prob(trt, nx=4, ny=3, order=4, ng=1)
    o = oracle_mod.Oracle(prob)
    _randomize(o, prob, seed=5)
    o.call("transport_scb")
    I = o["I_crn"]
    Iavg = 0.25 * (I[0::2, 0::2] + I[1::2, 0::2] + I[1::2, 1::2] + I[0::2, 1::2])
    assert np.allclose(o["I_avg"], Iavg, rtol=1e-14, atol=0)
    nm = prob.nm
    for m in range(nm):
        if prob.ox[m] > 0:   # right face of cell i = mean of its right corners
            ref = 0.5 * (I[1::2, 0::2, m] + I[1::2, 1::2, m])
            assert np.allclose(o["I_edgV"][1:, :, m], ref, rtol=1e-14)
        else:
            ref = 0.5 * (I[0::2, 0::2, m] + I[0::2, 1::2, m])
            assert np.allclose(o["I_edgV"][:-1, :, m], ref, rtol=1e-14)


def test_constant_solution_is_preserved_on_any_grid(trt, oracle_mod):
    """Uniform T, I_old = B_g and Planckian boundaries at T: the sweep returns B_g exactly (any dx, dy)."""
    prob = _prob(trt, nx=6, ny=5, order=8, ng=3)
    prob.tini = 37.0
    prob.bcT = (37.0,) * 4
    o = oracle_mod.Oracle(prob)
    rng = np.random.default_rng(1)
    dx, dy = rng.uniform(0.01, 0.5, prob.nx), rng.uniform(0.01, 0.5, prob.ny)
    o["Delx"][...] = dx
    o["Dely"][...] = dy
    o["A"][...] = np.outer(dx, dy)
    o.call("transport_scb")
    bg = o["Bg"][0, 0, :]
    assert np.max(np.abs(o["I_crn"] / bg[None, None, None, :] - 1)) < 1e-14


def test_moments_at_equilibrium(trt, oracle_mod):
    """At equilibrium f = sum w Ox^2 / sum w, f_xy = sum w OxOy / sum w = 0, C = half-range ratios."""
    prob = _prob(trt, nx=3, ny=3, order=12, ng=2)
    prob.bcT = (prob.tini,) * 4
    o = oracle_mod.Oracle(prob)
    o.call("transport_scb")
    o.call("collapse_intensities")
    o.call("fg_calc")
    o.call("cg_calc")
    w, ox, oy = prob.w, prob.ox, prob.oy
    fxx = (w * ox ** 2).sum() / w.sum()
    assert np.allclose(o["fg_avg_xx"], fxx, rtol=1e-13)
    assert np.allclose(o["fg_edgV_xx"], fxx, rtol=1e-13)
    assert np.max(np.abs(o["fg_avg_xy"])) < 1e-14
    nm = prob.nm
    left = (ox * w)[nm // 4: 3 * nm // 4].sum() / w[nm // 4: 3 * nm // 4].sum()
    assert np.allclose(o["Cg_L"], left, rtol=1e-13)
    assert np.allclose(o["Cg_R"], -left, rtol=1e-13)
    assert fxx == pytest.approx(1 / 3, rel=3e-6)
    E = o["HO_Eg_avg"]
    assert np.allclose(E, 4 * math.pi * o["Bg"] / prob.c, rtol=1e-13)


def test_incoming_boundary_sums(trt, oracle_mod):
    """MGQD_In_Calc: E_in = sum_in w I / c, F_in = sum_in Omega_n w I with I = B_g(bcT) on the left, 0 elsewhere."""
    prob = _prob(trt, nx=3, ny=4, order=8, ng=2)
    o = oracle_mod.Oracle(prob)
    nm, w, ox = prob.nm, prob.w, prob.ox
    inc = np.r_[0:nm // 4, 3 * nm // 4:nm]
    for g in range(prob.ng):
        b = oracle_mod.Bg_planck(prob.bcT[0], prob.nu[g], prob.nu[g + 1])
        assert np.allclose(o["Eg_in_L"][:, g], w[inc].sum() * b / prob.c, rtol=1e-14)
        assert np.allclose(o["Fg_in_L"][:, g], (ox * w)[inc].sum() * b, rtol=1e-14)
    assert np.all(o["Eg_in_R"] == 0) and np.all(o["Fg_in_T"] == 0)

"""Reflective boundaries (BC_Type 1) on the GPU against the oracle, through the C ABI:
RT_BC_UPDATE (MainSrc/UPDATES.f08:12-105) after each sweep, the sweep reading the reflected
inflow, the boundary-face moments of COLLAPSE_INTENSITIES, Cg_Calc and MGQD_In_Calc (zero on a
reflective side), the whole time step (ALGORITHMS.f08:541-572), the Tgen_QDf step (no
update: the initial Planckian inflow stays), and a restart (the inflow back to RT_INIT's).

The tile engine keeps a reflective side's corner intensities in boundary slices laid out
in the quadrant's swept frame; the grids below have partial tiles on both ends so that the
frame flips (the mirror quadrant's rows and columns counted from the other end) are covered.
At a larger size the GPU solution of a half problem with a reflective side equals the GPU
solution of the doubled, mirrored problem (the mirror pin of the oracle, on the GPU path)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

SIDES = [(1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1), (1, 1, 1, 1)]
VIEWS = ("I_edgV", "I_edgH", "Ic_edgV", "Ic_edgH")
MOMENTS = ("Hg_edgV_xx", "Hg_edgV_xy", "HO_Eg_edgV", "HO_Fxg_edgV", "Hg_edgH_yy", "Hg_edgH_xy", "HO_Eg_edgH",
           "HO_Fyg_edgH", "Hg_avg_xx", "HO_Eg_avg")


def _pair(trt, oracle_mod, bc_type, tile, mode, precision=0, nx=13, ny=11, order=8, ng=3, seed=3):
    prob = trt.problem.small(nx=nx, ny=ny, order=order, ng=ng, seed=seed)
    prob = prob.with_(bc_type=tuple(bc_type), bcT=(1000.0, 300.0, 500.0, 200.0))
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob, precision=precision, tile=tile, sweep_mode=mode)
    rng = np.random.default_rng(seed)
    T = rng.uniform(0.5, 800.0, size=(nx, ny))
    o["Temp"][...] = T
    ctx.push("Temp", T)
    o.call("material_update")
    ctx.call("material_update")
    bg = o["Bg"]
    Iold = np.asfortranarray(rng.uniform(0.2, 2.0, size=(2 * nx, 2 * ny, prob.nm, ng)))
    Iold *= np.repeat(np.repeat(bg, 2, axis=0), 2, axis=1)[:, :, None, :]
    o["I_crn_old"][...] = Iold
    ctx.push("I_crn_old", Iold)
    return prob, o, ctx


@pytest.mark.parametrize("bc_type", SIDES)
@pytest.mark.parametrize("tile,mode", [((2, 4), 2), ((4, 2), 3), ((2, 2), 4)])
def test_reflective_sweeps_match_oracle(trt, oracle_mod, bc_type, tile, mode):
    """Three sweeps with RT_BC_UPDATE after each: the second and third read reflected inflow."""
    prob, o, ctx = _pair(trt, oracle_mod, bc_type, tile, mode)
    for sweep in range(3):
        o.call("transport_scb")
        ctx.call("transport_scb")
        o.call("collapse_intensities")
        ctx.call("collapse_intensities")
        assert rel_err(ctx.pull("I_crn", o["I_crn"].shape), o["I_crn"]) < 1e-12, sweep
        for name in MOMENTS:
            assert rel_err(ctx.pull(name), o[name].ravel(order="F")) < 1e-12, (sweep, name)
        o.call("rt_bc_update")
        ctx.call("rt_bc_update")
        for name in VIEWS:
            assert rel_err(ctx.pull(name), o[name].ravel(order="F")) < 1e-12, (sweep, name)
        for name in ("mgqd_in_calc", "cg_calc"):
            o.call(name)
            ctx.call(name)
        for name in ("Cg_L", "Cg_B", "Cg_R", "Cg_T", "Eg_in_L", "Eg_in_B", "Eg_in_R", "Eg_in_T", "Fg_in_L", "Fg_in_B",
                     "Fg_in_R", "Fg_in_T"):
            g, r = ctx.pull(name), o[name].ravel(order="F")
            assert np.max(np.abs(g - r)) < 1e-12 * max(1.0, np.max(np.abs(r))), (sweep, name)
    for k, side in enumerate(("Cg_L", "Cg_B", "Cg_R", "Cg_T")):
        if bc_type[k]:
            assert not np.any(ctx.pull(side)), side
    ctx.close()


@pytest.mark.parametrize("precision,tile", [(1, (4, 4)), (1, (8, 2)), (2, (2, 2))])
def test_reflective_sweeps_fp32_within_tolerance(trt, oracle_mod, precision, tile):
    prob, o, ctx = _pair(trt, oracle_mod, (1, 1, 1, 1), tile, 0, precision=precision)
    for sweep in range(3):
        for x in (o, ctx):
            x.call("transport_scb")
            x.call("collapse_intensities")
            x.call("rt_bc_update")
        assert rel_err(ctx.pull("I_crn", o["I_crn"].shape), o["I_crn"]) < 2e-5, sweep
        for name in MOMENTS + VIEWS:
            assert rel_err(ctx.pull(name), o[name].ravel(order="F")) < 2e-5, (sweep, name)
    ctx.close()


def _l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("bc_type", [(0, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 1)])
def test_reflective_steps_match_oracle(trt, oracle_mod, bc_type):
    prob = trt.problem.mapped(nx=14, order=4, ng=4, sigma=1.0)
    prob = prob.with_(bc_type=bc_type, bcT=(800.0, 600.0, 700.0, 500.0), use_line_search=0,
                      conv_ho=1e-11, conv_lo=1e-12, conv_gr1=1e-13)
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob)
    for s in range(3):
        it_o = o.step()
        it_g = ctx.step()
        for n in ("Temp", "E_avg", "HO_E_avg"):
            err = _l2(ctx.pull(n), np.asarray(o[n]).ravel(order="F"))
            assert err < 1e-8, (bc_type, s, n, err, it_o, it_g)
        for n in VIEWS:      # the intensities carry the step's iteration tolerances
            g, r = ctx.pull(n), o[n].ravel(order="F")
            assert _l2(g, r) < 1e-8 and rel_err(g, r) < 1e-7, (s, n, _l2(g, r), rel_err(g, r))
    ctx.close()


def test_tgen_steps_keep_the_initial_inflow(trt):
    """Tgen_QDf calls no RT_BC_UPDATE (ALGORITHMS.f08:1031-1095): the incoming faces of a
    reflective side hold Bg_planck(bcT) of that side after any number of steps."""
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=2).with_(bc_type=(0, 0, 1, 0), bcT=(900.0, 0.0, 400.0, 0.0))
    ctx = trt.native.Context(prob)
    bcin = None
    for _ in range(3):
        ctx.tgen_step()
        v = ctx.pull("I_edgV").reshape(prob.nx + 1, prob.ny, prob.nm, prob.ng, order="F")
        K = prob.nm // 4
        inc = v[prob.nx, :, K:3 * K, :]          # right side, Omega_x < 0
        if bcin is None:
            bcin = inc[0, 0, :].copy()
        assert np.all(inc == bcin[None, None, :]) and np.all(bcin > 0)
    assert not np.any(ctx.pull("Cg_R"))
    ctx.close()


def test_restart_resets_the_reflected_inflow(trt, tmp_path):
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=2).with_(bc_type=(1, 0, 0, 0), bcT=(300.0, 900.0, 0.0, 0.0))
    fresh = trt.native.Context(prob)
    planck = fresh.pull("I_edgV").reshape(prob.nx + 1, prob.ny, prob.nm, prob.ng, order="F")
    ctx = trt.native.Context(prob)
    ctx.step()
    ctx.step()
    K = prob.nm // 4
    moved = ctx.pull("I_edgV").reshape(planck.shape, order="F")
    inc = np.r_[0:K, 3 * K:4 * K]                 # left side, Omega_x > 0
    assert not np.allclose(moved[0, :, inc, :], planck[0, :, inc, :])
    path = str(tmp_path / "r.bin")
    ctx.restart_out(path)
    again = trt.native.Context(prob)
    again.restart_in(path)
    back = again.pull("I_edgV").reshape(planck.shape, order="F")
    np.testing.assert_array_equal(back[0, :, inc, :], planck[0, :, inc, :])
    for c in (fresh, ctx, again):
        c.close()


@pytest.mark.parametrize("side", [2, 3])
def test_reflective_half_equals_mirrored_whole_on_gpu(trt, side):
    """A 40x36 problem (8 groups, S8) with a reflective right (top) side against the 80x36
    (40x72) problem mirrored about it, both on the GPU, converged tightly."""
    n, ny = 40, 36
    base = trt.problem.mapped(nx=n, order=8, ng=8, sigma=1.0)
    base = base.with_(ny=ny, mat=base.mat[:, :ny], dt=2e-3, maxit_rte=300, conv_ho=1e-9, conv_lo=1e-10,
                      conv_gr1=1e-11)
    bcT = [0.0, 0.0, 0.0, 0.0]
    bcT[side - 2] = 1000.0
    bct = [0, 0, 0, 0]
    bct[side] = 1
    half = base.with_(bc_type=tuple(bct), bcT=tuple(bcT))
    fbcT = list(bcT)
    fbcT[side] = bcT[side - 2]
    if side == 2:
        whole = base.with_(nx=2 * n, xlen=2.0, mat=np.concatenate([base.mat, base.mat[::-1]], axis=0),
                           bcT=tuple(fbcT))
        cells = (slice(0, n), slice(None))
    else:
        whole = base.with_(ny=2 * ny, ylen=2.0 * base.ylen, mat=np.concatenate([base.mat, base.mat[:, ::-1]], axis=1),
                           bcT=tuple(fbcT))
        cells = (slice(None), slice(0, ny))
    a = trt.native.Context(half)
    b = trt.native.Context(whole)
    for _ in range(2):
        a.step()
        b.step()
    for name in ("Temp", "HO_E_avg"):
        ta = a.pull(name).reshape(half.nx, half.ny, order="F")
        tb = b.pull(name).reshape(whole.nx, whole.ny, order="F")[cells]
        assert np.max(np.abs(ta / tb - 1.0)) < 1e-7, name
    a.close()
    b.close()

"""GPU parity of the sweep (TRANSPORT_SCB), the moments (COLLAPSE_INTENSITIES),
fg_Calc, Cg_Calc and MATERIAL_UPDATE against the FP64 oracle."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = [
    # (nx, ny, S_N, ng, tile, sweep_mode): 1 = per-diagonal launches, per-cell moment reduction;
    # 2 = one persistent launch; 3 = per-diagonal launches of the tile engine
    (12, 10, 4, 3, (8, 8), 1),
    (13, 11, 8, 5, (4, 4), 1),
    (9, 17, 12, 2, (16, 8), 1),
    (20, 20, 16, 3, (8, 4), 1),
    (12, 10, 4, 3, (4, 2), 2),
    (13, 11, 8, 5, (2, 4), 2),
    (9, 17, 12, 2, (2, 2), 2),
    (20, 20, 16, 3, (4, 2), 2),
    (11, 9, 16, 9, (2, 4), 2),     # two group blocks of 7 groups
    (10, 7, 4, 40, (4, 2), 2),     # 320 (cell, group) pairs per tile: two reduction passes
]
CASES += [c[:5] + (3,) for c in CASES if c[5] == 2]
# tile shapes only the FP32 tile engine has
FP32_CASES = [(13, 11, 8, 5, (4, 4), 3), (20, 20, 16, 3, (8, 2), 2), (10, 7, 4, 40, (2, 8), 3)]


def _randomized_pair(trt, oracle_mod, nx, ny, order, ng, tile, precision=0, seed=1, mode=0):
    prob = trt.problem.small(nx=nx, ny=ny, order=order, ng=ng, seed=seed)
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


@pytest.mark.parametrize("nx,ny,order,ng,tile,mode", CASES)
def test_sweep_fp64_matches_oracle(trt, oracle_mod, nx, ny, order, ng, tile, mode):
    prob, o, ctx = _randomized_pair(trt, oracle_mod, nx, ny, order, ng, tile, mode=mode)
    for name in ("Bg", "KapB", "KapE", "MGQD_Src", "RT_Src"):
        assert rel_err(ctx.pull(name), o[name].ravel(order="F")) < 1e-13, name
    o.call("transport_scb")
    ctx.call("transport_scb")
    Ig = ctx.pull("I_crn", o["I_crn"].shape)
    assert rel_err(Ig, o["I_crn"]) < 1e-12
    o.call("collapse_intensities")
    ctx.call("collapse_intensities")
    for name in ("Hg_avg_xx", "Hg_avg_yy", "Hg_avg_xy", "HO_Eg_avg", "Hg_edgV_xx", "Hg_edgV_xy", "HO_Eg_edgV",
                 "HO_Fxg_edgV", "Hg_edgH_yy", "Hg_edgH_xy", "HO_Eg_edgH", "HO_Fyg_edgH", "HO_E_avg", "HO_E_edgV",
                 "HO_E_edgH", "HO_Fx_edgV", "HO_Fy_edgH"):
        assert rel_err(ctx.pull(name), o[name].ravel(order="F")) < 1e-12, name
    o.call("fg_calc")
    ctx.call("fg_calc")
    o.call("cg_calc")
    ctx.call("cg_calc")
    for name in ("fg_avg_xx", "fg_avg_yy", "fg_avg_xy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy",
                 "Cg_L", "Cg_B", "Cg_R", "Cg_T"):
        g = ctx.pull(name)
        r = o[name].ravel(order="F")
        assert np.max(np.abs(g - r)) < 1e-12 * max(1.0, np.max(np.abs(r))), name
    ctx.close()


@pytest.mark.parametrize("mode,tile", [(1, (4, 8)), (2, (2, 4)), (3, (2, 4))])
def test_two_sweeps_same_step_reuse_old(trt, oracle_mod, mode, tile):
    """Two sweeps in one step both read I_crn_old (ALGORITHMS.f08:541 inside the RT loop)."""
    prob, o, ctx = _randomized_pair(trt, oracle_mod, 10, 12, 8, 2, tile, mode=mode)
    o.call("transport_scb")
    ctx.call("transport_scb")
    o.call("transport_scb")
    ctx.call("transport_scb")
    assert rel_err(ctx.pull("I_crn", o["I_crn"].shape), o["I_crn"]) < 1e-12
    assert rel_err(ctx.pull("I_crn_old", o["I_crn_old"].shape), o["I_crn_old"]) == 0.0
    ctx.close()


@pytest.mark.parametrize("nx,ny,order,ng,tile,mode", CASES[:2] + CASES[4:6] + CASES[-2:] + FP32_CASES)
def test_sweep_fp32_within_tolerance(trt, oracle_mod, nx, ny, order, ng, tile, mode):
    prob, o, ctx = _randomized_pair(trt, oracle_mod, nx, ny, order, ng, tile, precision=1, mode=mode)
    o.call("transport_scb")
    ctx.call("transport_scb")
    assert rel_err(ctx.pull("I_crn", o["I_crn"].shape), o["I_crn"]) < 2e-6
    o.call("collapse_intensities")
    ctx.call("collapse_intensities")
    o.call("fg_calc")
    ctx.call("fg_calc")
    for name in ("HO_Eg_avg", "fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgH_yy"):
        assert rel_err(ctx.pull(name), o[name].ravel(order="F")) < 2e-6, name
    ctx.close()


@pytest.mark.parametrize("tile,mode", [((2, 4), 2), ((4, 2), 3), ((2, 2), 2)])
def test_sweep_fp32_storage_fp64_arithmetic(trt, oracle_mod, tile, mode):
    """precision 2: intensities stored in FP32, corner solve in FP64.  The only
    rounding is the storage of I_crn and of the tile faces: the FP32 relative
    rounding (6e-8) bounds the error of every stored value."""
    prob, o, ctx = _randomized_pair(trt, oracle_mod, 13, 11, 8, 5, tile, precision=2, mode=mode)
    o.call("transport_scb")
    ctx.call("transport_scb")
    assert rel_err(ctx.pull("I_crn", o["I_crn"].shape), o["I_crn"]) < 3e-7
    for c in (o, ctx):
        c.call("collapse_intensities")
        c.call("fg_calc")
    for name in ("HO_Eg_avg", "fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgH_yy"):
        assert rel_err(ctx.pull(name), o[name].ravel(order="F")) < 3e-7, name
    ctx.close()


def test_equilibrium_sweep_is_exact(trt):
    """T = Tini everywhere with Planckian boundaries at Tini: the sweep returns B_g in every corner."""
    prob = trt.problem.small(nx=8, ny=8, order=8, ng=3)
    prob.tini = 300.0
    prob.bcT = (prob.tini,) * 4
    ctx = trt.native.Context(prob)
    ctx.call("transport_scb")
    I = ctx.pull("I_crn", (16, 16, prob.nm, prob.ng))
    bg = ctx.pull("Bg", (8, 8, prob.ng))[0, 0, :]
    assert np.all(bg > 0)
    assert np.all(np.abs(I - bg[None, None, None, :]) <= 1e-13 * bg[None, None, None, :])
    ctx.close()


def test_derived_intensity_views(trt, oracle_mod):
    """I_avg, I_edgV, I_edgH, Ic_edgV, Ic_edgH pulled from the device equal the reference arrays."""
    prob, o, ctx = _randomized_pair(trt, oracle_mod, 7, 6, 8, 2, (0, 0))
    o.call("transport_scb")
    ctx.call("transport_scb")
    for name in ("I_avg", "I_edgV", "I_edgH", "Ic_edgV", "Ic_edgH"):
        ref = o[name]
        assert rel_err(ctx.pull(name, ref.shape), ref) < 1e-12, name
    ctx.close()


def test_push_kape_and_source_reaches_the_sweep(trt, oracle_mod):
    """Pushing KapE and RT_Src (as the Fortran drop-in does) feeds the next sweep."""
    prob, o, ctx = _randomized_pair(trt, oracle_mod, 6, 5, 4, 2, (0, 0))
    rng = np.random.default_rng(9)
    ke = o["KapE"] * rng.uniform(0.5, 2.0, o["KapE"].shape)
    rs = o["RT_Src"] * rng.uniform(0.5, 2.0, o["RT_Src"].shape)
    o["KapE"][...] = ke
    o["RT_Src"][...] = rs
    ctx.push("KapE", ke)
    ctx.push("RT_Src", rs)
    o.call("transport_scb")
    ctx.call("transport_scb")
    assert rel_err(ctx.pull("I_crn", o["I_crn"].shape), o["I_crn"]) < 1e-12
    ctx.close()


@pytest.mark.parametrize("precision,tile", [(0, (4, 2)), (0, (2, 2)), (1, (4, 4)), (1, (2, 8)), (2, (2, 2))])
def test_persistent_equals_per_diagonal_bitwise(trt, precision, tile):
    """The persistent launch (tile items), the column-segment launch and the
    per-diagonal launches run the same tile body on the same data, so any stale or
    early read of an exchanged face (or of a face carried in registers) would change
    bits."""
    prob = trt.problem.small(nx=37, ny=29, order=8, ng=9, seed=5)
    rng = np.random.default_rng(5)
    T = rng.uniform(0.5, 800.0, size=(prob.nx, prob.ny))
    outs = []
    for mode in (2, 3, 4):
        ctx = trt.native.Context(prob, precision=precision, tile=tile, sweep_mode=mode)
        ctx.push("Temp", T)
        ctx.call("material_update")
        res = []
        for _ in range(3):
            ctx.call("transport_scb")
            ctx.call("collapse_intensities")
            res.append([ctx.pull(n) for n in ("I_crn", "HO_Eg_avg", "HO_Fxg_edgV", "HO_Fyg_edgH", "Hg_avg_xy")])
        outs.append(res)
        ctx.close()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", [1, 2])
def test_fp32_storage_keeps_tiny_intensities(trt, precision):
    """FP32-stored intensities carry a power-of-two factor: values down to 1e-40
    (cold high groups) survive push and pull with FP32's relative rounding, where
    plain FP32 would hold them as denormals with a few bits."""
    prob = trt.problem.small(nx=6, ny=5, order=4, ng=3, seed=4)
    ctx = trt.native.Context(prob, precision=precision)
    shape = (2 * prob.nx, 2 * prob.ny, prob.nm, prob.ng)
    rng = np.random.default_rng(4)
    vals = np.asfortranarray(10.0 ** rng.uniform(-40.0, 3.0, size=shape))
    ctx.push("I_crn_old", vals)
    back = ctx.pull("I_crn_old", shape)
    assert np.max(np.abs(back - vals) / vals) < 1e-7
    ctx.close()


def test_spin_limit_exceeded_is_reported(trt, monkeypatch):
    """A dependency wait of the persistent sweep that exceeds its spin limit (forced
    from the host: limit 0, so waits end at their first check) lets the sweep finish
    and is reported as an error; a context with the normal limit sweeps cleanly."""
    prob = trt.problem.small(nx=64, ny=64, order=8, ng=3, seed=2)
    monkeypatch.setenv("TRT_SPIN_LIMIT", "0")
    ctx = trt.native.Context(prob, precision=1, tile=(4, 2), sweep_mode=2)
    with pytest.raises(trt.native.TrtError, match="spin limit"):
        ctx.call("transport_scb")
    ctx.close()
    monkeypatch.delenv("TRT_SPIN_LIMIT")
    ctx = trt.native.Context(prob, precision=1, tile=(4, 2), sweep_mode=2)
    ctx.call("transport_scb")
    ctx.close()

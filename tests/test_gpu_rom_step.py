"""The ROM runs' time step through the C ABI (trt_step_given_qdf): QD factors and C_g
pushed by the caller to the *_next arrays, no sweep, GPU against the oracle's same step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QDF = ("fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy", "Cg_L", "Cg_B", "Cg_R",
       "Cg_T")


def _l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("name", ["small", "cfg0"])
def test_rom_step_matches_oracle(trt, oracle_mod, name):
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3) if name == "small" else trt.problem.cfg0()
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob)
    o.step()
    ctx.step()
    rng = np.random.default_rng(5)
    for s in range(2):
        icrn = ctx.pull("I_crn")
        # "reconstructed" factors: the current ones with a smooth relative error, as a
        # truncated basis would leave
        for n in QDF:
            a = np.array(o[n])
            a = a * (1.0 + 1e-3 * np.sin(np.arange(a.size).reshape(a.shape, order="F") * 0.37 + s))
            o[n + "_next"][...] = a
            ctx.push(n + "_next", a)
        it_o = o.step_given_qdf()
        it_g = ctx.step_given_qdf()
        assert it_g[0] == 1 and it_g[3] == 0 and it_o[0] == 1, (it_o, it_g)
        assert np.array_equal(ctx.pull("I_crn"), icrn)                     # no sweep
        for n in QDF:
            assert np.array_equal(ctx.pull(n), np.asarray(o[n]).ravel(order="F")), n
        for n in ("Temp", "E_avg", "MGQD_E_avg", "Eg_avg", "Fxg_edgV"):
            err = _l2(ctx.pull(n), np.asarray(o[n]).ravel(order="F"))
            assert err < 1e-8, (s, n, err, it_o, it_g)
    assert ctx.time == pytest.approx(o.time)
    ctx.close()

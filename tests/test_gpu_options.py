"""Options beyond the reference's default mlqd path, GPU against the oracle through the C ABI:
use_grey = 0 (MEB_SOLVE in the MLOQD loop, UPDATES.f08:189-231, ALGORITHMS.f08:767-770), the
positivity safety search of EGP_FV_NEWT (GLOQD_SOLVES.f08:643-752, run once per Newton
iteration), and the grey emission consistent with the group sources (grey_emission 1, a
labelled deviation from GLOQD_SOLVES.f08:439-440)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def _tight(prob):
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    return prob


def test_meb_solve_matches_oracle(trt, oracle_mod):
    prob = trt.problem.mapped(nx=12, order=4, ng=5, sigma=1.0)
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob)
    rng = np.random.default_rng(8)
    eg = rng.uniform(0.5, 2.0, o["Eg_avg"].shape) * np.asarray(o["Eg_avg"]).max()
    kap = rng.uniform(1.0, 50.0, o["KapE"].shape)
    told = rng.uniform(0.8, 3.0, o["Temp_Old"].shape)
    for name, a in (("Eg_avg", eg), ("KapE", kap), ("Temp_Old", told)):
        o[name][...] = a
        ctx.push(name, a)
    o.call("meb_solve")
    ctx.L.trt_meb_solve(ctx.h, b"Eg_avg")
    ref = np.asarray(o["Temp"]).ravel(order="F")
    got = ctx.pull("Temp")
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-14
    assert ctx.L.trt_meb_solve(ctx.h, b"Temp") != 0        # not a group energy array
    ctx.close()


@pytest.mark.parametrize("case", ["no_grey", "safety", "grey_emission", "grey_emission_map15"])
def test_option_steps_match_oracle(trt, oracle_mod, case):
    if case == "grey_emission_map15":
        # the cfg1 map's multiplier spread, lognormal(0, 1.5), on a 24x24 grid
        prob = _tight(trt.problem.mapped(nx=24, order=4, ng=6, sigma=1.5))
    else:
        prob = _tight(trt.problem.mapped(nx=16, order=4, ng=4, sigma=1.0))
    if case == "no_grey":
        prob.no_grey = 1
    elif case == "safety":
        prob.use_safety_search, prob.T_bound_low = 1, 0.99    # binds in the second step
    else:
        prob.grey_emission = 1
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob)
    for s in range(3):
        it_o = o.step()
        it_g = ctx.step()
        for n in ("Temp", "E_avg" if case != "no_grey" else "MGQD_E_avg", "HO_E_avg"):
            err = _l2(ctx.pull(n), np.asarray(o[n]).ravel(order="F"))
            assert err < 1e-8, (case, s, n, err, it_o, it_g)
    if case == "safety":
        assert np.min(ctx.pull("Temp")) >= 0.99
    ctx.close()


def test_safety_search_rejects_unreachable_bound(trt):
    prob = _tight(trt.problem.mapped(nx=8, order=4, ng=3, sigma=1.0))
    prob.use_safety_search, prob.T_bound_low = 1, 2.0          # every temperature starts below it
    ctx = trt.native.Context(prob)
    with pytest.raises(trt.native.TrtError, match="already below its lower bound"):
        ctx.step()
    ctx.close()

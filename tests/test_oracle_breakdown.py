"""Where the reference algorithm stops on the 8x8 lognormal(0, 1.5) opacity-multiplier map
(multipliers 0.04..53), shown on the oracle (CPU only).

* The analytic grey emission of EGP_FV_NEWT, 15 kap0 c aR T / pi^4 (GLOQD_SOLVES.f08:439-440),
  ignores the multipliers, so the grey equations emit too little where the multiplier is
  above 1: cells cool below the initial temperature that nothing in the problem lies below.
  The emission consistent with the group sources (grey_emission 1) keeps them at it.
* Next to region boundaries with opacity jumps of up to 1e3 the reduced QD solution has
  negative group energies (E_g < 0).  GREY_COEFS then weights by sign-changing E_g
  (GLOQD_SOLVES.f08:89-143) and the MLOQD iteration of the next step drifts: the
  temperature of those cells falls from iteration to iteration until the grey Newton
  iteration returns NaN, with either emission."""
import numpy as np


def test_analytic_grey_emission_cools_below_the_initial_temperature(trt, oracle_mod):
    tmin = {}
    for ge in (0, 1):
        prob = trt.problem.mapped(nx=32, order=8, ng=8, sigma=1.5)
        prob.grey_emission = ge
        o = oracle_mod.Oracle(prob, threads=8)
        for _ in range(3):
            o.step()
        assert np.all(np.isfinite(o["Temp"]))
        tmin[ge] = float(np.min(o["Temp"]))
    assert tmin[0] < 0.99 * prob.tini            # 0.981 eV after 3 steps
    assert tmin[1] > 0.9999 * prob.tini


def test_negative_group_energies_then_nan(trt, oracle_mod):
    prob = trt.problem.mapped(nx=48, order=8, ng=16, sigma=1.5)
    o = oracle_mod.Oracle(prob, threads=8)
    o.step()
    assert np.all(np.isfinite(o["Temp"]))
    eg = np.asarray(o["Eg_avg"])
    i, j, g = np.unravel_index(np.argmin(eg), eg.shape)
    assert eg[i, j, g] < 0
    # the most negative group energy sits in a cell at a region boundary of the map
    mat = np.asarray(prob.mat)
    nb = [mat[a, b] for a, b in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)) if 0 <= a < prob.nx and 0 <= b < prob.ny]
    assert any(m != mat[i, j] for m in nb)
    o.step()
    assert not np.all(np.isfinite(o["Temp"]))

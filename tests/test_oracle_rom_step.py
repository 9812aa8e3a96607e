"""The oracle's ROM-run step (orc_step_ex, ho_form 1: QD factors and C_g from the caller,
no sweep), pinned by a fixed point: a full step converges the low-order iteration for the
factors of its last sweep, so a fresh state handed exactly those factors reaches the same
temperature and energies in one RT iteration without sweeping."""
import numpy as np
import pytest

QDF = ("fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy", "Cg_L", "Cg_B", "Cg_R",
       "Cg_T")


def _l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("name", ["small", "cfg0"])
def test_given_factors_reach_the_full_steps_fixed_point(trt, oracle_mod, name):
    prob = trt.problem.small(nx=8, ny=6, order=4, ng=3) if name == "small" else trt.problem.cfg0()
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-12, 1e-13, 1e-14
    full = oracle_mod.Oracle(prob, threads=4)
    its_full = full.step()
    assert its_full[3] >= 1
    rom = oracle_mod.Oracle(prob, threads=4)
    icrn0 = np.array(rom["I_crn"])
    for n in QDF:
        rom[n + "_next"][...] = full[n]
    its = rom.step_given_qdf()
    assert its[0] == 1 and its[3] == 0, its
    assert np.array_equal(rom["I_crn"], icrn0)                  # no sweep
    for n in QDF:
        assert np.array_equal(rom[n], full[n]), n              # the given factors are the ones used
    for n in ("Temp", "E_avg", "MGQD_E_avg", "Eg_avg"):
        assert _l2(rom[n], full[n]) < 1e-9, (n, _l2(rom[n], full[n]))
    # the old level's factors were the initial ones, as in the full step
    for n in ("fg_avg_xx_old", "fg_edgV_xy_old"):
        assert np.array_equal(rom[n], full[n]), n
    assert rom.time == full.time

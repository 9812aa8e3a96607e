"""Pins the oracle's temperature functions (TEMP_FUNCTIONS.f08:11-141) and the
level-symmetric quadrature with known answers, closed forms and identities."""
import math

import numpy as np
import pytest

H = 6.62613e-19
C = 2.99792458e2
ERG = 6.24150934e11
CU = 1e13
NORM = H ** 3 * C ** 2 * ERG ** 4 * CU


def test_sigint_known_answers(oracle_mod):
    # lower branch z**3*(1/3 - z/8 + z**2/62.4) (TF:127-128)
    assert oracle_mod.sigint(1.0) == pytest.approx(1 / 3 - 1 / 8 + 1 / 62.4, rel=1e-15)
    assert oracle_mod.sigint(0.0) == 0.0
    # upper branch pi**4/15 - exp(-z)*(z**3+3z**2+6z+7.28) (TF:129-130)
    z = 3.0
    assert oracle_mod.sigint(z) == pytest.approx(math.pi ** 4 / 15 - math.exp(-z) * (27 + 27 + 18 + 7.28), rel=1e-15)
    assert oracle_mod.sigint(800.0) == pytest.approx(math.pi ** 4 / 15, rel=1e-15)
    # band (1.95, 2.05]: linear between the branch end points (unvendored Interpolator; parity unpinned)
    lo = oracle_mod.sigint(1.95)
    hi = math.pi ** 4 / 15 - math.exp(-2.05) * (2.05 ** 3 + 3 * 2.05 ** 2 + 6 * 2.05 + 7.28)
    assert oracle_mod.sigint(2.0) == pytest.approx(0.5 * (lo + hi), rel=1e-14)


def test_sigint_tracks_the_exact_planck_integral(oracle_mod):
    """sigint approximates int_0^z x^3/(e^x-1) dx; its fit error is below 1.5e-2 relative."""
    from scipy.integrate import quad
    for z in (0.05, 0.5, 1.0, 1.9, 2.2, 3.0, 5.0, 10.0, 30.0):
        exact = quad(lambda x: x ** 3 / math.expm1(x) if x > 0 else 0.0, 0.0, z, limit=200)[0]
        assert abs(oracle_mod.sigint(z) - exact) / exact < 1.5e-2, z


def test_planck_sum_telescopes(oracle_mod):
    """Contiguous groups spanning [0, inf) sum to 2 T^4 (pi^4/15) / (h^3 c^2 erg^4 comp_unit)."""
    edges = np.array([0.0, 3.0, 17.0, 90.0, 410.0, 2500.0, 1e10])
    for T in (0.7, 12.0, 1000.0):
        s = sum(oracle_mod.Bg_planck(T, edges[g], edges[g + 1], CU) for g in range(len(edges) - 1))
        assert s == pytest.approx(2 * T ** 4 * (math.pi ** 4 / 15) / NORM, rel=1e-13)
    assert oracle_mod.Bg_planck(1e-16, 1.0, 2.0, CU) == 0.0


def test_kapB_times_Bg_cancels_sigint(oracle_mod):
    """For nu_low/T <= 10, kapB*Bg = 2 sig0 T (e^-nl/T - e^-nh/T)/(h^3 c^2 erg^4 cu): sigint cancels (TF:28,102-105)."""
    sig0 = 27e9
    for T, nl, nh in ((1.0, 0.0, 3.0), (50.0, 10.0, 56.2), (1000.0, 316.0, 1778.0), (5.0, 2.0, 9.0)):
        lhs = oracle_mod.kapB(T, nl, nh) * oracle_mod.Bg_planck(T, nl, nh, CU)
        rhs = 2 * sig0 * T * (math.exp(-nl / T) - math.exp(-nh / T)) / NORM
        assert lhs == pytest.approx(rhs, rel=1e-12)


def test_kapB_high_branch_matches_closed_form(oracle_mod):
    """nu_low/T > 10 uses the asymptotic form (TF:22-24)."""
    T, nl, nh = 1.0, 12.0, 15.0
    z1, z2 = nh / T, nl / T
    sig = -math.exp(-(z1 - z2)) * (z1 ** 3 + 3 * z1 ** 2 + 6 * z1 + 7.28) + (z2 ** 3 + 3 * z2 ** 2 + 6 * z2 + 7.28)
    assert oracle_mod.kapB(T, nl, nh) == pytest.approx(27e9 * (1 - math.exp(-(nh - nl) / T)) / (T ** 3 * sig), rel=1e-14)


def test_midpoint_opacity(oracle_mod):
    """sigma_g(T) = 27 keV^3/cm (1-exp(-nu_g/T))/nu_g^3 with nu_g the arithmetic midpoint; 27 keV^3 = 27e9 eV^3."""
    assert oracle_mod.kap_midpoint(1000.0, 1000.0, 3000.0) == pytest.approx(27.0 * (1 - math.exp(-2.0)) / 8.0, rel=1e-14)
    # optically thin, hot limit: sigma -> 27e9 / (nu^2 T)
    assert oracle_mod.kap_midpoint(1e6, 10.0, 30.0) == pytest.approx(27e9 / (20.0 ** 2 * 1e6), rel=1e-4)


@pytest.mark.parametrize("order,nm", [(4, 12), (8, 40), (12, 84), (16, 144)])
def test_level_symmetric_moments(trt, order, nm):
    ox, oy, w = trt.problem.level_symmetric(order)
    assert len(w) == nm
    assert w.sum() == pytest.approx(4 * math.pi, rel=1e-15)          # INPUTS.f08:1669
    assert abs((w * ox).sum()) < 1e-13 and abs((w * oy).sum()) < 1e-13
    assert abs((w * ox * oy).sum()) < 1e-13
    # level-symmetric sets integrate mu^2 (published weights carry 7 digits)
    assert (w * ox ** 2).sum() == pytest.approx(4 * math.pi / 3, rel=3e-6)
    assert (w * oy ** 2).sum() == pytest.approx(4 * math.pi / 3, rel=3e-6)
    assert np.all(ox[: nm // 4] > 0) and np.all(oy[: nm // 4] > 0)         # q1 (+,+)
    assert np.all(ox[nm // 4: nm // 2] < 0) and np.all(oy[nm // 4: nm // 2] > 0)
    assert np.all(ox[nm // 2: 3 * nm // 4] < 0) and np.all(oy[nm // 2: 3 * nm // 4] < 0)
    assert np.all(ox[3 * nm // 4:] > 0) and np.all(oy[3 * nm // 4:] < 0)
    assert np.all(ox ** 2 + oy ** 2 < 1.0)


def test_splitmix_reproducible(trt):
    a = trt.problem.SplitMix64(42)
    b = trt.problem.SplitMix64(42)
    xs = [a.next_u64() for _ in range(5)]
    assert xs == [b.next_u64() for _ in range(5)]
    # reference value of SplitMix64 from seed 0 (Vigna's published generator)
    assert trt.problem.SplitMix64(0).next_u64() == 0xE220A8397B1DCDAF
    m = trt.problem.cfg1(nx=16)
    assert m.kmult.shape == (64,) and np.all(m.kmult > 0)
    assert m.mat.shape == (16, 16) and m.mat.max() == 63

"""Pins of oracle functions that the golden vectors leave open, by what the mathematics or a hand computation fixes:

* normalise_profiles (xpg/integrate.py:546-556): unit mass; no data gives the Gaussian prior, which on the 41 x 41
  grid is itself a normalised density; overwhelming data gives the data's own density; negative densities are clipped
  (the prior alone remains); three pseudo-counts per node give the even mix of data and prior;
* the adaptive fits' pieces (xpg/integrate.py:404-461): the strong-row rule on rows just over and under every
  threshold, the 300-reflection shrinkage and the lambda = 200 / n weight mixing against hand values, sigma_D from
  second moments generated as those of a Gaussian truncated at FG_R sigma (by quadrature, not the closed form);
* the profile-learning strong rule of the block loop (xpg/integrate.py:530-533) on rows over and under each bound;
* the box side and super-pixel rule (xpg/integrate.py:263-287) at every XY_BUCKETS edge and at the super-pixel
  switches, for the oracle and the product's vectorised planning alike.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import adapt as oa
from oracle import kernels as ok
from oracle import planning as op
from oracle import sweep as osw

G = 41
DU = (9.0 / 40) ** 2


def _grid():
    u = np.linspace(-4.5, 4.5, G)
    return np.meshgrid(u, u)


# ---- normalise_profiles -----------------------------------------------------------------------------------------

def test_normalised_profiles_have_unit_mass():
    rng = np.random.default_rng(0)
    den = rng.uniform(0, 50, (9, G, G)) * (rng.random((9, G, G)) < 0.7)
    num = den * rng.normal(0.05, 0.05, (9, G, G))
    prof = ok.normalise_profiles(num, den)
    assert prof.dtype == np.float32 and prof.shape == (9, G, G)
    np.testing.assert_allclose(prof.sum((1, 2)) * DU, 1.0, rtol=1e-6)
    assert (prof >= 0).all()


def test_no_data_gives_the_gaussian_prior():
    # the unit 2D Gaussian sampled on the grid is a normalised density to ~1e-5 (trapezoid rule on a smooth,
    # vanishing integrand; the tails beyond 4.5 sigma hold 1.4e-5 of the mass)
    U1, U2 = _grid()
    gauss = np.exp(-0.5 * (U1 ** 2 + U2 ** 2)) / (2 * np.pi)
    assert abs(gauss.sum() * DU - 1) < 2e-5
    prof = ok.normalise_profiles(np.zeros((2, G, G)), np.zeros((2, G, G)))
    np.testing.assert_allclose(prof[0], gauss, rtol=3e-5, atol=1e-9)
    np.testing.assert_array_equal(prof[0], prof[1])


def test_overwhelming_data_gives_its_own_density():
    U1, U2 = _grid()
    target = np.exp(-0.5 * ((U1 - 0.4) ** 2 / 1.7 + U2 ** 2 / 0.8))          # an elliptical, off-centre spot
    den = np.full((1, G, G), 1e12)
    prof = ok.normalise_profiles(den * target, den)[0]
    np.testing.assert_allclose(prof, target / (target.sum() * DU), rtol=1e-6, atol=1e-10)


def test_negative_densities_are_clipped():
    # every node's data density is negative: only the prior's 3 pseudo-counts remain, in the Gaussian's shape
    U1, U2 = _grid()
    gauss = np.exp(-0.5 * (U1 ** 2 + U2 ** 2)) / (2 * np.pi)
    den = np.full((1, G, G), 1e6)
    prof = ok.normalise_profiles(-den, den)[0]
    np.testing.assert_allclose(prof, gauss / (gauss.sum() * DU), rtol=1e-5)


def test_three_pseudo_counts_mix_data_and_prior_evenly():
    U1, U2 = _grid()
    gauss = ok.gaussian_profile().astype(np.float64)
    target = np.exp(-0.5 * ((U1 + 1.0) ** 2 + (U2 - 0.5) ** 2 / 2.0))
    den = np.full((1, G, G), 3.0)
    prof = ok.normalise_profiles(den * target, den)[0]
    mix = 0.5 * (target + gauss)
    np.testing.assert_allclose(prof, mix / (mix.sum() * DU), rtol=1e-6)


# ---- adaptive fits ------------------------------------------------------------------------------------------------

def test_adapt_strong_rule_edges():
    eps = 1e-9
    # isum / sqrt(vsum) exactly 15 fails, just above passes; vsum below 1e-6 counts as 1e-6
    isum = np.array([15.0, 15.0 + eps, 15.0, 15.0, 15.0 + eps, 0.0, -1.0, 1.6e-2, 1.4e-2, 30.0, 30.0, 30.0])
    vsum = np.array([1.0, 1.0, 1.0 - 1e-6, 1.0 + 1e-6, 1.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0])
    n_over = np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0])
    frac = np.array([0.99] * 10 + [0.98, 0.98 + eps])
    fb = np.ones(12, int)
    got = oa.strong_rows(isum, vsum, n_over, frac, fb)
    assert got.tolist() == [False, True, True, False, True, False, False, True, False, False, False, True]
    fb2 = fb.copy()
    fb2[1] = 2
    assert not oa.strong_rows(isum, vsum, n_over, frac, fb2)[1]


def test_shrinkage_by_hand():
    assert oa.shrink(4.0, 1.0, 300) == pytest.approx(2.0, rel=1e-15)              # equal weights: geometric mean
    assert oa.shrink(0.3, 0.1, 900) == pytest.approx(0.3 ** 0.75 * 0.1 ** 0.25, rel=1e-14)
    assert oa.shrink(5.0, 0.1, 0) == pytest.approx(0.1, rel=1e-15)                # no data: the run's value


def test_weight_mixing_by_hand():
    g = ok.rc_default_weights().astype(np.float64)
    w = np.zeros(21)
    w[3] = 2.0
    w[10] = 6.0
    got = oa.mix_weights(w, 200)                                                  # lambda = 1: the mean
    want = 0.5 * g
    want[3] += 0.5 * 0.25
    want[10] += 0.5 * 0.75
    np.testing.assert_allclose(got, want, rtol=1e-6)
    got = oa.mix_weights(w, 800)                                                  # lambda = 1/4
    want = (0.25 * g + np.where(np.arange(21) == 3, 0.25, 0) + np.where(np.arange(21) == 10, 0.75, 0)) / 1.25
    np.testing.assert_allclose(got, want, rtol=1e-6)
    # an all-zero fit keeps only the prior's share, lambda / (1 + lambda) = 2/3 at n = 100 (not renormalised)
    np.testing.assert_allclose(oa.mix_weights(np.zeros(21), 100), 2.0 / 3.0 * g, rtol=1e-6)


def _truncated_moment_factor(R=3.0):
    num = integrate.quad(lambda r: 0.5 * r ** 3 * np.exp(-0.5 * r ** 2), 0, R)[0]
    return num / integrate.quad(lambda r: r * np.exp(-0.5 * r ** 2), 0, R)[0]


@pytest.mark.parametrize("ratio,want", [(1.3, 1.3), (0.7, 0.7), (3.0, 2.0), (0.3, 0.5)])
def test_sigma_d_from_truncated_moments(ratio, want):
    rng = np.random.default_rng(5)
    n, qx, sd = 501, 0.075, np.radians(0.02)
    Xlen = rng.uniform(150, 240, n)
    cosa = 150.0 / Xlen
    pix = (qx / Xlen) ** 2 * (1 + cosa ** 2) / 24.0
    sig = ratio * sd
    MD = _truncated_moment_factor() * (sig ** 2 + pix)
    MD[:200] *= rng.uniform(1.5, 40, 200)           # contaminated spots above the median do not move it
    MD[200:230] *= rng.uniform(0.0, 0.5, 30)
    assert oa.sd_from_moments(MD, Xlen, cosa, qx, sd) == pytest.approx(want * sd, rel=1e-9)


# ---- profile-learning strong rule ---------------------------------------------------------------------------------

def test_profile_strong_rule_edges():
    eps = 1e-9
    rows = [  # isum, vsum, n_over, frac_rec, n_fg_nb, n_fg_masked, strong
        (10.0, 1.0, 0, 0.99, 0, 0, False),
        (10.0 + eps, 1.0, 0, 0.99, 0, 0, True),
        (20.0 + eps, 4.0, 0, 0.99, 0, 0, True),
        (20.0, 4.0 + 1e-6, 0, 0.99, 0, 0, False),
        (50.0, 1.0, 1, 0.99, 0, 0, False),
        (50.0, 1.0, 0, 0.95, 0, 0, False),
        (50.0, 1.0, 0, 0.95 + eps, 0, 0, True),
        (50.0, 1.0, 0, 0.99, 1, 0, False),
        (50.0, 1.0, 0, 0.99, 0, 1, False),
        (0.0, 0.0, 0, 0.99, 0, 0, False),
        (1.1e-2, 0.0, 0, 0.99, 0, 0, True),         # vsum floored at 1e-6: I / sigma = 11
    ]
    o1 = {k: np.array([r[i] for r in rows]) for i, k in enumerate(osw.STRONG_KEYS)}
    assert osw.profile_strong(o1).tolist() == [r[6] for r in rows]


# ---- box side and super-pixels --------------------------------------------------------------------------------------

# (spx, global box, want bxs, want pxs): need = 2 max(ceil(5 spx) + 1, 2) + 1 mapped up onto XY_BUCKETS (capped at the
# global box); pxs = clip(floor(spx / 2.5), 1, 4), and with pxs > 1 the need is taken in super-pixels
SIDE_CASES = [
    (0.10, 97, 9, 1), (0.59, 97, 9, 1), (0.61, 97, 13, 1), (0.99, 97, 13, 1), (1.01, 97, 17, 1), (1.39, 97, 17, 1),
    (1.41, 97, 23, 1), (1.99, 97, 23, 1), (2.01, 97, 31, 1), (2.79, 97, 31, 1), (2.81, 97, 41, 1), (3.79, 97, 41, 1),
    (3.81, 97, 55, 1), (4.99, 97, 55, 1),
    (5.00, 97, 31, 2), (5.59, 97, 31, 2), (5.61, 97, 41, 2), (7.49, 97, 41, 2), (7.50, 97, 31, 3),
    (9.99, 97, 41, 3), (10.00, 97, 31, 4), (14.00, 97, 41, 4), (25.00, 97, 73, 4), (40.0, 97, 97, 4),
    (60.0, 97, 97, 4), (3.90, 41, 41, 1), (3.90, 31, 31, 1), (0.30, 13, 9, 1),
]


@pytest.mark.parametrize("spx,box,bxs,pxs", SIDE_CASES)
def test_box_side_hand_cases(spx, box, bxs, pxs):
    got = op.box_side(100.0, 200.0, spx, box, False)
    assert got[:2] == (bxs, pxs)
    # the origin: floor(x - PIX0 + 0.5) in super-pixels, less half the side
    assert got[2] == math.floor(100.0 - 0.5) // pxs - bxs // 2
    assert got[3] == math.floor(200.0 - 0.5) // pxs - bxs // 2


def test_box_side_origin_at_negative_coordinates():
    # floor division towards minus infinity for pixels left of the detector's first super-pixel
    assert op.box_side(0.4, 0.6, 5.2, 97, False)[2:] == ((-1) // 2 - 15, 0 // 2 - 15)
    assert op.box_side(0.4, 0.6, 0.2, 97, True)[:2] == (97, 1)


def test_product_plan_sides_agrees_with_the_hand_cases():
    from xpg_amd import integrate as I
    from xpg_amd.geometry import Geometry
    # a detector on which spx is a known function of the pixel: flat panel normal to the beam, sigma_D chosen per case
    g = Geometry(wavelength=1.0, beam=np.array([0, 0, 1.0]), axis=np.array([1.0, 0, 0]), orgx=1000.5, orgy=1000.5,
                 distance=100.0, ed1=np.array([1.0, 0, 0]), ed2=np.array([0, 1.0, 0]), qx=0.1, qy=0.1, nx=2000,
                 ny=2000, phi0=0.0, osc=0.1)
    x = np.array([1000.5])
    y = np.array([1000.5])
    L = 100.0            # |X| at the beam centre; cos = 1, so spx = sqrt((L sd / q)^2 + 1/12) with sd in radians
    for spx, box, bxs, pxs in SIDE_CASES:
        # (a hair above the hand value: the degree round trip must not land just below a super-pixel switch)
        s_rad = math.sqrt(max((spx * (1 + 1e-10)) ** 2 - 1.0 / 12.0, 1e-12)) * g.qx / L
        model = I.IntegrationModel(sigma_d=math.degrees(s_rad) / 1.3, sigma_m=0.1)    # adaptive boxes use 1.3 sd
        got = I.plan_sides(dict(x=x, y=y), g, model, box, adaptive=True)
        if spx ** 2 > 1.0 / 12.0 + 1e-9:
            assert (int(got[0][0]), int(got[1][0])) == (bxs, pxs), (spx, box)

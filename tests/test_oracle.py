"""Pins of the FP64 oracle (oracle/kernels.py) before anything is compared against it.

1. Golden vectors: the reference's own pass1 / accumulate_profiles / pass2 evaluated on float64 inputs
   (tests/golden/make_golden.py) -- agreement to ~1e-10 and identical integer masks.
2. What the mathematics fixes, chosen so that a dropped term, a wrong sign or index, or a transposed operand fails:
   numpy.gradient, erf integrals summing to one, bilinear mass conservation, the clipped mean of a flat background,
   exact recovery of a noise-free model by the Poisson fit (single and 2x2 partner fit by Cramer's rule),
   numpy.corrcoef, the closed-form flat-panel lab position.
"""
import numpy as np
import pytest
from scipy.special import erf

from golden_io import cases, load
from oracle import kernels as ok

P1_FLOAT = ("bg", "nbg", "isum", "vsum", "vsum0", "xo", "yo", "zo", "fz", "img2", "frac_rec", "zeta", "J", "u1", "u2",
            "seff")
P1_INT = ("n_over", "n_fg_masked", "n_fg_nb", "nfg", "all_valid2")
P2_FLOAT = ("iprf", "vprf", "vprf0", "corr", "p_valid", "p_total", "ipartner", "rho", "p_fg")


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = np.maximum(np.abs(b), np.max(np.abs(b)) * 1e-6 + 1e-300)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def _args(c):
    return (c["sb"], c["rc"], c["det"], c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"])


@pytest.mark.parametrize("name", cases())
def test_oracle_matches_reference_fp64(name):
    c, out = load(name)
    g = out[64]
    r1 = ok.pass1(*_args(c), c["nb"], c["rcw"])
    for k in P1_FLOAT:
        assert _rel(r1[k], g["p1_" + k]) < 1e-9, k
    # moments are differences of nearly equal terms: compare on their own scale
    for k in ("m_d", "m_m"):
        assert np.max(np.abs(r1[k] - g["p1_" + k])) < 1e-9 * np.max(np.abs(g["p1_" + k])) + 1e-16, k
    for k in P1_INT:
        assert np.array_equal(np.asarray(r1[k]), g["p1_" + k]), k
    num, den = ok.accumulate_profiles(r1["img2"], r1["all_valid2"], r1["J"], r1["u1"], r1["u2"], c["rc"]["region"],
                                      r1["isum"], g["acc_sel"], r1["seff"], 9)
    assert np.max(np.abs(num - g["acc_num"])) < 1e-10 * np.max(np.abs(g["acc_num"]))
    assert np.max(np.abs(den - g["acc_den"])) < 1e-10 * np.max(np.abs(g["acc_den"]))
    pix = {k: r1[k] for k in ("zeta", "J", "u1", "u2", "seff")}
    r2 = ok.pass2(*_args(c), c["prof"], r1["bg"], r1["nbg"], c["nb"], c["rcw"], pix)
    for k in P2_FLOAT:
        assert _rel(r2[k], g["p2_" + k]) < 1e-9, k
    dI = np.abs(r2["iprf"] - g["p2_iprf"]) / np.sqrt(r2["vprf"])
    assert dI.max() < 1e-9


@pytest.mark.parametrize("name", cases())
def test_reference_fp32_drift_is_bounded(name):
    """The reference's own float32 run against its FP64 algorithm: recorded, not a parity target (it exceeds the
    1e-3 sigma budget for strong spots because eps = e.(S' - S) cancels in float32)."""
    c, out = load(name)
    r1 = ok.pass1(*_args(c), c["nb"], c["rcw"])
    pix = {k: r1[k] for k in ("zeta", "J", "u1", "u2", "seff")}
    r2 = ok.pass2(*_args(c), c["prof"], r1["bg"], r1["nbg"], c["nb"], c["rcw"], pix)
    dI = np.abs(r2["iprf"] - out[32]["p2_iprf"]) / np.sqrt(r2["vprf"])
    assert dI.max() < 2e-2


def test_irls_rules_agree_within_budget():
    """Per-reflection stop rule against the reference's chunk-global one: within 1e-3 sigma."""
    for name in cases():
        c, _ = load(name)
        r1 = ok.pass1(*_args(c), c["nb"], c["rcw"])
        pix = {k: r1[k] for k in ("zeta", "J", "u1", "u2", "seff")}
        a = ok.pass2(*_args(c), c["prof"], r1["bg"], r1["nbg"], c["nb"], c["rcw"], pix, irls_rule="chunk")
        b = ok.pass2(*_args(c), c["prof"], r1["bg"], r1["nbg"], c["nb"], c["rcw"], pix, irls_rule="reflection")
        assert np.max(np.abs(a["iprf"] - b["iprf"]) / np.sqrt(a["vprf"])) < 1e-3


def test_helpers_match_reference():
    z = np.load(ok.__file__.replace("oracle/kernels.py", "tests/golden/helpers.npz"))
    R = ok.phi_rc(z["lo"], z["hi"], z["phic"], z["zeta"], float(z["sm"]), z["rcw"])
    assert np.max(np.abs(R - z["R_mix"])) < 1e-14
    Rg = ok.phi_rc(z["lo"], z["hi"], z["phic"], z["zeta"], float(z["sm"]), None)
    assert np.max(np.abs(Rg - z["R_gauss"])) < 1e-14
    B = ok.rc_basis(z["lo"], z["hi"], z["phic"][:, 0], z["zeta"][:, 0], float(z["sm"]))
    assert np.max(np.abs(B - z["basis"])) < 1e-14
    inside, i1, i2, t1, t2 = ok.bilinear(z["u1"], z["u2"])
    assert np.array_equal(inside, z["inside"]) and np.array_equal(i1, z["i1"]) and np.array_equal(i2, z["i2"])
    assert np.max(np.abs(t1 - z["t1"])) < 1e-13 and np.max(np.abs(t2 - z["t2"])) < 1e-13
    nb = dict(x=z["nb_x"], y=z["nb_y"], z=z["nb_z"], sx=z["nb_sx"], sz=z["nb_sz"])
    assert np.array_equal(ok.neighbour_mask(z["xx"], z["yy"], z["zz"], nb, z["nbm"].shape), z["nbm"])
    assert np.array_equal(ok.neighbour_mask(z["xx"], z["yy"], z["zz"], nb, z["nbm"].shape, z["skip"]), z["nbm_skip"])
    assert np.array_equal(ok.gaussian_profile(), z["gauss"])
    assert np.array_equal(ok.rc_default_weights(), z["rcw_default"])


# ---------------------------------------------------------------------------------------------------------------------
# mathematics
# ---------------------------------------------------------------------------------------------------------------------

def test_gradient_is_numpy_gradient():
    f = np.random.default_rng(0).normal(size=(3, 7, 9))
    for ax in (1, 2):
        assert np.array_equal(ok.gradient2(f, ax), np.gradient(f, axis=ax, edge_order=2))


def test_rocking_curve_sums_to_one():
    rng = np.random.default_rng(1)
    sm, zeta, phic = np.radians(0.2), 0.6, 0.01
    edges = phic + np.linspace(-0.05, 0.05, 400)
    lo, hi = edges[:-1][None], edges[1:][None]
    for rcw in (None, ok.rc_default_weights(), rng.dirichlet(np.ones(21)).astype(np.float32)):
        R = ok.phi_rc(lo, hi, np.full_like(lo, phic), np.full_like(lo, zeta), sm, rcw)
        w = 1.0 if rcw is None else float(np.sum(np.asarray(rcw, float)))
        assert abs(R.sum() - w) < 1e-7
    # closed form of one frame of a Gaussian
    R = ok.phi_rc(np.array([0.0]), np.array([0.001]), np.array([0.0005]), np.array([1.0]), 0.001, None)
    assert abs(R[0] - erf(0.5 / np.sqrt(2))) < 1e-15


def test_default_mixture_is_unit_gaussian():
    """21 components of width 0.25 with the default weights: variance 1 (0.25^2 + weighted variance of centres)."""
    w = ok.rc_default_weights().astype(float)
    var = ok.RC_W ** 2 + np.sum(w * ok.RC_C ** 2) - np.sum(w * ok.RC_C) ** 2
    assert abs(var - 1.0) < 2e-3


def test_bilinear_conserves_mass():
    rng = np.random.default_rng(2)
    n = 5
    u1 = rng.uniform(-4, 4, (n, 9, 9))
    u2 = rng.uniform(-4, 4, (n, 9, 9))
    img2 = rng.uniform(0, 10, (n, 9, 9))
    J = np.full((n, 9, 9), 0.3)
    seff = np.ones(n)
    num, den = ok.accumulate_profiles(img2, np.ones((n, 9, 9), bool), J, u1, u2, np.zeros(n, int), np.ones(n),
                                      np.ones(n, bool), seff, 1)
    assert abs(den.sum() - n * 81) < 1e-9
    assert abs(num.sum() - img2.sum() / 0.3) < 1e-9 * img2.sum() / 0.3


def test_flat_background_is_clipped_mean():
    """No spot, no neighbours, flat template: b is the iterated clipped mean over the background ring."""
    from chunks import make_chunk
    c = make_chunk(n=6, seed=5, partners=False, neighbours=False, codes=False, intensity=(1e-9, 2e-9))
    r1 = ok.pass1(*_args(c), c["nb"], c["rcw"])
    gt = ok.geometry_terms(c["det"], c["axis"], c["s0"], c["rc"], c["frames_phi0"], c["osc"], c["sd"], c["sm"],
                           c["rcw"], c["Dz"], c["Dy"], c["Dx"])
    for r in range(6):
        r2 = gt["u1"][r, 0] ** 2 + gt["u2"][r, 0] ** 2
        fgz = gt["R"][r] >= 1e-3 * gt["R"][r].max() + 1e-6
        vals = []
        for j in range(c["Dz"]):
            for iy in range(c["Dy"]):
                for ix in range(c["Dx"]):
                    fg = r2[iy, ix] <= 9 and fgz[j]
                    if not fg and r2[iy, ix] >= 12.25:
                        vals.append(c["sb"][r, j, iy, ix])
        vals = np.array(vals)
        b = vals.mean()
        for _ in range(4):
            thr = b + 3 * np.sqrt(max(b, 0.5)) + 1
            b = vals[vals <= thr].mean()
        assert abs(r1["bg"][r] - b) < 1e-12 * max(b, 1)


def _noise_free(partner):
    from chunks import make_chunk
    c = make_chunk(n=8, seed=9, partners=partner, neighbours=False, codes=False)
    if partner:
        c["rc"]["has_p"][:] = 1.0
        x, y = ok.pixel_coords(c["rc"], c["Dy"], c["Dx"])
        c["rc"]["pS"] = ok.pixel_S(c["det"], x[:, c["Dx"] // 2] + 2.5, y[:, c["Dy"] // 2] - 1.5)
        c["rc"]["pphi"] = c["rc"]["phi"] + 0.7 * c["osc"]
    pix_in = ok.pass1(*_args(c), c["nb"], c["rcw"])
    pix = {k: pix_in[k] for k in ("zeta", "J", "u1", "u2", "seff")}
    gt = ok.geometry_terms(c["det"], c["axis"], c["s0"], c["rc"], c["frames_phi0"], c["osc"], c["sd"], c["sm"],
                           c["rcw"], c["Dz"], c["Dy"], c["Dx"], pix)
    p = ok.profile_values(c["prof"], c["rc"]["region"], gt) * gt["R"][:, :, None, None]
    bg = np.full(8, 3.0)
    I = np.linspace(10, 1e4, 8)
    sb = bg[:, None, None, None] + I[:, None, None, None] * p
    Ip = np.zeros(8)
    if partner:
        rcp = dict(c["rc"], S=c["rc"]["pS"], phi=c["rc"]["pphi"])
        gtp = ok.geometry_terms(c["det"], c["axis"], c["s0"], rcp, c["frames_phi0"], c["osc"], c["sd"], c["sm"],
                                c["rcw"], c["Dz"], c["Dy"], c["Dx"])
        pp = ok.profile_values(c["prof"], c["rc"]["region"], gtp) * gtp["R"][:, :, None, None]
        Ip = np.linspace(5e3, 50, 8)
        sb = sb + Ip[:, None, None, None] * pp
    c["sb"] = sb
    r2 = ok.pass2(*_args(c), c["prof"], bg, np.full(8, 100.0), c["nb"], c["rcw"], pix)
    return r2, I, Ip, p, sb, bg


def test_fit_recovers_noise_free_intensity():
    r2, I, _, p, sb, bg = _noise_free(False)
    assert np.max(np.abs(r2["iprf"] - I) / I) < 1e-10
    # a11 = sum p^2 / (b + I p) over the fit mask; var = 1/a11 + background term
    use = r2["_use"]
    a11 = np.sum(np.where(use, p * p / (bg[:, None, None, None] + I[:, None, None, None] * p), 0), axis=(1, 2, 3))
    assert np.all(r2["vprf"] >= 1 / a11 * (1 - 1e-9))
    # corr is numpy.corrcoef of p and the net counts on the fit mask
    for r in range(3):
        m = use[r]
        cc = np.corrcoef(p[r][m], (sb[r] - bg[r])[m])[0, 1]
        assert abs(cc - r2["corr"][r]) < 1e-9


def test_partner_fit_recovers_both_intensities():
    r2, I, Ip, *_ = _noise_free(True)
    assert np.max(np.abs(r2["iprf"] - I) / I) < 1e-8
    assert np.max(np.abs(r2["ipartner"] - Ip) / Ip) < 1e-8


def test_pixel_S_flat_closed_form():
    from chunks import flat_detector
    det = flat_detector(orgx=500.25, orgy=480.75, distance=150.0)
    rng = np.random.default_rng(3)
    x, y = rng.uniform(1, 1000, 50), rng.uniform(1, 1000, 50)
    X = np.c_[(x - 500.25) * 0.075, (y - 480.75) * 0.075, np.full(50, 150.0)]
    S = X / np.linalg.norm(X, axis=1, keepdims=True) / det["wl"]
    assert np.max(np.abs(ok.pixel_S(det, x, y) - S)) < 1e-15

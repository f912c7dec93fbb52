"""The oracle of the post-integration corrections and SCALE's combination (oracle/finalise.py), pinned by closed forms
and hand-computed values; the product's host pieces (attenuation tables, space-group numbers, group medians, the
combination, the INTEGRATE.HKL record formatter) against them.  No GPU needed."""
import numpy as np
import pytest

from oracle import finalise as of


def _beam_at(two_theta, azimuth):
    """Unit diffracted beam at scattering angle 2 theta, azimuth measured from +x about the beam (0, 0, 1)."""
    return np.stack([np.sin(two_theta) * np.cos(azimuth), np.sin(two_theta) * np.sin(azimuth), np.cos(two_theta)], -1)


def test_lorentz_is_sin_two_theta_normal_to_the_axis_and_zero_along_it():
    tt = np.radians(np.linspace(1, 80, 40))
    S = _beam_at(tt, np.pi / 2) / 0.9                     # the plane normal to the axis (1, 0, 0)
    z, _ = of.lp_factor(S, [0, 0, 1.0], [1.0, 0, 0], [0, 1.0, 0], 0.5)
    np.testing.assert_allclose(z, np.sin(tt), rtol=1e-14)
    z0, _ = of.lp_factor(_beam_at(tt, 0.0), [0, 0, 1.0], [1.0, 0, 0], [0, 1.0, 0], 0.5)
    np.testing.assert_allclose(z0, 0.0, atol=1e-15)


def test_polarisation_unpolarised_and_fully_polarised_limits():
    rng = np.random.default_rng(3)
    tt, az = rng.uniform(0.01, 1.4, 200), rng.uniform(0, 2 * np.pi, 200)
    S = _beam_at(tt, az)
    _, P = of.lp_factor(S, [0, 0, 1.0], [1.0, 0, 0], [0, 1.0, 0], 0.5)
    np.testing.assert_allclose(P, (1 + np.cos(tt) ** 2) / 2, rtol=1e-14)
    # fully polarised along x (normal y): no loss in the y-z plane, cos^2 2 theta in the x-z plane
    _, P1 = of.lp_factor(_beam_at(tt, np.pi / 2), [0, 0, 1.0], [1.0, 0, 0], [0, 1.0, 0], 1.0)
    np.testing.assert_allclose(P1, 1.0, rtol=1e-14)
    _, P2 = of.lp_factor(_beam_at(tt, 0.0), [0, 0, 1.0], [1.0, 0, 0], [0, 1.0, 0], 1.0)
    np.testing.assert_allclose(P2, np.cos(tt) ** 2, rtol=1e-12, atol=1e-15)
    # a normal not orthogonal to the beam is made so first: (0, 1, 1) acts as (0, 1, 0)
    _, P3 = of.lp_factor(S, [0, 0, 1.0], [1.0, 0, 0], [0, 1.0, 1.0], 0.8)
    _, P4 = of.lp_factor(S, [0, 0, 1.0], [1.0, 0, 0], [0, 1.0, 0.0], 0.8)
    np.testing.assert_allclose(P3, P4, rtol=1e-14)


def test_absorption_closed_forms():
    mu, t, mua, D = 3.7, 0.45, 3.3e-4, 150.0
    S = np.array([[0, 0, 1.0], [0.6, 0, 0.8], [1.0, 0, 1e-5]])
    f = of.absorption_factors(S, [0, 0, 1.0], D, t, mu, mua)
    c = np.array([1.0, 0.8, 1e-3])                         # the last one clipped at grazing incidence
    np.testing.assert_allclose(f, (1 - np.exp(-mu * t / c)) * np.exp(-mua * D / c), rtol=1e-13)
    f0 = of.absorption_factors(S, [0, 0, 1.0], D, 0.0, mu, mua)          # no sensor model: air only
    np.testing.assert_allclose(f0, np.exp(-mua * D / c), rtol=1e-13)
    # segmented: the segment's normal and the lab path of the detector point
    nrm = np.array([[0, 0, 1.0], [0, 0.6, 0.8], [0, 0, 1.0]])
    lab = np.array([[0, 0, 150.0], [10, 20, 160.0], [0, 0, 1.0]])
    e = S / np.linalg.norm(S, axis=1, keepdims=True)
    cs = np.clip((e * nrm).sum(1), 1e-3, 1)
    fs = of.absorption_factors(S, None, None, t, mu, mua, normals=nrm, lab=lab)
    np.testing.assert_allclose(fs, (1 - np.exp(-mu * t / cs)) * np.exp(-mua * np.linalg.norm(lab, axis=1)),
                               rtol=1e-13)


def test_mu_linear_table_nodes_and_log_log_midpoint():
    from xpg_amd.geometry import HC_KEV_A, mu_linear
    lam = lambda e: HC_KEV_A / e
    assert mu_linear(lam(10.0), "Si") == pytest.approx(32.87 * 2.33 / 10, rel=1e-13)
    assert mu_linear(lam(10.0), "silicon") == pytest.approx(32.87 * 2.33 / 10, rel=1e-13)
    assert mu_linear(lam(20.0), "AIR") == pytest.approx(0.7779 * 1.205e-3 / 10, rel=1e-13)
    assert mu_linear(lam(np.sqrt(150.0)), "Si") == pytest.approx(np.sqrt(32.87 * 9.76) * 2.33 / 10, rel=1e-12)
    assert mu_linear(1.0, "unobtainium") == 0.0


def test_columns_by_hand():
    S = np.array([[0, 1.0, 1.0], [0, 0, 1.0], [0, 1.0, 1.0], [0, 1.0, 1.0]])
    z, P = np.array([0.5, 0.0, 0.5, -0.5]), np.array([0.8, 1.0, 0.8, 0.8])
    ab = np.array([0.9, 0.9, 0.9, 0.9])
    c = of.columns(S, iprf=np.array([100.0, 50.0, np.nan, 10.0]), vprf=np.array([25.0, 4.0, 1.0, -1.0]),
                   isum=np.array([90.0, 1.0, 2.0, 3.0]), vsum=np.array([16.0, -2.0, 1.0, 1.0]),
                   p_valid=np.array([0.5, 1.0, 0.0, 2.0]), p_total=np.array([1.0, 0.0, 0.0, 1.0]),
                   corr=np.array([0.9, np.nan, 0.1, -0.25]), coverage=np.array([1.0, 1.0, 1.0, 0.5]), zeta=z, P=P,
                   absorb=ab)
    rlp = np.abs(z) / P / ab
    np.testing.assert_allclose(c["rlp"], rlp)
    assert c["I"][0] == pytest.approx(100 * rlp[0]) and c["sig"][0] == pytest.approx(5 * rlp[0])
    assert c["Isum"][0] == pytest.approx(90 * rlp[0]) and c["sigsum"][0] == pytest.approx(4 * rlp[0])
    assert c["sigsum"][1] == 0.0                            # negative variance -> 0
    np.testing.assert_allclose(c["peak"], [50.0, 100.0, 0.0, 50.0])   # p_total 0 -> 1e-9; clipped at 1; coverage
    np.testing.assert_allclose(c["corr"], [90.0, 0.0, 10.0, -25.0])
    assert c["good"].tolist() == [True, False, False, False]   # rlp 0, NaN, sigma 0


def _mates():
    """Two unique reflections observed twice each in frame block 0, one observed once."""
    r = dict(I=np.array([100.0, 110.0, 50.0, 70.0, 30.0]), sig=np.ones(5), Isum=np.array([104.0, 106.0, 60.0, 60.0,
                                                                                         30.0]),
             sigsum=np.ones(5) * 2, peak=np.full(5, 100.0), n_over=np.zeros(5, int), z=np.array([3.0, 40, 5, 60, 7]),
             x=np.full(5, 10.0), y=np.full(5, 10.0))
    return r, np.array([7, 7, 9, 9, 11])


def test_rmeas_by_hand():
    r, uid = _mates()
    _, inv = np.unique(uid, return_inverse=True)
    n = np.bincount(inv)
    # group 7: mean 105, |dev| 5 each, x sqrt(2); group 9: mean 60, |dev| 10 each; the single one is left out
    expect = (2 * 5 * np.sqrt(2) + 2 * 10 * np.sqrt(2)) / (100 + 110 + 50 + 70)
    assert of.rmeas(r["I"], inv, n) == pytest.approx(expect, rel=1e-14)


def test_combination_weights_and_limits():
    r, uid = _mates()
    I, sig, imid, rm0, rm = of.combine_prf_sum(r, uid, 100, 100)
    # summation agrees better within the groups: some finite I_mid wins over profile fitting alone
    assert np.isfinite(imid) and rm < rm0
    w = 1 / (1 + np.maximum(r["Isum"], 0) / imid)
    np.testing.assert_allclose(I, w * r["I"] + (1 - w) * r["Isum"], rtol=1e-14)
    np.testing.assert_allclose(sig, np.sqrt((w * 1) ** 2 + ((1 - w) * 2) ** 2), rtol=1e-14)
    # profile-fitted intensities that agree exactly within their groups: any summation weight only adds spread, so
    # profile fitting alone is kept (I_mid = inf, w = 1)
    r2 = dict(r, I=np.array([100.0, 100.0, 50.0, 50.0, 30.0]))
    I2, sig2, imid2, rm02, rm2 = of.combine_prf_sum(r2, uid, 100, 100)
    assert imid2 == np.inf and rm02 == rm2 == 0.0
    np.testing.assert_array_equal(I2, r2["I"])
    np.testing.assert_array_equal(sig2, r2["sig"])


def test_product_combination_and_group_medians_equal_the_oracle():
    from xpg_amd import finalise as pf
    rng = np.random.default_rng(11)
    n = 20000
    uid = rng.integers(0, 3000, n)
    Itrue = rng.exponential(500.0, 3000)[uid]
    r = dict(I=Itrue + rng.normal(0, 20, n), sig=np.full(n, 20.0), Isum=Itrue * 0.97 + rng.normal(0, 15, n),
             sigsum=np.full(n, 25.0), peak=rng.uniform(60, 100, n), n_over=(rng.uniform(size=n) < 0.01).astype(int),
             z=rng.uniform(0, 900, n), x=rng.uniform(1, 2000, n), y=rng.uniform(1, 2000, n),
             p_fg=rng.uniform(0.7, 1.0, n), zeta_k=rng.uniform(-1, 1, n))
    r["p_fg"][rng.uniform(size=n) < 0.05] = np.nan
    I_o, s_o, imid_o, rm0_o, rm_o = of.combine_prf_sum(r, uid, 2000, 2000)
    I_p, s_p, info = pf.combine_prf_sum(r, uid, 2000, 2000)
    assert info["i_mid"] == imid_o and info["rmeas"] == pytest.approx(rm_o, rel=1e-14)
    np.testing.assert_array_equal(I_p, I_o)
    np.testing.assert_array_equal(s_p, s_o)
    v = rng.normal(size=1001)
    g = rng.integers(0, 37, 1001)
    _, inv = np.unique(g, return_inverse=True)
    med = pf._group_medians(v, inv, inv.max() + 1)
    assert np.array_equal(med, [np.median(v[inv == k]) for k in range(inv.max() + 1)])


def test_space_group_numbers():
    from xpg_amd.finalise import space_group_number
    assert space_group_number("P 21 21 21") == 19 and space_group_number("P4(3)2(1)2") == 96
    assert space_group_number("C 2 2 21") == 20 and space_group_number("I 41 3 2") == 214
    assert space_group_number("") == 1 and space_group_number("P -1") == 1


def test_record_formatter_prints_what_python_formats():
    """xpg_format_integrate_rows (library host code) against the record the f-string of xpg/io/legacy.py:168-174
    writes, on values that exercise rounding ties, signs, NaN / inf, huge and tiny numbers."""
    from xpg_amd.finalise import integrate_hkl_records
    from xpg_amd.geometry import Geometry
    rng = np.random.default_rng(5)
    n = 3000
    g = Geometry(wavelength=1.0, beam=np.array([0, 0, 1.0]), axis=np.array([1.0, 0, 0]), orgx=1000, orgy=1000,
                 distance=150, ed1=np.array([1.0, 0, 0]), ed2=np.array([0, 1.0, 0]), qx=0.075, qy=0.075, nx=2000,
                 ny=2000, phi0=0, osc=0.1)
    vals = lambda: np.concatenate([rng.normal(0, 1e4, n - 12), [0.05, 0.15, 0.25, -0.05, -0.0, 1e300, -1e-300,
                                                                 np.nan, np.inf, -np.inf, 2.5, 99.995]])
    refl = dict(hkl=rng.integers(-60, 60, (n, 3)), I=vals(), sig=np.abs(vals()), x=vals(), y=vals(),
                z=rng.uniform(0, 100, n), rlp=np.abs(vals()), peak=rng.uniform(0, 100, n), corr=vals(),
                xo=np.where(rng.uniform(size=n) < 0.3, np.nan, vals()), yo=vals(), zo=rng.uniform(0, 100, n),
                S=rng.normal(size=(n, 3)))
    got = integrate_hkl_records(g, refl).split("\n")[:-1]
    u0 = g.s0 / np.linalg.norm(g.s0)
    alf0, bet0 = np.degrees(np.arctan2(u0[1], u0[0])), np.degrees(np.arccos(np.clip(u0[2], -1, 1)))
    u1 = refl["S"] / np.linalg.norm(refl["S"], axis=1, keepdims=True)
    alf1, bet1 = np.degrees(np.arctan2(u1[:, 1], u1[:, 0])), np.degrees(np.arccos(np.clip(u1[:, 2], -1, 1)))
    fs = lambda z: np.minimum(z, np.floor(z) + 0.94)
    xo = np.where(np.isfinite(refl["xo"]), refl["xo"], refl["x"])
    h = refl["hkl"]
    assert len(got) == n
    for i in range(n):
        want = (f" {h[i, 0]} {h[i, 1]} {h[i, 2]} {refl['I'][i]:.3E} {refl['sig'][i]:.3E} {refl['x'][i]:.1f} "
                f"{refl['y'][i]:.1f} {fs(refl['z'][i]):.1f} {refl['rlp'][i]:.5f} {float(refl['peak'][i]):.5f} "
                f"{int(np.nan_to_num(refl['corr'][i])):d} 0 {xo[i]:.1f} {refl['yo'][i]:.1f} {fs(refl['zo'][i]):.1f} "
                f"{alf0:.2f} {bet0:.2f} {alf1[i]:.2f} {bet1[i]:.2f} 0.00 1")
        assert got[i] == want, (i, got[i], want)

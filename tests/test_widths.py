"""Host logic of the spot-width estimate (xpg_amd/widths.py; xpg/integrate.py:1002-1154), pinned by hand cases,
a per-row statement of the resolved-curve rule, noise-free recovery of the rocking width and a numerical integral of
the truncated 2D Gaussian."""
import numpy as np
import pytest
from scipy import integrate
from scipy.special import erf

from xpg_amd import widths as W


def test_sample_ranges_hand_cases():
    # 100 frames of 0.1 deg: span 50 frames for 5 deg, capped at 100 // 5 = 20
    assert W.sample_ranges(1, 100, 0.1, 30) == [(1, 20), (21, 40), (41, 60), (61, 80), (81, 100)]
    # 1 deg frames: 5 frames for 5 deg, 30 / 3 = 10 wins; starts 1 + int(i * 90 / 4) = 1, 23, 46, 68, 91
    assert W.sample_ranges(1, 100, 1.0, 30) == [(1, 10), (23, 32), (46, 55), (68, 77), (91, 100)]
    # a short sweep: span at least 3 frames, ranges clipped at the last frame
    # (starts 5 + int(i * 2 / 4) = 5, 5, 6, 6, 7)
    assert W.sample_ranges(5, 9, 0.5, 30) == [(5, 7), (5, 7), (6, 8), (6, 8), (7, 9)]
    # negative oscillation counts by its magnitude
    assert W.sample_ranges(1, 100, -0.1, 30) == W.sample_ranges(1, 100, 0.1, 30)


def _resolved_row(fz, dz, z, iz0, fb, max_edge=0.5):
    """The rule for one row, written with explicit loops."""
    dz = min(int(dz), len(fz))
    if dz < 3:
        return True
    rec = [k for k in range(dz) if fz[k] != 0]
    first, last = (rec[0], rec[-1]) if rec else (0, 0)
    c = int(np.floor((z - iz0) / max(fb, 1)))
    c = min(max(c, 0), max(dz - 1, 0))
    centre = fz[c]
    edge = 0.5 * (fz[first] + fz[last])
    return bool(centre > 0 and edge < max_edge * centre and last - first >= 2)


def test_rocking_resolved_hand_cases():
    rows = [
        ([0, 1, 5, 10, 5, 1, 0], 7, 3.5, True),      # a falling-off curve
        ([5, 5, 5, 5, 5, 0, 0], 5, 2.5, False),      # flat: the edges equal the centre
        ([0, 0, 2, 8, 2, 0, 0], 7, 3.5, True),       # zeros outside the sweep: first / last recorded frames 2 and 4
        ([0, 0, 0, 8, 2, 0, 0], 7, 3.5, False),      # only two recorded frames
        ([1, -4, 1, 0, 0, 0, 0], 3, 1.5, False),     # negative centre
        ([9, 9, 0, 0, 0, 0, 0], 2, 1.0, True),       # fewer than three frames always pass
        ([0, 0, 0, 0, 0, 0, 0], 7, 3.5, False),      # nothing recorded
        ([1, 2, 10, 2, 9, 0, 0], 5, 2.5, False),     # the last recorded frame is high
    ]
    fz = np.array([r[0] for r in rows], np.float32)
    dz = np.array([r[1] for r in rows])
    z = np.array([r[2] for r in rows])
    got = W.rocking_resolved(fz, dz, z, np.zeros(len(rows), int), np.ones(len(rows), int))
    assert got.tolist() == [r[3] for r in rows]


def test_rocking_resolved_matches_the_row_rule():
    rng = np.random.default_rng(4)
    n, width = 3000, 17
    fz = rng.integers(-3, 12, (n, width)).astype(np.float32)
    fz[rng.random((n, width)) < 0.4] = 0
    dz = rng.integers(1, width + 3, n)
    iz0 = rng.integers(-5, 5, n)
    fb = rng.integers(1, 4, n)
    z = iz0 + rng.uniform(-2, width * 3, n)
    got = W.rocking_resolved(fz, dz, z, iz0, fb)
    want = [_resolved_row(fz[i], dz[i], z[i], iz0[i], fb[i]) for i in range(n)]
    assert got.tolist() == want
    assert 0.05 < got.mean() < 0.95


def _rocking(sm, phi, zeta, iz0, dz, z_to_phi):
    k = iz0[:, None] + np.arange(dz)[None]
    lo, hi = z_to_phi(k.astype(float)), z_to_phi(k + 1.0)
    s = sm / np.abs(zeta)[:, None] * np.sqrt(2)
    return 0.5 * (erf((hi - phi[:, None]) / s) - erf((lo - phi[:, None]) / s))


def test_fit_rocking_width_recovers_a_grid_width():
    rng = np.random.default_rng(7)
    n, dz = 400, 15
    z_to_phi = lambda z: 12.0 + 0.1 * np.asarray(z)
    iz0 = rng.integers(0, 500, n)
    phi = z_to_phi(iz0 + dz / 2 + rng.uniform(-1, 1, n))
    zeta = rng.uniform(0.2, 1.0, n) * rng.choice([-1, 1], n)
    guess = 0.07
    grid = guess * np.exp(np.linspace(np.log(0.2), np.log(3.0), 61))
    for k in (5, 37, 60):
        fz = _rocking(grid[k], phi, zeta, iz0, dz, z_to_phi) * rng.uniform(50, 5000, n)[:, None]
        assert W.fit_rocking_width(fz, iz0, phi, zeta, z_to_phi, guess) == pytest.approx(grid[k], rel=1e-12)
        # grid points evaluated in small groups give the same answer
        assert W.fit_rocking_width(fz, iz0, phi, zeta, z_to_phi, guess, max_cells=dz * n * 3) == \
            pytest.approx(grid[k], rel=1e-12)


def test_fit_rocking_width_edge_cases():
    z_to_phi = lambda z: 0.1 * np.asarray(z)
    # no row with a positive total: every cost is 0, the first grid point wins
    assert W.fit_rocking_width(np.zeros((3, 5)), np.zeros(3, int), np.zeros(3), np.ones(3), z_to_phi, 0.1) == \
        pytest.approx(0.02)


def test_strong_selection_tiers():
    n = 400
    ok = np.ones(n, bool)
    # >= 150 above 20: that tier
    isg = np.r_[np.full(160, 25.0), np.full(240, 12.0)]
    (a,) = W.strong_selection([(ok, isg)])
    assert a.sum() == 160
    # < 150 above 20: the tier above 10
    isg = np.r_[np.full(100, 25.0), np.full(50, 12.0), np.full(250, 1.0)]
    (b,) = W.strong_selection([(ok, isg)])
    assert b.sum() == 150
    # fewer than 20 above 10 in every range, >= 30 above 5 pooled: every range above 5
    r1 = (ok, np.r_[np.full(15, 12.0), np.full(10, 6.0), np.full(375, 1.0)])
    r2 = (ok, np.r_[np.full(5, 7.0), np.full(395, 1.0)])
    p = W.strong_selection([r1, r2])
    assert [x.sum() for x in p] == [25, 5]
    # ... and fewer than 30 above 5 pooled: nothing
    r2 = (ok, np.r_[np.full(4, 7.0), np.full(396, 1.0)])
    assert W.strong_selection([(ok, np.r_[np.full(25, 6.0), np.full(375, 1.0)]), r2]) == [None, None]
    # a range with fewer than 20 is dropped while another is kept
    p = W.strong_selection([(ok, np.r_[np.full(30, 30.0), np.full(370, 1.0)]), (ok, np.full(n, 11.0) * (np.arange(n) < 10))])
    assert p[0].sum() == 30 and p[1] is None


def test_truncated_gaussian_moment_factor():
    # <r^2 / 2> / sigma^2 over the disc r <= FG_R of a unit 2D Gaussian, by quadrature
    num = integrate.quad(lambda r: 0.5 * r ** 2 * r * np.exp(-0.5 * r ** 2), 0, W.FG_R)[0]
    den = integrate.quad(lambda r: r * np.exp(-0.5 * r ** 2), 0, W.FG_R)[0]
    assert W.TRUNC_T2 == pytest.approx(num / den, rel=1e-12)


def test_grows_with_box():
    class M:
        def __init__(self, d, m):
            self.sigma_d, self.sigma_m = d, m
    assert W.grows_with_box(M(0.13, 0.1), M(0.11, 0.1), M(0.09, 0.1))
    assert not W.grows_with_box(M(0.13, 0.1), M(0.11, 0.1), M(0.10, 0.1))     # no growth before
    assert not W.grows_with_box(M(0.12, 0.1), M(0.11, 0.1), M(0.09, 0.1))     # no growth now
    assert W.grows_with_box(M(0.1, 0.3), M(0.1, 0.25), M(0.1, 0.2))

"""The forward-model oracle (oracle/forward.py) against brute force written the other way round: a loop over the
window's voxels that finds, per reflection, the support pixel a voxel is (xpg/forward/kernels.py:34-134).  CPU only."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import forward as FO
from oracle import kernels as ok


def _case(seed, T, n=7, Dz=3, b=4):
    rng = np.random.default_rng(seed)
    win = SimpleNamespace(f0=5, f1=9, y0=10, ny=6, x0=20, nx=7)
    nv = (win.f1 - win.f0) * win.ny * win.nx
    R = rng.uniform(0.1, 1.0, (T, n, Dz)).astype(np.float32)
    P = rng.uniform(0.0, 1.0, (T, n, b, b)).astype(np.float32)
    # origins that clip the supports on every side of the window (and one entirely outside)
    iz0 = rng.integers(win.f0 - Dz + 1, win.f1, n)
    iy0 = rng.integers(win.y0 - b + 1, win.y0 + win.ny, n)
    ix0 = rng.integers(win.x0 - b + 1, win.x0 + win.nx, n)
    iz0[0], iy0[0], ix0[0] = win.f0, win.y0, win.x0
    iz0[1] = win.f1 + 3
    I = rng.uniform(1, 50, n).astype(np.float32)
    beta = rng.uniform(0, 2, n).astype(np.float32)
    A = rng.normal(0, 3, nv).astype(np.float32)
    M = rng.uniform(0.5, 9, nv).astype(np.float32)
    NS = rng.integers(-2, 20, nv).astype(np.int32)
    return SimpleNamespace(win=win, nv=nv, R=R, P=P, iz0=iz0, iy0=iy0, ix0=ix0, I=I, beta=beta, A=A, M=M, NS=NS,
                           n=n, Dz=Dz, b=b, T=T)


def _brute_terms(c):
    """{(row, voxel): f} by a loop over window voxels and reflections."""
    w = c.win
    out = {}
    for kz in range(w.f1 - w.f0):
        for ky in range(w.ny):
            for kx in range(w.nx):
                v = (kz * w.ny + ky) * w.nx + kx
                for r in range(c.n):
                    z = kz + w.f0 - int(c.iz0[r])
                    y = ky + w.y0 - int(c.iy0[r])
                    x = kx + w.x0 - int(c.ix0[r])
                    if 0 <= z < c.Dz and 0 <= y < c.b and 0 <= x < c.b:
                        out[(r, v)] = sum(float(c.R[t, r, z]) * float(c.P[t, r, y, x]) for t in range(c.T))
    return out


@pytest.mark.parametrize("T", [1, 2])
@pytest.mark.parametrize("with_beta", [True, False])
def test_scatter_matches_brute_force(T, with_beta):
    c = _case(3 + T, T)
    mu0 = np.linspace(0, 1, c.nv)
    want = mu0.copy()
    for (r, v), f in _brute_terms(c).items():
        want[v] += f * float(c.I[r]) + (float(c.beta[r]) if with_beta else 0.0)
    got = FO.scatter(mu0, c.R, c.P, c.I, c.beta if with_beta else None, c.iz0, c.iy0, c.ix0, c.win)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    assert not np.array_equal(got, mu0)


@pytest.mark.parametrize("T", [1, 2])
@pytest.mark.parametrize("mode", [0, 1])
def test_gather_matches_brute_force(T, mode):
    c = _case(11 + T, T)
    c.NS[::5] = -1
    rows = np.array([4, 0, 2, 2, 1, 3, 0])            # repeated output rows add
    n_out = 5
    NSf = c.NS.astype(np.float32)
    NSf[1::7] = np.nan                                  # NaN counts: no term
    g0, fd0 = np.zeros(n_out), np.zeros(n_out)
    g_w, fd_w = g0.copy(), fd0.copy()
    for (r, v), f in _brute_terms(c).items():
        ns = float(NSf[v])
        if not ns >= 0:
            continue
        m = float(c.M[v])
        if mode == 0:
            g_w[rows[r]] += f * float(c.A[v]) / m
        else:
            g_w[rows[r]] += f * (1.0 - max(ns, 0.0) / m)
            fd_w[rows[r]] += f * f / max(m, 1e-6)
    g, fd = FO.gather(g0, fd0 if mode == 1 else None, c.R, c.P, rows, c.iz0, c.iy0, c.ix0, c.win,
                      c.A if mode == 0 else None, c.M, NSf, mode)
    np.testing.assert_allclose(g, g_w, rtol=1e-12, atol=1e-12)
    if mode == 1:
        np.testing.assert_allclose(fd, fd_w, rtol=1e-12, atol=1e-12)
    assert np.count_nonzero(g_w) >= 4


def test_fraction_voxel_ids_orientation():
    """A single support at the window origin: voxel ids run x fastest, then y, then z, and f is R[z] P[y, x]."""
    win = SimpleNamespace(f0=2, f1=4, y0=1, ny=3, x0=0, nx=5)
    R = np.array([[[1.0, 10.0]]], np.float32)
    P = np.arange(9, dtype=np.float32).reshape(1, 1, 3, 3)
    v = FO.voxel_ids([2], [1], [0], 2, 3, win)
    assert v[0, 0, 0, 0] == 0 and v[0, 0, 0, 1] == 1 and v[0, 0, 1, 0] == 5 and v[0, 1, 0, 0] == 15
    f = FO.fractions(R, P)
    assert f[0, 1, 2, 1] == 10.0 * 7.0
    mu = FO.scatter(np.zeros(30), R, P, [1.0], None, [2], [1], [0], win)
    assert mu[(1 * 3 + 2) * 5 + 1] == 70.0


# ---- fraction_terms (the separable factors themselves) ---------------------------------------------------------------

def _fm_geom(**kw):
    from xpg_amd.geometry import Geometry
    base = dict(wavelength=1.0, beam=np.array([0, 0, 1.0]), axis=np.array([1.0, 0, 0]), orgx=1024.5, orgy=1024.5,
                distance=150, ed1=np.array([1.0, 0, 0]), ed2=np.array([0, 1.0, 0]), qx=0.075, qy=0.075, nx=2048,
                ny=2048, phi0=0, osc=0.1)
    base.update(kw)
    return Geometry(**base)


def _fm_rows(g, n, b, seed=0, pxs=1):
    """n reflections on the detector with supports of b x b (super-)pixels centred on their predicted position."""
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(300, 1700, n), rng.uniform(300, 1700, n)
    S = g.pixel_to_s(x, y)
    phi = np.radians(rng.uniform(2.0, 3.0, n))
    z = np.degrees(phi) / g.osc
    ix0 = np.floor(x / pxs - b / 2).astype(int)
    iy0 = np.floor(y / pxs - b / 2).astype(int)
    return S, phi, ix0, iy0, z


def test_fraction_terms_gaussian_sums_to_one_and_rocking_curve_covers_the_range():
    from xpg_amd.geometry import detector_consts
    g = _fm_geom()
    det = detector_consts(g)
    n, b, Dz = 40, 31, 41
    S, phi, ix0, iy0, z = _fm_rows(g, n, b)
    iz0 = np.floor(z - Dz / 2).astype(int)
    sd, sm = np.radians(0.03), np.radians(0.03)
    for rcw in (None, ok.rc_default_weights()):
        (R, P), = FO.fraction_terms(det, g.axis, g.s0, 0.0, np.radians(g.osc), sd, sm, S, phi, ix0, iy0, iz0, Dz, b,
                                    rcw=rcw)
        # the pixel-integrated Gaussian of a spot of ~3 pixels inside 31 pixels: sum 1 (midpoint rule)
        np.testing.assert_allclose(P.sum((1, 2)), 1.0, atol=2e-3)
        # the rocking curve over +-20 frames of 0.1 degree (sigma_M / |zeta| below 0.2 degree here): all of it
        np.testing.assert_allclose(R.sum(1), 1.0, atol=1e-9)
    # a box narrower than 3: J is the constant (qx pxs)^2
    (_, P2), = FO.fraction_terms(det, g.axis, g.s0, 0.0, np.radians(g.osc), sd, sm, S, phi, ix0 + 14, iy0 + 14, iz0,
                                 Dz, 2, pxs=2)
    e = 2 * 0.075
    assert np.all(P2 > 0) and np.isfinite(P2).all()
    (_, P3), = FO.fraction_terms(det, g.axis, g.s0, 0.0, np.radians(g.osc), 1.0, sm, S, phi, ix0, iy0, iz0, Dz, 2)
    np.testing.assert_allclose(P3.max(), e * e / 4 / (2 * np.pi * (1 + e * e / 4 / 12.0)), rtol=0.02)


def test_fraction_terms_halo_and_learnt_profile_limits():
    from xpg_amd.geometry import detector_consts
    g = _fm_geom()
    det = detector_consts(g)
    n, b, Dz = 30, 15, 21
    S, phi, ix0, iy0, z = _fm_rows(g, n, b, seed=2)
    iz0 = np.floor(z - Dz / 2).astype(int)
    sd, sm = np.radians(0.03), np.radians(0.08)
    args = (det, g.axis, g.s0, 0.0, np.radians(g.osc), sd, sm, S, phi, ix0, iy0, iz0, Dz, b)
    (R, P), = FO.fraction_terms(*args)
    # kappa = 1 without mixture weights: the halo is the core itself, so the terms add up to the core's fraction
    (R1, P1), (Rh, Ph) = FO.fraction_terms(*args, halo=(0.3, 1.0))
    f = R1[:, :, None, None] * P1[:, None] + Rh[:, :, None, None] * Ph[:, None]
    np.testing.assert_allclose(f, R[:, :, None, None] * P[:, None], rtol=1e-12, atol=1e-300)
    # w = 0: the first term is the core
    (R0, P0), (Rz, _) = FO.fraction_terms(*args, halo=(0.0, 2.0))
    assert np.array_equal(R0, R) and np.array_equal(P0, P) and not Rz.any()
    # a learnt profile that is the unit Gaussian on the grid reproduces the Gaussian factor to interpolation error
    prof = np.repeat(ok.gaussian_profile()[None], 9, 0).astype(np.float32)
    (Rp, Pp), = FO.fraction_terms(*args, prof=prof, region=np.arange(n) % 9, umax=ok.UMAX)
    assert np.array_equal(Rp, R)
    inner = P > 0.3 * P.max(axis=(1, 2), keepdims=True)
    assert np.abs(Pp[inner] / P[inner] - 1).max() < 1.5e-2

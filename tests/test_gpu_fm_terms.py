"""fraction_terms on the device (xpg_fm_terms) against the float64 oracle (oracle/forward.py): parametric and learnt
profiles, the mixture rocking curve and the plain Gaussian, the halo, per-reflection widths, boxes narrower than 3,
super-pixels, a two-segment detector with parallax; the factors feed the fused scatter unchanged."""
import numpy as np
import pytest
import torch

from oracle import forward as FO
from oracle import kernels as ok

pytestmark = pytest.mark.gpu


def _geom(**kw):
    from xpg_amd.geometry import Geometry
    base = dict(wavelength=1.0, beam=np.array([0, 0, 1.0]), axis=np.array([1.0, 0, 0]), orgx=1024.5, orgy=1024.5,
                distance=150, ed1=np.array([1.0, 0, 0]), ed2=np.array([0, 1.0, 0]), qx=0.075, qy=0.075, nx=2048,
                ny=2048, phi0=0, osc=0.1)
    base.update(kw)
    return Geometry(**base)


def _rows(g, n, b, seed, pxs=1):
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(300, 1700, n), rng.uniform(300, 1700, n)
    S = g.pixel_to_s(x, y)
    phi = np.radians(rng.uniform(2.0, 3.0, n))
    z = np.degrees(phi) / g.osc
    return S, phi, np.floor(x / pxs - b / 2).astype(int), np.floor(y / pxs - b / 2).astype(int), z


CASES = dict(
    gauss=dict(),
    mixture=dict(rcw=True),
    profile=dict(prof=True),
    halo=dict(halo=(0.2, 2.5), rcw=True, prof=True),
    per_row_widths=dict(rows_sd=True, halo=(0.1, 3.0)),
    narrow=dict(b=2),
    superpixel=dict(pxs=2, b=7),
    segmented=dict(segmented=True, prof=True),
)


@pytest.mark.parametrize("case", sorted(CASES))
def test_fraction_terms_match_oracle(case):
    from xpg_amd import forward as F
    from xpg_amd.geometry import Segment, detector_consts
    c = CASES[case]
    kw = dict(sensor_thickness=0.45, mu=3.7)
    if c.get("segmented"):
        th = np.radians(0.4)
        kw["segments"] = [Segment(1, 2048, 1, 1000),
                          Segment(1, 2048, 1040, 2048, orgx=0.4, orgy=-0.3, distance=0.2,
                                  ex=np.array([np.cos(th), np.sin(th), 0.02]),
                                  ey=np.array([-np.sin(th), np.cos(th), 0.0]))]
    g = _geom(**kw)
    det = detector_consts(g)
    n, b, Dz, pxs = 3000, c.get("b", 11), 9, c.get("pxs", 1)
    S, phi, ix0, iy0, z = _rows(g, n, b, seed=hash(case) % 1000, pxs=pxs)
    iz0 = np.floor(z - Dz / 2).astype(int)
    rng = np.random.default_rng(5)
    sd, sm = np.radians(0.03), np.radians(0.08)
    if c.get("rows_sd"):
        sd = sd * rng.uniform(0.7, 1.4, n)
        sm = sm * rng.uniform(0.7, 1.4, n)
    rcw = None
    if c.get("rcw"):
        rcw = ok.rc_default_weights() * rng.uniform(0.2, 1.8, 21).astype(np.float32)
        rcw = (rcw / rcw.sum()).astype(np.float32)
    prof = region = None
    if c.get("prof"):
        prof = np.abs(ok.gaussian_profile()[None] * rng.uniform(0.5, 1.5, (9, 41, 41))).astype(np.float32)
        region = rng.integers(0, 9, n)
    halo = c.get("halo")
    args = (g.axis, g.s0, 0.0, np.radians(g.osc), sd, sm, S, phi, ix0, iy0, iz0, Dz, b)
    print(f"fraction terms {case}: {n} reflections, b {b}, Dz {Dz}", flush=True)
    want = FO.fraction_terms(det, *args, prof=prof, region=region, umax=ok.UMAX, halo=halo, rcw=rcw, pxs=pxs)
    dev = torch.device("cuda:0")
    dargs = (torch.as_tensor(g.axis, device=dev), torch.as_tensor(g.s0, device=dev), 0.0, np.radians(g.osc),
             torch.as_tensor(sd, device=dev) if np.ndim(sd) else sd,
             torch.as_tensor(sm, device=dev) if np.ndim(sm) else sm, torch.as_tensor(S, device=dev),
             torch.as_tensor(phi, device=dev), torch.as_tensor(ix0, device=dev), torch.as_tensor(iy0, device=dev),
             torch.as_tensor(iz0, device=dev), Dz, b)
    got = F.fraction_terms(det, *dargs, prof=None if prof is None else torch.as_tensor(prof, device=dev),
                           region=None if region is None else torch.as_tensor(region, device=dev), umax=ok.UMAX,
                           halo=halo, rcw=rcw, pxs=pxs)
    assert len(got) == len(want)
    for (Rg, Pg), (Rw, Pw) in zip(got, want):
        assert Rg.dtype == torch.float32 and Rg.shape == Rw.shape and Pg.shape == Pw.shape
        for a, w in ((Rg.cpu().numpy(), Rw), (Pg.cpu().numpy(), Pw)):
            np.testing.assert_allclose(a, w.astype(np.float32), rtol=3e-7, atol=1e-37)


def test_factors_feed_the_scatter():
    """fraction_factors' stacked (T, n, Dz) / (T, n, b, b) factors are what scatter takes: the scattered fraction of
    each reflection over a window holding its whole support sums to sum_t (sum R_t)(sum P_t)."""
    from types import SimpleNamespace
    from xpg_amd import forward as F
    from xpg_amd.geometry import detector_consts
    g = _geom()
    n, b, Dz = 200, 9, 7
    S, phi, ix0, iy0, z = _rows(g, n, b, seed=9)
    iz0 = np.floor(z - Dz / 2).astype(int)
    dev = torch.device("cuda:0")
    R, P = F.fraction_factors(detector_consts(g), g.axis, g.s0, 0.0, np.radians(g.osc), np.radians(0.03),
                              np.radians(0.08), torch.as_tensor(S, device=dev), torch.as_tensor(phi, device=dev), ix0,
                              iy0, iz0, Dz, b, halo=(0.2, 2.0), rcw=ok.rc_default_weights())
    assert R.shape == (2, n, Dz) and P.shape == (2, n, b, b)
    win = SimpleNamespace(f0=int(iz0.min()), f1=int(iz0.max()) + Dz, y0=int(iy0.min()), ny=int(iy0.max() - iy0.min()) + b,
                          x0=int(ix0.min()), nx=int(ix0.max() - ix0.min()) + b)
    want = (R.double().sum(2) * P.double().sum((2, 3))).sum(0)
    for r in range(0, n, 37):                   # one reflection at a time: no overlaps between supports
        mu = torch.zeros((win.f1 - win.f0) * win.ny * win.nx, dtype=torch.float32, device=dev)
        sel = slice(r, r + 1)
        F.scatter(mu, R[:, sel].contiguous(), P[:, sel].contiguous(), torch.ones(1, device=dev), None,
                  torch.as_tensor(iz0[sel], device=dev), torch.as_tensor(iy0[sel], device=dev),
                  torch.as_tensor(ix0[sel], device=dev), win)
        assert float(mu.double().sum()) == pytest.approx(float(want[r]), rel=1e-5)

"""Seeded synthetic kernel inputs (shoebox chunks) for the kernel parity tests and the golden-vector script.

A chunk is what xpg/integrate.py:505-519 hands to one kernel call: value-coded shoeboxes ``sb`` (n, Dz, Dy, Dx), the
per-reflection dict ``rc``, the neighbour columns ``nb`` and the call scalars.  Spots are drawn from the kernels' own
model (Gaussian in the Kabsch frame times a Gaussian rocking curve) so that the profile fit has something to fit, with
Poisson noise, masked / overloaded pixels, unrecorded frames, neighbours inside the box and overlapping partners.
"""
from __future__ import annotations

import numpy as np

from oracle import kernels as ok


def flat_detector(nx=1024, ny=1024, orgx=512.5, orgy=512.5, distance=200.0, qx=0.075, qy=0.075, wl=1.0,
                  bkg=None, sensor_thickness=0.0, mu=0.0):
    """Detector constants of a flat single panel, float64 (xpg/integrate_kernels.py:48-65 with xpg/detector.py:84-100:
    P0 = distance*ed3, Bx = qx*ed1, By = qy*ed2 for ed1 = x, ed2 = y)."""
    P0 = np.array([[0.0, 0.0, distance]])
    Bx = np.array([[qx, 0.0, 0.0]])
    By = np.array([[0.0, qy, 0.0]])
    U1, U2 = Bx / qx, By / qy
    plx = float(sensor_thickness > 0 and mu > 0)
    return dict(nx=nx, ny=ny, orgx=orgx, orgy=orgy, P0=P0, Bx=Bx, By=By, U1=U1, U2=U2, UN=np.cross(U1, U2), qx=qx,
                qy=qy, wl=wl, tsens=float(max(sensor_thickness, 1e-6)), mu=float(max(mu, 1e-6)), plx=plx,
                segmap=None, gx=None, gy=None, bkg=bkg)


def analytic_template(nx=1024, ny=1024):
    """A smooth static background template of mean ~1 (rings, polarisation), float32 (xpg/init.py:203-235 makes such
    a template from the data); analytic so that fixtures can rebuild it instead of storing it."""
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    r = np.hypot(x - nx / 2, y - ny / 2) / max(nx, ny)
    return (1.0 + 0.35 * np.cos(2 * np.pi * r * 6.0) * np.exp(-r) + 0.1 * (x - nx / 2) / nx).astype(np.float32)


def make_chunk(n=48, Dz=7, Dy=11, Dx=11, seed=0, fb=1, pxs=1, partners=True, neighbours=True, codes=True,
               bkg_template=False, intensity=(50.0, 5e4), background=2.0, sd_deg=0.02, sm_deg=0.1, osc_deg=0.1,
               rcw="default", det=None):
    rng = np.random.default_rng(seed)
    det = det or flat_detector(bkg=(analytic_template() if bkg_template else None))
    axis = np.array([1.0, 0.0, 0.0])
    s0 = np.array([0.0, 0.0, 1.0]) / det["wl"]
    osc = np.radians(osc_deg)
    frames_phi0 = 0.0
    sd, sm = np.radians(sd_deg), np.radians(sm_deg)
    nfr = 400
    # centres away from the rotation axis (|zeta| not tiny)
    x = rng.uniform(60, det["nx"] - 60, n)
    y = np.where(rng.random(n) < 0.5, rng.uniform(60, 380, n), rng.uniform(640, det["ny"] - 60, n))
    z = rng.uniform(60, nfr - 60, n)
    S = ok.pixel_S(det, x, y)
    phi = frames_phi0 + osc * z
    # box origin: xpg/integrate.py:250-287 (pxs super-pixels, fb frames per bin)
    hx, hy = Dx // 2, Dy // 2
    ix0 = np.floor(x - ok.PIX0 + 0.5).astype(np.int64) // pxs - hx
    iy0 = np.floor(y - ok.PIX0 + 0.5).astype(np.int64) // pxs - hy
    hz = Dz // 2
    iz0 = np.floor(z - (fb - 1) / 2.0).astype(np.int64) - hz * fb
    region = rng.integers(0, 9, n)
    rc = dict(S=S, phi=phi, ix0=ix0, iy0=iy0, iz0=iz0, fb=np.full(n, float(fb)), region=region)
    if pxs > 1:
        rc["pxs"] = np.full(n, float(pxs))
    if isinstance(rcw, str) and rcw == "default":
        rcw = ok.rc_default_weights()
    elif isinstance(rcw, str) and rcw == "random":
        w = rng.gamma(2.0, 1.0, ok.RC_C.size) * ok.rc_default_weights()
        rcw = (w / w.sum()).astype(np.float32)
    gt = ok.geometry_terms(det, axis, s0, rc, frames_phi0, osc, sd, sm, rcw, Dz, Dy, Dx)
    gauss = np.exp(-0.5 * (gt["u1"] ** 2 + gt["u2"] ** 2)) / (2 * np.pi)
    ptrue = gauss * (gt["J"] / gt["seff"][:, None, None] ** 2)[:, None] * gt["R"][:, :, None, None]
    I = np.exp(rng.uniform(np.log(intensity[0]), np.log(intensity[1]), n))
    bref = ok.bkg_ref(det, gt["x"], gt["y"], Dy, Dx)
    lam = background * bref * (fb * pxs * pxs) + I[:, None, None, None] * ptrue
    sb = rng.poisson(lam).astype(np.float64)
    has_p = np.zeros(n)
    pS, pphi = S.copy(), phi.copy()
    if partners:
        hp = rng.random(n) < 0.3
        has_p[hp] = 1.0
        xp_ = x + rng.uniform(-2.5, 2.5, n) * pxs
        yp_ = y + rng.uniform(-2.5, 2.5, n) * pxs
        pS = np.where(hp[:, None], ok.pixel_S(det, xp_, yp_), S)
        pphi = np.where(hp, phi + osc * rng.uniform(-1.5, 1.5, n), phi)
        # the partner's spot is in the data too
        rcp = dict(rc, S=pS, phi=pphi)
        gtp = ok.geometry_terms(det, axis, s0, rcp, frames_phi0, osc, sd, sm, rcw, Dz, Dy, Dx)
        gp = np.exp(-0.5 * (gtp["u1"] ** 2 + gtp["u2"] ** 2)) / (2 * np.pi)
        pp = gp * (gtp["J"] / gtp["seff"][:, None, None] ** 2)[:, None] * gtp["R"][:, :, None, None]
        Ip = np.exp(rng.uniform(np.log(intensity[0]), np.log(intensity[1]), n))
        sb = sb + np.where(hp[:, None, None, None], rng.poisson(Ip[:, None, None, None] * pp), 0)
    rc.update(has_p=has_p, pS=pS, pphi=pphi)
    if codes:
        u = rng.random(sb.shape)
        sb[u < 0.02] = -1.0
        sb[(u >= 0.02) & (u < 0.025)] = -2.0
        # a few boxes lose frames (sweep ends / excluded frames: unrecorded) and one loses a whole detector row
        for r in rng.choice(n, max(1, n // 8), replace=False):
            sb[r, : rng.integers(1, Dz)] = -3.0
        sb[0, :, 0, :] = -1.0
    # neighbour columns: K = 8 (xpg/integrate.py:288-302), padding -1e6 / width 1
    K = 8
    nbx = np.full((n, K), -1e6, np.float32)
    nby = np.full((n, K), -1e6, np.float32)
    nbz = np.full((n, K), -1e6, np.float32)
    nbsx = np.ones((n, K), np.float32)
    nbsz = np.ones((n, K), np.float32)
    if neighbours:
        for r in range(n):
            m = rng.integers(0, 4)
            for k in range(m):
                nbx[r, k] = x[r] + rng.uniform(-1.2, 1.2) * Dx * pxs
                nby[r, k] = y[r] + rng.uniform(-1.2, 1.2) * Dy * pxs
                nbz[r, k] = z[r] + rng.uniform(-1.0, 1.0) * Dz * fb
                nbsx[r, k] = rng.uniform(1.0, 3.0)
                nbsz[r, k] = rng.uniform(0.8, 3.0)
    nb = dict(x=nbx, y=nby, z=nbz, sx=nbsx, sz=nbsz)
    prof = np.repeat(ok.gaussian_profile()[None], 9, 0).astype(np.float64)
    prof = prof * rng.uniform(0.85, 1.15, prof.shape)
    du = (2 * ok.UMAX / (ok.PROFILE_GRID - 1)) ** 2
    prof = (prof / (prof.sum(axis=(1, 2), keepdims=True) * du)).astype(np.float32)
    return dict(sb=sb, rc=rc, det=det, axis=axis, s0=s0, frames_phi0=frames_phi0, osc=osc, sd=sd, sm=sm, gain=1.0,
                nb=nb, rcw=rcw, prof=prof, Dz=Dz, Dy=Dy, Dx=Dx, I_true=I)

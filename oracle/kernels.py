"""FP64 oracle of the integration kernels (pass 1, profile accumulation, pass 2) -- TEST INFRASTRUCTURE ONLY.

Restates xpg/integrate_kernels.py in float64 numpy.  Every elementary operation whose rounding decides a mask
(norms, dot products, the neighbour-ellipsoid test, the foreground radius, the fit threshold) is written out in a fixed
left-to-right order, which the HIP kernels repeat operation for operation, so that masks agree bit for bit.

Functions and the reference lines they follow:

  pixel_S              xpg/integrate_kernels.py:78-105   (flat or segmented panel, geometric correction, parallax)
  eps12                xpg/integrate_kernels.py:108-115
  phi_rc               xpg/integrate_kernels.py:130-141  (mixture centres rounded to float32 as the reference does)
  rc_basis             xpg/integrate_kernels.py:144-151
  bilinear             xpg/integrate_kernels.py:154-166
  neighbour_mask       xpg/integrate_kernels.py:186-213
  geometry_terms       xpg/integrate_kernels.py:219-276
  profile_values       xpg/integrate_kernels.py:278-285
  pass1                xpg/integrate_kernels.py:287-365
  accumulate_profiles  xpg/integrate_kernels.py:367-380
  pass2                xpg/integrate_kernels.py:382-483  (IRLS stop rule: "chunk" as the reference, or "reflection")
  gaussian_profile     xpg/integrate_kernels.py:491-495
  normalise_profiles   xpg/integrate.py:546-556

Pinning: tests/test_oracle.py checks every function here against (1) golden vectors written by running the
reference's own make_kernels on float64 inputs (tests/golden/), which must agree to ~1e-12 with identical masks,
and (2) properties the mathematics fixes: numpy.gradient for the pixel-area derivatives, closed-form erf integrals
that sum to one, bilinear weights that conserve mass, the clipped mean of a flat background, exact recovery of a
noise-free b*T + I*p model by the Poisson fit, Cramer's rule for the 2x2 partner fit, numpy.corrcoef for corr.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf, erfc

PIX0 = 1.0
N_FG = 3.0
N_BOX = 5.0
FG_R = N_FG
BG_R = N_FG + 0.5
PROFILE_GRID = 41
N_IRLS = 6
IRLS_TOL = 1e-3
VAR_FLOOR = 1e-3
RC_C = np.linspace(-3.5, 3.5, 21)
RC_W = 0.25
UMAX = N_BOX * 0.9
# the reference builds the mixture centres of its device rocking curve as a float32 tensor
RC_C32 = RC_C.astype(np.float32).astype(np.float64)
# working precision of the per-voxel arithmetic (geometry, masks, sums, fits): float64 for the oracle.  The CPU
# throughput baseline sets float32, the precision of the reference's tensor program; profile accumulation and
# normalisation stay float64 either way.
FT = np.float64


def _dot3(a, b):
    """(a0*b0 + a1*b1) + a2*b2 over the last axis, in this order."""
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _norm3(a):
    return np.sqrt(_dot3(a, a))


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def rc_default_weights():
    """xpg/integrate_kernels.py:125-127: mixture weights that make a unit Gaussian."""
    w = np.exp(-0.5 * RC_C ** 2 / (1 - RC_W ** 2))
    return (w / w.sum()).astype(np.float32)


def gaussian_profile(G=PROFILE_GRID):
    """xpg/integrate_kernels.py:491-495."""
    u = np.linspace(-UMAX, UMAX, G)
    U1, U2 = np.meshgrid(u, u)
    return (np.exp(-0.5 * (U1 ** 2 + U2 ** 2)) / (2 * np.pi)).astype(np.float32)


# ---------------------------------------------------------------------------------------------------------------------
# detector
# ---------------------------------------------------------------------------------------------------------------------

def _lookup_segment(segmap, x, y):
    """xpg/detector.py:134-138."""
    ix = np.clip(np.floor(x - 0.5).astype(np.int64), 0, segmap.shape[1] - 1)
    iy = np.clip(np.floor(y - 0.5).astype(np.int64), 0, segmap.shape[0] - 1)
    return segmap[iy, ix]


def _bilinear_table(tab, x, y):
    """xpg/detector.py:141-153."""
    ny, nx = tab.shape
    fx = np.clip(x - 1.0, 0, nx - 1.000001)
    fy = np.clip(y - 1.0, 0, ny - 1.000001)
    ix = np.floor(fx).astype(np.int64)
    iy = np.floor(fy).astype(np.int64)
    ix1 = np.minimum(ix + 1, nx - 1)
    iy1 = np.minimum(iy + 1, ny - 1)
    tx, ty = fx - ix, fy - iy
    return (((tab[iy, ix] * (1 - tx) * (1 - ty) + tab[iy, ix1] * tx * (1 - ty))
             + tab[iy1, ix] * (1 - tx) * ty) + tab[iy1, ix1] * tx * ty)


def pixel_S(det, x, y):
    """S of detector points (x, y): xpg/integrate_kernels.py:78-105.  ``det`` holds float64 arrays P0, Bx, By, U1, U2,
    UN (K,3), scalars orgx, orgy, qx, qy, wl, tsens, mu, plx, and optional segmap / gx, gy / bkg tables (None)."""
    x, y = np.broadcast_arrays(np.asarray(x, FT), np.asarray(y, FT))
    if det.get("segmap") is None:
        k = np.zeros(x.shape, np.int64)
    else:
        k = np.maximum(_lookup_segment(det["segmap"], x, y), 0).astype(np.int64)
    P0, Bx, By = det["P0"][k], det["Bx"][k], det["By"][k]
    U1, U2, UN = det["U1"][k], det["U2"][k], det["UN"][k]
    if det.get("gx") is None:
        xc, yc = x, y
    else:
        xc = x + _bilinear_table(det["gx"], x, y)
        yc = y + _bilinear_table(det["gy"], x, y)

    def plane(xx, yy):
        return (P0 + (xx - det["orgx"])[..., None] * Bx) + (yy - det["orgy"])[..., None] * By
    X = plane(xc, yc)
    for _ in range(2 if det["plx"] else 0):
        e = X / _norm3(X)[..., None]
        cosa = np.clip(_dot3(e, UN), 1e-3, 1.0)
        L = det["tsens"] / cosa
        muL = det["mu"] * L
        lbar = 1.0 / det["mu"] - L * np.exp(-muL) / np.maximum(-np.expm1(-muL), 1e-12)
        d1, d2 = lbar * _dot3(e, U1), lbar * _dot3(e, U2)
        X = plane(xc - det["plx"] * d1 / det["qx"], yc - det["plx"] * d2 / det["qy"])
    return X / (_norm3(X) * det["wl"])[..., None]


def eps12(det, S, e1, e2, x, y):
    """xpg/integrate_kernels.py:108-115: x (n,Dx), y (n,Dy) -> eps1, eps2 (n,Dy,Dx)."""
    Sp = pixel_S(det, x[:, None, :], y[:, :, None])
    d = Sp - S[:, None, None, :]
    invS = 1.0 / _norm3(S)[:, None, None]
    return _dot3(d, e1[:, None, None, :]) * invS, _dot3(d, e2[:, None, None, :]) * invS


def _erfc6(x):
    """erfc, 0 beyond 6 (erfc < 2.2e-17 there; float64 erf rounds to 1, so erf differences are 0 as well)."""
    return np.where(x > 6.0, 0.0, erfc(np.minimum(x, 6.0)))


def erf_diff(a, b):
    """erf(a) - erf(b) for a >= b, from ea = erfc|a| and eb = erfc|b|: eb - ea where both arguments are >= 0 and
    ea - eb where both are <= 0 -- the same number as the reference's erf(a) - erf(b) without the cancellation of two
    erf values rounded towards +-1, so that the tails of a rocking curve keep their relative precision (the
    reference's keeps an absolute 1e-16) -- and (2 - ea) - eb across zero."""
    a, b = np.broadcast_arrays(np.asarray(a), np.asarray(b))
    ea, eb = _erfc6(np.abs(a)), _erfc6(np.abs(b))
    return np.where(b >= 0, eb - ea, np.where(a <= 0, ea - eb, (2.0 - ea) - eb))


def phi_rc(phi_lo, phi_hi, phi_c, zeta, sm, rcw=None):
    """Recorded fraction of the rocking curve between phi_lo and phi_hi: xpg/integrate_kernels.py:130-141 (the erf
    differences by erf_diff)."""
    s = sm / np.maximum(np.abs(zeta), 1e-3)
    if rcw is None:
        a = (phi_hi - phi_c) / (s * np.sqrt(2.0))
        b = (phi_lo - phi_c) / (s * np.sqrt(2.0))
        return 0.5 * erf_diff(a, b)
    sw = s * RC_W * np.sqrt(2.0)
    out = np.zeros(np.shape(phi_lo), FT)
    rcw = np.asarray(rcw, FT)
    for j in range(RC_C32.size):
        a = ((phi_hi - phi_c) - RC_C32[j] * s) / sw
        b = ((phi_lo - phi_c) - RC_C32[j] * s) / sw
        out = out + (rcw[j] * 0.5) * erf_diff(a, b)
    return out


def rc_basis(phi_lo, phi_hi, phi_c, zeta, sm):
    """Frame integrals of each mixture component (N, Dz, K): xpg/integrate_kernels.py:144-151."""
    s = (sm / np.maximum(np.abs(zeta), 1e-3))[:, None, None]
    sw = s * RC_W * np.sqrt(2.0)
    a = (phi_hi[..., None] - phi_c[:, None, None] - RC_C * s) / sw
    b = (phi_lo[..., None] - phi_c[:, None, None] - RC_C * s) / sw
    return 0.5 * (erf(a) - erf(b))


def bilinear(u1, u2, G=PROFILE_GRID, umax=UMAX):
    """xpg/integrate_kernels.py:154-166."""
    f1 = (u1 + umax) / (2 * umax) * (G - 1)
    f2 = (u2 + umax) / (2 * umax) * (G - 1)
    inside = (f1 >= 0) & (f1 <= G - 1) & (f2 >= 0) & (f2 <= G - 1)
    f1 = np.clip(f1, 0, G - 1 - 1e-4)
    f2 = np.clip(f2, 0, G - 1 - 1e-4)
    i1 = np.floor(f1).astype(np.int64)
    i2 = np.floor(f2).astype(np.int64)
    return inside, i1, i2, f1 - i1, f2 - i2


def neighbour_mask(xx, yy, zz, nb, shape, skip_first=None):
    """Voxels inside any neighbour's foreground ellipsoid: xpg/integrate_kernels.py:186-213.  The reference's per-column
    skip of columns that reach no box is exact (monotone rounding), so every column is evaluated here."""
    nbm = np.zeros(shape, bool)
    for kk in range(nb["x"].shape[1]):
        sx = nb["sx"][:, kk].astype(FT)[:, None, None, None]
        sz = nb["sz"][:, kk].astype(FT)[:, None, None, None]
        dx = (xx - nb["x"][:, kk].astype(FT)[:, None, None, None]) / sx
        dy = (yy - nb["y"][:, kk].astype(FT)[:, None, None, None]) / sx
        dz = (zz - nb["z"][:, kk].astype(FT)[:, None, None, None]) / sz
        m = ((dx * dx + dy * dy) + dz * dz) < 1.0
        if kk == 0 and skip_first is not None:
            m = m & ~np.asarray(skip_first, bool)[:, None, None, None]
        nbm |= m
    return nbm


def bkg_ref(det, x, y, Dy, Dx):
    """Template at shoebox pixel centres: xpg/integrate_kernels.py:68-75.  x (n,Dx), y (n,Dy) -> (n,1,Dy,Dx)."""
    n = x.shape[0]
    if det.get("bkg") is None:
        return np.ones((n, 1, Dy, Dx))
    ny, nx = det["bkg"].shape
    ix = np.clip(np.round(x - 1.0).astype(np.int64), 0, nx - 1)
    iy = np.clip(np.round(y - 1.0).astype(np.int64), 0, ny - 1)
    return det["bkg"].astype(FT)[iy[:, :, None], ix[:, None, :]][:, None]


# ---------------------------------------------------------------------------------------------------------------------
# per-reflection geometry
# ---------------------------------------------------------------------------------------------------------------------

def gradient2(f, axis):
    """Second-order differences of unit-spaced samples, one-sided second order at the ends (numpy.gradient with
    edge_order=2), written out in numpy's operation order."""
    f = np.moveaxis(f, axis, -1)
    g = np.empty_like(f)
    g[..., 1:-1] = (f[..., 2:] - f[..., :-2]) / 2.0
    g[..., 0] = (-1.5 * f[..., 0] + 2.0 * f[..., 1]) + -0.5 * f[..., 2]
    g[..., -1] = (0.5 * f[..., -3] + -2.0 * f[..., -2]) + 1.5 * f[..., -1]
    return np.moveaxis(g, -1, axis)


def pixel_coords(rc, Dy, Dx):
    """Pixel (or super-pixel) centres of the box in detector pixels: xpg/integrate_kernels.py:224-230."""
    s_ = np.asarray(rc.get("pxs", np.ones(len(rc["ix0"]))), FT)[:, None]
    x = (np.asarray(rc["ix0"], FT)[:, None] + np.arange(Dx, dtype=FT)[None]) * s_ + (s_ - 1) * 0.5 + PIX0
    y = (np.asarray(rc["iy0"], FT)[:, None] + np.arange(Dy, dtype=FT)[None]) * s_ + (s_ - 1) * 0.5 + PIX0
    return x, y


def geometry_terms(det, axis, s0, rc, frames_phi0, osc, sd, sm, rcw, Dz, Dy, Dx, pix=None):
    """xpg/integrate_kernels.py:219-276 in float64."""
    S = np.asarray(rc["S"], FT)
    x, y = pixel_coords(rc, Dy, Dx)
    axis = np.asarray(axis, FT)
    s0 = np.asarray(s0, FT)
    if pix is not None:
        zeta = pix["zeta"]
    else:
        c1 = _cross(S, np.broadcast_to(s0, S.shape))
        e1 = c1 / _norm3(c1)[:, None]
        c2 = _cross(S, e1)
        e2 = c2 / _norm3(c2)[:, None]
        zeta = _dot3(e1, np.broadcast_to(axis, e1.shape))
        eps1, eps2 = eps12(det, S, e1, e2, x, y)
    if pix is not None:
        J = pix["J"]
    elif Dx >= 3 and Dy >= 3:
        g1x, g1y = gradient2(eps1, 2), gradient2(eps1, 1)
        g2x, g2y = gradient2(eps2, 2), gradient2(eps2, 1)
        J = np.abs(g1x * g2y - g2x * g1y)
    else:
        a1, a2 = eps12(det, S, e1, e2, x + 0.5, y)
        b1, b2 = eps12(det, S, e1, e2, x - 0.5, y)
        c1_, c2_ = eps12(det, S, e1, e2, x, y + 0.5)
        d1, d2 = eps12(det, S, e1, e2, x, y - 0.5)
        J = np.abs((a1 - b1) * (c2_ - d2) - (a2 - b2) * (c1_ - d1))
    fb = np.asarray(rc["fb"], FT)[:, None]
    k = np.asarray(rc["iz0"], FT)[:, None] + fb * np.arange(Dz, dtype=FT)[None]
    phi_lo = frames_phi0 + osc * k
    phi_hi = phi_lo + osc * fb
    lo, hi = np.minimum(phi_lo, phi_hi), np.maximum(phi_lo, phi_hi)
    phi = np.asarray(rc["phi"], FT)
    R = phi_rc(lo, hi, np.broadcast_to(phi[:, None], lo.shape), np.broadcast_to(zeta[:, None], lo.shape), sm, rcw)
    if pix is not None:
        seff, u1c, u2c = pix["seff"], pix["u1"], pix["u2"]
    else:
        psz = np.sqrt(J[:, Dy // 2, Dx // 2])
        seff = sd * np.sqrt(1.0 + (psz / sd) ** 2 / 12.0)
        u1c = eps1 / seff[:, None, None]
        u2c = eps2 / seff[:, None, None]
    phi_c = 0.5 * (lo + hi)
    return dict(x=x, y=y, u1=u1c[:, None], u2=u2c[:, None], u1c=u1c, u2c=u2c, J=J, R=R, zeta=zeta, phi_c=phi_c,
                seff=seff)


def profile_values(prof, region, gt):
    """p2 = P2(u) J / seff^2: xpg/integrate_kernels.py:278-285.  prof (R,G,G) float32 values."""
    prof = np.asarray(prof, FT)
    inside, i1, i2, t1, t2 = bilinear(gt["u1"], gt["u2"], prof.shape[-1])
    r = np.asarray(region)[:, None, None, None]
    v = ((((1 - t1) * (1 - t2)) * prof[r, i2, i1] + (t1 * (1 - t2)) * prof[r, i2, i1 + 1])
         + ((1 - t1) * t2) * prof[r, i2 + 1, i1]) + (t1 * t2) * prof[r, i2 + 1, i1 + 1]
    return np.where(inside, v, 0.0) * (gt["J"] / (gt["seff"] ** 2)[:, None, None])[:, None]


def _zz(rc, Dz):
    fb = np.asarray(rc["fb"], FT)[:, None]
    return (np.asarray(rc["iz0"], FT)[:, None] + fb * (np.arange(Dz, dtype=FT)[None] + 0.5))[:, :, None, None]


# ---------------------------------------------------------------------------------------------------------------------
# pass 1
# ---------------------------------------------------------------------------------------------------------------------

def pass1(sb, rc, det, axis, s0, frames_phi0, osc, sd, sm, gain, nb, rcw):
    """Background, summation, centroids, moments and projected images: xpg/integrate_kernels.py:287-365.
    sb: (n,Dz,Dy,Dx) value codes (-3 unrecorded, -2 overloaded, -1 masked, >= 0 counts)."""
    sb = np.asarray(sb, FT)
    n, Dz, Dy, Dx = sb.shape
    gt = geometry_terms(det, axis, s0, rc, frames_phi0, osc, sd, sm, rcw, Dz, Dy, Dx)
    valid = sb >= 0
    recorded = sb > -2.5
    over = sb == -2
    r2 = gt["u1"] ** 2 + gt["u2"] ** 2
    fg2 = r2 <= FG_R ** 2
    Rmax = np.max(gt["R"], axis=1, keepdims=True)
    fgz = gt["R"] >= 1e-3 * Rmax + 1e-6
    fg = fg2 & fgz[:, :, None, None]
    xx = gt["x"][:, None, None, :]
    yy = gt["y"][:, None, :, None]
    zz = _zz(rc, Dz)
    nbm = neighbour_mask(xx, yy, zz, nb, sb.shape)
    bgm = valid & ~fg & ~nbm & (r2 >= BG_R ** 2)
    c = np.where(valid, sb, 0.0)
    bref = bkg_ref(det, gt["x"], gt["y"], Dy, Dx)
    ax = (1, 2, 3)
    w = bgm.astype(FT)
    b = np.sum(w * c, axis=ax) / np.maximum(np.sum(w * bref, axis=ax), 1e-6)
    for _ in range(4):
        bp = b[:, None, None, None] * bref
        thr = (bp + 3.0 * np.sqrt(gain * np.maximum(bp, 0.5))) + 1.0
        w = (bgm & (c <= thr)).astype(FT)
        b = np.sum(w * c, axis=ax) / np.maximum(np.sum(w * bref, axis=ax), 1e-6)
    nbg = np.sum(w * bref, axis=ax)
    fgv = fg & valid & ~nbm
    fgw = fgv.astype(FT)
    bpix = b[:, None, None, None] * bref
    net = c - bpix
    nfg = np.sum(fgw, axis=ax)
    sfb = np.sum(fgw * bref, axis=ax)
    isum = np.sum(fgw * net, axis=ax)
    vb = sfb ** 2 * gain * np.maximum(b, VAR_FLOOR) / np.maximum(nbg, 1.0)
    vsum = gain * np.sum(fgw * c, axis=ax) + vb
    vsum0 = gain * np.sum(fgw * bpix, axis=ax) + vb
    wpos = np.maximum(net, 0.0) * fgw
    sw = np.maximum(np.sum(wpos, axis=ax), 1e-6)
    xo = np.sum(wpos * xx, axis=ax) / sw
    yo = np.sum(wpos * yy, axis=ax) / sw
    zo = np.sum(wpos * zz, axis=ax) / sw
    e1v = gt["u1"] * gt["seff"][:, None, None, None]
    e2v = gt["u2"] * gt["seff"][:, None, None, None]
    wsg = net * fgw
    sws = np.maximum(np.sum(wsg, axis=ax), 1e-6)
    mu1 = np.sum(wsg * e1v, axis=ax) / sws
    mu2 = np.sum(wsg * e2v, axis=ax) / sws
    m_d = (np.sum(wsg * (e1v ** 2 + e2v ** 2), axis=ax) / sws - mu1 ** 2 - mu2 ** 2) / 2.0
    e3 = gt["zeta"][:, None] * (gt["phi_c"] - np.asarray(rc["phi"], FT)[:, None])
    mu3 = np.sum(wsg * e3[:, :, None, None], axis=ax) / sws
    m_m = np.sum(wsg * (e3 ** 2)[:, :, None, None], axis=ax) / sws - mu3 ** 2
    clean = valid & ~nbm
    fgz4 = fgz[:, :, None, None]
    use_net = np.where(fgz4 & clean, net, 0.0)
    img2 = np.sum(use_net, axis=1)
    all_valid2 = np.all(clean | ~fgz4, axis=1)
    frac_rec = np.sum(np.where(np.any(recorded, axis=(2, 3)), gt["R"], 0.0), axis=1)
    n_over = np.sum(over & fg, axis=ax)
    n_fg_masked = np.sum(fg & ~valid & recorded, axis=ax)
    n_fg_nb = np.sum(fg & nbm, axis=ax)
    fz = np.sum(np.where(fg2 & valid & ~nbm, net, 0.0), axis=(2, 3))
    return dict(bg=b, nbg=nbg, isum=isum, vsum=vsum, vsum0=vsum0, xo=xo, yo=yo, zo=zo, m_d=m_d, m_m=m_m, fz=fz,
                img2=img2, all_valid2=all_valid2, frac_rec=frac_rec, n_over=n_over, n_fg_masked=n_fg_masked,
                nfg=nfg, zeta=gt["zeta"], J=gt["J"], u1=gt["u1c"], u2=gt["u2c"], seff=gt["seff"], n_fg_nb=n_fg_nb,
                # masks, for bit-exact comparison with the device kernels
                _nbm=nbm, _fg=fg, _bgm=bgm, _w=w.astype(bool), _R=gt["R"])


# ---------------------------------------------------------------------------------------------------------------------
# profile accumulation and normalisation
# ---------------------------------------------------------------------------------------------------------------------

def accumulate_profiles(img2, ok2, J, u1, u2, region, isum, sel, seff, n_regions, G=PROFILE_GRID):
    """Scatter normalised projected spots into per-region grids: xpg/integrate_kernels.py:367-380.  float64 grids."""
    img2, J, u1, u2 = (np.asarray(a, np.float64) for a in (img2, J, u1, u2))
    isum, seff = np.asarray(isum, np.float64), np.asarray(seff, np.float64)
    dens = img2 / np.maximum(isum, 1e-6)[:, None, None] / np.maximum(J / (seff ** 2)[:, None, None], 1e-12)
    w = (np.asarray(sel, bool)[:, None, None] & np.asarray(ok2, bool)).astype(np.float64)
    inside, i1, i2, t1, t2 = bilinear(u1, u2, G)
    w = w * inside
    r = np.broadcast_to(np.asarray(region)[:, None, None], u1.shape)
    num = np.zeros((n_regions, G, G))
    den = np.zeros((n_regions, G, G))
    for di, dj, ww in ((0, 0, (1 - t1) * (1 - t2)), (1, 0, t1 * (1 - t2)), (0, 1, (1 - t1) * t2), (1, 1, t1 * t2)):
        np.add.at(num, (r, i2 + dj, i1 + di), w * ww * dens)
        np.add.at(den, (r, i2 + dj, i1 + di), w * ww)
    return num, den


def normalise_profiles(num, den, G=PROFILE_GRID, k=3.0):
    """Empirical density regularised towards the Gaussian and normalised to unit mass: xpg/integrate.py:546-556."""
    gauss = gaussian_profile(G).astype(np.float64)
    dens = np.where(den > 0, num / np.maximum(den, 1e-12), 0.0)
    du = (2 * UMAX / (G - 1)) ** 2
    prof = np.empty(num.shape, np.float32)
    for r in range(num.shape[0]):
        emp = np.clip(dens[r], 0, None)
        pr = (den[r] * emp + k * gauss) / (den[r] + k)
        pr = pr / max(pr.sum() * du, 1e-12)
        prof[r] = pr
    return prof


# ---------------------------------------------------------------------------------------------------------------------
# pass 2
# ---------------------------------------------------------------------------------------------------------------------

def pass2(sb, rc, det, axis, s0, frames_phi0, osc, sd, sm, gain, prof, bg, nbm_x, nb, rcw, pix=None,
          irls_rule="chunk"):
    """Poisson-weighted profile fit with the 2x2 partner deconvolution: xpg/integrate_kernels.py:382-483.
    irls_rule "chunk": stop when no reflection of the call moved by more than IRLS_TOL sigma (the reference,
    xpg/integrate_kernels.py:464-467); "reflection": each reflection stops on its own test."""
    sb = np.asarray(sb, FT)
    n, Dz, Dy, Dx = sb.shape
    gain = float(gain)
    bg, nbm_x = np.asarray(bg, FT), np.asarray(nbm_x, FT)
    gt = geometry_terms(det, axis, s0, rc, frames_phi0, osc, sd, sm, rcw, Dz, Dy, Dx, pix)
    valid = sb >= 0
    region = np.asarray(rc["region"])
    p = profile_values(prof, region, gt) * gt["R"][:, :, None, None]
    has_p = np.asarray(rc["has_p"], FT)
    anyp = bool(np.any(has_p > 0.5))
    if anyp:
        rc_p = dict(S=rc["pS"], phi=rc["pphi"], ix0=rc["ix0"], iy0=rc["iy0"], iz0=rc["iz0"], region=rc["region"],
                    fb=rc["fb"])
        if "pxs" in rc:
            rc_p["pxs"] = rc["pxs"]
        gtp = geometry_terms(det, axis, s0, rc_p, frames_phi0, osc, sd, sm, rcw, Dz, Dy, Dx)
        pp = (profile_values(prof, region, gtp) * gtp["R"][:, :, None, None]) * has_p[:, None, None, None]
    else:
        pp = np.zeros_like(p)
    xx = gt["x"][:, None, None, :]
    yy = gt["y"][:, None, :, None]
    zz = _zz(rc, Dz)
    nbm = neighbour_mask(xx, yy, zz, nb, sb.shape, skip_first=has_p > 0.5)
    c = np.where(valid, sb, 0.0)
    bref = bkg_ref(det, gt["x"], gt["y"], Dy, Dx)
    b = bg[:, None, None, None] * bref
    net = c - b
    ax = (1, 2, 3)
    h4 = has_p[:, None, None, None]

    pmax = np.max(p, axis=ax, keepdims=True)
    ppmax = np.max(pp, axis=ax, keepdims=True)
    use = valid & ~nbm & ((p > 0.01 * pmax) | (pp > 0.01 * np.maximum(ppmax, 1e-12)))
    wuse = use.astype(FT)
    wp = wuse * p
    I = np.sum(wuse * net, axis=ax) / np.maximum(np.sum(wp + wuse * pp, axis=ax), 1e-6)
    Ip = I * has_p
    wpp_, wpn = wp * p, wp * net
    wq = wuse * pp
    wqq, wqn, wpq = wq * pp, wq * net, wp * pp
    done = np.zeros(n, bool)              # "reflection" rule: reflections whose own test passed keep their values
    st = {}
    for it_ in range(N_IRLS):
        I_old = I
        mu = b + np.maximum(I, 0.0)[:, None, None, None] * p
        mu = mu + np.maximum(Ip, 0.0)[:, None, None, None] * pp
        iv = 1.0 / (gain * np.maximum(mu, VAR_FLOOR))
        a11 = np.sum(wpp_ * iv, axis=ax)
        b1 = np.sum(wpn * iv, axis=ax)
        a12 = np.sum(wpq * iv, axis=ax)
        a22 = np.sum(wqq * iv, axis=ax) + (1.0 - has_p) * 1.0
        b2 = np.sum(wqn * iv, axis=ax)
        det2 = np.maximum(a11 * a22 - a12 * a12, 1e-30)
        In = (a22 * b1 - a12 * b2) / det2
        Ipn = (a11 * b2 - a12 * b1) / det2 * has_p
        new = dict(I=In, Ip=Ipn, a11=a11, a12=a12, a22=a22, det2=det2, iv=iv)
        for k_, v_ in new.items():
            if k_ not in st:
                st[k_] = v_.copy()
            else:
                st[k_][~done] = v_[~done]
        if IRLS_TOL > 0 and 1 <= it_ < N_IRLS - 1:
            dz = np.abs(In - I_old) * np.sqrt(np.maximum(det2 / np.maximum(a22, 1e-30), 0.0))
            if irls_rule == "chunk":
                if float(np.max(dz)) < IRLS_TOL:
                    break
            else:
                done |= dz < IRLS_TOL
                if done.all():
                    break
        I, Ip = st["I"].copy(), st["Ip"].copy()
    I, Ip, a11, a12, a22, det2, iv = (st[k_] for k_ in ("I", "Ip", "a11", "a12", "a22", "det2", "iv"))
    # tail: xpg/integrate_kernels.py:405-429
    C = np.sum(wp * bref * iv, axis=ax)
    var_b = gain * np.maximum(bg, VAR_FLOOR) / np.maximum(nbm_x, 1.0)
    a11i = a22 / det2
    varI = a11i + (C * a11i) ** 2 * var_b
    iv0 = 1.0 / (gain * np.maximum(b, VAR_FLOOR))
    a110 = np.maximum(np.sum(wpp_ * iv0, axis=ax), 1e-12)
    C0 = np.sum(wp * bref * iv0, axis=ax)
    infl = a11i * a11
    var0 = (1.0 / a110 + (C0 / a110) ** 2 * var_b) * np.maximum(infl, 1.0)
    nuse = np.maximum(np.sum(wuse, axis=ax), 1)
    pm = np.sum(wp, axis=ax) / nuse
    resid = net - np.maximum(Ip, 0.0)[:, None, None, None] * pp
    nm = np.sum(wuse * resid, axis=ax) / nuse
    dp = (p - pm[:, None, None, None]) * wuse
    dn = (resid - nm[:, None, None, None]) * wuse
    corr = np.sum(dp * dn, axis=ax) / np.sqrt(np.maximum(np.sum(dp * dp, axis=ax) * np.sum(dn * dn, axis=ax), 1e-20))
    p_valid = np.sum(np.where(valid & ~nbm, p, 0.0), axis=ax)
    p_total = np.sum(p, axis=ax)
    rho = np.where(has_p > 0.5, a12 / np.sqrt(np.maximum(a11 * a22, 1e-30)), 0.0)
    r2 = gt["u1"] ** 2 + gt["u2"] ** 2
    Rmax = np.max(gt["R"], axis=1, keepdims=True)
    fgz = gt["R"] >= 1e-3 * Rmax + 1e-6
    fgm = (r2 <= FG_R ** 2) & fgz[:, :, None, None]
    p_fg = np.sum(np.where(fgm, p, 0.0), axis=ax) / np.maximum(np.sum(p, axis=ax), 1e-9)
    del h4
    return dict(iprf=I, vprf=varI, vprf0=var0, corr=corr, p_valid=p_valid, p_total=p_total, ipartner=Ip, rho=rho,
                p_fg=p_fg, _use=use, _nbm=nbm, _p=p, _pp=pp)

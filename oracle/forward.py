"""Oracle (test infrastructure only): float64 numpy restatement of the forward (window) model's scatter and gather
passes, xpg/forward/kernels.py:34-134 (the eager forms they fuse: WindowModel.mu / grad_ratio / grad_newton,
xpg/forward/model.py:477-537).  Only the tests use it, as the checker.

Pinning: tests/test_forward_oracle.py checks every function against brute force written the other way round -- a
loop over the window's voxels that asks, per reflection, whether the voxel lies in its support and which support
pixel it is -- on small windows that clip supports on every side, with T = 1 and 2, NS < 0 and NaN voxels, repeated
rows, and beta absent.  A transposed y / x, a dropped halo term, a wrong window origin or a missing NS mask fails it.
"""
from __future__ import annotations

import numpy as np


def fractions(R, P):
    """f (n, Dz, b, b) = sum_t R[t][:, z] P[t][:, y, x] (the separable support of each reflection,
    xpg/forward/kernels.py:54-57)."""
    R = np.asarray(R, np.float64)
    P = np.asarray(P, np.float64)
    return np.einsum("tnz,tnyx->nzyx", R, P)


def voxel_ids(iz0, iy0, ix0, Dz, b, win):
    """Window voxel index of every support voxel (n, Dz, b, b), -1 outside the window (xpg/forward/kernels.py:38-53):
    frames [f0, f1), rows [y0, y0 + ny), columns [x0, x0 + nx), index (z - f0) ny nx + (y - y0) nx + (x - x0)."""
    nf = int(win.f1) - int(win.f0)
    kz = np.asarray(iz0, np.int64)[:, None] + np.arange(Dz)[None, :] - int(win.f0)
    ky = np.asarray(iy0, np.int64)[:, None] + np.arange(b)[None, :] - int(win.y0)
    kx = np.asarray(ix0, np.int64)[:, None] + np.arange(b)[None, :] - int(win.x0)
    KZ, KY, KX = kz[:, :, None, None], ky[:, None, :, None], kx[:, None, None, :]
    ok = (KZ >= 0) & (KZ < nf) & (KY >= 0) & (KY < int(win.ny)) & (KX >= 0) & (KX < int(win.nx))
    return np.where(ok, (KZ * int(win.ny) + KY) * int(win.nx) + KX, -1)


def scatter(mu, R, P, I, beta, iz0, iy0, ix0, win):
    """mu (float64 copy) + sum over supports of I_row f + beta_row at the in-window voxels."""
    out = np.array(mu, np.float64, copy=True)
    n, Dz = np.shape(R)[1:]
    if n == 0:
        return out
    b = np.shape(P)[-1]
    f = fractions(R, P)
    val = f * np.asarray(I, np.float64)[:, None, None, None]
    if beta is not None:
        val = val + np.asarray(beta, np.float64)[:, None, None, None]
    v = voxel_ids(iz0, iy0, ix0, Dz, b, win)
    ok = v >= 0
    np.add.at(out, v[ok], val[ok])
    return out


def gather(g, fd, R, P, rows, iz0, iy0, ix0, win, A, M, NS, mode):
    """(g, fd) float64 copies plus, per reflection over its in-window voxels with NS >= 0: mode 0 sum f A / M; mode 1
    sum f (1 - max(NS, 0) / M) into g and sum f^2 / max(M, 1e-6) into fd (xpg/forward/kernels.py:63-104)."""
    g = np.array(g, np.float64, copy=True)
    fd = None if fd is None else np.array(fd, np.float64, copy=True)
    n, Dz = np.shape(R)[1:]
    if n == 0:
        return g, fd
    b = np.shape(P)[-1]
    f = fractions(R, P)
    v = voxel_ids(iz0, iy0, ix0, Dz, b, win)
    vc = np.maximum(v, 0)
    ns = np.asarray(NS).astype(np.float64)[vc]
    ok = (v >= 0) & (ns >= 0)                      # NaN compares false
    m = np.asarray(M, np.float64)[vc]
    with np.errstate(divide="ignore", invalid="ignore"):
        if mode == 0:
            w = np.asarray(A, np.float64)[vc] / m
        else:
            w = 1.0 - np.maximum(ns, 0.0) / m
            d = np.where(ok, f * f / np.maximum(m, 1e-6), 0.0).sum(axis=(1, 2, 3))
        s = np.where(ok, f * w, 0.0).sum(axis=(1, 2, 3))
    rows = np.asarray(rows, np.int64)
    np.add.at(g, rows, s)
    if mode == 1:
        np.add.at(fd, rows, d)
    return g, fd


def term_scale(R, P, rows, iz0, iy0, ix0, win, A, M, NS, mode, n_out):
    """Per output row, the sum of |terms| a gather adds (the scale a float32 rounding bound is taken against)."""
    n, Dz = np.shape(R)[1:]
    b = np.shape(P)[-1]
    f = np.abs(fractions(R, P))
    v = voxel_ids(iz0, iy0, ix0, Dz, b, win)
    vc = np.maximum(v, 0)
    ns = np.asarray(NS).astype(np.float64)[vc]
    ok = (v >= 0) & (ns >= 0)
    m = np.abs(np.asarray(M, np.float64)[vc])
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.abs(np.asarray(A, np.float64)[vc]) / m if mode == 0 else 1.0 + np.maximum(ns, 0.0) / m
        s = np.where(ok, f * w, 0.0).sum(axis=(1, 2, 3))
    out = np.zeros(n_out)
    np.add.at(out, np.asarray(rows, np.int64), s)
    return out


def fraction_terms(det, axis, s0, frames_phi0, osc, sd, sm, S, phi, ix0, iy0, iz0, Dz, Dx, prof=None, region=None,
                   umax=None, halo=None, rcw=None, pxs=1):
    """The separable factors [(R (n, Dz), P (n, Dx, Dx)), ...] of xpg/forward/model.py:206-275 in float64: Kabsch frame,
    eps at the (super-)pixel centres, J by second-order differences (Dx >= 3; else (qx pxs)^2), seff from sigma_D and
    the centre pixel, the rocking curve over frames iz0 + j, and P the learnt profile sampled bilinearly at
    u = eps / seff times J / seff^2 (or the pixel-integrated Gaussian); with halo = (w, kappa) the factors
    ((1 - w) R, P) and (w R_h, Gaussian of sqrt((kappa sd)^2 + pix^2 / 12)).  ``det`` as oracle.kernels.pixel_S takes
    it; sd, sm scalars or (n,) arrays.  The geometry pieces are oracle.kernels' (pinned there)."""
    from . import kernels as ok
    S = np.asarray(S, np.float64)
    n = S.shape[0]
    Dy = Dx
    rc = dict(ix0=np.asarray(ix0), iy0=np.asarray(iy0), pxs=np.full(n, float(pxs)))
    x, y = ok.pixel_coords(rc, Dy, Dx)
    s0 = np.asarray(s0, np.float64)
    c1 = np.cross(S, np.broadcast_to(s0, S.shape))
    e1 = c1 / np.linalg.norm(c1, axis=1, keepdims=True)
    c2 = np.cross(S, e1)
    e2 = c2 / np.linalg.norm(c2, axis=1, keepdims=True)
    zeta = e1 @ np.asarray(axis, np.float64)
    eps1, eps2 = ok.eps12(det, S, e1, e2, x, y)
    if Dx >= 3:
        g1x, g1y = ok.gradient2(eps1, 2), ok.gradient2(eps1, 1)
        g2x, g2y = ok.gradient2(eps2, 2), ok.gradient2(eps2, 1)
        J = np.abs(g1x * g2y - g2x * g1y)
    else:
        J = np.full_like(eps1, (float(det["qx"]) * pxs) ** 2)
    k = np.asarray(iz0, np.float64)[:, None] + np.arange(Dz, dtype=np.float64)[None]
    lo = frames_phi0 + osc * k
    hi = lo + osc
    lo, hi = np.minimum(lo, hi), np.maximum(lo, hi)
    phb = np.broadcast_to(np.asarray(phi, np.float64)[:, None], lo.shape)
    zb = np.broadcast_to(zeta[:, None], lo.shape)
    sd = np.broadcast_to(np.asarray(sd, np.float64), (n,))
    sm = np.broadcast_to(np.asarray(sm, np.float64), (n,))
    R = ok.phi_rc(lo, hi, phb, zb, sm[:, None], rcw)
    pix = np.sqrt(J[:, Dy // 2, Dx // 2])
    seff = sd * np.sqrt(1.0 + (pix / sd) ** 2 / 12.0)
    r2 = eps1 ** 2 + eps2 ** 2

    def gauss2(se):
        return np.exp(-0.5 * r2 / se[:, None, None] ** 2) / (2 * np.pi) * J / se[:, None, None] ** 2

    if prof is None:
        P = gauss2(seff)
    else:
        prof = np.asarray(prof, np.float64)
        inside, i1, i2, t1, t2 = ok.bilinear(eps1 / seff[:, None, None], eps2 / seff[:, None, None], prof.shape[-1],
                                            umax)
        r = np.asarray(region)[:, None, None]
        v = ((1 - t1) * (1 - t2) * prof[r, i2, i1] + t1 * (1 - t2) * prof[r, i2, i1 + 1]
             + (1 - t1) * t2 * prof[r, i2 + 1, i1] + t1 * t2 * prof[r, i2 + 1, i1 + 1])
        P = np.where(inside, v, 0.0) * J / seff[:, None, None] ** 2
    if halo is None:
        return [(R, P)]
    w, kap = halo
    sh = np.sqrt((kap * sd) ** 2 + pix ** 2 / 12.0)
    Rh = ok.phi_rc(lo, hi, phb, zb, kap * sm[:, None])
    return [((1 - w) * R, P), (w * Rh, gauss2(sh))]

"""Oracle -- test infrastructure only (tests/, the smoke check and the bench's CPU baseline use it; the product never
does).  Float64 numpy restatement of the geometry refinement before integration:

- the spot geometry it reads: pixel -> lab -> S with segments, geometric correction and four parallax iterations
  (xpg/geometry.py:186-268, xpg/detector.py:36-160), spot_to_p0 (xpg/geometry.py:313-319), predict_phi
  (xpg/geometry.py:482-502), assign_indices (xpg/indexing.py:530-535);
- the parameterisation, residuals, forward-difference Jacobian, normal equations and the Levenberg-Marquardt
  refinement with Cauchy weights, outlier rejection, Gaussian priors and ESDs (xpg/refine.py:26-443);
- the per-phi-block models with orientation tracking and the empirical-Bayes position prior, the prior itself, and
  the robust Whittaker smoothing of the orientation trajectory (xpg/stage_integrate.py:28-229).

Geometry, segment, crystal and experiment objects are read by attribute and rebuilt with dataclasses.replace, so the
oracle runs on the product's objects without importing the product.  Pinned in tests/test_refine_oracle.py.
"""
from __future__ import annotations

import copy
from dataclasses import replace

import numpy as np
from scipy.spatial.transform import Rotation

TWO_PI = 2 * np.pi


# ---------------------------------------------------------------------------------------------------------------------
# small linear algebra
# ---------------------------------------------------------------------------------------------------------------------

def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def skew(u):
    z = np.zeros_like(u[..., 0])
    return np.stack([np.stack([z, -u[..., 2], u[..., 1]], -1), np.stack([u[..., 2], z, -u[..., 0]], -1),
                     np.stack([-u[..., 1], u[..., 0], z], -1)], -2)


def rotvec_to_matrix(r):
    """xpg/geometry.py:52-65 (Rodrigues with the small-angle series)."""
    r = np.asarray(r, float)
    th2 = np.sum(r * r, -1)[..., None, None]
    th = np.sqrt(th2 + 1e-30)
    k = skew(r)
    small = th2 < 1e-12
    with np.errstate(invalid="ignore", divide="ignore"):
        a = np.where(small, 1.0 - th2 / 6.0, np.sin(th) / th)
        b = np.where(small, 0.5 - th2 / 24.0, (1.0 - np.cos(th)) / (th2 + 1e-30))
    return np.eye(3) + a * k + b * (k @ k)


def rotation_matrices(axis, angles):
    """xpg/geometry.py:410-416."""
    u = unit(axis)
    a = np.asarray(angles, float)
    c, s = np.cos(a)[..., None, None], np.sin(a)[..., None, None]
    ux = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    return c * np.eye(3) + s * ux + (1 - c) * np.outer(u, u)


def b_matrix(cell):
    """xpg/models.py:64-78, batched over leading dimensions."""
    cell = np.asarray(cell, float)
    a, b, c = cell[..., 0], cell[..., 1], cell[..., 2]
    al, be, ga = (np.radians(cell[..., i]) for i in (3, 4, 5))
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    cz = np.sqrt(np.maximum(c * c - cx * cx - cy * cy, 1e-12))
    z = np.zeros_like(a)
    real = np.stack([np.stack([a, z, z], -1), np.stack([b * np.cos(ga), b * np.sin(ga), z], -1),
                     np.stack([cx, cy, cz], -1)], -2)
    return np.linalg.inv(real)


CELL_PARAMS = {"a": ["a", "b", "c", "al", "be", "ga"], "m": ["a", "b", "c", "be"], "o": ["a", "b", "c"],
               "t": ["a", "c"], "h": ["a", "c"], "c": ["a"]}


def cell_from_free(system, v):
    """xpg/models.py:21-35."""
    v = np.asarray(v, float)
    e = [v[..., i] for i in range(v.shape[-1])]
    n90, n120 = np.full(v.shape[:-1], 90.0), np.full(v.shape[:-1], 120.0)
    full = {"a": lambda: e[:6], "m": lambda: [e[0], e[1], e[2], n90, e[3], n90],
            "o": lambda: [e[0], e[1], e[2], n90, n90, n90], "t": lambda: [e[0], e[0], e[1], n90, n90, n90],
            "h": lambda: [e[0], e[0], e[1], n90, n90, n120], "c": lambda: [e[0], e[0], e[0], n90, n90, n90]}
    return np.stack(full[system](), -1)


def free_from_cell(system, cell):
    a, b, c, al, be, ga = (float(x) for x in cell)
    return np.array({"a": [a, b, c, al, be, ga], "m": [a, b, c, be], "o": [a, b, c], "t": [(a + b) / 2, c],
                     "h": [(a + b) / 2, c], "c": [(a + b + c) / 3]}[system], float)


def perp_basis(v):
    """xpg/refine.py:20-25."""
    v = unit(v)
    t = np.array([1.0, 0, 0]) if abs(v[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = unit(np.cross(v, t))
    return e1, np.cross(v, e1)


# ---------------------------------------------------------------------------------------------------------------------
# detector and spot geometry
# ---------------------------------------------------------------------------------------------------------------------

def seg_local(seg, qx, qy):
    """(c, ex, ey) of a segment in the detector system (xpg/detector.py:49-56)."""
    ex = unit(seg.ex)
    ey = np.asarray(seg.ey, float)
    ey = unit(ey - np.dot(ey, ex) * ex)
    ez = np.cross(ex, ey)
    return -qx * seg.orgx * ex - qy * seg.orgy * ey + seg.distance * ez, ex, ey


def seg_from_local(seg, c, ex, ey, qx, qy):
    """xpg/detector.py:57-63: the segment with offset c and axes ex, ey."""
    ex = unit(ex)
    ey = unit(ey - np.dot(ey, ex) * ex)
    ez = np.cross(ex, ey)
    return replace(seg, orgx=float(-(c @ ex) / qx), orgy=float(-(c @ ey) / qy), distance=float(c @ ez), ex=ex, ey=ey)


def segment_rows(g):
    """(rects, C, EX, EY, refine flags) of the geometry's segments; one full-detector segment when it has none."""
    if g.segments:
        segs = list(g.segments)
        loc = [seg_local(s, g.qx, g.qy) for s in segs]
        rects = [(s.x1, s.x2, s.y1, s.y2) for s in segs]
        ref = [tuple(getattr(s, "refine", ()) or ()) for s in segs]
    else:
        loc = [(np.zeros(3), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]))]
        rects = [(1, max(g.nx, 1), 1, max(g.ny, 1))]
        ref = [()]
    C, EX, EY = (np.array([l[i] for l in loc]) for i in range(3))
    return rects, C, EX, EY, ref


def panel_affine(C, EX, EY, orgx, orgy, dist, ed1, ed2, qx, qy):
    """xpg/detector.py:84-110, batched: orgx / orgy / dist (...), ed1 / ed2 (..., 3), C / EX / EY (..., K, 3) or (K, 3)
    -> P0, Bx, By (..., K, 3)."""
    ed3 = np.cross(ed1, ed2)
    E = np.stack([ed1, ed2, ed3], -2)
    Bx = qx * (EX @ E)
    By = qy * (EY @ E)
    orgx, orgy, dist = (np.asarray(v, float) for v in (orgx, orgy, dist))
    origin = np.stack([-qx * orgx, -qy * orgy, dist * np.ones_like(orgx)], -1)
    A = (C + origin[..., None, :]) @ E
    P0 = A + orgx[..., None, None] * Bx + orgy[..., None, None] * By
    return P0, Bx, By


def panels(g):
    _, C, EX, EY, _ = segment_rows(g)
    return panel_affine(C, EX, EY, g.orgx, g.orgy, g.distance, unit(g.ed1), unit(g.ed2), g.qx, g.qy)


def segment_of(g, x, y):
    x, y = np.asarray(x, float), np.asarray(y, float)
    if not g.segments:
        return np.zeros(x.shape, int)
    rects, *_ = segment_rows(g)
    m = np.full((g.ny, g.nx), -1, np.int64)
    for k, (x1, x2, y1, y2) in enumerate(rects):
        m[max(y1 - 1, 0):min(y2, g.ny), max(x1 - 1, 0):min(x2, g.nx)] = k
    ix = np.clip(np.floor(x - 0.5).astype(np.int64), 0, g.nx - 1)
    iy = np.clip(np.floor(y - 0.5).astype(np.int64), 0, g.ny - 1)
    return m[iy, ix]


def bilinear(tab, x, y):
    """xpg/detector.py:133-144."""
    ny, nx = tab.shape
    fx = np.clip(x - 1.0, 0, nx - 1.000001)
    fy = np.clip(y - 1.0, 0, ny - 1.000001)
    ix, iy = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
    ix1, iy1 = np.minimum(ix + 1, nx - 1), np.minimum(iy + 1, ny - 1)
    tx, ty = fx - ix, fy - iy
    t = np.asarray(tab, float)
    return (t[iy, ix] * (1 - tx) * (1 - ty) + t[iy, ix1] * tx * (1 - ty) + t[iy1, ix] * (1 - tx) * ty
            + t[iy1, ix1] * tx * ty)


def corrected_xy(g, x, y):
    x, y = np.asarray(x, float), np.asarray(y, float)
    if g.geo_corr is None:
        return x, y
    return x + bilinear(np.asarray(g.geo_corr[0], float), x, y), y + bilinear(np.asarray(g.geo_corr[1], float), x, y)


def has_parallax(g):
    return g.sensor_thickness > 0 and g.mu > 0


def pixel_to_lab(g, x, y):
    """xpg/geometry.py:235-262: the segment's affine map of the corrected position, then four parallax iterations."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    k = np.maximum(segment_of(g, x, y), 0)
    xc, yc = corrected_xy(g, x, y)
    P0, Bx, By = (a[k] for a in panels(g))
    lab = lambda a, b: P0 + (a - g.orgx)[..., None] * Bx + (b - g.orgy)[..., None] * By
    X = lab(xc, yc)
    if not has_parallax(g):
        return X
    u1, u2 = Bx / g.qx, By / g.qy
    un = np.cross(u1, u2)
    for _ in range(4):
        e = X / np.linalg.norm(X, axis=-1, keepdims=True)
        cosa = np.clip(np.sum(e * un, -1), 1e-3, 1.0)
        L = g.sensor_thickness / cosa
        lbar = 1.0 / g.mu - L * np.exp(-g.mu * L) / np.maximum(-np.expm1(-g.mu * L), 1e-12)
        d1, d2 = lbar * np.sum(e * u1, -1), lbar * np.sum(e * u2, -1)
        X = lab(xc - d1 / g.qx, yc - d2 / g.qy)
    return X


def s0_of(g):
    return unit(g.beam) / g.wavelength


def pixel_to_s(g, x, y):
    X = pixel_to_lab(g, x, y)
    return X / (np.linalg.norm(X, axis=-1, keepdims=True) * g.wavelength)


def z_to_phi(g, z):
    return g.phi0 + g.osc * (np.asarray(z, float) - (g.starting_frame - 1))


def spot_to_p0(g, x, y, z):
    """xpg/geometry.py:313-319: the reciprocal-lattice vector at phi = 0 of an observed spot."""
    p = pixel_to_s(g, x, y) - s0_of(g)
    R = rotation_matrices(g.axis, -np.radians(z_to_phi(g, z)))
    return np.einsum("nij,nj->ni", R, p)


def predict_phi(g, p0):
    """xpg/geometry.py:482-502: both roots of S0 . R(phi) p0 = -|p0|^2 / 2 and where they exist."""
    m, s0 = np.asarray(g.axis, float), s0_of(g)
    par = (p0 @ m)[:, None] * m
    perp = p0 - par
    a = perp @ s0
    b = np.cross(m, p0) @ s0
    c = -0.5 * np.einsum("ij,ij->i", p0, p0) - par @ s0
    r = np.hypot(a, b)
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = c / r
    base, delta = np.arctan2(b, a), np.arccos(np.clip(ratio, -1, 1))
    return base + delta, base - delta, np.abs(ratio) <= 1.0


def choose_branch(g, Astar, hkl, phi_obs):
    """xpg/refine.py:298-304."""
    pa, pb, valid = predict_phi(g, hkl @ Astar.T)
    da = np.abs(np.angle(np.exp(1j * (pa - phi_obs))))
    db = np.abs(np.angle(np.exp(1j * (pb - phi_obs))))
    return np.where(da <= db, 1.0, -1.0), valid, np.minimum(da, db)


def assign_indices(Astar, p0, tol=0.3):
    """xpg/indexing.py:530-535."""
    h = p0 @ np.linalg.inv(Astar).T
    hr = np.round(h)
    ok = np.all(np.abs(h - hr) < tol, axis=1) & np.any(hr != 0, axis=1)
    return hr.astype(int), ok


# ---------------------------------------------------------------------------------------------------------------------
# parameterisation (xpg/refine.py:28-150)
# ---------------------------------------------------------------------------------------------------------------------

class Param:
    """A flat parameter vector <-> experiment quantities; angles in mrad, positions px / mm, cell A / degree."""

    def __init__(self, exp, flags):
        flags = [f.upper() for f in flags]
        g, c = exp.geom, exp.crystal
        self.exp, self.flags, self.blocks = exp, flags, []
        self.beam0, self.axis0 = unit(g.beam), unit(g.axis)
        self.be1, self.be2 = perp_basis(self.beam0)
        self.ae1, self.ae2 = perp_basis(self.axis0)
        self.ed1_0, self.ed2_0 = np.array(g.ed1, float), np.array(g.ed2, float)
        self.U0, self.system = np.array(c.U, float), c.system
        x0, names = [], []
        if "BEAM" in flags:
            self.blocks.append(("beam", len(x0), 2)); x0 += [0.0, 0.0]; names += ["beam_r1", "beam_r2"]
        if "AXIS" in flags:
            self.blocks.append(("axis", len(x0), 2)); x0 += [0.0, 0.0]; names += ["axis_r1", "axis_r2"]
        if "POSITION" in flags or "DISTANCE" in flags or "ORIGIN" in flags:
            self.blocks.append(("origin", len(x0), 3)); x0 += [g.orgx, g.orgy, g.distance]
            names += ["orgx", "orgy", "distance"]
        if "DETECTOR_ORIENTATION" in flags or ("POSITION" in flags and "SEGMENT" in flags):
            self.blocks.append(("dettilt", len(x0), 3)); x0 += [0.0, 0.0, 0.0]; names += ["det_r1", "det_r2", "det_r3"]
        if "ORIENTATION" in flags:
            self.blocks.append(("orient", len(x0), 3)); x0 += [0.0, 0.0, 0.0]; names += ["U_r1", "U_r2", "U_r3"]
        if "CELL" in flags:
            fc = free_from_cell(self.system, c.cell)
            self.blocks.append(("cell", len(x0), fc.size)); x0 += list(fc); names += CELL_PARAMS[self.system]
        self.rects, self.C, self.EX, self.EY, ref = segment_rows(g)
        K = len(self.rects)
        self.segmask = np.zeros((K, 6))
        if "SEGMENT" in flags and K > 1:
            for k in range(K):
                r = ref[k] or ("POSITION", "ORIENTATION")
                self.segmask[k, :3] = "POSITION" in r
                self.segmask[k, 3:] = "ORIENTATION" in r
            self.blocks.append(("segment", len(x0), 6 * K)); x0 += [0.0] * (6 * K)
            names += [f"seg{k + 1}_{n}" for k in range(K) for n in ("d1", "d2", "d3", "r1", "r2", "r3")]
        self.n_extra = 6 if any(b[0] == "segment" for b in self.blocks) else 0
        self.cen = np.array([(0.5 * (r[0] + r[1]), 0.5 * (r[2] + r[3])) for r in self.rects], float)
        self.x0, self.names = np.array(x0, float), names

    def unpack(self, X):
        """Quantities of parameter vectors X (M, p): each (M, ...)."""
        X = np.atleast_2d(np.asarray(X, float))
        M = X.shape[0]
        g, c = self.exp.geom, self.exp.crystal
        rep = lambda a: np.broadcast_to(a, (M,) + np.shape(a)).copy()
        beam, axis = rep(self.beam0), rep(self.axis0)
        orgx, orgy, dist = np.full(M, g.orgx), np.full(M, g.orgy), np.full(M, g.distance)
        ed1, ed2 = rep(self.ed1_0), rep(self.ed2_0)
        U, cell = rep(self.U0), rep(np.asarray(c.cell, float))
        C, EX, EY = rep(self.C), rep(self.EX), rep(self.EY)
        seg_v = None
        for name, i, n in self.blocks:
            v = X[:, i:i + n]
            if name == "beam":
                beam = (rotvec_to_matrix(1e-3 * (v[:, 0:1] * self.be1 + v[:, 1:2] * self.be2)) @ beam[..., None])[..., 0]
            elif name == "axis":
                axis = (rotvec_to_matrix(1e-3 * (v[:, 0:1] * self.ae1 + v[:, 1:2] * self.ae2)) @ axis[..., None])[..., 0]
            elif name == "origin":
                orgx, orgy, dist = v[:, 0], v[:, 1], v[:, 2]
            elif name == "dettilt":
                e3 = np.cross(self.ed1_0, self.ed2_0)
                R = rotvec_to_matrix(1e-3 * (v[:, 0:1] * self.ed1_0 + v[:, 1:2] * self.ed2_0 + v[:, 2:3] * e3))
                ed1, ed2 = (R @ ed1[..., None])[..., 0], (R @ ed2[..., None])[..., 0]
            elif name == "orient":
                U = rotvec_to_matrix(1e-3 * v) @ U
            elif name == "cell":
                cell = cell_from_free(self.system, v)
            elif name == "segment":
                # each segment: shift (mm) and rotation (mrad) about its centre, in the detector system
                seg_v = v.reshape(M, -1, 6) * self.segmask
                qx, qy = g.qx, g.qy
                cx, cy = self.cen[:, 0:1], self.cen[:, 1:2]
                mid = C + qx * cx * EX + qy * cy * EY
                R = rotvec_to_matrix(1e-3 * seg_v[..., 3:])
                EX = (R @ EX[..., None])[..., 0]
                EY = (R @ EY[..., None])[..., 0]
                C = mid + seg_v[..., :3] - qx * cx * EX - qy * cy * EY
        Astar = U @ b_matrix(cell)
        P0, Bx, By = panel_affine(C, EX, EY, orgx, orgy, dist, ed1, ed2, g.qx, g.qy)
        return dict(beam=beam, axis=axis, orgx=orgx, orgy=orgy, dist=dist, ed1=ed1, ed2=ed2, U=U, cell=cell,
                    Astar=Astar, C=C, EX=EX, EY=EY, P0=P0, Bx=Bx, By=By, seg_v=seg_v)

    def gauge(self, x):
        """The six restraint rows of segment refinement (mean shift / 1 um, mean rotation / 0.01 mrad), or none."""
        if not self.n_extra:
            return np.zeros(0)
        sv = self.unpack(x)["seg_v"][0]
        return np.concatenate([sv[:, :3].mean(0) / 1e-3, sv[:, 3:].mean(0) / 1e-2])

    def apply(self, x):
        """xpg/refine.py:97-108."""
        d = {k: (v[0] if v is not None else None) for k, v in self.unpack(x).items()}
        g = self.exp.geom
        segs = None
        if g.segments:
            segs = [seg_from_local(copy.deepcopy(s), d["C"][k], d["EX"][k], d["EY"][k], g.qx, g.qy)
                    for k, s in enumerate(g.segments)]
        g2 = replace(g, beam=d["beam"] / np.linalg.norm(d["beam"]), axis=d["axis"] / np.linalg.norm(d["axis"]),
                     orgx=float(d["orgx"]), orgy=float(d["orgy"]), distance=float(d["dist"]), ed1=d["ed1"],
                     ed2=d["ed2"], segments=segs if segs is not None else g.segments)
        return replace(self.exp, geom=g2, crystal=replace(self.exp.crystal, U=d["U"], cell=d["cell"]))


# ---------------------------------------------------------------------------------------------------------------------
# residuals, Jacobian, normal equations (xpg/refine.py:151-280)
# ---------------------------------------------------------------------------------------------------------------------

def predict_xyphi(q, wl, hkl, phi_obs, branch):
    """xpg/refine.py:253-280 for models (M, ...) and spots (n): phi (M, n), S (M, n, 3)."""
    m = q["axis"] / np.linalg.norm(q["axis"], axis=-1, keepdims=True)
    s0 = q["beam"] / np.linalg.norm(q["beam"], axis=-1, keepdims=True) / wl
    p0 = hkl @ np.swapaxes(q["Astar"], -1, -2)                   # (M, n, 3)
    m_, s0_ = m[:, None, :], s0[:, None, :]
    par = np.sum(p0 * m_, -1)[..., None] * m_
    perp = p0 - par
    mxp = np.cross(np.broadcast_to(m_, p0.shape), p0)
    a, b = np.sum(perp * s0_, -1), np.sum(mxp * s0_, -1)
    c = -0.5 * np.sum(p0 * p0, -1) - np.sum(par * s0_, -1)
    r = np.sqrt(a * a + b * b)
    ratio = np.clip(c / r, -1 + 1e-9, 1 - 1e-9)
    phi = np.arctan2(b, a) + branch * np.arccos(ratio)
    phi = phi + TWO_PI * np.round((phi_obs - phi) / TWO_PI)
    cph, sph = np.cos(phi)[..., None], np.sin(phi)[..., None]
    S = par + cph * perp + sph * mxp + s0_
    return phi, S


def residuals(par, X, data):
    """Normalised residuals (M, 3 n) of parameter vectors X (M, p) (xpg/refine.py:154-183), without restraint rows."""
    q = par.unpack(X)
    phi, S = predict_xyphi(q, data["wl"], data["H"], data["PH"], data["BR"])
    seg = data["SEG"]
    P0, Bx, By = q["P0"][:, seg], q["Bx"][:, seg], q["By"][:, seg]
    n = np.cross(Bx, By)
    tt = np.sum(P0 * n, -1) / np.sum(S * n, -1)
    P = tt[..., None] * S - P0
    xp = q["orgx"][:, None] + np.sum(P * Bx, -1) / np.sum(Bx * Bx, -1)
    yp = q["orgy"][:, None] + np.sum(P * By, -1) / np.sum(By * By, -1)
    e = S / np.linalg.norm(S, axis=-1, keepdims=True)
    u1, u2 = Bx / data["qx"], By / data["qy"]
    un = np.cross(u1, u2)
    cosa = np.clip(np.sum(e * un, -1), 1e-3, 1.0)
    L = data["tsens"] / cosa
    lbar = 1.0 / data["mu"] - L * np.exp(-data["mu"] * L) / np.maximum(-np.expm1(-data["mu"] * L), 1e-12)
    xp = xp + data["plx"] * lbar * np.sum(e * u1, -1) / data["qx"]
    yp = yp + data["plx"] * lbar * np.sum(e * u2, -1) / data["qy"]
    return np.concatenate([(xp - data["XO"]) / data["sxy"], (yp - data["YO"]) / data["sxy"],
                           (phi - data["PH"]) / data["sphi"]], -1)


def full_residuals(par, x, data):
    """Residual rows of one vector x: 3 n spot rows, then the restraint rows."""
    return np.concatenate([residuals(par, x[None], data)[0], par.gauge(x)])


def jacobian(par, x, data):
    """Forward differences, h = 1e-7 max(|x|, 1) (xpg/refine.py:185-210): (J, r)."""
    h = 1e-7 * np.maximum(np.abs(x), 1.0)
    X = x + np.diag(h)
    R = residuals(par, np.vstack([x[None], X]), data)
    r0 = np.concatenate([R[0], par.gauge(x)])
    if par.n_extra:
        Rg = np.array([par.gauge(X[k]) for k in range(x.size)])
        R = np.hstack([R[1:], Rg])
    else:
        R = R[1:]
    return ((R - r0) / h[:, None]).T, r0


def normal(par, x, data, w):
    """(J^T W J, J^T W r) (xpg/refine.py:218-222)."""
    J, r = jacobian(par, x, data)
    JW = J * w[:, None]
    return JW.T @ J, JW.T @ r


# ---------------------------------------------------------------------------------------------------------------------
# the refinement (xpg/refine.py:307-443)
# ---------------------------------------------------------------------------------------------------------------------

REFINE_SUB, PREFILTER_N = 40000, 20000


class Result:
    def __init__(self, exp, rmsd_xy, rmsd_phi, n_used, inliers, x, names, esd=None):
        self.exp, self.rmsd_xy, self.rmsd_phi, self.n_used = exp, rmsd_xy, rmsd_phi, n_used
        self.inliers, self.x, self.names, self.esd = inliers, x, names, esd


def spot_data(exp, hkl, xobs, yobs, zobs, sigma_xy, sigma_phi_deg):
    g = exp.geom
    phi_obs = np.radians(z_to_phi(g, zobs))
    branch, valid, _ = choose_branch(g, exp.crystal.U @ b_matrix(exp.crystal.cell), hkl, phi_obs)
    seg = segment_of(g, xobs, yobs)
    valid = valid & (seg >= 0)
    xc, yc = corrected_xy(g, xobs, yobs)
    data = dict(H=np.asarray(hkl, float), PH=phi_obs, BR=branch, XO=xc, YO=yc, SEG=np.maximum(seg, 0),
                wl=g.wavelength, qx=g.qx, qy=g.qy, sxy=float(sigma_xy), sphi=float(np.radians(sigma_phi_deg)),
                tsens=float(max(g.sensor_thickness, 1e-6)), mu=float(max(g.mu, 1e-6)), plx=float(has_parallax(g)))
    return data, valid


def refine(exp, hkl, xobs, yobs, zobs, flags, sigma_xy=0.5, sigma_phi_deg=None, max_iter=30, reject=True,
           n_reject_cycles=4, robust=True, fixed_inliers=None, max_err_xy=3.0, max_err_phi=2.0, prior=None):
    """xpg/refine.py:307-430."""
    g = exp.geom
    hkl = np.asarray(hkl, float)
    xo, yo, zo = (np.asarray(v, float) for v in (xobs, yobs, zobs))
    if reject and fixed_inliers is None and len(hkl) > PREFILTER_N and flags:
        pre = refine(exp, hkl, xo, yo, zo, [], sigma_xy, sigma_phi_deg, reject=True, max_err_xy=3.0 * max_err_xy,
                     max_err_phi=3.0 * max_err_phi).inliers
        filtered = 0 < pre.sum() < 0.9 * len(hkl)
        idx = np.flatnonzero(pre) if filtered else np.arange(len(hkl))
        if REFINE_SUB and idx.size > REFINE_SUB:
            k = int(np.ceil(idx.size / REFINE_SUB))
            ss = idx[np.argsort(zo[idx], kind="stable")[::k]]
            sub = refine(exp, hkl[ss], xo[ss], yo[ss], zo[ss], flags, sigma_xy, sigma_phi_deg, max_iter, reject,
                         n_reject_cycles, robust, None, max_err_xy, max_err_phi, prior)
            sub = refine(sub.exp, hkl[idx], xo[idx], yo[idx], zo[idx], flags, sigma_xy, sigma_phi_deg, 4, reject, 1,
                         robust, np.ones(idx.size, bool), max_err_xy, max_err_phi, prior)
        elif filtered:
            sub = refine(exp, hkl[idx], xo[idx], yo[idx], zo[idx], flags, sigma_xy, sigma_phi_deg, max_iter, reject,
                         n_reject_cycles, robust, None, max_err_xy, max_err_phi, prior)
        else:
            sub = None
        if sub is not None:
            full = refine(sub.exp, hkl, xo, yo, zo, [], sigma_xy, sigma_phi_deg, reject=True, n_reject_cycles=1,
                          max_err_xy=max_err_xy, max_err_phi=max_err_phi, robust=False)
            return Result(sub.exp, full.rmsd_xy, full.rmsd_phi, full.n_used, full.inliers, sub.x, sub.names, sub.esd)
    if sigma_phi_deg is None:
        sigma_phi_deg = max(abs(g.osc) * 0.5, 0.02)
    data, valid = spot_data(exp, hkl, xo, yo, zo, sigma_xy, sigma_phi_deg)
    par = Param(exp, flags)
    n, ne = hkl.shape[0], par.n_extra
    res = lambda xx: full_residuals(par, xx, data)

    def normal_j(xx, w1):
        return normal(par, xx, data, np.concatenate([np.tile(np.asarray(w1, float), 3), np.ones(ne)]))

    inl = valid.copy() if fixed_inliers is None else (np.asarray(fixed_inliers) & valid)
    x = par.x0.copy()
    if x.size == 0:
        rr = res(x)[:3 * n].reshape(3, n)
        if reject:
            rn = np.sqrt((rr ** 2).sum(0))
            med = np.median(rn[inl]) if inl.any() else 1.0
            inl = inl & (rn < max(6.0 * med, 3.0)) & (np.hypot(rr[0], rr[1]) * sigma_xy <= max_err_xy) \
                & (np.abs(rr[2]) * sigma_phi_deg <= max_err_phi)
        return result(par, x, rr, inl, sigma_xy, sigma_phi_deg, None)
    pidx = np.array([par.names.index(k) for k in (prior or {}) if k in par.names], int)
    pc = np.array([prior[par.names[i]][0] for i in pidx], float)
    ps = np.array([prior[par.names[i]][1] for i in pidx], float)
    prior_cost = lambda xx: float(np.sum(((xx[pidx] - pc) / ps) ** 2)) if pidx.size else 0.0
    cost_w = lambda r, w, xx: float(np.sum(w * r * r)) + prior_cost(xx)
    for cycle in range(n_reject_cycles + 1 if reject else 1):
        lam = 1e-3
        r = res(x)
        for it in range(max_iter):
            rr = r[:3 * n].reshape(3, n)
            rn = np.sqrt((rr ** 2).sum(0))
            w1 = inl.astype(float)
            if robust:
                w1 = w1 / (1.0 + (rn / 3.0) ** 2)
            w = np.concatenate([np.tile(w1, 3), np.ones(ne)])
            A, gr = normal_j(x, w1)
            if pidx.size:
                A[pidx, pidx] += 1.0 / ps ** 2
                gr[pidx] += (x[pidx] - pc) / ps ** 2
            D = np.diag(A).copy()
            D[D <= 0] = 1.0
            c0 = cost_w(r, w, x)
            improved = False
            for _ in range(10):
                try:
                    dx = -np.linalg.solve(A + lam * np.diag(D), gr)
                except np.linalg.LinAlgError:
                    lam *= 10
                    continue
                xn = x + dx
                rn_ = res(xn)
                c1 = cost_w(rn_, w, xn)
                if np.isfinite(c1) and c1 < c0:
                    x, r = xn, rn_
                    lam = max(lam / 3, 1e-9)
                    improved = True
                    break
                lam *= 4
            if not improved or (c0 - c1) < 1e-7 * c0:
                break
        if not reject:
            break
        rr = r[:3 * n].reshape(3, n)
        rn = np.sqrt((rr ** 2).sum(0))
        med = np.median(rn[inl]) if inl.any() else 1.0
        new_inl = valid & (rn < max(6.0 * med, 3.0)) & (np.hypot(rr[0], rr[1]) * sigma_xy <= max_err_xy) \
            & (np.abs(rr[2]) * sigma_phi_deg <= max_err_phi)
        if fixed_inliers is not None:
            new_inl &= fixed_inliers
        if np.array_equal(new_inl, inl):
            break
        inl = new_inl
    try:
        w = np.concatenate([np.tile(inl.astype(float), 3), np.ones(ne)])
        A, _ = normal_j(x, inl.astype(float))
        dof = max(int(3 * inl.sum()) - x.size, 1)
        s2 = float(np.sum((w * r * r)[:3 * n])) / dof
        esd = np.sqrt(np.clip(np.diag(np.linalg.pinv(A)) * s2, 0, None))
    except Exception:
        esd = None
    return result(par, x, r[:3 * n].reshape(3, n), inl, sigma_xy, sigma_phi_deg, esd)


def result(par, x, rr, inl, sigma_xy, sigma_phi, esd):
    exp = par.apply(x)
    if inl.any():
        rxy = float(np.sqrt(np.mean(rr[0, inl] ** 2 + rr[1, inl] ** 2))) * sigma_xy
        rphi = float(np.sqrt(np.mean(rr[2, inl] ** 2))) * sigma_phi
    else:
        rxy = rphi = np.inf
    return Result(exp, rxy, rphi, int(inl.sum()), inl, x, par.names, esd)


# ---------------------------------------------------------------------------------------------------------------------
# per-block models and smoothing (xpg/stage_integrate.py:28-229)
# ---------------------------------------------------------------------------------------------------------------------

def pos_prior(flags, g, sds=None, explicit=None, learnt=None):
    """xpg/stage_integrate.py:214-229: explicit stands for XPG_POS_PRIOR, learnt for results['pos_prior_sd']."""
    if "POSITION" not in flags:
        return None
    if sds is None:
        sds = explicit if explicit is not None else (learnt if learnt is not None else [0.5, 0.5, 0.1])
    if not all(v > 0 for v in sds):
        return None
    return dict(orgx=(g.orgx, sds[0]), orgy=(g.orgy, sds[1]), distance=(g.distance, sds[2]))


def block_edges(g, data_range, delphi):
    f0, f1 = data_range
    z0, z1 = f0 - 1.0, float(f1)
    nfr = max(int(round(delphi / abs(g.osc))), 1)
    edges = np.append(np.arange(z0, z1, nfr, dtype=float), z1)
    if len(edges) > 2 and edges[-1] - edges[-2] < 0.5 * nfr:
        edges = np.delete(edges, -2)
    return edges


def block_models(exp, spots, data_range, delphi, flags, exclude_spots=None, explicit_prior=None, state=None):
    """xpg/stage_integrate.py:28-153: (blocks, edges).  spots = (x, y, z) arrays; state receives pos_prior_sd."""
    state = {} if state is None else state
    g = exp.geom
    edges = block_edges(g, data_range, delphi)
    sx, sy, sz = (np.asarray(v, float) for v in spots)
    flags = [f for f in flags if f != "CELL"] if len(edges) > 2 else list(flags)
    auto = "POSITION" in flags and explicit_prior is None
    cap_mm = 0.1
    cap = [cap_mm / g.qx, cap_mm / g.qy, cap_mm]
    a_max = float(np.max(exp.crystal.cell[:3]))
    spacing = g.wavelength * g.distance / (a_max * min(g.qx, g.qy))
    cap = [min(cap[0], 0.05 * spacing), min(cap[1], 0.05 * spacing), cap[2]]
    prior = pos_prior(flags, g, sds=[min(5.0, 3 * cap[0]), min(5.0, 3 * cap[1]), min(1.0, 3 * cap[2])]
                      if auto else None, explicit=explicit_prior, learnt=state.get("pos_prior_sd"))
    nb = len(edges) - 1
    allowed = np.ones(len(sx), bool) if exclude_spots is None else ~exclude_spots
    in_blk = [np.flatnonzero((sz >= edges[i]) & (sz < edges[i + 1]) & allowed) for i in range(nb)]
    esd_pos = {}

    def fit_block(i, start, tols=(0.3, 0.2, 0.12)):
        idx = in_blk[i]
        if idx.size < 50:
            return start, False
        cur = start
        for tol in tols:
            p0 = spot_to_p0(cur.geom, sx[idx], sy[idx], sz[idx])
            hkl, ok = assign_indices(cur.crystal.U @ b_matrix(cur.crystal.cell), p0, tol=tol)
            if ok.sum() < 30:
                return start, False
            r = refine(cur, hkl[ok], sx[idx][ok], sy[idx][ok], sz[idx][ok], flags, max_iter=15, n_reject_cycles=2,
                       prior=prior)
            if not np.isfinite(r.rmsd_xy) or r.rmsd_xy > 3:
                return start, False
            cur = r.exp
            if r.esd is not None:
                esd_pos[i] = [r.esd[r.names.index(k)] if k in r.names else np.nan for k in ("orgx", "orgy", "distance")]
        return cur, True

    A0 = exp.crystal.U @ b_matrix(exp.crystal.cell)
    frac = []
    for i in range(nb):
        idx = in_blk[i]
        if idx.size < 50:
            frac.append(0.0)
            continue
        _, ok = assign_indices(A0, spot_to_p0(g, sx[idx], sy[idx], sz[idx]), tol=0.2)
        frac.append(ok.mean())
    seed = int(np.argmax(frac))
    blocks, n_fallback = [exp] * nb, 0
    m, okb = fit_block(seed, exp)
    blocks[seed] = m
    n_fallback += not okb
    for direction in (1, -1):
        prev = blocks[seed]
        i = seed + direction
        while 0 <= i < nb:
            m, okb = fit_block(i, prev)
            blocks[i] = m
            n_fallback += not okb
            prev = m
            i += direction
    if auto and len(esd_pos) >= 4:
        ii = sorted(esd_pos)
        est = np.array([[blocks[i].geom.orgx, blocks[i].geom.orgy, blocks[i].geom.distance] for i in ii])
        esd = np.array([esd_pos[i] for i in ii])
        spread = (1.4826 * np.median(np.abs(est - np.median(est, 0)), 0)) ** 2
        tau2 = spread - np.nanmedian(esd ** 2, 0)
        tau = np.sqrt(np.maximum(np.nan_to_num(tau2, nan=0.0), np.array([0.05, 0.05, 0.01]) ** 2))
        tau = np.minimum(tau, np.array(cap))
        state["pos_prior_sd"] = [float(v) for v in tau]
        prior = pos_prior(flags, g, explicit=explicit_prior, learnt=state["pos_prior_sd"])
        esd_pos.clear()
        first, n_fallback, blocks = blocks, 0, []
        for i in range(nb):
            m, okb = fit_block(i, first[i], tols=(0.12,))
            blocks.append(m)
            n_fallback += not okb
    state.update(seed=seed, n_fallback=n_fallback)
    return blocks, edges


def smooth_blocks(blocks, ref, lam=4.0, n_iter=4):
    """xpg/stage_integrate.py:156-212: robust Whittaker smoothing of the blocks' rotation vectors relative to the
    reference orientation (second-difference penalty, strength by GCV, floor 4) and of their beam directions."""
    nb = len(blocks)
    if nb < 5:
        return blocks
    U0 = np.asarray(ref.crystal.U, float)
    rv = np.array([Rotation.from_matrix(np.asarray(b.crystal.U) @ U0.T).as_rotvec() for b in blocks])
    bm = np.array([np.asarray(b.geom.beam, float) / np.linalg.norm(b.geom.beam) for b in blocks])
    D = np.diff(np.eye(nb), 2, axis=0)
    DtD = D.T @ D

    def gcv_lambda(w):
        best = None
        n_eff = w.sum()
        for lm in np.logspace(-1, 3, 17):
            H = np.linalg.solve(np.diag(w) + lm * DtD, np.diag(w))
            z = H @ rv
            rss = float(np.sum(w[:, None] * (rv - z) ** 2))
            dof = n_eff - np.trace(H)
            if dof <= 1:
                continue
            gg = n_eff * rss / dof ** 2
            if best is None or gg < best[0]:
                best = (gg, lm)
        return best[1] if best else lam

    w = np.ones(nb)
    for _ in range(n_iter):
        lam = max(gcv_lambda(w), 4.0)
        zr = np.linalg.solve(np.diag(w) + lam * DtD, w[:, None] * rv)
        res = np.linalg.norm(rv - zr, axis=1)
        mad = 1.4826 * np.median(res) + 1e-9
        w_new = (res < 3.0 * mad + np.radians(0.05)).astype(float)
        if np.array_equal(w_new, w) or w_new.sum() < 0.5 * nb:
            break
        w = w_new
    zb = np.linalg.solve(np.diag(w) + lam * DtD, w[:, None] * bm)
    out = []
    for i, b in enumerate(blocks):
        g = replace(b.geom, beam=zb[i] / np.linalg.norm(zb[i]))
        out.append(replace(b, geom=g, crystal=replace(b.crystal, U=Rotation.from_rotvec(zr[i]).as_matrix() @ U0)))
    return out

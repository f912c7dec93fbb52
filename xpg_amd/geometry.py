"""Host-side diffraction geometry consumed by the integration path (float64 numpy).

Follows the conventions of xpg/geometry.py:1-14 (Kabsch 1988/2010): right-handed lab frame, S0 = beam / lambda,
pixel (x, y) continuous with array element ix centred at x = ix + 1 (PIX0), lab position
X = QX (x - ORGX) ED1 + QY (y - ORGY) ED2 + F ED3, frame i covering z in [i - 1, i), phi(z) = phi0 + osc (z - (first - 1)),
A* columns = a*, b*, c* at phi = 0.  Only what the integrator reads is provided (the Geometry fields of
xpg/geometry.py:129-407 it uses): one flat panel or several segments (the SEGMENT= model of xpg/detector.py:1-19),
optional geometric-correction tables (X-/Y-GEO_CORR, pixel offsets added to raw coordinates) and sensor parallax.
``detector_consts`` hands all of it to the kernels (per-segment affine maps, segment map, correction tables).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PIX0 = 1.0


def unit(v):
    v = np.asarray(v, dtype=float)
    return v / np.linalg.norm(v)


def rotation_matrices(axis, angles):
    """Right-handed rotations about unit ``axis`` by ``angles`` (radians), shape angles.shape + (3, 3)."""
    u = unit(axis)
    a = np.asarray(angles, float)
    c, s = np.cos(a)[..., None, None], np.sin(a)[..., None, None]
    K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    return c * np.eye(3) + s * K + (1 - c) * np.outer(u, u)


def cell_to_B(cell):
    """Reciprocal basis (columns a*, b*, c*) of a cell in the Busing-Levy frame: a along x, b in the xy plane."""
    a, b, c, al, be, ga = cell
    al, be, ga = np.radians([al, be, ga])
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    real = np.array([[a, 0.0, 0.0], [b * np.cos(ga), b * np.sin(ga), 0.0],
                     [cx, cy, np.sqrt(max(c * c - cx * cx - cy * cy, 1e-12))]])
    return np.linalg.inv(real)


def random_rotation(rng):
    """Uniform random rotation from a unit quaternion."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@dataclass
class Segment:
    """One detector segment (module) of the SEGMENT= model (xpg/detector.py:36-80): raw pixels x1..x2, y1..y2
    (1-based, inclusive) map to detector-system points qx (x - orgx) ex + qy (y - orgy) ey + distance (ex x ey), with
    ex, ey the segment's axes in the detector system (ey made orthogonal to ex)."""
    x1: int
    x2: int
    y1: int
    y2: int
    orgx: float = 0.0
    orgy: float = 0.0
    distance: float = 0.0
    ex: np.ndarray = field(default_factory=lambda: np.array([1.0, 0.0, 0.0]))
    ey: np.ndarray = field(default_factory=lambda: np.array([0.0, 1.0, 0.0]))
    refine: tuple = ()          # REFINE_SEGMENT= (POSITION / ORIENTATION); empty: both

    def axes(self, qx, qy):
        """(offset c, unit ex, unit ey) in the detector system: a raw pixel (x, y) is at qx x ex + qy y ey + c."""
        ex = unit(self.ex)
        ey = np.asarray(self.ey, float)
        ey = unit(ey - (ey @ ex) * ex)
        ez = np.cross(ex, ey)
        return -qx * self.orgx * ex - qy * self.orgy * ey + self.distance * ez, ex, ey

    @property
    def rect(self):
        return (self.x1, self.x2, self.y1, self.y2)

    @property
    def centre(self):
        return 0.5 * (self.x1 + self.x2), 0.5 * (self.y1 + self.y2)

    def with_axes(self, c, ex, ey, qx, qy):
        """This segment (same pixels) placed at detector-system offset c with axes ex, ey (xpg/detector.py:57-63)."""
        ex = unit(ex)
        ey = unit(ey - np.dot(ey, ex) * ex)
        ez = np.cross(ex, ey)
        return Segment(self.x1, self.x2, self.y1, self.y2, orgx=float(-(c @ ex) / qx), orgy=float(-(c @ ey) / qy),
                       distance=float(c @ ez), ex=ex, ey=ey, refine=tuple(self.refine))


def bilinear_table(tab, x, y):
    """Bilinear interpolation of a per-pixel table whose element ix sits at x = ix + 1 (coordinates clamped to the
    table, xpg/detector.py:141-153)."""
    ny, nx = tab.shape
    fx = np.clip(np.asarray(x, float) - 1.0, 0, nx - 1.000001)
    fy = np.clip(np.asarray(y, float) - 1.0, 0, ny - 1.000001)
    ix, iy = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
    ix1, iy1 = np.minimum(ix + 1, nx - 1), np.minimum(iy + 1, ny - 1)
    tx, ty = fx - ix, fy - iy
    t = np.asarray(tab, float)
    return ((t[iy, ix] * (1 - tx)) * (1 - ty) + (t[iy, ix1] * tx) * (1 - ty)) + (t[iy1, ix] * (1 - tx)) * ty + \
        (t[iy1, ix1] * tx) * ty


@dataclass
class Geometry:
    wavelength: float
    beam: np.ndarray
    axis: np.ndarray
    orgx: float
    orgy: float
    distance: float
    ed1: np.ndarray
    ed2: np.ndarray
    qx: float
    qy: float
    nx: int
    ny: int
    phi0: float                  # degrees at the start of starting_frame
    osc: float                   # degrees per frame
    starting_frame: int = 1
    polarization_fraction: float = 0.5
    polarization_normal: np.ndarray = field(default_factory=lambda: np.array([0.0, 1.0, 0.0]))
    sensor_thickness: float = 0.0
    mu: float = 0.0
    segments: list | None = None
    geo_corr: tuple | None = None

    @property
    def ed3(self):
        return np.cross(self.ed1, self.ed2)

    @property
    def s0(self):
        return unit(self.beam) / self.wavelength

    @property
    def has_parallax(self):
        return self.sensor_thickness > 0 and self.mu > 0

    @property
    def segmented(self):
        return bool(self.segments)

    def z_to_phi(self, z):
        return self.phi0 + self.osc * (np.asarray(z) - (self.starting_frame - 1))

    def phi_to_z(self, phi):
        return (np.asarray(phi) - self.phi0) / self.osc + (self.starting_frame - 1)

    def segment_list(self):
        return list(self.segments) if self.segments else [Segment(1, max(self.nx, 1), 1, max(self.ny, 1))]

    def panels(self):
        """Affine lab maps (P0, Bx, By), each (K, 3), one per segment: X = P0 + (x - orgx) Bx + (y - orgy) By.  The
        detector system's axes are ED1, ED2 and ED3 = ED1 x ED2, its origin -qx orgx ED1 - qy orgy ED2 + F ED3 from
        the crystal (xpg/detector.py:84-95)."""
        E = np.stack([unit(self.ed1), unit(self.ed2)])
        E = np.vstack([E, np.cross(E[0], E[1])])
        origin = np.array([-self.qx * self.orgx, -self.qy * self.orgy, self.distance])
        P0, Bx, By = [], [], []
        for seg in self.segment_list():
            c, ex, ey = seg.axes(self.qx, self.qy)
            bx, by = self.qx * (ex @ E), self.qy * (ey @ E)
            P0.append((c + origin) @ E + self.orgx * bx + self.orgy * by)
            Bx.append(bx)
            By.append(by)
        return np.array(P0), np.array(Bx), np.array(By)

    def segmap(self):
        """(ny, nx) int16 segment index of every raw pixel, -1 in the gaps between segments (None: one panel)."""
        if not self.segmented:
            return None
        m = np.full((self.ny, self.nx), -1, np.int16)
        for k, seg in enumerate(self.segments):
            m[max(seg.y1 - 1, 0):min(seg.y2, self.ny), max(seg.x1 - 1, 0):min(seg.x2, self.nx)] = k
        return m

    def segment_of(self, x, y):
        """Segment of raw continuous coordinates (the pixel containing them), -1 in gaps; 0 for one panel."""
        x = np.asarray(x, float)
        if not self.segmented:
            return np.zeros(x.shape, np.int64)
        m = self.segmap()
        ix = np.clip(np.floor(x - 0.5).astype(np.int64), 0, m.shape[1] - 1)
        iy = np.clip(np.floor(np.asarray(y, float) - 0.5).astype(np.int64), 0, m.shape[0] - 1)
        return m[iy, ix].astype(np.int64)

    def corrected_xy(self, x, y):
        """Raw coordinates with the geometric-correction offsets added (unchanged without tables)."""
        x, y = np.asarray(x, float), np.asarray(y, float)
        if self.geo_corr is None:
            return x, y
        return x + bilinear_table(self.geo_corr[0], x, y), y + bilinear_table(self.geo_corr[1], x, y)

    def _parallax(self, e, u1=None, u2=None, un=None):
        """Shift (panel axes u1, u2, mm) of the absorption-weighted mean interaction depth along unit rays e (sensor of
        thickness t, attenuation mu): lbar = 1/mu - L e^{-mu L} / (1 - e^{-mu L}), L = t / cos(incidence)."""
        if u1 is None:
            u1, u2, un = unit(self.ed1), unit(self.ed2), self.ed3
        cosa = np.clip(np.sum(e * un, axis=-1), 1e-3, 1.0)
        L = self.sensor_thickness / cosa
        muL = self.mu * L
        lbar = 1.0 / self.mu - L * np.exp(-muL) / np.maximum(-np.expm1(-muL), 1e-12)
        return lbar * np.sum(e * u1, axis=-1), lbar * np.sum(e * u2, axis=-1)

    def _plane(self, x, y):
        x, y = np.asarray(x, float), np.asarray(y, float)
        return (self.qx * (x[..., None] - self.orgx) * unit(self.ed1) + self.qy * (y[..., None] - self.orgy)
                * unit(self.ed2) + self.distance * self.ed3)

    def pixel_to_lab(self, x, y):
        """Lab position (mm) where the ray recorded at raw pixel (x, y) enters the sensor: segment map, geometric
        correction, then four parallax iterations (xpg/geometry.py:246-262)."""
        x, y = np.asarray(x, float), np.asarray(y, float)
        if not self.segmented and self.geo_corr is None:
            X = self._plane(x, y)
            if self.has_parallax:
                for _ in range(4):
                    d1, d2 = self._parallax(X / np.linalg.norm(X, axis=-1, keepdims=True))
                    X = self._plane(x - d1 / self.qx, y - d2 / self.qy)
            return X
        k = np.maximum(self.segment_of(x, y), 0)
        xc, yc = self.corrected_xy(x, y)
        P0, Bx, By = (a[k] for a in self.panels())
        lab = lambda a, b: P0 + (a - self.orgx)[..., None] * Bx + (b - self.orgy)[..., None] * By
        X = lab(xc, yc)
        if self.has_parallax:
            u1, u2 = Bx / self.qx, By / self.qy
            un = np.cross(u1, u2)
            for _ in range(4):
                d1, d2 = self._parallax(X / np.linalg.norm(X, axis=-1, keepdims=True), u1, u2, un)
                X = lab(xc - d1 / self.qx, yc - d2 / self.qy)
        return X

    def spot_to_p0(self, x, y, z):
        """Reciprocal-lattice vector at phi = 0 of an observed spot (x, y, z) (xpg/geometry.py:313-319)."""
        p = self.pixel_to_s(x, y) - self.s0
        R = rotation_matrices(self.axis, -np.radians(self.z_to_phi(z)))
        return np.einsum("nij,nj->ni", R, p)

    def pixel_to_s(self, x, y):
        X = self.pixel_to_lab(x, y)
        return X / (np.linalg.norm(X, axis=-1, keepdims=True) * self.wavelength)

    def s_to_pixel(self, S):
        """Raw pixel coordinates where rays S are recorded (nan when they miss the plane, or every segment): the
        intersection with each segment's plane, the parallax shift, the inverse geometric correction (fixed-point
        iteration), the first segment whose pixels contain the point (xpg/geometry.py:268-304)."""
        S = np.asarray(S, float)
        if not self.segmented and self.geo_corr is None:
            n = self.ed3
            sn = S @ n
            with np.errstate(divide="ignore", invalid="ignore"):
                t = self.distance / sn
            P = t[..., None] * S - self.distance * n
            x = self.orgx + (P @ unit(self.ed1)) / self.qx
            y = self.orgy + (P @ unit(self.ed2)) / self.qy
            if self.has_parallax:
                d1, d2 = self._parallax(S / np.linalg.norm(S, axis=-1, keepdims=True))
                x, y = x + d1 / self.qx, y + d2 / self.qy
            ok = np.isfinite(t) & (t > 0)
            return np.where(ok, x, np.nan), np.where(ok, y, np.nan)
        P0, Bx, By = self.panels()
        x = np.full(S.shape[:-1], np.nan)
        y = np.full(S.shape[:-1], np.nan)
        for k, seg in enumerate(self.segment_list()):
            nrm = np.cross(Bx[k], By[k])
            with np.errstate(divide="ignore", invalid="ignore"):
                t = (P0[k] @ nrm) / (S @ nrm)
            P = t[..., None] * S - P0[k]
            xk = self.orgx + (P @ Bx[k]) / (Bx[k] @ Bx[k])
            yk = self.orgy + (P @ By[k]) / (By[k] @ By[k])
            if self.has_parallax:
                u1, u2 = Bx[k] / self.qx, By[k] / self.qy
                with np.errstate(invalid="ignore"):
                    d1, d2 = self._parallax(S / np.linalg.norm(S, axis=-1, keepdims=True), u1, u2, np.cross(u1, u2))
                xk, yk = xk + d1 / self.qx, yk + d2 / self.qy
            if self.geo_corr is not None:              # raw (x, y) whose corrected position is (xk, yk)
                xr, yr = xk, yk
                for _ in range(4):
                    xr = xk - bilinear_table(self.geo_corr[0], xr, yr)
                    yr = yk - bilinear_table(self.geo_corr[1], xr, yr)
                xk, yk = xr, yr
            ok = np.isfinite(t) & (t > 0) & np.isnan(x)
            if self.segmented:
                ok &= (xk >= seg.x1 - 0.5) & (xk < seg.x2 + 0.5) & (yk >= seg.y1 - 0.5) & (yk < seg.y2 + 0.5)
            x, y = np.where(ok, xk, x), np.where(ok, yk, y)
        return x, y

    def lp_factor(self, S):
        """(Lorentz zeta, polarisation P) of diffracted beams S."""
        e = S / np.linalg.norm(S, axis=-1, keepdims=True)
        e0 = unit(self.beam)
        zeta = np.cross(e, e0) @ unit(self.axis)
        n = self.polarization_normal - np.dot(self.polarization_normal, e0) * e0
        n = n / np.linalg.norm(n)
        emain = np.cross(n, e0)
        f = self.polarization_fraction
        return zeta, f * (1 - (e @ emain) ** 2) + (1 - f) * (1 - (e @ n) ** 2)

    def resolution_of_pixel(self, x, y):
        return 1.0 / np.linalg.norm(self.pixel_to_s(x, y) - self.s0, axis=-1)

    def panel_normal_at(self, x, y):
        """Unit sensor normal (lab) of the segment containing raw pixel (x, y) (gaps: segment 0)."""
        if not self.segmented:
            return np.broadcast_to(self.ed3, np.shape(x) + (3,))
        P0, Bx, By = self.panels()
        un = np.cross(Bx / self.qx, By / self.qy)
        return un[np.maximum(self.segment_of(x, y), 0)]


# Photoelectric (sensors) and total (air) mass attenuation coefficients (cm^2/g) against energy (keV), with densities
# (g/cm^3): the NIST XCOM values the reference tabulates (xpg/geometry.py:422-431), interpolated log-log.
MU_TABLES = {
    "SI": (2.33, ((5, 242.0), (6, 145.5), (8, 62.0), (10, 32.87), (15, 9.76), (20, 4.06), (30, 1.15), (40, 0.476),
                  (50, 0.240), (60, 0.137), (80, 0.0566), (100, 0.0287))),
    "CDTE": (5.85, ((8, 250.0), (10, 132.0), (15, 44.0), (20, 21.0), (25, 12.0), (26.70, 10.1), (26.72, 36.0),
                    (31.80, 23.5), (31.82, 40.5), (40, 22.5), (50, 12.5), (60, 7.6), (80, 3.5), (100, 1.95))),
    "GAAS": (5.32, ((8, 37.0), (10, 20.5), (11.0, 16.0), (11.1, 95.0), (11.8, 77.0), (11.9, 150.0), (15, 80.0),
                    (20, 37.0), (30, 12.0), (40, 5.3), (60, 1.7), (100, 0.4))),
    "AIR": (1.205e-3, ((5, 39.31), (6, 22.70), (8, 9.921), (10, 5.120), (15, 1.614), (20, 0.7779), (30, 0.3538),
                       (40, 0.2485), (50, 0.2080), (60, 0.1875), (80, 0.1662), (100, 0.1541))),
}
HC_KEV_A = 12.398419843320026           # h c in keV Angstrom


def mu_linear(wavelength, material="Si"):
    """Linear attenuation coefficient (1/mm) of a material at a wavelength (Angstrom): log-log interpolation of the
    mass attenuation table at E = hc / lambda, times the density (xpg/geometry.py:434-446); 0 for a material not
    tabulated."""
    key = str(material).upper().replace(" ", "")
    key = "SI" if key == "SILICON" else key
    if key not in MU_TABLES:
        return 0.0
    rho, tab = MU_TABLES[key]
    e = np.log([t[0] for t in tab])
    m = np.log([t[1] for t in tab])
    return float(np.exp(np.interp(np.log(HC_KEV_A / wavelength), e, m))) * rho / 10.0


@dataclass
class Crystal:
    Astar: np.ndarray
    space_group: str = "P 1"


@dataclass
class Experiment:
    geom: Geometry
    crystal: Crystal


def detector_consts(g: Geometry, bkg_template=None):
    """Detector constants for the kernels (float64 host arrays), the fields of xpg/integrate_kernels.py:48-65: one
    affine map per segment, the segment map of a segmented detector and the geometric-correction tables."""
    P0, Bx, By = g.panels()
    U1, U2 = Bx / g.qx, By / g.qy
    gx = gy = None
    if g.geo_corr is not None:
        gx, gy = (np.asarray(t, np.float32) for t in g.geo_corr)
    return dict(orgx=g.orgx, orgy=g.orgy, P0=P0, Bx=Bx, By=By, U1=U1, U2=U2, UN=np.cross(U1, U2), qx=g.qx, qy=g.qy,
                wl=g.wavelength, tsens=float(max(g.sensor_thickness, 1e-6)), mu=float(max(g.mu, 1e-6)),
                plx=float(g.has_parallax), segmap=g.segmap(), gx=gx, gy=gy,
                bkg=None if bkg_template is None else np.asarray(bkg_template, np.float32))


# ---------------------------------------------------------------------------------------------------------------------
# prediction (rotation method)
# ---------------------------------------------------------------------------------------------------------------------

def generate_hkl(Astar, dmin, dmax=1e9):
    """All Miller indices with dmin <= d < dmax (not 000)."""
    R = np.linalg.inv(Astar)
    hmax = [int(np.ceil(np.linalg.norm(R[i]) / dmin)) for i in range(3)]
    H = np.stack(np.meshgrid(*[np.arange(-m, m + 1) for m in hmax], indexing="ij"), -1).reshape(-1, 3)
    p = H @ Astar.T
    d2 = np.einsum("ij,ij->i", p, p)
    keep = (d2 <= 1.0 / dmin ** 2) & (d2 > 1.0 / dmax ** 2)
    return H[keep].astype(np.int32)


def centring_allowed(hkl, space_group):
    """Reflections allowed by the lattice centring of the space-group symbol's first letter (h.t integral for every
    centring vector t): P, A, B, C, I, F, R (hexagonal axes, obverse)."""
    c = (space_group.strip()[:1] or "P").upper()
    h, k, l = (hkl[:, i].astype(np.int64) for i in range(3))
    if c == "P":
        return np.ones(len(hkl), bool)
    if c == "A":
        return (k + l) % 2 == 0
    if c == "B":
        return (h + l) % 2 == 0
    if c == "C":
        return (h + k) % 2 == 0
    if c == "I":
        return (h + k + l) % 2 == 0
    if c == "F":
        return ((h + k) % 2 == 0) & ((h + l) % 2 == 0) & ((k + l) % 2 == 0)
    if c == "R":
        return (-h + k + l) % 3 == 0
    return np.ones(len(hkl), bool)


def predict_reflections(g: Geometry, Astar, dmin, dmax=1e9, z_range=None, hkl=None):
    """Rotation-method predictions with z in z_range: hkl, phi (deg), S, x, y, z, zeta (Lorentz), polarization.
    Both roots of S0 . R(phi) p0 = -|p0|^2 / 2 are kept."""
    if hkl is None:
        hkl = generate_hkl(Astar, dmin, dmax)
    p0 = hkl @ Astar.T
    m, s0 = unit(g.axis), g.s0
    par = (p0 @ m)[:, None] * m
    perp = p0 - par
    a = perp @ s0
    b = np.cross(m, p0) @ s0
    c = -0.5 * np.einsum("ij,ij->i", p0, p0) - par @ s0
    r = np.hypot(a, b)
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = c / r
    valid = np.abs(ratio) <= 1.0
    base = np.arctan2(b, a)
    delta = np.arccos(np.clip(ratio, -1, 1))
    if z_range is None:
        z_range = (g.starting_frame - 1, g.starting_frame - 1 + 360.0 / abs(g.osc))
    lo, hi = sorted(g.z_to_phi(np.array(z_range, float)))
    out = {k: [] for k in ("hkl", "phi", "S")}
    for root in (base + delta, base - delta):
        ph = lo + np.mod(np.degrees(root[valid]) - lo, 360.0)
        for shift in (0.0, 360.0):
            phs = ph + shift
            ok = (phs >= lo) & (phs < hi)
            if not ok.any():
                continue
            R = rotation_matrices(g.axis, np.radians(phs[ok]))
            out["hkl"].append(hkl[valid][ok])
            out["phi"].append(phs[ok])
            out["S"].append(np.einsum("nij,nj->ni", R, p0[valid][ok]) + s0)
    if not out["hkl"]:
        z = np.zeros(0)
        return dict(hkl=np.zeros((0, 3), np.int32), phi=z, S=np.zeros((0, 3)), x=z, y=z, z=z, zeta=z, polarization=z)
    res = {k: np.concatenate(v) for k, v in out.items()}
    res["x"], res["y"] = g.s_to_pixel(res["S"])
    res["z"] = g.phi_to_z(res["phi"])
    res["zeta"], res["polarization"] = g.lp_factor(res["S"])
    ok = np.isfinite(res["x"])
    return {k: v[ok] for k, v in res.items()}

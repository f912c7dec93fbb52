"""Seeded synthetic rotation datasets (the benchmark workloads), generated on the device.

The physics follows xpg/simulate.py:228-398: reflections predicted for a seeded orientation, Wilson intensities,
Lorentz-polarisation scaled counts, spots as elliptical Gaussians on the detector (radial width stretched by the
obliquity) times an erf-integrated Gaussian rocking curve, Poisson background, static dead / saturated pixels.
Reflection parameters are computed on the host; the frames (uint32 counts) are written by HIP kernels
(include/xpg_synth.h) straight into the device frame store.  Randomness is counter-based, keyed by (seed, frame,
pixel) and (seed, reflection, frame, pixel), so any frame range can be regenerated independently.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np
import torch

from . import _lib as L
from .frames import DeviceFrames
from .geometry import Crystal, Experiment, Geometry, cell_to_B, generate_hkl, predict_reflections, random_rotation

FLAG = 0xFFFFFFFF          # dead pixel (value code -1)
SAT = 0xFFFFFFFE           # saturated pixel (>= overload: value code -2)
OVERLOAD = 0xFFFFFFFE


class SynthBg(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nx", C.c_int32), ("ny", C.c_int32), ("f0", C.c_int32), ("n", C.c_int32),
                ("cx", C.c_double), ("cy", C.c_double), ("bg_centre", C.c_double), ("bg_edge", C.c_double),
                ("dead_frac", C.c_double), ("sat_frac", C.c_double), ("flag_value", C.c_uint32),
                ("sat_value", C.c_uint32)]


class SynthSpots(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nx", C.c_int32), ("ny", C.c_int32), ("f0", C.c_int32), ("nf", C.c_int32),
                ("n", C.c_int32), ("half", C.c_int32)] + \
               [(k, C.c_void_p) for k in ("x", "y", "z", "counts", "sig_r", "sig_t", "sig_z", "rdx", "rdy")] + \
               [("flag_value", C.c_uint32), ("sat_value", C.c_uint32), ("sid", C.c_void_p)]


@dataclass
class SynthSpec:
    name: str = "custom"
    nx: int = 1024
    ny: int = 1024
    nframes: int = 100
    osc: float = 0.1                  # deg
    phi0: float = 0.0
    wavelength: float = 1.0
    distance: float = 200.0
    qx: float = 0.075
    cell: tuple = (79.1, 79.1, 38.0, 90.0, 90.0, 90.0)
    space_group: str = "P 43 21 2"
    wilson_b: float = 20.0
    mean_i0: float = 2.0e4
    bg_centre: float = 2.0            # counts / pixel at the beam centre
    bg_edge: float = 2.0              # ... at the detector corner
    dead_frac: float = 0.0
    sat_frac: float = 0.0
    sigma_d: float = 0.02             # deg
    sigma_m: float = 0.1              # deg (or per-region range below)
    sigma_m_range: tuple | None = None   # (lo, hi): sigma_M drawn per detector region, uniform
    peak_range: tuple | None = None   # (lo, hi): spot peak counts log-uniform (overrides Wilson intensities)
    box_xy: int = 0
    box_z: int = 0
    box_draw: tuple | None = None     # ((side lo, hi), (depth lo, hi)): per-reflection boxes log-uniform, odd
    block_frames: int = 50
    seed: int = 1
    dmin: float | None = None

    @property
    def orgx(self):
        return self.nx / 2 + 0.5

    @property
    def orgy(self):
        return self.ny / 2 + 0.5


def spec_small():
    """100 frames of 1024 x 1024, 0.1 deg, lambda 1.0 A, 200 mm, 0.075 mm pixels, beam at the centre; P4(3)2(1)2 cell
    79.1 x 79.1 x 38.0 A; mosaicity 0.1 deg, divergence 0.02 deg; Wilson B 20 A^2; background 2 counts / pixel;
    fixed 9 x 9 x 5 shoeboxes."""
    return SynthSpec(name="small_1024x1024x100", seed=1001, box_xy=9, box_z=5)


def spec_full():
    """1800 frames of 4148 x 4362 uint32 at 0.1 deg, 150 mm; P2(1)2(1)2(1) 55 x 70 x 95 A; Wilson B 25 A^2;
    background 1-5 counts / pixel varying radially; 1% dead pixels; 11 x 11 x 7 shoeboxes."""
    return SynthSpec(name="full_4148x4362x1800", nx=4148, ny=4362, nframes=1800, distance=150.0,
                     cell=(55.0, 70.0, 95.0, 90.0, 90.0, 90.0), space_group="P 21 21 21", wilson_b=25.0,
                     bg_centre=5.0, bg_edge=1.0, dead_frac=0.01, box_xy=11, box_z=7, seed=1002)


def spec_uneven():
    """The full-size stream with sigma_M per detector region uniform in 0.05-0.8 deg (uneven shoebox sizes, wide
    reflections binned in phi, partials over frame blocks); boxes from the reference's own rules."""
    return replace(spec_full(), name="uneven_4148x4362x1800", sigma_m_range=(0.05, 0.8), box_xy=0, box_z=0,
                   box_draw=((5, 21), (3, 15)), seed=1003)


def _odd_log_uniform(u, lo, hi):
    """Odd integers in [lo, hi] (lo, hi odd) from uniforms u: log-uniform over [lo, hi], rounded to the nearest odd
    (so E[v] before rounding is (hi - lo) / ln(hi / lo))."""
    v = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u)
    return np.clip(2 * np.floor((v - 1.0) / 2.0 + 0.5) + 1, lo, hi).astype(int)


def draw_boxes(spec: SynthSpec, refl):
    """Per-reflection shoebox (depth, side) of a workload with drawn boxes: side and depth independent, log-uniform
    over spec.box_draw's ranges and odd, keyed on the Miller index (the same reflection gets the same box in every
    run, whatever else is predicted)."""
    (s_lo, s_hi), (d_lo, d_hi) = spec.box_draw
    h = np.asarray(refl["hkl"]).astype(np.int64)
    key = ((h[:, 0] + 512) * 1048576 + (h[:, 1] + 512) * 1024 + (h[:, 2] + 512)).astype(np.uint64)
    side = _odd_log_uniform(_hash_uniform(key, spec.seed + 11), s_lo, s_hi)
    depth = _odd_log_uniform(_hash_uniform(key, spec.seed + 12), d_lo, d_hi)
    return depth, side


def box_draw_for(spec: SynthSpec):
    """The per-reflection box override of spec (a callable of the prediction table), or None."""
    if spec.box_draw is None:
        return None
    return lambda refl: draw_boxes(spec, refl)


def spec_stress():
    """900 frames of 4148 x 4362, spot peaks log-uniform 1 .. 2^20 on a 0.5 counts / pixel background, 2% saturated
    pixels flagged."""
    return replace(spec_full(), name="stress_4148x4362x900", nframes=900, bg_centre=0.5, bg_edge=0.5, dead_frac=0.0,
                   sat_frac=0.02, peak_range=(1.0, 2.0 ** 20), seed=1004)


def make_geometry(spec: SynthSpec) -> Geometry:
    return Geometry(wavelength=spec.wavelength, beam=np.array([0.0, 0.0, 1.0]), axis=np.array([1.0, 0.0, 0.0]),
                    orgx=spec.orgx, orgy=spec.orgy, distance=spec.distance, ed1=np.array([1.0, 0, 0]),
                    ed2=np.array([0, 1.0, 0]), qx=spec.qx, qy=spec.qx, nx=spec.nx, ny=spec.ny, phi0=spec.phi0,
                    osc=spec.osc, starting_frame=1, polarization_fraction=0.99)


def make_experiment(spec: SynthSpec) -> Experiment:
    rng = np.random.default_rng(spec.seed)
    U = random_rotation(rng)
    return Experiment(make_geometry(spec), Crystal(U @ cell_to_B(spec.cell), spec.space_group))


def edge_resolution(g: Geometry):
    c = np.array([[0.5, 0.5], [g.nx + .5, 0.5], [0.5, g.ny + .5], [g.nx + .5, g.ny + .5]])
    return float(g.resolution_of_pixel(c[:, 0], c[:, 1]).min())


def _hash_uniform(key, salt):
    x = np.asarray(key, np.uint64) * np.uint64(0x9E3779B97F4A7C15) + np.uint64((salt * 0xBF58476D1CE4E5B9) % 2 ** 64)
    x ^= x >> np.uint64(31)
    x = x * np.uint64(0x94D049BB133111EB)
    x ^= x >> np.uint64(29)
    return (x >> np.uint64(11)).astype(np.float64) / float(2 ** 53) * (1 - 2e-12) + 1e-12


def region_sigma_m(spec: SynthSpec, x, y, n_regions_xy=3):
    """sigma_M (deg) of each reflection: constant, or drawn per detector region (3 x 3) uniformly in sigma_m_range."""
    if spec.sigma_m_range is None:
        return np.full(np.shape(x), spec.sigma_m)
    rx = np.clip((np.asarray(x) / spec.nx * n_regions_xy).astype(int), 0, n_regions_xy - 1)
    ry = np.clip((np.asarray(y) / spec.ny * n_regions_xy).astype(int), 0, n_regions_xy - 1)
    lo, hi = spec.sigma_m_range
    per = lo + (hi - lo) * np.random.default_rng(spec.seed + 77).random(n_regions_xy * n_regions_xy)
    return per[ry * n_regions_xy + rx]


def spot_table(spec: SynthSpec, exp: Experiment | None = None):
    """Host parameters of every simulated spot (xpg/simulate.py:255-300 physics)."""
    exp = exp or make_experiment(spec)
    g = exp.geom
    dmin = spec.dmin or edge_resolution(g)
    hkl = generate_hkl(exp.crystal.Astar, dmin)
    pr = predict_reflections(g, exp.crystal.Astar, dmin, z_range=(-5.0, spec.nframes + 5.0), hkl=hkl)
    e1 = np.cross(pr["S"], g.s0)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    zeta_u = np.abs(e1 @ g.axis)
    zeta = np.abs(pr["zeta"])
    keep = (zeta > 1e-3) & (zeta_u > 0.03)
    pr = {k: v[keep] for k, v in pr.items()}
    zeta, zeta_u = zeta[keep], zeta_u[keep]
    h = pr["hkl"].astype(np.int64)
    key = ((h[:, 0] + 512) * 1048576 + (h[:, 1] + 512) * 1024 + (h[:, 2] + 512)).astype(np.uint64)
    s2 = np.einsum("ij,ij->i", *(2 * [h @ exp.crystal.Astar.T]))
    if spec.peak_range is not None:
        lo, hi = spec.peak_range
        peak = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * _hash_uniform(key, spec.seed + 1))
    else:
        I = spec.mean_i0 * np.exp(-0.5 * spec.wilson_b * s2) * -np.log(_hash_uniform(key, spec.seed + 1))
    Xv = g.pixel_to_lab(pr["x"], pr["y"])
    Xlen = np.linalg.norm(Xv, axis=1)
    cosa = (Xv @ g.ed3) / Xlen
    sig_t = Xlen * np.radians(spec.sigma_d) / spec.qx
    sig_r = sig_t / cosa
    sm = region_sigma_m(spec, pr["x"], pr["y"])
    sig_z = sm / zeta_u / spec.osc
    if spec.peak_range is not None:
        # intensity from the peak pixel count: peak = I * fz_max * g_max, g_max = 1 / (2 pi s_r s_t)
        fz_max = np.minimum(1.0, 1.0 / (np.sqrt(2 * np.pi) * sig_z))
        counts = peak * 2 * np.pi * sig_r * sig_t / fz_max
    else:
        counts = I * pr["polarization"] / zeta
    rd = np.c_[pr["x"] - g.orgx, pr["y"] - g.orgy]
    rn = np.linalg.norm(rd, axis=1, keepdims=True)
    rd = np.where(rn > 1e-6, rd / np.maximum(rn, 1e-6), np.array([[1.0, 0.0]]))
    return dict(x=pr["x"], y=pr["y"], z=pr["z"], counts=counts, sig_r=sig_r, sig_t=sig_t, sig_z=sig_z, rdx=rd[:, 0],
                rdy=rd[:, 1], hkl=pr["hkl"], exp=exp, dmin=dmin)


def _lib_synth():
    return L.lib()


def spot_half(spots):
    """Half-width (pixels) of the square each spot is rendered over, from the spread of the whole table."""
    return int(np.clip(np.ceil(4.0 * np.quantile(spots["sig_r"], 0.99)), 6, 40)) if len(spots["sig_r"]) else 6


def _select(lo_sorted, order, hi, span, f0, nf):
    a = np.searchsorted(lo_sorted, f0 - span, side="left")
    b = np.searchsorted(lo_sorted, f0 + nf, side="left")
    cand = order[a:b]
    return np.sort(cand[hi[cand] >= f0])


def batch_spot_index(spots, nframes, batch):
    """Spots that reach each batch of ``batch`` frames, as one index array and per-batch offsets (host)."""
    z, sz = spots["z"], spots["sig_z"]
    lo, hi = z - 5 * sz, z + 5 * sz
    order = np.argsort(lo, kind="stable")
    span = float(np.max(hi - lo)) if z.size else 0.0
    parts = [_select(lo[order], order, hi, span, f0, min(batch, nframes - f0)) for f0 in range(0, nframes, batch)]
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([p_.size for p_ in parts])
    return (np.concatenate(parts) if parts else np.zeros(0, np.int64)).astype(np.int64), off


class DeviceSpots:
    """A dataset's spot table on the device, with the host index to pick the spots that reach a range of frames."""
    KEYS = ("x", "y", "z", "counts", "sig_r", "sig_t", "sig_z", "rdx", "rdy")

    def __init__(self, spots, device):
        self.spots = spots
        self.t = {k: torch.as_tensor(np.ascontiguousarray(spots[k]), dtype=torch.float64, device=device)
                  for k in self.KEYS}
        self.half = spot_half(spots)
        self.n = int(len(spots["z"]))
        self._index = None

    def select(self, f0, nf):
        """Table indices of the spots whose +-5 sigma frame range meets frames [f0, f0 + nf)."""
        if self._index is None:
            z, sz = self.spots["z"], self.spots["sig_z"]
            lo, hi = z - 5 * sz, z + 5 * sz
            order = np.argsort(lo, kind="stable")
            self._index = (lo[order], order, hi, float(np.max(hi - lo)) if z.size else 0.0)
        lo_sorted, order, hi, span = self._index
        return _select(lo_sorted, order, hi, span, f0, nf)

    def nbytes(self):
        return 8 * len(self.KEYS) * self.n


def render(spec: SynthSpec, spots, out: torch.Tensor, f0: int, nf: int, sync=True, sel=None):
    """Frames [f0, f0 + nf) (0-based) of the dataset into ``out`` (nf, ny, nx) int32/uint32 device tensor, on the
    current stream.  ``spots``: the spot table (dict) or its DeviceSpots.  Every spot keys its random draws by its
    table index, so a frame's values do not depend on the range a call renders.  ``sync`` = False leaves the launches
    in flight (temporaries are allocated on the current stream).  ``sel``: the spots reaching these frames as an int64
    device tensor of table indices (batch_spot_index), else selected here."""
    h = _lib_synth()
    dev = out.device
    bg = SynthBg(spec.seed, spec.nx, spec.ny, f0, nf, spec.orgx, spec.orgy, spec.bg_centre, spec.bg_edge,
                 spec.dead_frac, spec.sat_frac, FLAG, SAT)
    st = L.stream_ptr(dev)
    if h.xpg_synth_background(C.byref(bg), L.ptr(out), st) != 0:
        raise L.XpgError("synthetic background launch failed")
    ds = spots if isinstance(spots, DeviceSpots) else DeviceSpots(spots, dev)
    if sel is None:
        sel_h = ds.select(f0, nf)
        idx = torch.from_numpy(sel_h.astype(np.int64)).pin_memory().to(dev, non_blocking=True) if sel_h.size else None
    else:
        idx = sel if sel.numel() else None
    if idx is not None:
        arrs = {k: ds.t[k].index_select(0, idx) for k in DeviceSpots.KEYS}
        sid = idx.to(torch.int32)
        sp = SynthSpots(spec.seed, spec.nx, spec.ny, f0, nf, int(idx.numel()), ds.half,
                        *[L.ptr(arrs[k]) for k in DeviceSpots.KEYS], FLAG, SAT, L.ptr(sid))
        if h.xpg_synth_spots_add(C.byref(sp), L.ptr(out), st) != 0:
            raise L.XpgError("synthetic spot launch failed")
    if sync:
        torch.cuda.current_stream(dev).synchronize()


def generate(spec: SynthSpec, device, batch=50, log=None):
    """The whole dataset on the device: (DeviceFrames, spot table)."""
    spots = spot_table(spec)
    data = torch.empty((spec.nframes, spec.ny, spec.nx), dtype=torch.int32, device=device)
    ds = DeviceSpots(spots, device)
    for f0 in range(0, spec.nframes, batch):
        nf = min(batch, spec.nframes - f0)
        render(spec, ds, data[f0:f0 + nf], f0, nf, sync=False)
        if log and (f0 // batch) % 6 == 0:
            log(f"  synth {spec.name}: frames {f0 + 1}-{f0 + nf}")
    torch.cuda.synchronize(device)
    return DeviceFrames(data, first=1, overload=OVERLOAD, flag_value=FLAG), spots


def block_edges(spec: SynthSpec):
    edges = np.arange(0, spec.nframes + spec.block_frames, spec.block_frames, dtype=float)
    edges[-1] = min(edges[-1], spec.nframes)
    return np.unique(np.minimum(edges, spec.nframes))

"""3D profile-fitting integration on MI355X: the call surface of xpg/integrate.py (integrate, estimate_sigmas,
IntegrationModel, IntegrationResult, N_BOX) over the HIP kernels.

Model (Kabsch 1988/2010), as xpg/integrate.py:1-21: each shoebox voxel maps to the reflection frame
eps1 = e1.(S'-S)/|S|, eps2 = e2.(S'-S)/|S|, eps3 = zeta (phi' - phi); the expected fraction of the reflection in voxel
(i, j, k) is p = P2(eps1, eps2) J_ij R_k with P2 a learnt 2D profile per detector region and phi block, J the pixel
area and R_k the frame-integrated rocking curve.

What runs where:
  host (numpy/scipy, before the device work): prediction, box planning, neighbour lists, block order, and the small
      per-block fits of the adaptive widths (xpg/integrate.py:404-461);
  device (HIP, libxpg_integrate.so): shoebox extraction from the frame store fused into pass 1 and pass 2, the masks,
      background, sums and moments, the profile accumulation (fused into pass 1), the float64 profile normalisation and
      the Poisson profile fit.  There is no CPU fallback: without a GPU and the built library the calls raise.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch
from scipy.spatial import cKDTree

from . import integrate_kernels as K
from ._lib import XpgError
from .finalise import FINAL_F64
from .frames import DeviceFrames, from_imageset
from .geometry import PIX0, Geometry, centring_allowed, detector_consts, generate_hkl, predict_reflections
from .integrate_kernels import (FG_R, N_BOX, N_FG, PROFILE_GRID, RC_C, gaussian_profile, rc_basis_np,
                                rc_default_weights)

N_BOX_Z = 4.0
DZ_BUCKETS = (3, 5, 9, 17, 33, 65)
BIN_SLOTS = 33
BIN_MAX = 8
XY_BUCKETS = (9, 13, 17, 23, 31, 41, 55, 73, 97)
SUPERPIXEL = 2.5
SUPERPIXEL_MAX = 4


@dataclass
class IntegrationModel:
    """xpg/integrate.py:60-79.  ``box_z`` (odd, 0 = the reference's depth rule) fixes the shoebox depth in frames, for
    workloads specified with fixed box shapes; it has no counterpart in the reference."""
    sigma_d: float
    sigma_m: float
    gain: float = 1.0
    box_xy: int = 0
    box_z: int = 0

    def box(self, geom: Geometry, dmin=None):
        if self.box_xy:
            return self.box_xy
        corners = geom.pixel_to_lab(np.array([0.5, geom.nx + .5, 0.5, geom.nx + .5]),
                                    np.array([0.5, 0.5, geom.ny + .5, geom.ny + .5]))
        L = np.linalg.norm(corners, axis=1).max()
        cos2t = geom.distance / L
        spx = L * np.radians(self.sigma_d) / min(geom.qx, geom.qy) / max(cos2t, 0.4)
        spx = math.sqrt(spx ** 2 + 1.0 / 12.0)
        h = int(math.ceil(N_BOX * spx)) + 1
        return 2 * max(h, 2) + 1


@dataclass
class IntegrationResult:
    refl: dict
    model: IntegrationModel
    profiles: dict = field(default_factory=dict)
    timing: dict = field(default_factory=dict)


def _bucket(dz):
    for b in DZ_BUCKETS:
        if dz <= b:
            return b
    return None


def _neighbours(x, y, z, sx, sz, k=8):
    """k nearest predicted centres in scaled (x/sx, y/sx, z/sz) space within 4 N_FG (xpg/integrate.py:164-172)."""
    pts = np.c_[x / sx, y / sx, z / sz]
    tree = cKDTree(pts)
    d, j = tree.query(pts, k=k + 1, distance_upper_bound=4.0 * N_FG)
    j, d = j[:, 1:], d[:, 1:]
    has = np.isfinite(d)
    return np.where(has, j, 0), has, np.where(has, d, np.inf)


def predict_for_integration(exps_by_block, block_edges, dmin, dmax, z_lo, z_hi, minimum_zeta=0.05, lattice_of=None):
    """Predictions per phi block (xpg/integrate.py:101-122).  Blocks sharing one Experiment object are predicted in
    one call over their union of z and split by the block edges (the same reflections)."""
    out = []
    nb0 = len(block_edges) - 1
    groups = {}
    for b, exp in enumerate(exps_by_block):
        groups.setdefault(id(exp), (exp, []))[1].append(b)
    hkl_cache = {}
    for exp, blocks in groups.values():
        key = (id(exp.crystal.Astar), exp.crystal.space_group)
        if key not in hkl_cache:
            hkl = generate_hkl(exp.crystal.Astar, dmin, dmax)
            hkl_cache[key] = hkl[centring_allowed(hkl, exp.crystal.space_group)]
        edges = [(block_edges[b % nb0], block_edges[b % nb0 + 1]) for b in blocks]
        za, zb = min(e[0] for e in edges), max(e[1] for e in edges)
        pr = predict_reflections(exp.geom, exp.crystal.Astar, dmin, dmax, z_range=(za, zb), hkl=hkl_cache[key])
        for b, (a, c) in zip(blocks, edges):
            m = (pr["z"] >= a) & (pr["z"] < c)
            q = {k: v[m] for k, v in pr.items()}
            q["block"] = np.full(int(m.sum()), b, np.int32)
            q["lattice"] = np.full(int(m.sum()), 0 if lattice_of is None else lattice_of[b], np.int32)
            out.append((b, q))
    out.sort(key=lambda t: t[0])
    res = {k: np.concatenate([o[k] for _, o in out]) for k in out[0][1]}
    g = exps_by_block[0].geom
    e1 = np.cross(res["S"], g.s0)
    e1 /= np.maximum(np.linalg.norm(e1, axis=1, keepdims=True), 1e-12)
    res["zeta_k"] = e1 @ g.axis
    ok = (np.abs(res["zeta_k"]) >= minimum_zeta) & (res["x"] >= 0.5) & (res["x"] < g.nx + 0.5) & \
         (res["y"] >= 0.5) & (res["y"] < g.ny + 0.5) & (res["z"] >= z_lo) & (res["z"] < z_hi)
    return {k: v[ok] for k, v in res.items()}


def plan_boxes(refl, g0, model, adaptive=True):
    """Pixel-to-reflection assignment (xpg/integrate.py:230-287): shoebox depth, phi-binning, side, super-pixels and
    origin of every reflection.  Integer results from float64 floor / ceil; returns (keep, plan dict)."""
    box = (IntegrationModel(model.sigma_d * (1.3 if adaptive else 1.0), model.sigma_m).box(g0)
           if not model.box_xy else model.box_xy)
    sz_frames = model.sigma_m / np.abs(refl["zeta_k"]) / abs(g0.osc)
    if model.box_z:
        n = refl["x"].size
        hz = np.full(n, model.box_z // 2)
        fb = np.ones(n, int)
        iz0 = np.floor(refl["z"]).astype(int) - hz
        dz = 2 * hz + 1
        bucket = dz.copy()
        cov = np.ones(n, np.float32)
        wide = trunc = np.zeros(n, bool)
    else:
        from scipy.special import erf as _erf
        hz = np.ceil(N_BOX_Z * sz_frames + 0.5).astype(int)
        hmax = (max(DZ_BUCKETS) - 1) // 2
        wide = hz > hmax
        fb = np.where(wide, np.clip(np.ceil((2 * hz + 1) / BIN_SLOTS), 2, BIN_MAX), 1).astype(int)
        hz = np.where(wide, np.minimum(np.ceil(N_BOX_Z * sz_frames / fb + 0.5).astype(int), (BIN_SLOTS - 1) // 2),
                      hz).astype(int)
        trunc = wide & (np.ceil(N_BOX_Z * sz_frames / fb + 0.5) > (BIN_SLOTS - 1) // 2)
        iz_c = np.floor(refl["z"] - (fb - 1) / 2.0).astype(int)
        iz0 = iz_c - hz * fb
        dz = 2 * hz + 1
        bucket = np.array([_bucket(d) or -1 for d in dz], int)
        zc = refl["z"]
        cov = 0.5 * (_erf((iz0 + dz * fb - zc) / (np.sqrt(2) * sz_frames)) - _erf((iz0 - zc) / (np.sqrt(2) * sz_frames)))
        cov = np.where(trunc, np.clip(cov, 0, 1), 1.0).astype(np.float32)
    keep = bucket > 0
    plan = dict(box=box, iz0=iz0, dz=dz, fb=fb, bucket=bucket, coverage=cov, wide=wide, trunc=trunc)
    return keep, plan


def plan_sides(refl, g0, model, box, adaptive=True):
    """Per-reflection shoebox side, super-pixel size and origin (xpg/integrate.py:263-287)."""
    n = refl["x"].size
    if model.box_xy:
        bxs = np.full(n, box, int)
        spx = None
    else:
        Xv = g0.pixel_to_lab(refl["x"], refl["y"])
        Lr = np.linalg.norm(Xv, axis=1)
        cosr = np.abs(Xv @ g0.ed3) / Lr
        spx = Lr * np.radians(model.sigma_d * (1.3 if adaptive else 1.0)) / min(g0.qx, g0.qy) / np.maximum(cosr, 0.4)
        spx = np.sqrt(spx ** 2 + 1.0 / 12.0)
        need = 2 * np.maximum(np.ceil(N_BOX * spx).astype(int) + 1, 2) + 1
        xyb = np.array([b_ for b_ in XY_BUCKETS if b_ < box] + [box])
        bxs = xyb[np.minimum(np.searchsorted(xyb, need), len(xyb) - 1)]
    pxs = np.ones(n, int)
    if SUPERPIXEL > 0 and not model.box_xy and n:
        pxs = np.clip(np.floor(spx / SUPERPIXEL), 1, SUPERPIXEL_MAX).astype(int)
        need_s = 2 * np.maximum(np.ceil(N_BOX * spx / pxs).astype(int) + 1, 2) + 1
        bxs = np.where(pxs > 1, xyb[np.minimum(np.searchsorted(xyb, need_s), len(xyb) - 1)], bxs)
    halfs = bxs // 2
    ix0 = np.floor(refl["x"] - PIX0 + 0.5).astype(int) // pxs - halfs
    iy0 = np.floor(refl["y"] - PIX0 + 0.5).astype(int) // pxs - halfs
    return bxs, pxs, ix0, iy0


def block_order(blocks_refl, need_lo, need_hi, all_frames):
    """Order in which the reference's frame loop hands completed blocks to the block worker (xpg/integrate.py:
    729-803): blocks open in need_lo order as the loop reaches them and are submitted, in open order, after the first
    processed frame >= need_hi; blocks still open at the end follow in open order."""
    pending = sorted((b for b in range(len(blocks_refl)) if blocks_refl[b].size), key=lambda b: need_lo[b])
    opened, order = [], []
    for f in all_frames:
        fi = f - 1
        while pending and need_lo[pending[0]] <= fi:
            opened.append(pending.pop(0))
        for b in list(opened):
            if need_hi[b] <= fi:
                opened.remove(b)
                order.append(b)
    return order + opened


def _kernel_device(images):
    dev = getattr(images, "device", None)
    if dev is not None and torch.device(dev).type == "cuda":
        return torch.device(dev)
    if not torch.cuda.is_available():
        raise XpgError("the MI355X integration path needs a GPU (torch.cuda / ROCm device)")
    return torch.device("cuda", torch.cuda.current_device())


def integrate(*args, **kw):
    """3D profile-fitting integration of the predicted reflections (see _integrate)."""
    return _integrate(*args, **kw)


def estimate_sigmas(*args, **kw):
    """sigma_D / sigma_M from strong spots (see _estimate_sigmas)."""
    return _estimate_sigmas(*args, **kw)


# ---------------------------------------------------------------------------------------------------------------------
# the sweep
# ---------------------------------------------------------------------------------------------------------------------

OUT_F64 = ("bg", "isum", "vsum", "vsum0", "iprf", "vprf", "vprf0", "corr", "xo", "yo", "zo", "p_valid", "p_total",
           "frac_rec", "rho", "m_d", "m_m", "p_fg")
P1_KEEP = ("bg", "nbg", "isum", "vsum", "vsum0", "xo", "yo", "zo", "frac_rec", "m_d", "m_m", "n_over", "n_fg_masked",
           "n_fg_nb", "nfg", "zeta", "seff")


# LDS bytes per wavefront that separate the size bins of the batched form (a launch covers one bin of one block, with
# mixed shoebox shapes; its wavefronts get the bin's largest slice)
LDS_BINS = (8192, 12288, 16384, 24576, 32768, 49152, 65536, 98304, 163840)

# copies of each block's fixed-point profile grids: pass 1's workgroups spread their integer increments over them
# (fewer atomics on one address; the normalisation adds them up, exactly)
ACC_COPIES = 8
PREP_LEAD = 2           # batched form: blocks whose preparation the host submits before the chain starts


class SweepPlan:
    """Everything the device needs for one integrate() call: the reflection table in processing order, uploaded once.
    Rows are ordered by phi block, then shoebox shape in order of its LDS bytes (so size bins are contiguous), with
    the rows that have an overlapping partner first in each shape group (xpg/integrate.py:372-383).  ``groups`` are the
    (block, shape) ranges (one launch each in the plain form), ``bins`` the (block, bin) ranges (one launch each in the
    batched form, mixed shapes)."""

    def __init__(self, refl, iz0, fb, bucket, bxs, pxs, ix0, iy0, region, has_partner, pj, nb, blocks_refl, order,
                 device, lds_bins=None):
        from . import _lib as L
        lds_bins = tuple(lds_bins or LDS_BINS)
        self.lds_bins = lds_bins
        self.wg_waves = 0
        self.groups = []                 # (block, (B, bx, sp), start, n)
        self.bins = []                   # (block, bin, start, n, (max B, max bx), LDS bytes per wavefront)
        perm = []
        lds = {}

        def lds_of(B, bx):
            if (B, bx) not in lds:
                lds[(B, bx)] = int(L.lib().xpg_lds_bytes(B, bx, bx))
            return lds[(B, bx)]

        bin_of = lambda k: int(np.searchsorted(lds_bins, lds_of(k[0], k[1])))

        for b in order:
            rows = blocks_refl[b]
            ck = np.unique((bucket[rows].astype(np.int64) * 4096 + bxs[rows]) * 4096 + pxs[rows])
            # shapes in order of their LDS need (then depth, side, super-pixel): any bin ladder cuts this order into
            # contiguous launches, so the row order -- and with it the order the fits reduce rows in -- does not
            # depend on the bins
            keys = sorted(((int(k // 4096 ** 2), int(k // 4096 % 4096), int(k % 4096)) for k in ck),
                          key=lambda k: (lds_of(k[0], k[1]), k[0], k[1], k[2]))
            cur = None
            for gk in keys:
                sel = rows[(bucket[rows] == gk[0]) & (bxs[rows] == gk[1]) & (pxs[rows] == gk[2])]
                sel = sel[np.argsort(~has_partner[sel], kind="stable")]
                if not sel.size:
                    continue
                bi = bin_of(gk)
                if cur is None or cur[1] != bi:
                    if cur is not None:
                        self.bins.append(tuple(cur))
                    cur = [b, bi, len(perm), 0, (0, 0), 0]
                cur[3] += int(sel.size)
                cur[4] = (max(cur[4][0], gk[0]), max(cur[4][1], gk[1]))
                cur[5] = max(cur[5], lds_of(gk[0], gk[1]))
                self.groups.append((b, gk, len(perm), int(sel.size)))
                perm.extend(sel.tolist())
            if cur is not None:
                self.bins.append(tuple(cur))
        self.perm = np.asarray(perm, np.int64)
        p = self.perm
        T = lambda a, dt: np.ascontiguousarray(a, dtype=dt)
        self.h = dict(S=T(refl["S"][p], np.float64), phi=T(np.radians(refl["phi"][p]), np.float64),
                      ix0=T(ix0[p], np.int32), iy0=T(iy0[p], np.int32), iz0=T(iz0[p], np.int32),
                      fb=T(fb[p], np.int32), pxs=T(pxs[p], np.int32), region=T(region[p], np.int32),
                      has_p=T(has_partner[p], np.float32), pS=T(refl["S"][pj[p]], np.float64),
                      pphi=T(np.radians(refl["phi"][pj[p]]), np.float64), x=T(refl["x"][p], np.float64),
                      y=T(refl["y"][p], np.float64),
                      coverage=T(refl["coverage"][p] if "coverage" in refl else np.ones(p.size), np.float32))
        for k in ("x", "y", "z", "sx", "sz"):
            self.h["nb" + k] = T(nb[k][p], np.float32)
        # per-row shapes, and each row's voxel / pixel offset inside its bin's buffers
        dzr, bxr = bucket[p].astype(np.int64), bxs[p].astype(np.int64)
        Vr, Pr = dzr * bxr * bxr, bxr * bxr
        voff = np.zeros(p.size, np.int64)
        poff = np.zeros(p.size, np.int64)
        self.bin_sizes = []
        for b, bi, start, n, mx, wl in self.bins:
            sl = slice(start, start + n)
            voff[sl] = np.cumsum(Vr[sl]) - Vr[sl]
            poff[sl] = np.cumsum(Pr[sl]) - Pr[sl]
            self.bin_sizes.append((int(Vr[sl].sum()), int(Pr[sl].sum())))
        # each bin's voxel / pixel / row offset inside its block (bins of a block back to back), and the largest block
        self.bin_arena = []
        acc = {}
        for bi, (b, lb, start, n, mx, wl) in enumerate(self.bins):
            v0, p0, r0 = acc.get(b, (0, 0, 0))
            self.bin_arena.append((v0, p0, r0))
            nV, nP = self.bin_sizes[bi]
            acc[b] = (v0 + nV, p0 + nP, r0 + n)
        self.block_arena = tuple(max((t[k] for t in acc.values()), default=0) for k in range(3))
        self.rows_dz, self.rows_bx = dzr, bxr
        self.h.update(rdz=T(dzr, np.int32), rdx=T(bxr, np.int32), voff=T(voff, np.int64), poff=T(poff, np.int64))
        self.K = int(nb["x"].shape[1])
        self.n = int(p.size)
        self.device = None
        self.t = {}
        if device is not None:
            self.t = {k: torch.as_tensor(v, device=device) for k, v in self.h.items()}
            self.device = device

    def add_fit_inputs(self, refl, exp_blocks):
        """Per-row inputs of the device adaptive fits, in table order: phi, zeta, |X| and |cos| of the detector
        normal angle of each reflection (host arrays; uploaded by to())."""
        perm = self.perm
        Xlen = np.empty(perm.size)
        cosa = np.empty(perm.size)
        starts = {}
        for b, gk, start, m in self.groups:
            s0, n0 = starts.get(b, (start, 0))
            starts[b] = (min(s0, start), n0 + m)
        for b, (s0, n_b) in starts.items():
            g = exp_blocks[b].geom
            rows = perm[s0:s0 + n_b]
            Xv = g.pixel_to_lab(refl["x"][rows], refl["y"][rows])
            Xlen[s0:s0 + n_b] = np.linalg.norm(Xv, axis=1)
            cosa[s0:s0 + n_b] = np.abs(Xv @ g.ed3) / Xlen[s0:s0 + n_b]
        self.h.update(ad_phi=np.ascontiguousarray(refl["phi"][perm], np.float64),
                      ad_zeta_k=np.ascontiguousarray(refl["zeta_k"][perm], np.float64), ad_Xlen=Xlen, ad_cosa=cosa)
        if self.device is not None:
            for k in ("ad_phi", "ad_zeta_k", "ad_Xlen", "ad_cosa"):
                self.t[k] = torch.as_tensor(self.h[k], device=self.device)
        return self

    def to(self, device):
        """A copy of the plan with its table uploaded to ``device`` (the host arrays are shared)."""
        import copy
        c = copy.copy(self)
        c.t = {k: torch.as_tensor(v, device=device) for k, v in self.h.items()}
        c.device = device
        return c

    def refl_view(self, start, n, shape, mixed=False, wave_lds=0):
        """ctypes table of rows [start, start + n) (pointer offsets into the uploaded table); ``mixed``: with the
        per-row shapes and bin offsets (``shape`` then gives the maxima, ``wave_lds`` the LDS bytes per wavefront its
        largest row needs)."""
        import ctypes as C
        from . import _lib as L
        s = L.Refl()
        s.n, (s.Dz, s.Dy, s.Dx) = n, shape
        for k in ("S", "pS"):
            setattr(s, k, C.c_void_p(self.t[k].data_ptr() + 24 * start))
        for k in ("phi", "pphi"):
            setattr(s, k, C.c_void_p(self.t[k].data_ptr() + 8 * start))
        for k in ("ix0", "iy0", "iz0", "fb", "pxs", "region", "has_p"):
            setattr(s, k, C.c_void_p(self.t[k].data_ptr() + 4 * start))
        s.K = self.K
        s.wg_waves = int(self.wg_waves)
        for k in ("x", "y", "z", "sx", "sz"):
            setattr(s, "nb" + k, C.c_void_p(self.t["nb" + k].data_ptr() + 4 * self.K * start))
        # pass 2 of register-kernel rows: the partner rows lead each shape group, so the LDS kernel needs to cover only
        # the rows up to the last one that has a partner or a box past the register kernel's (8 frames, 128 pixels)
        need = (self.h["has_p"][start:start + n] > 0.5) | (self.rows_dz[start:start + n] > 8) | \
            (self.rows_bx[start:start + n] ** 2 > 128)
        s.lds_rows = int(np.flatnonzero(need)[-1]) + 1 if need.any() else -1
        if mixed:
            s.rdz = C.c_void_p(self.t["rdz"].data_ptr() + 4 * start)
            s.rdy = s.rdx = C.c_void_p(self.t["rdx"].data_ptr() + 4 * start)
            s.voff = C.c_void_p(self.t["voff"].data_ptr() + 8 * start)
            s.poff = C.c_void_p(self.t["poff"].data_ptr() + 8 * start)
            s.wave_lds = int(wave_lds)
        v = K.DeviceRefl.__new__(K.DeviceRefl)
        v.s, v.n, v.t = s, n, {}
        return v


def _integrate(images, exp_blocks, block_edges, model: IntegrationModel, dmin, dmax=1e9, frames=None, mask=None,
               minimum_zeta=0.05, verbose=True, log=print, n_regions_xy=3, learn_profiles=True, exclude_frames=(),
               adaptive=True, deconvolve_overlaps=True, extra_lattices=None, bkg_template=None, foreign_spots=None,
               n_foreign=4, moments_only=False, plan_only=False, timing=None, hbm_budget=None, _box_draw=None):
    """Integrate a sweep (xpg/integrate.py:209-999).  ``images``: a DeviceFrames store, or an ImageSet-like object
    (first, last, nx, ny, iter_frames) whose frames are uploaded through pinned buffers -- all of them when they and
    the batched form's prepared data fit ``hbm_budget`` (bytes; default 85% of the device's free memory), else
    streamed through a ring of frames block by block (xpg_amd/passes.py; bit-identical results).  ``plan_only``
    returns the host plan without device work (planning tests); ``timing`` (dict) receives host / device times."""
    lattice_of = [0] * len(exp_blocks)
    if extra_lattices:
        exp_blocks = list(exp_blocks)
        for li, lat in enumerate(extra_lattices, start=1):
            exp_blocks += list(lat)
            lattice_of += [li] * len(lat)
    t0 = time.time()
    g0 = exp_blocks[0].geom
    if frames is None:
        frames = (images.first, images.last)
    f_first, f_last = frames
    z_lo, z_hi = f_first - 1.0, float(f_last)
    refl = predict_for_integration(exp_blocks, block_edges, dmin, dmax, z_lo, z_hi, minimum_zeta, lattice_of)
    keep, pl = plan_boxes(refl, g0, model, adaptive)
    box = pl["box"]
    sd = np.radians(model.sigma_d)
    sm = np.radians(model.sigma_m)
    n0 = refl["x"].size
    if verbose:
        log(f"  INTEGRATE: {n0} predicted reflections, box {box}x{box} px, sigma_D {model.sigma_d:.4f} "
            f"sigma_M {model.sigma_m:.4f} deg; {int(pl['wide'].sum())} wide in phi (> {max(DZ_BUCKETS)} frames): "
            f"integrated in bins of 2-{BIN_MAX} frames, {int(pl['trunc'].sum())} of them still truncated")
    if _box_draw is not None:
        keep = np.ones(n0, bool)
        pl["coverage"] = np.ones(n0, np.float32)
    refl["coverage"] = pl["coverage"]
    refl = {k: v[keep] for k, v in refl.items()}
    iz0, dz, bucket, fb = (pl[k][keep] for k in ("iz0", "dz", "bucket", "fb"))
    n = refl["x"].size
    if _box_draw is None:
        bxs, pxs, ix0, iy0 = plan_sides(refl, g0, model, box, adaptive)
    else:
        # a synthetic workload's drawn shapes (odd depth and side per reflection; no phi-binning, no super-pixels),
        # placed as the fixed-size boxes are
        dz, bxs = (np.asarray(a, int) for a in _box_draw(refl))
        fb = np.ones(n, int)
        pxs = np.ones(n, int)
        iz0 = np.floor(refl["z"]).astype(int) - dz // 2
        bucket = dz.copy()
        ix0 = np.floor(refl["x"] - PIX0 + 0.5).astype(int) - bxs // 2
        iy0 = np.floor(refl["y"] - PIX0 + 0.5).astype(int) - bxs // 2
    # neighbours for overlap masking (xpg/integrate.py:288-302)
    Xlen = np.linalg.norm(g0.pixel_to_lab(refl["x"], refl["y"]), axis=1)
    sx_px = np.maximum(Xlen * sd / g0.qx, 0.5) * N_FG
    sz_f = np.sqrt((model.sigma_m / np.abs(refl["zeta_k"])) ** 2 + (g0.osc ** 2) / 12.0) / abs(g0.osc) * N_FG
    if n:
        nbj, nbh, nbd = _neighbours(refl["x"], refl["y"], refl["z"], sx_px.mean(), sz_f.mean())
    else:
        nbj, nbh, nbd = np.zeros((0, 8), int), np.zeros((0, 8), bool), np.zeros((0, 8))
    has_partner = (nbh[:, 0] & (nbd[:, 0] < 2.0)) if deconvolve_overlaps else np.zeros(len(nbd), bool)
    nb = dict(x=np.where(nbh, refl["x"][nbj], -1e6).astype(np.float32),
              y=np.where(nbh, refl["y"][nbj], -1e6).astype(np.float32),
              z=np.where(nbh, refl["z"][nbj], -1e6).astype(np.float32),
              sx=np.where(nbh, sx_px[nbj], 1.0).astype(np.float32),
              sz=np.where(nbh, sz_f[nbj], 1.0).astype(np.float32))
    if foreign_spots is not None and len(foreign_spots[0]) and n:
        # spots of other crystals masked at 2 sigma unless within 2.5 sigma of the own centre (xpg/integrate.py:303-327)
        fx, fy, fz_ = (np.asarray(a, float) for a in foreign_spots)
        sxm, szm = float(sx_px.mean()), float(sz_f.mean())
        ft = cKDTree(np.c_[fx / sxm, fy / sxm, fz_ / szm])
        d_f, j_f = ft.query(np.c_[refl["x"] / sxm, refl["y"] / sxm, refl["z"] / szm], k=n_foreign + 2,
                            distance_upper_bound=4.0 * N_FG)
        d_f, j_f = np.atleast_2d(d_f), np.atleast_2d(j_f)
        hf = np.isfinite(d_f) & (d_f > 2.5 / N_FG)
        order_f = np.argsort(np.where(hf, d_f, np.inf), axis=1)[:, :n_foreign]
        rows_f = np.arange(d_f.shape[0])[:, None]
        hf, j_f = hf[rows_f, order_f], np.where(np.isfinite(d_f), j_f, 0)[rows_f, order_f]
        rad = 2.0 / N_FG
        nb["x"] = np.concatenate([nb["x"], np.where(hf, fx[j_f], -1e6).astype(np.float32)], 1)
        nb["y"] = np.concatenate([nb["y"], np.where(hf, fy[j_f], -1e6).astype(np.float32)], 1)
        nb["z"] = np.concatenate([nb["z"], np.where(hf, fz_[j_f], -1e6).astype(np.float32)], 1)
        nb["sx"] = np.concatenate([nb["sx"], np.full(hf.shape, sxm * rad, np.float32)], 1)
        nb["sz"] = np.concatenate([nb["sz"], np.full(hf.shape, szm * rad, np.float32)], 1)
    rx = np.clip((refl["x"] / g0.nx * n_regions_xy).astype(int), 0, n_regions_xy - 1)
    ry = np.clip((refl["y"] / g0.ny * n_regions_xy).astype(int), 0, n_regions_xy - 1)
    region = (ry * n_regions_xy + rx).astype(np.int32)
    nreg = n_regions_xy * n_regions_xy
    nblocks = len(exp_blocks)
    blk = refl["block"]
    zstart = np.maximum(iz0, f_first - 1)
    zend = np.minimum(iz0 + dz * fb - 1, f_last - 1)
    order_b = np.argsort(blk, kind="stable")
    blocks_refl = [order_b[blk[order_b] == b] for b in range(nblocks)]
    need_lo = [int(zstart[r].min()) if r.size else 10 ** 9 for r in blocks_refl]
    need_hi = [int(zend[r].max()) if r.size else -1 for r in blocks_refl]
    excl = set(exclude_frames)
    all_frames = [f for f in range(f_first, f_last + 1) if f not in excl]
    border = block_order(blocks_refl, need_lo, need_hi, all_frames)
    pj = nbj[:, 0] if n else np.zeros(0, int)
    host_plan = dict(refl=refl, iz0=iz0, dz=dz, fb=fb, bucket=bucket, bxs=bxs, pxs=pxs, ix0=ix0, iy0=iy0,
                     region=region, has_partner=has_partner, nb=nb, block_order=border, blocks_refl=blocks_refl,
                     need_lo=need_lo, need_hi=need_hi, pj=pj)
    if plan_only:
        return host_plan
    opts = dict(frames=frames, mask=mask, exclude_frames=exclude_frames, learn_profiles=learn_profiles,
                adaptive=adaptive, bkg_template=bkg_template, moments_only=moments_only, nreg=nreg, sd=sd, sm=sm)
    mp = None
    if not isinstance(images, DeviceFrames):
        from . import passes
        device = _kernel_device(images)
        host_plan["sweep_plan"] = sweep_plan(host_plan, exp_blocks)
        budget = hbm_budget if hbm_budget is not None else int(0.85 * torch.cuda.mem_get_info(device)[0])
        mp = passes.plan_memory(host_plan, host_plan["sweep_plan"], frames, images.ny, images.nx, budget,
                                frame_offset=f_first - 1)
        host_plan["memory_plan"] = mp
    if mp is not None and mp.streamed:
        sw = Sweep(passes.stream_store(images, mp, frames, device), exp_blocks, model, host_plan, opts)
        sw.t_plan = time.time() - t0
        run = passes.run_streamed(sw, images, mp, frames)
    else:
        sw = Sweep(images, exp_blocks, model, host_plan, opts)
        sw.t_plan = time.time() - t0
        run = sw.run()
    res = sw.finish(run, t0)
    if timing is not None:
        timing.update(res.timing)
    if verbose:
        ok = np.isfinite(res.refl["iprf"])
        log(f"  INTEGRATE: {ok.sum()} reflections integrated in {res.timing['total']:.1f}s (plan "
            f"{res.timing['plan']:.1f}s, device {res.timing['gpu']:.2f}s)")
    return res


def sweep_plan(hp, exp_blocks, lds_bins=None, wg_waves=None):
    """The host part of a sweep's device table (SweepPlan with the fits' inputs), from a host plan of
    _integrate(plan_only=True); a Sweep uploads it.  lds_bins / wg_waves: launch parameters (default: the tuned
    ones, xpg_amd/tuned.json, else LDS_BINS and as many wavefronts per workgroup as fit)."""
    tb, tw = tuned_launch()
    sp = SweepPlan(hp["refl"], hp["iz0"], hp["fb"], hp["bucket"], hp["bxs"], hp["pxs"], hp["ix0"], hp["iy0"],
                   hp["region"], hp["has_partner"], hp["pj"], hp["nb"], hp["blocks_refl"], hp["block_order"],
                   None, lds_bins=lds_bins or tb)
    sp.wg_waves = tw if wg_waves is None else int(wg_waves)
    return sp.add_fit_inputs(hp["refl"], exp_blocks)


def tuned_launch():
    """(LDS bin bounds, wavefronts per workgroup) chosen by the autotuner (xpg_amd/autotune.py) for this device
    type, from xpg_amd/tuned.json; defaults when absent."""
    import json
    from pathlib import Path
    f = Path(__file__).with_name("tuned.json")
    if f.exists():
        rec = json.loads(f.read_text())
        return tuple(rec.get("lds_bins") or LDS_BINS), int(rec.get("wg_waves", 0))
    return LDS_BINS, 0


class Sweep:
    """A planned sweep on the device: the reflection table uploaded once (SweepPlan), the frame store view, and the
    per-block loop of pass 1 -> adaptive widths -> profile normalisation -> pass 2 (xpg/integrate.py:463-574), which
    ``run`` executes (repeatably: benchmarks time it with the inputs resident in HBM).

    The adaptive chain stays on the device: pass 1 of a block reads its widths record (sigma_D, sigma_M, rocking-curve
    weights) from device memory, xpg_adapt_block writes the records for the block's pass 2 and for the next block, so
    a whole sweep is a sequence of launches with no host round trip.

    Two execution forms (``opts["form"]``):
      plain    one stream; per block: pass 1 reading its voxels from the frame store, fits, normalisation, pass 2
               (also from the frame store)
      batched  the widths-independent work of every block (shoebox gather, Kabsch frame, eps and J, neighbour masks:
               xpg_prepare) runs ahead on a third stream into contiguous buffers; the serial chain pass 1 -> fits ->
               normalisation of consecutive blocks runs on the caller's stream reading them;
               each block's pass 2 depends only on its own chain and runs on a second stream, overlapping the next
               blocks' chains.
    Both give identical results (the same kernels on the same values)."""

    def __init__(self, images, exp_blocks, model, hp, opts, uploaded=None):
        """``uploaded``: the device table (``sp``) and detectors (``dets``) already uploaded by the caller."""
        from . import _lib as L
        self.images, self.exp_blocks, self.model, self.hp, self.o = images, exp_blocks, model, hp, opts
        device = _kernel_device(images)
        self.device = device
        src = images if isinstance(images, DeviceFrames) else from_imageset(images, opts["frames"], device)
        self.frames_src = src
        self.store = src.store(mask=opts["mask"], exclude_frames=opts["exclude_frames"], frames=opts["frames"])
        if uploaded is not None:
            self.sp = uploaded["sp"]
        else:
            sp = hp.get("sweep_plan")          # built ahead on the host (sweep_plan())
            if sp is None or "lds_bins" in opts or "wg_waves" in opts:
                sp = sweep_plan(hp, exp_blocks, opts.get("lds_bins"), opts.get("wg_waves"))
            self.sp = sp.to(device)
        self.by_block = {}
        for gi, (b, gk, start, m) in enumerate(self.sp.groups):
            self.by_block.setdefault(b, []).append((gi, gk, start, m))
        self.bins_by_block = {}
        self.bin_of_group = {}
        for bi, (b, lb, start, m, mx, wl) in enumerate(self.sp.bins):
            self.bins_by_block.setdefault(b, []).append((bi, start, m, (mx[0], mx[1], mx[1]), wl))
        for gi, (b, gk, start, m) in enumerate(self.sp.groups):
            self.bin_of_group[gi] = next(bi for bi, (b2, lb, s2, m2, mx, wl) in enumerate(self.sp.bins)
                                         if b2 == b and s2 <= start < s2 + m2)
        self.order = [b for b in hp["block_order"] if b in self.by_block]
        self.dets = dict(uploaded["dets"]) if uploaded is not None else {}
        self._views = {}
        self._call_cache = {}
        self.t_plan = 0.0
        t = self.sp.t
        self.ad = dict(phi=t["ad_phi"], zeta_k=t["ad_zeta_k"], Xlen=t["ad_Xlen"], cosa=t["ad_cosa"])
        nmax = max([sum(m for *_, m in self.by_block[b]) for b in self.order] or [1])
        self.ws_bytes = int(L.lib().xpg_adapt_workspace_bytes(int(nmax)))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=device)
        self.wsz = C.sizeof(L.Widths)
        nb = len(self.order)
        self.widths = torch.zeros((2 * nb + 2) * self.wsz, dtype=torch.uint8, device=device)
        self.fz = torch.zeros((self.sp.n, max(DZ_BUCKETS)), dtype=torch.float32, device=device)

    def det_for(self, g):
        key = id(g)
        if key not in self.dets:
            self.dets[key] = (g, K.DeviceDetector(detector_consts(g, self.o["bkg_template"]), self.device))
        return self.dets[key][1]

    def launches(self):
        """(kernel, block, first row, n rows, (Dz, Dy, Dx) or its maxima) of every pass launch of the current form,
        in order."""
        out = []
        batched = self.o.get("form", "batched") == "batched"
        for b in self.order:
            units = ([(start, m, shape) for bi, start, m, shape, wl in self.bins_by_block[b]] if batched else
                     [(start, m, (B, bx, bx)) for gi, (B, bx, _), start, m in self.by_block[b]])
            for kind in ("pass1",) + (() if self.o["moments_only"] else ("pass2",)):
                out += [(kind, b, start, m, shape) for start, m, shape in units]
        return out

    def _prep_for(self, bi):
        """Prepared-data buffers of bin launch bi (allocated on first use): values and neighbour masks at each row's
        voxel offset, eps1 / eps2 / J (and the partner's) at three times its pixel offset, zeta per row."""
        from . import _lib as L
        if bi not in self.prep:
            b, lb, start, m, mx, wl = self.sp.bins[bi]
            nV, nP = self.sp.bin_sizes[bi]
            hasp = bool(self.hp["has_partner"][self.sp.perm[start:start + m]].any())
            dev = self.device
            bufs = dict(sb=torch.empty(nV, dtype=torch.float32, device=dev),
                        geo=torch.empty(3 * nP, dtype=torch.float64, device=dev),
                        zeta=torch.empty(m, dtype=torch.float64, device=dev),
                        nbm=torch.empty(nV, dtype=torch.uint8, device=dev),
                        geo_p=torch.empty(3 * nP if hasp else 1, dtype=torch.float64, device=dev),
                        zeta_p=torch.empty(m if hasp else 1, dtype=torch.float64, device=dev),
                        R=torch.empty(m * mx[0], dtype=torch.float64, device=dev))
            ps = L.Prepared()
            for k, t in bufs.items():
                setattr(ps, k, L.ptr(t))
            ps.r_stride = int(mx[0])
            self.prep[bi] = (ps, bufs)
        return self.prep[bi][0]

    def _prep_group(self, gi):
        """The prepared data of shape group gi inside its bin's buffers (per-row arrays offset to the group's rows)."""
        from . import _lib as L
        bi = self.bin_of_group[gi]
        ps = self._prep_for(bi)
        d = self.sp.groups[gi][2] - self.sp.bins[bi][2]
        g = L.Prepared()
        for k in ("sb", "geo", "nbm", "geo_p"):
            setattr(g, k, getattr(ps, k))
        g.zeta = ps.zeta + 8 * d
        g.zeta_p = ps.zeta_p + 8 * d if self.prep[bi][1]["zeta_p"].numel() > 1 else ps.zeta_p
        g.R, g.r_stride = ps.R + 8 * d * ps.r_stride, ps.r_stride
        return g

    def _widths_ptr(self, slot):
        return self.widths.data_ptr() + slot * self.wsz

    def _init_widths(self):
        """Every widths slot to the run's starting widths (slot 2i: pass 1 of the i-th block, 2i + 1: its pass 2),
        from a device copy made once (device-to-device: capturable in a graph)."""
        from . import _lib as L
        if not hasattr(self, "widths0"):
            w = L.Widths()
            w.sd, w.sm = float(self.o["sd"]), float(self.o["sm"])
            w.rcw[:] = [float(v) for v in rc_default_weights()]
            raw = torch.frombuffer(bytearray(bytes(w)), dtype=torch.uint8)
            self.widths0 = raw.repeat(2 * len(self.order) + 2).to(self.device)
            self.gauss_t = torch.as_tensor(gaussian_profile(), device=self.device)
            self.gauss_rep = self.gauss_t[None].repeat(self.o["nreg"], 1, 1).contiguous()
        self.widths.copy_(self.widths0)

    def run(self, timer=None):
        """Device execution of the whole sweep; ``timer(kind, b, gi, stream)`` may return a context that brackets
        launches on ``stream``."""
        t_dev0 = time.time()
        out = self.enqueue(timer)
        torch.cuda.synchronize(self.device)
        out["t_gpu"] = time.time() - t_dev0
        return out

    def capture(self, host=None):
        """Capture the batched form of the whole sweep in a HIP graph (after one eager run that allocates and
        uploads everything the launches need); ``replay()`` then reruns it.  Returns the graph's output tensors.
        ``host``: the per-block copies of the results to these pinned tensors are captured too (enqueue)."""
        self.o["form"] = "batched"
        self.run()
        if host is not None:
            self.copy_stream()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.graph_out = self.enqueue(None, host)
        return self.graph_out

    RESULT_F64 = OUT_F64 + FINAL_F64
    RESULT_I32 = ("n_over", "n_fg_masked")
    RESULT_KEYS = RESULT_F64 + RESULT_I32

    def host_results(self):
        """Pinned host rows for the per-reflection results (copy_results' destination): a dict of per-key views of
        one (len(RESULT_F64), N) float64 and one (len(RESULT_I32), N) int32 block, the blocks as ``.blocks``."""
        N = self.sp.n
        F = torch.empty((len(self.RESULT_F64), N), dtype=torch.float64, pin_memory=True)
        I = torch.empty((len(self.RESULT_I32), N), dtype=torch.int32, pin_memory=True)
        h = _Rows({k: F[i] for i, k in enumerate(self.RESULT_F64)})
        h.update({k: I[i] for i, k in enumerate(self.RESULT_I32)})
        h.blocks = (F, I)
        return h

    def copy_results(self, run, host, rows=None, stream=None):
        """Rows ``rows`` = (first, n) (default: all) of every per-reflection result of ``run`` (table order) into
        ``host`` (host_results()): one asynchronous 2-D copy per dtype on ``stream`` (default: the current one),
        complete at its next synchronisation."""
        from . import _lib as L
        s0, n = rows if rows is not None else (0, self.sp.n)
        st = stream.cuda_stream if stream is not None else torch.cuda.current_stream(self.device).cuda_stream
        for dst, src in zip(host.blocks, run["blocks"]):
            es, N = dst.element_size(), self.sp.n
            L.check(L.lib().xpg_copy_rows(C.c_void_p(dst.data_ptr() + es * s0), es * N,
                                          C.c_void_p(src.data_ptr() + es * s0), es * N, es * n, dst.shape[0],
                                          C.c_void_p(st)))
        return host

    def replay(self):
        t_dev0 = time.time()
        self.graph.replay()
        torch.cuda.synchronize(self.device)
        return dict(self.graph_out, t_gpu=time.time() - t_dev0)

    def _setup(self, batched):
        from . import _lib as L
        sp = self.sp
        if not hasattr(self, "prep"):
            # prepared-data buffers per bin launch, allocated on first use: every bin in the batched form; in the plain
            # form the bins of the shape groups too large for the fused kernels' LDS (gathered first, on the same
            # stream)
            lim = L.lib().xpg_lds_limit()
            self.big = {gi for gi, (b, (B, bx, _), start, m) in enumerate(sp.groups)
                        if L.lib().xpg_lds_bytes(B, bx, bx) > lim}
            self.prep = {}
            self.prep_g = {}
        if batched:
            for bi in range(len(sp.bins)):
                self._prep_for(bi)
        for gi in self.big:
            if gi not in self.prep_g:
                self.prep_g[gi] = self._prep_group(gi)

    def out_keys(self):
        """Row keys of the float64 and the int32 output block (alloc_outputs), results first."""
        i32 = list(self.RESULT_I32) + [k for k in P1_KEEP if k.startswith("n_") and k not in self.RESULT_I32]
        f64 = list(self.RESULT_F64) + [k for k in P1_KEEP + K.P2_F64 if k not in i32 and k not in self.RESULT_F64]
        assert set(FINAL_F64) <= set(f64)
        return f64, i32

    def alloc_outputs(self):
        """Per-reflection output tensors of one run (NaN until written): (o1, o2, blocks), the per-key tensors being
        rows of one float64 and one int32 block (``blocks``), the results first, so that the rows of a phi block move
        to the host in one 2-D copy per dtype (copy_results)."""
        N, device = self.sp.n, self.device
        f64, i32 = self.out_keys()
        F = torch.full((len(f64), N), float("nan"), dtype=torch.float64, device=device)
        I = torch.zeros((len(i32), N), dtype=torch.int32, device=device)
        rows = {k: F[i] for i, k in enumerate(f64)}
        rows.update({k: I[i] for i, k in enumerate(i32)})
        return {k: rows[k] for k in P1_KEEP}, {k: rows[k] for k in K.P2_F64 + FINAL_F64}, (F, I)

    def enqueue(self, timer=None, host=None):
        """Launches of the whole sweep on the current stream (and the form's side streams), with no synchronisation
        and no host-to-device copy.  ``host`` (host_results()): each block's results are copied there as soon as its
        pass 2 has run, on a copy stream, while later blocks compute; the current stream waits for the last copy."""
        from . import _lib as L
        o, sp, device, model = self.o, self.sp, self.device, self.model
        main = torch.cuda.current_stream(device)
        batched = o.get("form", "batched") == "batched"
        self._setup(batched)
        if batched and not hasattr(self, "side"):
            self.side = torch.cuda.Stream(device)      # default priorities
            self.xs = torch.cuda.Stream(device)
            self.aux1 = torch.cuda.Stream(device)
        side = self.side if batched else main
        rb = RunBufs(self)
        self._init_widths()
        nullctx = _Null()
        st = L.stream_ptr(device)
        ready = {}
        # the widths-independent work of a block (shoebox gather, Kabsch frame, eps and J, neighbour masks) on the
        # preparation stream: every block ahead, or (opts "prep_ahead" = D) D blocks ahead of the chain, each after
        # the previous block's pass 1, so that its output is still in the last-level cache when its pass 1 reads it.
        # In the first form the host submits PREP_LEAD blocks' preparations, then one more after each block of the
        # chain: the device still runs them as early as it can, and the chain's first launches are not held behind
        # the host's submission of every block's preparation.
        ahead = o.get("prep_ahead")

        def prepare(b):
            with torch.cuda.stream(self.xs):
                g = self.exp_blocks[b].geom
                det = self.det_for(g)
                callp = K.make_call(g.axis, g.s0, 0.0, 1.0, o["sd"], o["sm"], model.gain, rc_default_weights())
                for bi, start, m, shape, wl in self.bins_by_block[b]:
                    with (timer("prepare", b, (start, m), self.xs) if timer else nullctx):
                        L.check(L.lib().xpg_prepare(C.byref(det.s), C.byref(callp),
                                                    C.byref(sp.refl_view(start, m, shape, True, wl).s),
                                                    C.byref(self.store), C.byref(self.prep[bi][0]),
                                                    L.stream_ptr(device)))
                ready[b] = torch.cuda.Event()
                ready[b].record(self.xs)

        p1_done = {}
        # batched: a block's pass 2 is launched behind the next block's pass 1, so that it fills the device while the
        # (small) adaptive-fit launches of that next block run, instead of sharing it with pass 1
        defer = batched and o.get("defer_pass2", True)
        held = []
        if batched:
            self.xs.wait_stream(main)
            for b in self.order[:(PREP_LEAD if ahead is None else ahead + 1)]:
                prepare(b)
        for i, b in enumerate(self.order):
            if batched and ahead is not None and i + ahead + 1 < len(self.order):
                if i >= 1:
                    self.xs.wait_event(p1_done[self.order[i - 1]])
                prepare(self.order[i + ahead + 1])
            g = self.exp_blocks[b].geom
            det = self.det_for(g)
            # launch units: (first row, rows, shape or shape maxima, mixed, prepared data or None); mixed: None for a
            # view of one shape with its own row layout, else the LDS bytes per wavefront of a view with per-row
            # shapes and bin offsets (-1: sized from the shape)
            if batched:
                main.wait_event(ready[b])
                units = [(start, m, shape, wl, self.prep[bi][0]) for bi, start, m, shape, wl in self.bins_by_block[b]]
            else:
                units = []
                for gi, (B, bx, _), start, m in self.by_block[b]:
                    big = gi in self.big
                    units.append((start, m, (B, bx, bx), -1 if big else None, self.prep_g[gi] if big else None))
                    if big:
                        call0 = K.make_call(g.axis, g.s0, 0.0, 1.0, o["sd"], o["sm"], model.gain, rc_default_weights())
                        L.check(L.lib().xpg_prepare(C.byref(det.s), C.byref(call0),
                                                    C.byref(sp.refl_view(start, m, (B, bx, bx), mixed=True).s),
                                                    C.byref(self.store), C.byref(self.prep_g[gi]), st))
            def after_pass1(b=b):
                p1_done[b] = _record(main)
                if held:                 # the previous block's pass 2, behind this block's pass 1 (see below)
                    held.pop()()

            f = self._block_chain(i, b, units, rb, main, side, timer, cross=batched, after_pass1=after_pass1,
                                  defer=defer, host=host)
            if f is not None:
                held.append(f)
            if batched and ahead is None and i + PREP_LEAD < len(self.order):
                prepare(self.order[i + PREP_LEAD])
        while held:
            held.pop()()
        if batched:
            main.wait_stream(side)
        if host is not None and not o["moments_only"]:
            main.wait_stream(self.copy_stream())
        return rb.result()

    def copy_stream(self):
        if not hasattr(self, "cs"):
            self.cs = torch.cuda.Stream(self.device)      # default priority
        return self.cs

    def _view(self, start, m, shape, mixed):
        """Cached ctypes table view: mixed None for one shape with its own row layout, else the LDS bytes per
        wavefront of a view with per-row shapes and bin offsets (-1: sized from the shape)."""
        key = (start, m, shape, mixed)
        v = self._views.get(key)
        if v is None:
            v = self._views[key] = self.sp.refl_view(start, m, shape, mixed is not None, max(mixed or 0, 0))
        return v

    def _calls(self, i, b):
        """Cached (pass-1 call, pass-2 call, adaptive-fit arguments without the outputs) of the i-th block."""
        c = self._call_cache.get(i)
        if c is None:
            from . import _lib as L
            o, model = self.o, self.model
            g = self.exp_blocks[b].geom
            frames_phi0 = float(np.radians(g.z_to_phi(0.0)))
            osc = float(np.radians(g.osc))
            call1 = K.make_call(g.axis, g.s0, frames_phi0, osc, o["sd"], o["sm"], model.gain, rc_default_weights())
            call1.widths = C.c_void_p(self._widths_ptr(2 * i))
            call2 = K.make_call(g.axis, g.s0, frames_phi0, osc, o["sd"], o["sm"], model.gain, rc_default_weights())
            call2.widths = C.c_void_p(self._widths_ptr(2 * i + 1 if o["adaptive"] else 2 * i))
            groups = self.by_block[b]
            s0, n_b = groups[0][2], sum(m for *_, m in groups)
            N = self.sp.n
            a = lambda t: C.c_void_p(t.data_ptr() + t.element_size() * s0 * (t.numel() // N))
            fit = (s0, n_b, a(self.fz), a(self.ad["phi"]), a(self.ad["zeta_k"]), a(self.sp.t["iz0"]),
                   a(self.sp.t["fb"]), a(self.ad["Xlen"]), a(self.ad["cosa"]),
                   (float(o["sd"]), float(o["sm"]), float(g.z_to_phi(0.0)), float(g.osc), float(g.qx)),
                   (C.c_void_p(self._widths_ptr(2 * i)), C.c_void_p(self._widths_ptr(2 * i + 1)),
                    C.c_void_p(self._widths_ptr(2 * i + 2)), L.ptr(self.ws), self.ws_bytes))
            c = self._call_cache[i] = (call1, call2, fit)
        return c

    def _block_chain(self, i, b, units, rb, main, side, timer=None, cross=False, after_pass1=None, defer=False,
                     host=None):
        """The i-th block (table block b): pass 1 of every launch unit, the adaptive fits and the profile
        normalisation on ``main``, pass 2 on ``side`` (after ``main`` when ``cross``); outputs into ``rb``
        (RunBufs), and with ``host`` the block's rows copied there after its pass 2.  ``defer``: pass 2 is not
        launched but returned as a callable, for the caller to launch later."""
        from . import _lib as L
        o, device = self.o, self.device
        nullctx = _Null()
        lib = L.lib()
        st = L.stream_ptr(device)
        det = self.det_for(self.exp_blocks[b].geom)
        call1, call2, fit = self._calls(i, b)
        acc = None
        if rb.learn:
            G2 = rb.acc[i]
            acc = L.ProfileAcc(G2[0].data_ptr(), G2[1].data_ptr(), o["nreg"], ACC_COPIES)
        fz_base, fz_row = self.fz.data_ptr(), 4 * self.fz.shape[1]
        # several size bins: split between main and a helper stream (largest first, each to the less loaded one), so
        # that the few wavefronts per CU of large boxes and the many of small boxes share the CUs
        aux = self.aux1 if (cross and len(units) > 1 and getattr(self, "aux1", None) is not None) else None
        where = [0] * len(units)
        if aux is not None:
            load = [0, 0]
            for u in sorted(range(len(units)), key=lambda u: -units[u][1] * int(np.prod(units[u][2]))):
                where[u] = int(load[1] < load[0])
                load[where[u]] += units[u][1] * int(np.prod(units[u][2]))
            aux.wait_stream(main)
        for u, (start, m, shape, mixed, pp) in enumerate(units):
            view = self._view(start, m, shape, mixed)
            out = L.Pass1Out()
            for k in K.P1_F64 + K.P1_I32:
                setattr(out, k, rb.p1(k, start))
            out.fz = fz_base + fz_row * start
            out.fz_stride = self.fz.shape[1]
            s1 = aux if where[u] else main
            sp1 = s1.cuda_stream if where[u] else st
            if pp is not None:          # the rows' rocking curves for pass 1's widths, ahead of it
                with (timer("rocking", b, (start, m), s1) if timer else nullctx):
                    L.check(lib.xpg_rocking(C.byref(call1), C.byref(view.s), C.byref(pp), sp1))
            with (timer("pass1", b, (start, m), s1) if timer else nullctx):
                a_ = C.byref(acc) if acc is not None else None
                if pp is not None:
                    L.check(lib.xpg_pass1_prepared(C.byref(det.s), C.byref(call1), C.byref(view.s), C.byref(pp),
                                                   C.byref(out), a_, sp1))
                else:
                    L.check(lib.xpg_pass1_frames(C.byref(det.s), C.byref(call1), C.byref(view.s),
                                                 C.byref(self.store), C.byref(out), a_, sp1))
        if aux is not None:
            main.wait_stream(aux)
        if after_pass1 is not None:
            after_pass1()
        if o["adaptive"]:
            s0, n_b, fz, phi, zeta_k, iz0, fb, Xlen, cosa, scal, tail = fit
            L.check(lib.xpg_adapt_block(
                n_b, rb.p1("isum", s0), rb.p1("vsum", s0), rb.p1("n_over", s0), rb.p1("frac_rec", s0),
                rb.p1("m_d", s0), fz, phi, zeta_k, iz0, fb, Xlen, cosa, *scal, *tail, st))
        if rb.learn:
            prof = rb.prof[i]
            L.check(lib.xpg_normalise_profiles(C.byref(acc), L.ptr(self.gauss_t), L.ptr(prof), L.ptr(rb.f64[0]),
                                               L.ptr(rb.f64[1]), st))
        else:
            prof = self.gauss_rep
        rb.prof_by_block[b] = prof
        if o["moments_only"]:
            return None

        def pass2():
            if cross:
                side.wait_stream(main)
            sst = L.stream_ptr(device) if side is main else side.cuda_stream
            for start, m, shape, mixed, pp in units:
                view = self._view(start, m, shape, mixed)
                out = L.Pass2Out()
                for k in K.P2_F64:
                    setattr(out, k, rb.p2(k, start))
                bg, nbg = rb.p1("bg", start), rb.p1("nbg", start)
                if pp is not None:      # the rows' rocking curves for pass 2's widths (pass 1 has read its own)
                    with (timer("rocking", b, (start, m), side) if timer else nullctx):
                        L.check(lib.xpg_rocking(C.byref(call2), C.byref(view.s), C.byref(pp), sst))
                with (timer("pass2", b, (start, m), side) if timer else nullctx):
                    if pp is not None:
                        L.check(lib.xpg_pass2_prepared(C.byref(det.s), C.byref(call2), C.byref(view.s), C.byref(pp),
                                                       L.ptr(prof), int(prof.shape[0]), bg, nbg, 0, C.byref(out),
                                                       sst))
                    else:
                        L.check(lib.xpg_pass2_frames(C.byref(det.s), C.byref(call2), C.byref(view.s),
                                                     C.byref(self.store), L.ptr(prof), int(prof.shape[0]), bg, nbg,
                                                     0, C.byref(out), sst))
            self._finalise_block(b, rb, fit[0], fit[1], sst)
            if host is not None:        # the block's rows (contiguous in the table) to the host, behind its pass 2
                cs = self.copy_stream()
                cs.wait_stream(side)
                self.copy_results(rb.result(), host, (fit[0], fit[1]), cs)

        if defer:
            return pass2
        pass2()
        return None

    def _finalise_block(self, b, rb, s0, n, stream):
        """The block's rows [s0, s0 + n) (table order) corrected on the device behind its pass 2: Lorentz-polarisation,
        absorption and the columns of xpg/stage_integrate.py:408-421 (xpg_finalise) into the FINAL_F64 rows."""
        from . import _lib as L
        from .finalise import lp_geom
        g = self.exp_blocks[b].geom
        key = id(g)
        if not hasattr(self, "_lp"):
            self._lp = {}
        if key not in self._lp:
            self._lp[key] = (g, lp_geom(g))
        geom = self._lp[key][1]
        t = self.sp.t
        fin = L.FinalIn()
        fin.S = C.c_void_p(t["S"].data_ptr() + 24 * s0)
        fin.x, fin.y = C.c_void_p(t["x"].data_ptr() + 8 * s0), C.c_void_p(t["y"].data_ptr() + 8 * s0)
        fin.coverage = C.c_void_p(t["coverage"].data_ptr() + 4 * s0)
        for k in ("iprf", "vprf", "p_valid", "p_total", "corr"):
            setattr(fin, k, C.c_void_p(rb.p2(k, s0)))
        for k in ("isum", "vsum"):
            setattr(fin, k, C.c_void_p(rb.p1(k, s0)))
        fo = L.FinalOut()
        for k in FINAL_F64:
            setattr(fo, k, C.c_void_p(rb.p2(k, s0)))
        det = self.det_for(g) if g.segmented else None
        L.check(L.lib().xpg_finalise(C.byref(geom), C.byref(det.s) if det is not None else None, int(n),
                                     C.byref(fin), C.byref(fo), stream))

    def arena(self):
        """Prepared-data buffers for one block at a time (the largest block's bins back to back): the single-stream
        block-by-block form (enqueue_block)."""
        from . import _lib as L
        nV, nP, nR = self.sp.block_arena
        dev = self.device
        bufs = dict(sb=torch.empty(max(nV, 1), dtype=torch.float32, device=dev),
                    geo=torch.empty(max(3 * nP, 1), dtype=torch.float64, device=dev),
                    zeta=torch.empty(max(nR, 1), dtype=torch.float64, device=dev),
                    nbm=torch.empty(max(nV, 1), dtype=torch.uint8, device=dev),
                    geo_p=torch.empty(max(3 * nP, 1), dtype=torch.float64, device=dev),
                    zeta_p=torch.empty(max(nR, 1), dtype=torch.float64, device=dev),
                    R=torch.empty(max(nR, 1) * self.rocking_stride(), dtype=torch.float64, device=dev))
        return bufs

    def rocking_stride(self):
        """Doubles per row of the arena's rocking-curve table: the deepest box of the sweep."""
        return int(max((mx[0] for b, lb, start, m, mx, wl in self.sp.bins), default=1))

    def arena_bytes(self):
        nV, nP, nR = self.sp.block_arena
        return 5 * nV + 48 * nP + 16 * nR + 8 * nR * self.rocking_stride()

    def enqueue_block(self, i, store, arena, rb):
        """The i-th block of the order on the current stream only, outputs into ``rb`` (RunBufs), reading frames
        through ``store`` (any xpg_frames view holding the block's frames, e.g. a ring) and its prepared data from
        ``arena`` (arena()): every bin's
        shoebox gather and geometry, pass 1, the adaptive fits, normalisation and pass 2, in stream order.  Gives the
        results of enqueue() (the same kernels on the same values)."""
        from . import _lib as L
        o, sp, model = self.o, self.sp, self.model
        b = self.order[i]
        g = self.exp_blocks[b].geom
        det = self.det_for(g)
        st = L.stream_ptr(self.device)
        callp = K.make_call(g.axis, g.s0, 0.0, 1.0, o["sd"], o["sm"], model.gain, rc_default_weights())
        units = []
        for bi, start, m, shape, wl in self.bins_by_block[b]:
            v0, p0, r0 = sp.bin_arena[bi]
            pp = L.Prepared()
            pp.sb = arena["sb"].data_ptr() + 4 * v0
            pp.nbm = arena["nbm"].data_ptr() + v0
            pp.geo = arena["geo"].data_ptr() + 24 * p0
            pp.geo_p = arena["geo_p"].data_ptr() + 24 * p0
            pp.zeta = arena["zeta"].data_ptr() + 8 * r0
            pp.zeta_p = arena["zeta_p"].data_ptr() + 8 * r0
            pp.r_stride = self.rocking_stride()
            pp.R = arena["R"].data_ptr() + 8 * r0 * pp.r_stride
            L.check(L.lib().xpg_prepare(C.byref(det.s), C.byref(callp), C.byref(sp.refl_view(start, m, shape, True, wl).s),
                                        C.byref(store), C.byref(pp), st))
            units.append((start, m, shape, wl, pp))
        main = torch.cuda.current_stream(self.device)
        self._block_chain(i, b, units, rb, main, main)

    def block_params(self):
        """Per-block widths (xpg/integrate.py:427-429, :458-460) from the device records of the last run."""
        from . import _lib as L
        raw = self.widths.cpu().numpy().tobytes()
        out = {}
        for i, b in enumerate(self.order):
            w1 = L.Widths.from_buffer_copy(raw[2 * i * self.wsz:(2 * i + 1) * self.wsz])
            w2 = L.Widths.from_buffer_copy(raw[(2 * i + 1) * self.wsz:(2 * i + 2) * self.wsz])
            rec = dict(sigma_d=float(np.degrees(w1.sd)), sigma_m=float(np.degrees(w2.sm)), n=int(w2.n_strong),
                       rcw=np.round(np.array(w2.rcw[:], np.float64), 3).tolist())
            if w2.sm_fit > 0:
                rec.update(sigma_d_fit=float(np.degrees(w2.sd_fit)), sigma_m_fit=float(np.degrees(w2.sm_fit)))
            out[b] = rec
        return out

    def finish(self, run, t0=None):
        """Host results in prediction order (xpg/integrate.py:967-999)."""
        sp, hp = self.sp, self.hp
        N = sp.n
        inv = np.empty(N, np.int64)
        inv[sp.perm] = np.arange(N)
        F, I = (b.cpu().numpy() for b in run["blocks"])      # two copies for every per-reflection output
        f64, i32 = self.out_keys()
        host = {k: F[i][inv] for i, k in enumerate(f64)}
        host.update({k: I[i][inv] for i, k in enumerate(i32)})
        refl = dict(hp["refl"])
        for k in OUT_F64 + FINAL_F64:
            refl[k] = host[k].astype(np.float64)
        refl["n_over"] = host["n_over"].astype(np.int32)
        refl["n_fg_masked"] = host["n_fg_masked"].astype(np.int32)
        refl["overlap"] = hp["has_partner"].copy()
        if not self.o["learn_profiles"]:
            refl["fz"] = self.fz.cpu().numpy()[inv]
        refl["iz0"], refl["dz"], refl["fb"] = hp["iz0"], hp["dz"], hp["fb"]
        refl["region"] = hp["region"]
        refl["ix0"], refl["iy0"] = hp["ix0"] * hp["pxs"], hp["iy0"] * hp["pxs"]
        if (hp["pxs"] > 1).any():
            refl["pxs"] = hp["pxs"]
        profiles = {b: p.cpu().numpy() for b, p in run["prof_by_block"].items()}
        tm = dict(total=(time.time() - t0) if t0 else float("nan"), gpu=run["t_gpu"], plan=self.t_plan)
        res = IntegrationResult(refl, self.model, profiles, tm)
        res.block_params = self.block_params() if self.o["adaptive"] else {}
        return res


class RunBufs:
    """Device buffers of one run of a sweep: per-reflection outputs (table order), the fixed-point profile grids and
    the normalised profiles of every block, with the raw pointers the launches take (no per-launch tensor views)."""

    def __init__(self, sweep):
        self.o1, self.o2, self.blocks = sweep.alloc_outputs()
        dev, G = sweep.device, K.PROFILE_GRID
        nb, nreg = max(len(sweep.order), 1), sweep.o["nreg"]
        self.learn = bool(sweep.o["learn_profiles"])
        if self.learn:
            self.acc = torch.zeros((nb, 2, ACC_COPIES * nreg, G, G), dtype=torch.int64, device=dev)
            self.prof = torch.empty((nb, nreg, G, G), dtype=torch.float32, device=dev)
            self.f64 = torch.empty((2, nreg, G, G), dtype=torch.float64, device=dev)
        self.ptr1 = {k: (v.data_ptr(), v.element_size()) for k, v in self.o1.items()}
        self.ptr2 = {k: (v.data_ptr(), v.element_size()) for k, v in self.o2.items()}
        self.prof_by_block = {}

    def p1(self, k, start):
        base, es = self.ptr1[k]
        return base + es * start

    def p2(self, k, start):
        base, es = self.ptr2[k]
        return base + es * start

    def result(self):
        return dict(o1=self.o1, o2=self.o2, blocks=self.blocks, prof_by_block=self.prof_by_block)


def _record(stream):
    e = torch.cuda.Event()
    e.record(stream)
    return e


class _Rows(dict):
    """Per-key views of result blocks (host_results); ``blocks`` holds the blocks."""
    blocks = ()


class _Null:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


def _estimate_sigmas(images, exp, model0: IntegrationModel, dmin, frames, n_sample_frames=30, verbose=True, log=print,
                     extra_exps=(), bkg_template=None, foreign_spots=None, max_iter=None, guard=None):
    """sigma_D / sigma_M from strong spots of five sampled phi ranges, iterated (xpg/integrate.py:1002-1092): each
    range runs as a moments-only device sweep; the selection and fits are in xpg_amd/widths.py."""
    from . import widths as W
    g = exp.geom
    f0, f1 = frames

    def run_range(model, a, b):
        res = integrate(images, [exp], np.array([a - 1.0, float(b)]), model, dmin,
                        frames=(max(f0, a - 5), min(f1, b + 5)), verbose=False, learn_profiles=False, adaptive=False,
                        bkg_template=bkg_template, extra_lattices=[[e] for e in extra_exps] or None,
                        foreign_spots=foreign_spots, moments_only=True)
        return res.refl

    return W.estimate(run_range, g, model0, lambda sd, sm: IntegrationModel(sigma_d=sd, sigma_m=sm, gain=model0.gain),
                      W.sample_ranges(f0, f1, g.osc, n_sample_frames), max_iter=max_iter, guard=guard,
                      log=log if verbose else None)

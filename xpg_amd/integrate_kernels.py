"""Integration kernels on MI355X: the call surface of xpg/integrate_kernels.py over the HIP library.

``make_kernels(Dz, Dy, Dx, n_regions)`` returns ``(pass1, accumulate_profiles, pass2)`` with the reference's
signatures (xpg/integrate_kernels.py:216-485); each runs one HIP kernel on the current stream and returns device
tensors.  Constants and host helpers (rocking-curve basis, Gaussian prior) keep the reference's names and values.

Arguments are taken as the reference's ``process_block`` passes them (xpg/integrate.py:466-519): device tensors
(float32 ``S``/``phi``/``fb``/``pxs``, integer ``ix0``/``iy0``/``iz0``/``region``, the ``_detector_consts`` dict of
device tensors and Python scalars, ``axis``/``s0``/``rcw`` device tensors, numpy float32 scalars) or numpy arrays.

Differences from the reference that a caller can observe, all deliberate:
  * geometry is evaluated in float64 and voxel sums are compensated, so results match the float64 restatement of
    the algorithm to ~1e-7 instead of carrying float32 cancellation errors (|dI|/sigma up to ~3e-3 in the reference);
  * pass2's IRLS stops per reflection (results do not depend on which reflections share a call); the reference's
    chunk-global stop (xpg/integrate_kernels.py:464-467) differs from it by < 1e-3 sigma;
  * per-reflection outputs are float64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

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
FX_SCALE = float(2 ** 44)          # fixed-point profile grids (xpg_common.h)


def rc_default_weights():
    """Mixture weights that make a unit Gaussian rocking curve (xpg/integrate_kernels.py:125-127)."""
    w = np.exp(-0.5 * RC_C ** 2 / (1 - RC_W ** 2))
    return (w / w.sum()).astype(np.float32)


def rc_basis_np(phi_lo, phi_hi, phi_c, zeta, sm):
    """Frame integrals of each mixture component (N, Dz, K), numpy (xpg/integrate_kernels.py:144-151)."""
    from scipy.special import erf
    s = (sm / np.maximum(np.abs(zeta), 1e-3))[:, None, None]
    sw = s * RC_W * np.sqrt(2.0)
    a = (phi_hi[..., None] - phi_c[:, None, None] - RC_C * s) / sw
    b = (phi_lo[..., None] - phi_c[:, None, None] - RC_C * s) / sw
    return 0.5 * (erf(a) - erf(b))


def gaussian_profile(G=PROFILE_GRID):
    """Gaussian prior of the reference profiles (xpg/integrate_kernels.py:491-495)."""
    umax = N_BOX * 0.9
    u = np.linspace(-umax, umax, G)
    U1, U2 = np.meshgrid(u, u)
    return (np.exp(-0.5 * (U1 ** 2 + U2 ** 2)) / (2 * np.pi)).astype(np.float32)


def H(r):
    """Host copies of a kernel result's per-reflection (1-D) outputs in one transfer, cached in the dict
    (xpg/integrate_kernels.py:169-183)."""
    h = r.get("_h")
    if h is None:
        keys = [k for k, v in r.items() if torch.is_tensor(v) and v.dim() == 1]
        h = {}
        if keys:
            st = torch.stack([r[k].to(torch.float64) for k in keys]).cpu().numpy()
            for i, k in enumerate(keys):
                h[k] = st[i].astype(str(r[k].dtype).replace("torch.", ""))
        r["_h"] = h
    return h


# ---------------------------------------------------------------------------------------------------------------------
# argument marshalling
# ---------------------------------------------------------------------------------------------------------------------

def _dev(x, dtype, device):
    if torch.is_tensor(x):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(x)), dtype=dtype, device=device)


def _host64(x):
    """float64 host copy of an array, a device or host tensor, or a scalar."""
    if torch.is_tensor(x):
        return x.detach().to("cpu", torch.float64).numpy()
    return np.asarray(x, np.float64)


def _shape(x):
    return tuple(x.shape) if hasattr(x, "shape") else tuple(np.shape(x))


def _table(x):
    """A per-pixel detector table, or None for an absent one (None or the reference's (1, 1) placeholder)."""
    return None if x is None or _shape(x) == (1, 1) else x


class DeviceDetector:
    """Detector constants on the device (xpg/integrate_kernels.py:48-65).  ``d`` is a dict with P0, Bx, By, U1, U2, UN
    (K,3), orgx, orgy, qx, qy, wl, tsens, mu, plx and optional segmap, gx, gy, bkg tables (None when absent)."""

    def __init__(self, d, device):
        self.device = device
        f64 = lambda k: _dev(_host64(d[k]).reshape(-1, 3), torch.float64, device)
        self.t = {k: f64(k) for k in ("P0", "Bx", "By", "U1", "U2", "UN")}
        s = L.Detector()
        s.n_seg = int(self.t["P0"].shape[0])
        for k in ("P0", "Bx", "By", "U1", "U2", "UN"):
            setattr(s, k, L.ptr(self.t[k]))
        for k in ("orgx", "orgy", "qx", "qy", "wl", "tsens", "mu", "plx"):
            setattr(s, k, float(d[k]))
        seg = _table(d.get("segmap"))
        if seg is not None:
            self.t["segmap"] = _dev(seg, torch.int16, device)
            if int(self.t["segmap"].max()) >= s.n_seg:
                raise L.XpgError(f"segment map names segment {int(self.t['segmap'].max())} of {s.n_seg} affine maps")
            s.segmap, (s.seg_ny, s.seg_nx) = L.ptr(self.t["segmap"]), self.t["segmap"].shape
        gx = _table(d.get("gx"))
        if gx is not None:
            self.t["gx"] = _dev(gx, torch.float32, device)
            self.t["gy"] = _dev(d["gy"], torch.float32, device)
            if self.t["gy"].shape != self.t["gx"].shape:
                raise L.XpgError("geometric-correction tables gx and gy differ in shape")
            s.gx, s.gy = L.ptr(self.t["gx"]), L.ptr(self.t["gy"])
            s.geo_ny, s.geo_nx = self.t["gx"].shape
        bkg = _table(d.get("bkg"))
        if bkg is not None:
            self.t["bkg"] = _dev(bkg, torch.float32, device)
            s.bkg, (s.bkg_ny, s.bkg_nx) = L.ptr(self.t["bkg"]), self.t["bkg"].shape
        self.s = s


def device_detector(det, device):
    if isinstance(det, DeviceDetector):
        return det
    return DeviceDetector(det, device)


def make_call(axis, s0, frames_phi0, osc, sd, sm, gain, rcw):
    """Scalars of one kernel call (xpg_call); vectors may be device tensors, scalars numpy / 0-d tensors."""
    c = L.Call()
    c.axis[:] = [float(v) for v in _host64(axis).reshape(3)]
    c.s0[:] = [float(v) for v in _host64(s0).reshape(3)]
    c.frames_phi0, c.osc, c.sd, c.sm, c.gain = (float(v) for v in (frames_phi0, osc, sd, sm, gain))
    if rcw is None:
        c.use_rcw = 0
    else:
        c.rcw[:] = [float(v) for v in _host64(rcw).astype(np.float32).reshape(-1)]
        c.use_rcw = 1
    return c


class DeviceRefl:
    """Per-reflection table of a chunk on the device (xpg/integrate.py:509-517), with its ctypes view."""

    def __init__(self, rc, nb, shape, device):
        Dz, Dy, Dx = shape
        n = int(np.shape(rc["ix0"])[0])
        self.t = t = {}
        t["S"] = _dev(rc["S"], torch.float64, device)
        t["phi"] = _dev(rc["phi"], torch.float64, device)
        for k in ("ix0", "iy0", "iz0", "fb", "region", "pxs"):
            if k in rc and rc[k] is not None:
                t[k] = _dev(np.asarray(rc[k]) if not torch.is_tensor(rc[k]) else rc[k], torch.int32, device)
        if "has_p" in rc and rc["has_p"] is not None:
            t["has_p"] = _dev(rc["has_p"], torch.float32, device)
            t["pS"] = _dev(rc["pS"], torch.float64, device)
            t["pphi"] = _dev(rc["pphi"], torch.float64, device)
        K = 0
        if nb is not None:
            for k in ("x", "y", "z", "sx", "sz"):
                t["nb" + k] = _dev(nb[k], torch.float32, device)
            K = int(t["nbx"].shape[1]) if t["nbx"].dim() == 2 else 0
        s = L.Refl()
        s.n, s.Dz, s.Dy, s.Dx = n, Dz, Dy, Dx
        for k in ("S", "phi", "ix0", "iy0", "iz0", "fb", "pxs", "region", "has_p", "pS", "pphi"):
            setattr(s, k, L.ptr(t.get(k)))
        s.K = K
        for k in ("x", "y", "z", "sx", "sz"):
            setattr(s, "nb" + k, L.ptr(t.get("nb" + k)))
        self.s = s
        self.n = n


# ---------------------------------------------------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------------------------------------------------

P1_F64 = ("bg", "nbg", "isum", "vsum", "vsum0", "xo", "yo", "zo", "m_d", "m_m", "frac_rec", "nfg", "zeta", "seff")
P1_I32 = ("n_over", "n_fg_masked", "n_fg_nb")
P2_F64 = ("iprf", "vprf", "vprf0", "corr", "p_valid", "p_total", "ipartner", "rho", "p_fg")


def new_profile_acc(n_regions, device, G=PROFILE_GRID):
    num = torch.zeros((n_regions, G, G), dtype=torch.int64, device=device)
    den = torch.zeros((n_regions, G, G), dtype=torch.int64, device=device)
    acc = L.ProfileAcc(L.ptr(num), L.ptr(den), n_regions, 1)
    return num, den, acc


def run_prepare(det, call, refl, frames, device):
    """xpg_prepare of a chunk: (ctypes xpg_prepared, dict of its device buffers)."""
    n, Dz, Dy, Dx = refl.n, refl.s.Dz, refl.s.Dy, refl.s.Dx
    bufs = dict(sb=torch.empty((n, Dz, Dy, Dx), dtype=torch.float32, device=device),
                geo=torch.empty((n, 3, Dy, Dx), dtype=torch.float64, device=device),
                zeta=torch.empty(n, dtype=torch.float64, device=device),
                nbm=torch.empty((n, Dz, Dy, Dx), dtype=torch.uint8, device=device),
                geo_p=torch.empty((n, 3, Dy, Dx), dtype=torch.float64, device=device),
                zeta_p=torch.empty(n, dtype=torch.float64, device=device))
    ps = L.Prepared()
    for k, t in bufs.items():
        setattr(ps, k, L.ptr(t))
    L.check(L.lib().xpg_prepare(C.byref(det.s), C.byref(call), C.byref(refl.s), C.byref(frames), C.byref(ps),
                                L.stream_ptr(device)))
    return ps, bufs


def prepared_from_pix(det, call, refl, sb, pix):
    """Prepared data of a chunk (sb, neighbour masks, the partners' geometry: xpg_prepare_sb) whose own per-pixel
    geometry is pass 1's ``pix`` (xpg/integrate_kernels.py:236-260): eps = u seff in float64 from u and seff rounded
    to float32 (the reference's pass-1 dtype), so that the kernel's eps / seff gives back u exactly."""
    n, Dz, Dy, Dx = refl.n, refl.s.Dz, refl.s.Dy, refl.s.Dx
    device = det.device
    g = lambda k, dt: _dev(pix[k], dt, device)
    seff = g("seff", torch.float32).reshape(n).to(torch.float64)
    geo = torch.empty((n, 3, Dy, Dx), dtype=torch.float64, device=device)
    bufs = dict(sb=sb, geo=geo, zeta=torch.empty(n, dtype=torch.float64, device=device),
                nbm=torch.empty((n, Dz, Dy, Dx), dtype=torch.uint8, device=device),
                geo_p=torch.empty((n, 3, Dy, Dx), dtype=torch.float64, device=device),
                zeta_p=torch.empty(n, dtype=torch.float64, device=device), seff=seff)
    ps = L.Prepared()
    for k, t in bufs.items():
        if k != "seff":
            setattr(ps, k, L.ptr(t))
    L.check(L.lib().xpg_prepare_sb(C.byref(det.s), C.byref(call), C.byref(refl.s), L.ptr(sb), C.byref(ps),
                                   L.stream_ptr(device)))
    u1 = g("u1", torch.float32).reshape(n, Dy, Dx).to(torch.float64)
    u2 = g("u2", torch.float32).reshape(n, Dy, Dx).to(torch.float64)
    geo[:, 0] = u1 * seff[:, None, None]
    geo[:, 1] = u2 * seff[:, None, None]
    geo[:, 2] = g("J", torch.float64).reshape(n, Dy, Dx)
    bufs["zeta"].copy_(g("zeta", torch.float64).reshape(n))
    ps.seff = L.ptr(seff)
    return ps, bufs


def run_pass1(det, call, refl, sb=None, frames=None, acc=None, pixel_outputs=True, prep=None, masks=False):
    """One pass-1 launch; ``frames`` (a FrameStore view) instead of ``sb`` extracts from the frame store in-kernel,
    ``prep`` (run_prepare) reads prepared data.  ``masks``: also the per-voxel classes (``vmask``, XPG_VM_* bits)."""
    n, Dz, Dy, Dx = refl.n, refl.s.Dz, refl.s.Dy, refl.s.Dx
    device = det.device
    out = {k: torch.empty(n, dtype=torch.float64, device=device) for k in P1_F64}
    out.update({k: torch.empty(n, dtype=torch.int32, device=device) for k in P1_I32})
    out["fz"] = torch.empty((n, Dz), dtype=torch.float32, device=device)
    if pixel_outputs:
        for k in ("img2", "J", "u1", "u2"):
            out[k] = torch.empty((n, Dy, Dx), dtype=torch.float32, device=device)
        out["all_valid2"] = torch.empty((n, Dy, Dx), dtype=torch.uint8, device=device)
    if masks:
        out["vmask"] = torch.zeros((n, Dz, Dy, Dx), dtype=torch.uint8, device=device)
    o = L.Pass1Out()
    for k, _ in L.Pass1Out._fields_:
        if k != "fz_stride":
            setattr(o, k, L.ptr(out.get(k)))
    st = L.stream_ptr(device)
    accp = C.byref(acc) if acc is not None else None
    if prep is not None:
        L.check(L.lib().xpg_pass1_prepared(C.byref(det.s), C.byref(call), C.byref(refl.s), C.byref(prep), C.byref(o),
                                           accp, st))
    elif frames is not None:
        L.check(L.lib().xpg_pass1_frames(C.byref(det.s), C.byref(call), C.byref(refl.s), C.byref(frames), C.byref(o),
                                         accp, st))
    else:
        L.check(L.lib().xpg_pass1(C.byref(det.s), C.byref(call), C.byref(refl.s), L.ptr(sb), C.byref(o), accp, st))
    if pixel_outputs:
        out["all_valid2"] = out["all_valid2"].bool()
    return out


def run_pass2(det, call, refl, prof, bg, nbg, sb=None, frames=None, irls_rule=0, prep=None, masks=False):
    """One pass-2 launch (sources as run_pass1); ``masks``: also the per-voxel fit mask and neighbour bits."""
    n = refl.n
    device = det.device
    out = {k: torch.empty(n, dtype=torch.float64, device=device) for k in P2_F64}
    if masks:
        out["vmask"] = torch.zeros((n, refl.s.Dz, refl.s.Dy, refl.s.Dx), dtype=torch.uint8, device=device)
    o = L.Pass2Out()
    for k, _ in L.Pass2Out._fields_:
        setattr(o, k, L.ptr(out.get(k)))
    prof = _dev(prof, torch.float32, device)
    bg = _dev(bg, torch.float64, device)
    nbg = _dev(nbg, torch.float64, device)
    st = L.stream_ptr(device)
    if prep is not None:
        L.check(L.lib().xpg_pass2_prepared(C.byref(det.s), C.byref(call), C.byref(refl.s), C.byref(prep),
                                           L.ptr(prof), int(prof.shape[0]), L.ptr(bg), L.ptr(nbg), int(irls_rule),
                                           C.byref(o), st))
    elif frames is not None:
        L.check(L.lib().xpg_pass2_frames(C.byref(det.s), C.byref(call), C.byref(refl.s), C.byref(frames), L.ptr(prof),
                                         int(prof.shape[0]), L.ptr(bg), L.ptr(nbg), int(irls_rule), C.byref(o), st))
    else:
        L.check(L.lib().xpg_pass2(C.byref(det.s), C.byref(call), C.byref(refl.s), L.ptr(sb), L.ptr(prof),
                                  int(prof.shape[0]), L.ptr(bg), L.ptr(nbg), int(irls_rule), C.byref(o), st))
    return out


def normalise_profiles(num_fx, den_fx, device, G=PROFILE_GRID, gauss=None):
    """Fixed-point grids -> float32 profiles (xpg/integrate.py:546-556), on the device; also the float64 grids.
    ``gauss``: the (G, G) float32 Gaussian prior already on the device (no host copy: capturable in a graph)."""
    R = int(num_fx.shape[0])
    acc = L.ProfileAcc(L.ptr(num_fx), L.ptr(den_fx), R, 1)
    if gauss is None:
        gauss = torch.as_tensor(gaussian_profile(G), device=device)
    prof = torch.empty((R, G, G), dtype=torch.float32, device=device)
    num64 = torch.empty((R, G, G), dtype=torch.float64, device=device)
    den64 = torch.empty((R, G, G), dtype=torch.float64, device=device)
    L.check(L.lib().xpg_normalise_profiles(C.byref(acc), L.ptr(gauss), L.ptr(prof), L.ptr(num64), L.ptr(den64),
                                           L.stream_ptr(device)))
    return prof, num64, den64


def make_kernels(Dz, Dy, Dx, n_regions, G=PROFILE_GRID):
    """(pass1, accumulate_profiles, pass2) for one shoebox shape, as xpg/integrate_kernels.py:216-485."""

    def _setup(sb, rc, det, axis, s0, frames_phi0, osc, sd, sm, gain, nb, rcw):
        L.require_device(sb)
        device = sb.device
        d = device_detector(det, device)
        call = make_call(axis, s0, frames_phi0, osc, sd, sm, gain, rcw)
        refl = DeviceRefl(rc, nb, (Dz, Dy, Dx), device)
        if tuple(sb.shape) != (refl.n, Dz, Dy, Dx):
            raise L.XpgError(f"sb shape {tuple(sb.shape)} != {(refl.n, Dz, Dy, Dx)}")
        return d, call, refl, sb.to(torch.float32).contiguous()

    def pass1(sb, rc, det, axis, s0, frames_phi0, osc, sd, sm, gain, prof, nb, rcw):
        """Background, summation, centroids, 2D projected spot images (xpg/integrate_kernels.py:287-365)."""
        d, call, refl, sbf = _setup(sb, rc, det, axis, s0, frames_phi0, osc, sd, sm, gain, nb, rcw)
        return run_pass1(d, call, refl, sb=sbf)

    def accumulate_profiles(img2, ok2, J, u1, u2, region, isum, sel, seff):
        """Per-region grids of normalised projected spots (xpg/integrate_kernels.py:367-380): float64 (R, G, G)."""
        L.require_device(img2)
        device = img2.device
        n = int(img2.shape[0])
        num_fx, den_fx, acc = new_profile_acc(n_regions, device, G)
        args = [_dev(img2, torch.float32, device), _dev(ok2, torch.uint8, device), _dev(J, torch.float32, device),
                _dev(u1, torch.float32, device), _dev(u2, torch.float32, device), _dev(region, torch.int32, device),
                _dev(isum, torch.float64, device), _dev(sel, torch.uint8, device), _dev(seff, torch.float64, device)]
        L.check(L.lib().xpg_accumulate_profiles(*[L.ptr(a) for a in args], n, int(img2.shape[1]),
                                                int(img2.shape[2]), C.byref(acc), L.stream_ptr(device)))
        return num_fx.to(torch.float64) / FX_SCALE, den_fx.to(torch.float64) / FX_SCALE

    def pass2(sb, rc, det, axis, s0, frames_phi0, osc, sd, sm, gain, prof, bg, nbm_x, nb, rcw, pix=None):
        """Profile fitting with fixed background and 2x2 overlap deconvolution (xpg/integrate_kernels.py:382-483).
        ``pix``: pass 1's per-pixel geometry of these reflections (zeta, J, u1, u2, seff), used instead of
        recomputing it, as the reference does; the overlapping partners' geometry is computed afresh."""
        d, call, refl, sbf = _setup(sb, rc, det, axis, s0, frames_phi0, osc, sd, sm, gain, nb, rcw)
        if pix is None:
            return run_pass2(d, call, refl, prof, bg, nbm_x, sb=sbf)
        ps, keep = prepared_from_pix(d, call, refl, sbf, pix)
        out = run_pass2(d, call, refl, prof, bg, nbm_x, prep=ps)
        out["_prepared"] = keep                  # the buffers stay alive until the launch has completed
        return out

    return pass1, accumulate_profiles, pass2


_KERNEL_CACHE: dict = {}


def run_extract(store, refl, device):
    """Value-coded shoeboxes (n, Dz, Dy, Dx) of a chunk from a device frame store (DeviceFrames.store() view)."""
    sb = torch.empty((refl.n, refl.s.Dz, refl.s.Dy, refl.s.Dx), dtype=torch.float32, device=device)
    L.check(L.lib().xpg_extract(C.byref(store), C.byref(refl.s), L.ptr(sb), L.stream_ptr(device)))
    return sb

"""Forward (window) model passes on MI355X: the call surface of xpg/forward/kernels.py over the HIP kernels of
include/xpg_forward.h.

    scatter(mu, R, P, I, beta, iz0, iy0, ix0, win)            mu[v] += I f + beta over the in-window voxels
    gather(g, fd, R, P, rows, iz0, iy0, ix0, win, A, M, NS, mode)
        mode 0: g[rows] += sum f A / M; mode 1: g[rows] += sum f (1 - NS / M), fd[rows] += sum f^2 / M (voxels with
        NS >= 0, inside the window); device rows outside g add nothing, host rows outside g raise

    fraction_terms(det, axis, s0, frames_phi0, osc, sd, sm, S, phi, ix0, iy0, iz0, Dz, Dx, prof, region, umax, halo,
                   rcw, pxs)                                   the separable factors [(R, P), ...] (xpg_fm_terms)

R: (T, n, Dz), P: (T, n, b, b) float32 (T = 1, or 2 with the halo); iz0 / iy0 / ix0 the supports' origins; ``win``
any object with f0, f1, y0, ny, x0, nx (xpg/forward/model.py's Window).  Asynchronous on the current stream; there is
no CPU fallback (``available`` is False for the CPU, and it and the calls raise on a GPU without the library).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

NS_TYPES = {torch.int16: 0, torch.int32: 1, torch.float32: 2}


def available(dev) -> bool:
    """The fused device passes apply (xpg/forward/kernels.py:29-31): True for a GPU device, False for the CPU.  On a
    GPU device the library must be there: its absence raises rather than falling back to the eager passes."""
    if getattr(dev, "type", str(dev)) != "cuda":
        return False
    L.lib()
    return True


def _f32(t, name):
    L.require_device(t)
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise L.XpgError(f"{name}: contiguous float32 device tensor expected")
    return t


def _supports(R, P, iz0, iy0, ix0):
    _f32(R, "R")
    _f32(P, "P")
    if R.dim() != 3 or P.dim() != 4 or P.shape[:2] != R.shape[:2] or P.shape[2] != P.shape[3]:
        raise L.XpgError("R (T, n, Dz) and P (T, n, b, b) expected")
    T, n, Dz = R.shape
    idx = [t.to(device=R.device, dtype=torch.int32).contiguous() for t in (iz0, iy0, ix0)]
    if any(t.numel() != n for t in idx):
        raise L.XpgError("iz0, iy0, ix0: one origin per reflection")
    s = L.FmSupports(L.ptr(R), L.ptr(P), L.ptr(idx[0]), L.ptr(idx[1]), L.ptr(idx[2]), T, n, Dz, P.shape[-1])
    return s, idx


def _window(win):
    return L.FmWindow(int(win.f0), int(win.f1 - win.f0), int(win.y0), int(win.ny), int(win.x0), int(win.nx))


def _n_vox(win):
    return (int(win.f1) - int(win.f0)) * int(win.ny) * int(win.nx)


def scatter(mu, R, P, I, beta, iz0, iy0, ix0, win):
    """mu[voxel] += I_row f + beta_row over the in-window voxels of each reflection's support (as WindowModel.mu)."""
    _f32(mu, "mu")
    if mu.numel() < _n_vox(win):
        raise L.XpgError("mu is smaller than the window")
    if R.shape[1] == 0:
        return mu
    s, keep = _supports(R, P, iz0, iy0, ix0)
    _f32(I, "I")
    if beta is not None:
        _f32(beta, "beta")
    if I.numel() != R.shape[1] or (beta is not None and beta.numel() != R.shape[1]):
        raise L.XpgError("I / beta: one value per reflection")
    w = _window(win)
    L.check(L.lib().xpg_fm_scatter(L.ptr(mu), C.byref(s), L.ptr(I), L.ptr(beta), C.byref(w),
                                   L.stream_ptr(mu.device)))
    return mu


def gather(g, fd, R, P, rows, iz0, iy0, ix0, win, A, M, NS, mode):
    """mode 0: g[rows] += sum f A / M; mode 1: g[rows] += sum f (1 - NS / M), fd[rows] += sum f^2 / M (voxels with
    NS >= 0, inside the window)."""
    _f32(g, "g")
    if R.shape[1] == 0:
        return g, fd
    s, keep = _supports(R, P, iz0, iy0, ix0)
    nv = _n_vox(win)
    _f32(M, "M")
    if mode == 0:
        _f32(A, "A")
    elif mode == 1:
        _f32(fd, "fd")
        if fd.numel() < g.numel():
            raise L.XpgError("fd is smaller than g")
    else:
        raise L.XpgError("mode 0 or 1")
    L.require_device(NS)
    if NS.dtype not in NS_TYPES or not NS.is_contiguous():
        raise L.XpgError("NS: contiguous int16, int32 or float32")
    if M.numel() < nv or NS.numel() < nv or (mode == 0 and A.numel() < nv):
        raise L.XpgError("A / M / NS are smaller than the window")
    if torch.is_tensor(rows) and rows.is_cuda:
        rows64 = rows.to(dtype=torch.int64).contiguous()          # the kernel skips rows outside g: no sync here
    else:
        rh = torch.as_tensor(rows, dtype=torch.int64)
        if rh.numel() and (int(rh.min()) < 0 or int(rh.max()) >= g.numel()):
            raise L.XpgError("rows: an index into g per reflection")
        rows64 = rh.to(g.device)
    if rows64.numel() != R.shape[1]:
        raise L.XpgError("rows: one index per reflection")
    w = _window(win)
    L.check(L.lib().xpg_fm_gather(L.ptr(g), L.ptr(fd) if mode == 1 else None, C.byref(s), L.ptr(rows64), g.numel(),
                                  L.ptr(A) if mode == 0 else None, L.ptr(M), L.ptr(NS), NS_TYPES[NS.dtype], mode,
                                  C.byref(w), L.stream_ptr(g.device)))
    return g, fd


def fraction_factors(det, axis, s0, frames_phi0, osc, sd, sm, S, phi, ix0, iy0, iz0, Dz, Dx, prof=None, region=None,
                     umax=None, halo=None, rcw=None, pxs=1, device=None):
    """The factors of fraction_terms stacked as the fused passes take them: R (T, n, Dz), P (T, n, Dx, Dx) float32
    device tensors (xpg_fm_terms, include/xpg_forward.h).  ``det``: the integration kernels' DeviceDetector or the
    reference's detector-constants dict (torch tensors or arrays); sd, sm scalars, 0-d or (n,) tensors / arrays
    (radians); vectors may be device tensors."""
    from .integrate_kernels import DeviceDetector, _dev, _host64
    device = torch.device(device) if device is not None else (S.device if torch.is_tensor(S) else torch.device("cuda"))
    if device.type != "cuda":
        raise L.XpgError("fraction_terms runs on the GPU")
    d = det if isinstance(det, DeviceDetector) else DeviceDetector(det, device)
    n = int(S.shape[0])
    Dz, b, pxs = int(Dz), int(Dx), int(pxs)
    m = L.FmModel()
    m.axis[:] = [float(v) for v in _host64(axis).reshape(3)]
    m.s0[:] = [float(v) for v in _host64(s0).reshape(3)]
    m.frames_phi0, m.osc = float(frames_phi0), float(osc)
    keep = {}
    rows = L.FmRows()
    for k, v in (("sd", sd), ("sm", sm)):
        a = _host64(v) if not torch.is_tensor(v) else v
        if (torch.is_tensor(a) and a.numel() > 1) or (not torch.is_tensor(a) and np.size(a) > 1):
            keep[k] = _dev(a, torch.float64, device).reshape(-1)
            if keep[k].numel() != n:
                raise L.XpgError(f"{k}: a scalar or one value per reflection")
            setattr(rows, k, L.ptr(keep[k]))
            setattr(m, k, 0.0)
        else:
            setattr(m, k, float(np.asarray(_host64(a)).reshape(-1)[0]))
    keep["S"] = _dev(S, torch.float64, device).reshape(n, 3)
    keep["phi"] = _dev(phi, torch.float64, device).reshape(n)
    for k, v in (("ix0", ix0), ("iy0", iy0), ("iz0", iz0)):
        keep[k] = _dev(v, torch.int32, device).reshape(n)
    rows.S, rows.phi = L.ptr(keep["S"]), L.ptr(keep["phi"])
    rows.ix0, rows.iy0, rows.iz0 = L.ptr(keep["ix0"]), L.ptr(keep["iy0"]), L.ptr(keep["iz0"])
    rows.n, rows.Dz, rows.b, rows.pxs = n, Dz, b, pxs
    if prof is not None:
        keep["prof"] = _dev(prof, torch.float32, device)
        G = int(keep["prof"].shape[-1])
        keep["prof"] = keep["prof"].reshape(-1, G, G).contiguous()
        if region is None:
            raise L.XpgError("learnt profiles need each reflection's region")
        keep["region"] = _dev(region, torch.int32, device).reshape(n)
        m.prof, m.n_regions, m.G = L.ptr(keep["prof"]), int(keep["prof"].shape[0]), G
        m.umax = float(umax) if umax is not None else 4.5
        rows.region = L.ptr(keep["region"])
    if halo is not None:
        m.halo, m.halo_w, m.halo_kappa = 1, float(halo[0]), float(halo[1])
    if rcw is not None:
        m.rcw[:] = [float(v) for v in _host64(rcw).astype(np.float32).reshape(-1)]
        m.use_rcw = 1
    T = 2 if halo is not None else 1
    R = torch.empty((T, n, Dz), dtype=torch.float32, device=device)
    P = torch.empty((T, n, b, b), dtype=torch.float32, device=device)
    with torch.cuda.device(device):
        L.check(L.lib().xpg_fm_terms(C.byref(d.s), C.byref(m), C.byref(rows), L.ptr(R), L.ptr(P),
                                     L.stream_ptr(device)))
    return R, P


def fraction_terms(det, axis, s0, frames_phi0, osc, sd, sm, S, phi, ix0, iy0, iz0, Dz, Dx, prof=None, region=None,
                   umax=None, halo=None, rcw=None, pxs=1):
    """fraction_terms of xpg/forward/model.py:206-275 on the device: the separable factors [(R (n, Dz),
    P (n, Dx, Dx)), ...] of each reflection's fraction in its support (one term; two with the halo), float32 device
    tensors, from the float64 device geometry of the integration kernels (xpg_fm_terms)."""
    R, P = fraction_factors(det, axis, s0, frames_phi0, osc, sd, sm, S, phi, ix0, iy0, iz0, Dz, Dx, prof, region, umax,
                            halo, rcw, pxs)
    return [(R[t], P[t]) for t in range(R.shape[0])]

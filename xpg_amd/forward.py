"""Forward (window) model passes on MI355X: the call surface of xpg/forward/kernels.py over the HIP kernels of
include/xpg_forward.h.

    scatter(mu, R, P, I, beta, iz0, iy0, ix0, win)            mu[v] += I f + beta over the in-window voxels
    gather(g, fd, R, P, rows, iz0, iy0, ix0, win, A, M, NS, mode)
        mode 0: g[rows] += sum f A / M; mode 1: g[rows] += sum f (1 - NS / M), fd[rows] += sum f^2 / M (voxels with
        NS >= 0, inside the window); device rows outside g add nothing, host rows outside g raise

R: (T, n, Dz), P: (T, n, b, b) float32 (T = 1, or 2 with the halo); iz0 / iy0 / ix0 the supports' origins; ``win``
any object with f0, f1, y0, ny, x0, nx (xpg/forward/model.py's Window).  Asynchronous on the current stream; there is
no CPU fallback (``available`` is False for the CPU, and it and the calls raise on a GPU without the library).
"""
from __future__ import annotations

import ctypes as C

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

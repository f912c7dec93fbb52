"""The integration kernels as PyTorch operators (``torch.ops.xint``), from ``libxint_ops.so`` (xpg_amd/csrc/xint_ops.cpp,
TORCH_LIBRARY over the C ABI of include/xpg_integrate.h).

Each op runs on the current stream, allocates its outputs through torch's caching allocator and raises RuntimeError
on bad arguments.  The helpers below build the op arguments from the dictionaries the reference's ``process_block``
passes to ``make_kernels``' closures (xpg/integrate.py:466-519): ``rc`` (S, phi, ix0, iy0, iz0, fb, region, pxs,
has_p, pS, pphi), ``nb`` (x, y, z, sx, sz) and the ``_detector_consts`` dict.

    torch.ops.xint.extract(frames, frame0, min_valid, flag_value, overload, use_flag_value, rows, box) -> sb
    torch.ops.xint.pass1(sb, rows, det, det_scalars, call, rcw, acc_num=None, acc_den=None) -> [P1_OUT...]
    torch.ops.xint.accumulate_profiles(img2, ok2, J, u1, u2, region, isum, sel, seff, n_regions) -> (num, den)
    torch.ops.xint.normalise_profiles(num_fx, den_fx, gauss) -> (prof, num64, den64)
    torch.ops.xint.pass2(sb, rows, det, det_scalars, call, rcw, prof, bg, nbg) -> [P2_OUT...]
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from . import _lib as L
from .integrate_kernels import P1_F64, P1_I32, P2_F64, _dev, _host64, _table

OPS_PATH = Path(__file__).resolve().parent / "libxint_ops.so"
ROW_KEYS = ("S", "phi", "ix0", "iy0", "iz0", "fb", "pxs", "region", "has_p", "pS", "pphi", "nbx", "nby", "nbz", "nbsx",
            "nbsz")
P1_OUT = P1_F64 + P1_I32 + ("fz", "img2", "J", "u1", "u2", "all_valid2")
P2_OUT = P2_F64
DET_SCALARS = ("orgx", "orgy", "qx", "qy", "wl", "tsens", "mu", "plx")

_loaded = False


def load():
    """Register the xint operators (raises when the library has not been built)."""
    global _loaded
    if not _loaded:
        if not OPS_PATH.exists():
            raise L.XpgError(f"{OPS_PATH} is missing: build it with __graft_entry__.build()")
        L.lib()                       # the kernel library first, so that both share one copy of it
        torch.ops.load_library(str(OPS_PATH))
        _loaded = True
    return torch.ops.xint


def rows(rc, nb, device):
    """The row list of an op (ROW_KEYS order) from the reference's rc / nb dictionaries."""
    t = {"S": _dev(rc["S"], torch.float64, device), "phi": _dev(rc["phi"], torch.float64, device)}
    for k in ("ix0", "iy0", "iz0", "fb", "region", "pxs"):
        if rc.get(k) is not None:
            t[k] = _dev(rc[k], torch.int32, device)
    if rc.get("has_p") is not None:
        t["has_p"] = _dev(rc["has_p"], torch.float32, device)
        t["pS"] = _dev(rc["pS"], torch.float64, device)
        t["pphi"] = _dev(rc["pphi"], torch.float64, device)
    if nb is not None:
        for k in ("x", "y", "z", "sx", "sz"):
            t["nb" + k] = _dev(nb[k], torch.float32, device)
    return [t.get(k) for k in ROW_KEYS]


def detector(d, device):
    """(det list, det_scalars) of an op from a _detector_consts dict."""
    aff = np.stack([_host64(d[k]).reshape(-1, 3) for k in ("P0", "Bx", "By", "U1", "U2", "UN")])
    tabs = [torch.as_tensor(aff, dtype=torch.float64, device=device).contiguous()]
    seg, gx, gy, bkg = (_table(d.get(k)) for k in ("segmap", "gx", "gy", "bkg"))
    tabs.append(None if seg is None else _dev(seg, torch.int16, device))
    tabs.append(None if gx is None else _dev(gx, torch.float32, device))
    tabs.append(None if gx is None else _dev(gy, torch.float32, device))
    tabs.append(None if bkg is None else _dev(bkg, torch.float32, device))
    return tabs, [float(d[k]) for k in DET_SCALARS]


def call(axis, s0, frames_phi0, osc, sd, sm, gain, rcw):
    """(call scalars, mixture weights or None) of an op."""
    c = [float(v) for v in _host64(axis).reshape(3)] + [float(v) for v in _host64(s0).reshape(3)]
    c += [float(frames_phi0), float(osc), float(sd), float(sm), float(gain)]
    w = None if rcw is None else [float(v) for v in _host64(rcw).astype(np.float32).reshape(-1)]
    return c, w

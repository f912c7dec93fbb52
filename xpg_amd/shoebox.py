"""Shoebox value codes of xpg/shoebox.py for the device path.

SB_OFF and the stored-value convention (value + 3, 0 = unrecorded) are xpg/shoebox.py:11; ``_decode_bins`` runs the
bin decoding of xpg/shoebox.py:264-268 on the device (HIP kernel ``xpg_decode_bins``).  The reference's host shoebox
buffers and their memory planner (xpg/shoebox.py:19-261) have no counterpart here: frames stay in HBM, and sweeps
larger than the HBM budget stream their frames through a ring on the device (xpg_amd/passes.py).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L

SB_OFF = np.float32(3.0)


def _decode_bins(arr, cnt):
    """Binned shoebox buffer from IEEE-encoded sums (+inf masked, NaN overloaded, 0 = no frame yet) to the stored
    convention, in place, on the device.  arr (n, B, bx, bx) float32 and cnt (n, B) int16 device tensors."""
    L.require_device(arr)
    L.require_device(cnt)
    n, B = int(arr.shape[0]), int(arr.shape[1])
    bxbx = int(arr.shape[2] * arr.shape[3])
    if tuple(cnt.shape) != (n, B) or arr.dtype != torch.float32 or cnt.dtype != torch.int16:
        raise L.XpgError("arr (n, B, bx, bx) float32 and cnt (n, B) int16 expected")
    L.check(L.lib().xpg_decode_bins(L.ptr(arr), L.ptr(cnt), n, B, bxbx, L.stream_ptr(arr.device)))
    return arr

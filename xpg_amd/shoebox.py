"""Shoebox value codes and the pass planner of xpg/shoebox.py, for the device path.

SB_OFF and the stored-value convention (value + 3, 0 = unrecorded) are xpg/shoebox.py:11; ``_decode_bins`` runs the
bin decoding of xpg/shoebox.py:264-268 on the device (HIP kernel ``xpg_decode_bins``).

``_plan_passes`` / ``_split_blocks`` keep the contract of xpg/shoebox.py:134-196: given per-reflection buffer bytes
held from the reflection's first frame until its block completes, group blocks into passes over the frames (and
blocks into parts of consecutive phi) that fit a memory budget.  The device path keeps whole frames in HBM and needs
no shoebox buffers, so it does not call them; they are provided for callers that plan host-side shoebox gathers with
the reference's planner, with the reference's behaviour (its tests are restated in tests/test_shoebox.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

SB_OFF = np.float32(3.0)
KERNEL_RESERVE = 2.0 * 2 ** 30


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


def _plan_passes(order, blocks_refl, lattice_of, need_hi, t_row, w_row, n_frames, budget, max_passes=16):
    """Passes over the frames within ``budget`` bytes.  A reflection costs w_row bytes from its first frame t_row
    until its block's last frame need_hi; one pass peaks at the largest open total over the frames plus twice its
    largest block plus KERNEL_RESERVE.  Candidates, in order: one pass; k passes by lattice; then, within each
    lattice, phi blocks m apart (m = 2 ..).  The first candidate under the budget wins; if none fits, the fewest passes
    within 5% of the best achievable peak.  Returns (passes, dict(peak, open, block, one, plan))."""
    size = {b: float(w_row[blocks_refl[b]].sum()) for b in order}

    def peak(blist):
        rows = np.concatenate([blocks_refl[b] for b in blist])
        delta = np.bincount(t_row[rows], w_row[rows], minlength=n_frames + 1)[:n_frames + 1]
        closes = np.clip([need_hi[b] + 1 for b in blist], 0, n_frames)
        delta = delta - np.bincount(closes, [size[b] for b in blist], minlength=n_frames + 1)
        open_max = float(np.cumsum(delta).max())
        biggest = max(size[b] for b in blist)
        return open_max + 2 * biggest + KERNEL_RESERVE, open_max, biggest

    lat = [lattice_of[b] for b in order]
    lats = sorted(set(lat))
    rank, count = [], {}
    for l_ in lat:
        rank.append(count.get(l_, 0))
        count[l_] = rank[-1] + 1
    cands = [("one pass", [0] * len(order))]
    cands += [(f"{k} passes by lattice", [lats.index(l_) % k for l_ in lat]) for k in range(2, len(lats) + 1)]
    for m in range(2, max_passes // len(lats) + 1):
        label = f"{len(lats) * m} passes ({'per lattice, ' if len(lats) > 1 else ''}phi blocks {m} apart)"
        cands.append((label, [lats.index(l_) * m + r_ % m for l_, r_ in zip(lat, rank)]))
    tried = []
    for name, key in cands:
        passes = [p for p in ([b for b, k_ in zip(order, key) if k_ == q] for q in range(max(key) + 1)) if p]
        tried.append((passes, max((peak(p) for p in passes), key=lambda v: v[0]), name))
        if budget is None or tried[-1][1][0] <= budget:
            break
    else:
        best = min(t[1][0] for t in tried)
        cut = best + 0.05 * (tried[0][1][0] - best)
        p_, w_, n_ = next(t for t in tried if t[1][0] <= cut)
        tried.append((p_, w_, n_ + " (still above the budget)"))
    passes, worst, name = tried[-1]
    return passes, dict(peak=worst[0], open=worst[1], block=worst[2], one=tried[0][1][0], plan=name)


def _split_blocks(blocks_refl, z, w_row, k):
    """Each block's rows in k parts of consecutive z with about equal bytes (rows in index order within a part).
    Returns (parts, block of each part); empty parts are dropped."""
    parts, parent = [], []
    for b, rows in enumerate(blocks_refl):
        o = rows[np.argsort(z[rows], kind="stable")]
        c = np.cumsum(w_row[o])
        cuts = np.searchsorted(c, c[-1] * np.arange(1, k) / k, side="right") if o.size else []
        for p_ in np.split(o, cuts):
            if p_.size:
                parts.append(np.sort(p_))
                parent.append(b)
    return parts, parent

"""CPU oracle of the integration's host planning -- TEST INFRASTRUCTURE ONLY.

Per-reflection, pure-Python restatements of the integer planning of xpg/integrate.py, written one reflection at a
time (the product code is vectorised numpy):

  box_depth      xpg/integrate.py:235-250  depth half-width hz, phi-binning fb, first frame iz0, bucket
  box_side       xpg/integrate.py:263-287  side bxs (XY_BUCKETS), super-pixel side pxs, origin ix0 / iy0
  neighbours     xpg/integrate.py:164-172  k nearest in scaled space within 4 N_FG (brute force)
  frame_loop     xpg/integrate.py:729-803  the order in which blocks reach the block worker, by walking the frames

Pinning: these restate the reference's formulas term by term; tests/test_planning.py checks the product's
vectorised planning against them on seeded reflection lists and against hand-computed cases (the 0.1 deg
mosaicity example of the depth rule: dz >= 11, bucket 17).
"""
from __future__ import annotations

import math

import numpy as np

N_BOX = 5.0
N_BOX_Z = 4.0
N_FG = 3.0
DZ_BUCKETS = (3, 5, 9, 17, 33, 65)
BIN_SLOTS = 33
BIN_MAX = 8
XY_BUCKETS = (9, 13, 17, 23, 31, 41, 55, 73, 97)
SUPERPIXEL = 2.5
SUPERPIXEL_MAX = 4


def box_depth(z, zeta_k, sigma_m, osc):
    sz = sigma_m / abs(zeta_k) / abs(osc)
    hz = math.ceil(N_BOX_Z * sz + 0.5)
    hmax = (max(DZ_BUCKETS) - 1) // 2
    fb = 1
    if hz > hmax:
        fb = int(min(max(math.ceil((2 * hz + 1) / BIN_SLOTS), 2), BIN_MAX))
        hz = min(math.ceil(N_BOX_Z * sz / fb + 0.5), (BIN_SLOTS - 1) // 2)
    iz0 = math.floor(z - (fb - 1) / 2.0) - hz * fb
    dz = 2 * hz + 1
    bucket = next((b for b in DZ_BUCKETS if dz <= b), -1)
    return hz, fb, iz0, dz, bucket


def box_side(x, y, spx, box, box_xy):
    """spx: footprint of sigma_D in pixels at the reflection (with the pixel term); box: global box side."""
    xyb = [b for b in XY_BUCKETS if b < box] + [box]

    def pick(need):
        for b in xyb:
            if b >= need:
                return b
        return xyb[-1]
    if box_xy:
        bxs, pxs = box, 1
    else:
        bxs = pick(2 * max(math.ceil(N_BOX * spx) + 1, 2) + 1)
        pxs = int(min(max(math.floor(spx / SUPERPIXEL), 1), SUPERPIXEL_MAX))
        if pxs > 1:
            bxs = pick(2 * max(math.ceil(N_BOX * spx / pxs) + 1, 2) + 1)
    half = bxs // 2
    ix0 = math.floor(x - 1.0 + 0.5) // pxs - half
    iy0 = math.floor(y - 1.0 + 0.5) // pxs - half
    return bxs, pxs, ix0, iy0


def neighbours(x, y, z, sx, sz, k=8):
    pts = np.c_[np.asarray(x) / sx, np.asarray(y) / sx, np.asarray(z) / sz]
    n = len(pts)
    out = []
    for i in range(n):
        d = np.sqrt(((pts - pts[i]) ** 2).sum(1))
        d[i] = np.inf
        order = sorted((j for j in range(n) if d[j] <= 4.0 * N_FG), key=lambda j: (d[j], j))[:k]
        out.append(order)
    return out


def frame_loop(need_lo, need_hi, has_rows, frames):
    """Blocks in the order the frame loop submits them (walks every processed frame)."""
    pending = sorted((b for b in range(len(need_lo)) if has_rows[b]), key=lambda b: need_lo[b])
    open_ = []
    order = []
    for f in frames:
        fi = f - 1
        while pending and need_lo[pending[0]] <= fi:
            open_.append(pending.pop(0))
        for b in list(open_):
            if need_hi[b] <= fi:
                open_.remove(b)
                order.append(b)
    return order + open_

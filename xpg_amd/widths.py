"""Spot-width estimate: beam divergence sigma_D and mosaicity sigma_M from the strong spots of a few sampled phi
ranges, on top of moments-only device sweeps.

The procedure is the one of xpg/integrate.py:1002-1092 (``estimate_sigmas``) with its helpers
``_rocking_resolved`` (:1107-1127) and ``_fit_rocking_width`` (:1130-1154):

  * five phi ranges spread over the sweep, each at least 5 degrees wide, integrated with pass 1 only;
  * per range the candidate spots (inside the range, first lattice, no overlapping partner, positive summation
    intensity, no overload, > 98% of the rocking curve recorded, a rocking curve that falls off inside the box);
  * strong spots by tiers of I/sigma (> 20, or > 10 when that leaves < 150; ranges with < 20 are dropped; all ranges
    pooled at > 5 when no range is left);
  * sigma_D from the median truncation-corrected second moment minus the pixel footprint, sigma_M from a grid search
    of frame-integrated Gaussian rocking curves against the spots' per-frame fractions;
  * up to three iterations, each re-integrating with the new widths; an estimate that keeps growing with the box is
    rejected (the previous one is kept).

The arithmetic here is host numpy: the grid search evaluates every grid point at once over the strong spots
(``fit_rocking_width``), and the resolved-curve test finds each row's first and last recorded frame from cumulative
counts (``rocking_resolved``).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

from .integrate_kernels import FG_R

N_RANGES = 5
GRID_POINTS = 61
# <r^2 / 2> / sigma^2 of a 2D Gaussian truncated at radius FG_R sigma (the foreground ellipse of pass 1)
TRUNC_T2 = 1.0 - (FG_R ** 2 / 2) * math.exp(-FG_R ** 2 / 2) / (1 - math.exp(-FG_R ** 2 / 2))


def sample_ranges(first, last, osc, n_sample_frames=30, n_ranges=N_RANGES):
    """Inclusive frame ranges (a, b) sampled for the estimate: ``n_ranges`` ranges of equal span from the first to the
    last frame, each at least 5 degrees (and ``n_sample_frames`` / 3 frames) wide, at most 1/n_ranges of the sweep."""
    n = last - first + 1
    span = max(int(round(n_sample_frames / 3)), int(np.ceil(5.0 / max(abs(osc), 1e-6))), 3)
    span = min(span, max(n // n_ranges, 3))
    steps = max(n_ranges - 1, 1)
    out = []
    for i in range(n_ranges):
        a = first + int(i * (n - span) / steps)        # truncation towards zero, as int() of the quotient
        out.append((a, min(a + span - 1, last)))
    return out


def rocking_resolved(fz, dz, z, iz0, fb, max_edge=0.5):
    """Rows whose per-frame net counts ``fz`` (n, >= max dz) fall off inside the box: the mean of the first and last
    recorded frames (frames with fz != 0) below ``max_edge`` times the frame of the predicted centre, which must be
    positive, and at least three frames between them.  Boxes shallower than three frames always pass."""
    fz = np.asarray(fz)
    n, width = fz.shape
    dz = np.minimum(np.asarray(dz, np.int64), width)
    frame = np.arange(width)[None, :]
    recorded = (fz != 0) & (frame < dz[:, None])
    seen = np.cumsum(recorded, axis=1)
    total = seen[:, -1]
    any_rec = total > 0
    first = np.where(any_rec, (seen == 0).sum(1), 0)                   # frames before the first recorded one
    last = np.where(any_rec, (seen < total[:, None]).sum(1), 0)        # frames before the last recorded one
    rows = np.arange(n)
    centre_frame = np.floor((np.asarray(z, float) - iz0) / np.maximum(fb, 1)).astype(np.int64)
    centre_frame = np.clip(centre_frame, 0, np.maximum(dz - 1, 0))
    centre = fz[rows, centre_frame]
    edge = 0.5 * (fz[rows, first] + fz[rows, last])
    falls_off = (centre > 0) & (edge < max_edge * centre) & (last - first >= 2)
    return (dz < 3) | falls_off


def fit_rocking_width(fz, iz0, phi, zeta, z_to_phi, sm_guess, lo_fac=0.2, hi_fac=3.0, n_grid=GRID_POINTS,
                      max_cells=4_000_000):
    """sigma_M (degrees) on a log grid of ``n_grid`` points from lo_fac to hi_fac times ``sm_guess``: the width whose
    frame-integrated Gaussian rocking curves R (normalised over each box) best fit the spots' per-frame fractions
    fz / sum(fz), weighted by each spot's total (least squares).  Rows: strong spots with fb = 1; phi in degrees;
    z_to_phi maps frame coordinates to degrees.  The first grid point of the smallest cost wins."""
    fz = np.asarray(fz, float)
    tot = fz.sum(1)
    use = tot > 0
    fz, tot = fz[use], tot[use]
    iz0 = np.asarray(iz0)[use]
    phi = np.asarray(phi, float)[use]
    zeta = np.abs(np.asarray(zeta, float)[use])
    grid = sm_guess * np.exp(np.linspace(np.log(lo_fac), np.log(hi_fac), n_grid))
    if not fz.size:
        return float(grid[0])                           # every cost is 0: the first grid point
    frac = fz / tot[:, None]
    k = iz0[:, None] + np.arange(fz.shape[1])[None, :]
    p0, p1 = z_to_phi(k.astype(float)), z_to_phi(k + 1.0)
    lo, hi = np.minimum(p0, p1) - phi[:, None], np.maximum(p0, p1) - phi[:, None]
    cost = np.empty(n_grid)
    step = max(1, max_cells // max(fz.size, 1))          # grid points evaluated together (bounded memory)
    for g0 in range(0, n_grid, step):
        scale = (grid[g0:g0 + step, None] / zeta[None, :] * np.sqrt(2.0))[:, :, None]   # (grid, rows, 1)
        R = 0.5 * (erf(hi[None] / scale) - erf(lo[None] / scale))
        R /= np.maximum(R.sum(2, keepdims=True), 1e-9)
        cost[g0:g0 + step] = np.sum(tot[None, :, None] * (frac[None] - R) ** 2, axis=(1, 2))
    cost[np.isnan(cost)] = np.inf                       # a NaN cost never wins
    if not np.isfinite(cost).any():
        return float(sm_guess)
    return float(grid[int(np.argmin(cost))])


def candidate_rows(r, a, b):
    """(candidates, I/sigma) of one sampled range's moments-only result ``r`` for frames a..b: spots centred inside
    the range, of the first lattice, without an overlapping partner, integrated with positive intensity, no overload,
    > 98% of the rocking curve recorded and a resolved rocking curve."""
    isum, vsum = r["isum"], r["vsum"]
    inside = (r["z"] >= a - 1) & (r["z"] < b) & (r.get("lattice", 0) == 0) & ~r["overlap"]
    ok = inside & np.isfinite(isum) & (isum > 0) & (r["n_over"] == 0) & (r["frac_rec"] > 0.98)
    if r.get("fz") is not None:
        ok &= rocking_resolved(r["fz"], r["dz"], r["z"], r["iz0"], r.get("fb", np.ones(len(isum), int)))
    return ok, isum / np.sqrt(np.maximum(vsum, 1e-6))


def strong_selection(samples):
    """Strong spots of each sampled range from (candidates, I/sigma) pairs: I/sigma > 20, or > 10 when fewer than 150
    pass; a range with fewer than 20 is left out (None).  When every range is left out but the ranges together hold
    at least 30 candidates above 5, every range uses I/sigma > 5."""
    picks = []
    for ok, isg in samples:
        sel = ok & (isg > 20)
        if sel.sum() < 150:
            sel = ok & (isg > 10)
        picks.append(sel if sel.sum() >= 20 else None)
    if all(p is None for p in picks):
        weak = [ok & (isg > 5.0) for ok, isg in samples]
        if sum(int(w.sum()) for w in weak) >= 30:
            picks = weak
    return picks


def width_terms(g, r, sel):
    """Per strong spot: (second moment m_d, pixel-footprint variance in eps units), the rocking moment m_m less the
    frame-width variance, and the rows usable for the rocking-curve fit (fb = 1) as (fz, iz0, phi, zeta_k)."""
    X = g.pixel_to_lab(r["x"][sel], r["y"][sel])
    L = np.linalg.norm(X, axis=1)
    cos_n = np.abs(X @ g.ed3) / L
    footprint = (g.qx / L) ** 2 * (1 + cos_n ** 2) / 24.0       # tangential q/|X|, radial q cos/|X|, over 2 axes
    if "pxs" in r:
        footprint = footprint * r["pxs"][sel] ** 2
    fb = r["fb"][sel] if "fb" in r else 1
    mm = r["m_m"][sel] - (np.radians(g.osc) * fb * r["zeta_k"][sel]) ** 2 / 12.0
    unbinned = sel & (r["fb"] == 1) if "fb" in r else sel
    fit = (r["fz"][unbinned], r["iz0"][unbinned], r["phi"][unbinned], r["zeta_k"][unbinned])
    return np.c_[r["m_d"][sel], footprint], mm, fit


def grows_with_box(new, cur, prev):
    """An estimate above 1.15 x the current widths after the current ones already grew 1.15 x over the previous:
    the spots' widths follow the box (contamination by neighbours), not the crystal."""
    grew_now = new.sigma_d > 1.15 * cur.sigma_d or new.sigma_m > 1.15 * cur.sigma_m
    grew_before = cur.sigma_d > 1.15 * prev.sigma_d or cur.sigma_m > 1.15 * prev.sigma_m
    return grew_now and grew_before


def estimate(run_range, g, model0, make_model, ranges, max_iter=None, guard=None, log=None):
    """The iteration of the estimate.  ``run_range(model, a, b)`` returns the moments-only result of frames a..b;
    ``make_model(sigma_d, sigma_m)`` builds the next model.  Returns the final model with ``.history``, a list of
    (sigma_D, sigma_M, spots) per iteration."""
    model = prev = model0
    history = []
    for it in range(max_iter or 3):
        results = [run_range(model, a, b) for a, b in ranges]
        samples = [candidate_rows(r, a, b) for r, (a, b) in zip(results, ranges)]
        picks = strong_selection(samples)
        md_parts, mm_parts, fit_parts = [], [], []
        for r, sel in zip(results, picks):
            if sel is None or not sel.any():
                continue
            md2, mm, fit = width_terms(g, r, sel)
            md_parts.append(md2)
            mm_parts.append(mm)
            fit_parts.append(fit)
        if not md_parts:
            break
        md2 = np.concatenate(md_parts)
        md = md2[:, 0] / TRUNC_T2 - md2[:, 1]
        sd_new = np.degrees(np.sqrt(max(np.median(md), 1e-12)))
        fz, iz0, phi, zk = (np.concatenate([f[i] for f in fit_parts]) for i in range(4))
        # the grid centre: the moment estimate, not below half the current width (a rocking curve much narrower than
        # a frame leaves almost no moment after the frame-width correction)
        guess = max(np.degrees(np.sqrt(max(np.median(np.concatenate(mm_parts)), 1e-12))), 0.5 * model.sigma_m)
        sm_new = fit_rocking_width(fz, iz0, phi, zk, g.z_to_phi, guess)
        new = make_model(float(np.clip(sd_new, 0.005, 1.0)), float(np.clip(sm_new, 0.005, 5.0)))
        history.append((float(new.sigma_d), float(new.sigma_m), int(md.size)))
        if log is not None:
            log(f"  INTEGRATE: estimated sigma_D {new.sigma_d:.4f} deg, sigma_M {new.sigma_m:.4f} deg "
                f"from {md.size} strong spots (iteration {it + 1})")
        if (guard is None or guard) and it >= 1 and grows_with_box(new, model, prev):
            if log is not None:
                log("  INTEGRATE: spot-width estimate keeps growing with the box (contamination): keeping iteration "
                    f"{it} values")
            break
        prev, model = model, new
        if max_iter and it >= 2 and abs(model.sigma_d / prev.sigma_d - 1) < 0.03 and \
                abs(model.sigma_m / prev.sigma_m - 1) < 0.03:
            break
    model.history = history
    return model

"""Oracle (test infrastructure only): float64 numpy restatement of what follows integration in the reference's
INTEGRATE stage -- the Lorentz-polarisation factor (Geometry.lp_factor, xpg/geometry.py:394-407), the sensor
efficiency and air absorption (Geometry.absorption_factors, xpg/geometry.py:372-392), the derived columns rlp, I,
sig, Isum, sigsum, peak, corr and the selection of finite reflections (xpg/stage_integrate.py:408-421) -- and of
SCALE's profile / summation combination (xpg/stage_scale.py:623-695, without the tNCS class correction).  Only the
tests use it, as the checker.

Pinning (tests/test_finalise_oracle.py): the Lorentz factor against sin 2 theta for beams in the plane normal to the
axis and 0 along it; the polarisation factor against (1 + cos^2 2 theta) / 2 for an unpolarised beam and against the
fully polarised limits; the absorption against closed forms at normal and oblique incidence (1 - exp(-mu t / cos),
exp(-mu_air D / cos)); mu_linear against the tabulated values at tabulated energies and the log-log midpoint between
them; the columns against hand-computed rows (NaN, negative variances, zero totals, coverage); the combination against
hand-computed Rmeas values and weights on a small set of symmetry mates, and its limits (I_mid = inf gives profile
fitting alone).
"""
from __future__ import annotations

import numpy as np


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def lp_factor(S, beam, axis, pol_normal, pol_fraction):
    """(zeta, P): zeta = (e x e0) . axis, P = f (1 - (e . emain)^2) + (1 - f) (1 - (e . n)^2) with n the polarisation
    normal made orthogonal to the beam and emain = n x e0 (xpg/geometry.py:394-407)."""
    S = np.asarray(S, np.float64)
    e = S / np.linalg.norm(S, axis=-1, keepdims=True)
    e0 = np.asarray(beam, np.float64)
    zeta = np.cross(e, e0) @ np.asarray(axis, np.float64)
    pn = np.asarray(pol_normal, np.float64)
    n = pn - np.dot(pn, e0) * e0
    n = n / np.linalg.norm(n)
    emain = np.cross(n, e0)
    f = float(pol_fraction)
    return zeta, f * (1 - (e @ emain) ** 2) + (1 - f) * (1 - (e @ n) ** 2)


def absorption_factors(S, ed3, distance, sensor_thickness, mu, mu_air, normals=None, lab=None):
    """Attenuation factor of recorded intensities (xpg/geometry.py:372-392): sensor efficiency
    1 - exp(-mu t / clip(cos, 1e-3, 1)) when thickness and mu are > 0, times air transmission exp(-mu_air path).
    Flat detector: cos = e . ED3, path = D / clip(cos); segmented: ``normals`` (n, 3) of the reflections' segments and
    the lab positions ``lab`` (n, 3) of their detector points (path = |lab|)."""
    S = np.asarray(S, np.float64)
    e = S / np.linalg.norm(S, axis=-1, keepdims=True)
    f = np.ones(e.shape[:-1])
    cosn = np.einsum("...k,...k->...", e, normals) if normals is not None else e @ np.asarray(ed3, np.float64)
    if sensor_thickness > 0 and mu > 0:
        f = f * -np.expm1(-mu * sensor_thickness / np.clip(cosn, 1e-3, 1.0))
    if lab is not None:
        path = np.linalg.norm(lab, axis=-1)
    else:
        path = distance / np.clip(e @ np.asarray(ed3, np.float64), 1e-3, 1.0)
    return f * np.exp(-mu_air * path)


def columns(S, iprf, vprf, isum, vsum, p_valid, p_total, corr, coverage, zeta, P, absorb):
    """rlp = |zeta| / P / absorption and the columns of xpg/stage_integrate.py:412-421; ``good`` the finite rows with
    sigma > 0."""
    rlp = np.abs(zeta) / P / absorb
    I = iprf * rlp
    sig = np.sqrt(np.maximum(vprf, 0)) * rlp
    out = dict(rlp=rlp, I=I, sig=sig, Isum=isum * rlp, sigsum=np.sqrt(np.maximum(vsum, 0)) * rlp,
               peak=100 * np.clip(p_valid / np.maximum(p_total, 1e-9), 0, 1) * (1.0 if coverage is None else coverage),
               corr=100 * np.nan_to_num(corr))
    out["good"] = np.isfinite(I) & np.isfinite(sig) & (sig > 0)
    return out


def rmeas(I, inv, n):
    """Rmeas of intensities grouped by ``inv`` (group sizes n): sum over groups with >= 2 members of
    sqrt(n / (n - 1)) |I - mean| over the sum of |I| of those members."""
    m = np.bincount(inv, I) / n
    mm = n[inv] >= 2
    dev = np.abs(I - m[inv]) * np.sqrt(n[inv] / np.maximum(n[inv] - 1, 1))
    return dev[mm].sum() / max(np.abs(I[mm]).sum(), 1e-30)


def combine_prf_sum(r, uid, nx, ny, block_frames=100):
    """Profile-fitted / summation combination (xpg/stage_scale.py:623-695 without tNCS): I = w I_prf + (1 - w) I_sum / f
    with w = 1 / (1 + max(I_sum / f, 0) / I_mid), f the recorded foreground fraction, I_mid chosen among inf and the
    50/75/90/97/99% quantiles of |I_sum / f| by the lowest Rmeas of symmetry mates within blocks of frames.
    Returns (I, sig, I_mid, Rmeas of profile fitting, Rmeas of the combination)."""
    rec = np.clip(r["peak"] / 100.0, 0.5, 1.0)
    frac = rec
    if "p_fg" in r and np.isfinite(r["p_fg"]).mean() >= 0.9:
        v = np.clip(np.nan_to_num(r["p_fg"], nan=1.0), 0.5, 1.0)
        reg = np.clip((r["x"] / nx * 3).astype(int), 0, 2) * 3 + np.clip((r["y"] / ny * 3).astype(int), 0, 2)
        zk = np.abs(r.get("zeta_k", np.ones(len(v))))
        zb = np.clip(np.searchsorted(np.quantile(zk, [0.2, 0.4, 0.6, 0.8]), zk), 0, 4)
        key_ = (reg * 5 + zb) * 100000 + (r["z"] // 100).astype(int)
        uk_, inv_ = np.unique(key_, return_inverse=True)
        med = np.array([np.median(v[inv_ == k]) for k in range(uk_.size)])
        frac = np.clip(rec * med[inv_], 0.5, 1.0)
    Ip, sp, Is_, ss_ = r["I"], r["sig"], r["Isum"] / frac, r["sigsum"] / frac
    okc = np.isfinite(Is_) & (ss_ > 0)
    if "sum_ok" in r:
        okc &= np.asarray(r["sum_ok"], bool)
    key = np.asarray(uid, np.int64) * 10000 + (r["z"] // block_frames).astype(np.int64)
    uk, inv = np.unique(key, return_inverse=True)
    n = np.bincount(inv)
    cands = [np.inf] + list(np.quantile(np.abs(Is_[okc]), [0.5, 0.75, 0.9, 0.97, 0.99]))
    strong = okc & (Is_ / np.maximum(ss_, 1e-9) > 20) & (r["peak"] > 95) & (Ip > 0) & (r["n_over"] == 0)
    if strong.sum() > 100:
        lr = float(np.median(np.log(np.maximum(Is_[strong], 1e-9) / Ip[strong])))
        if abs(np.expm1(lr)) > 0.1:
            cands = cands[1:]
    best = None
    for imid in cands:
        w = np.where(okc, 1.0 / (1.0 + np.maximum(Is_, 0) / imid), 1.0)
        rm = rmeas(w * Ip + (1 - w) * np.where(okc, Is_, 0.0), inv, n)
        if best is None or rm < best[0] - 1e-5:
            best = (rm, imid, w)
    rm, imid, w = best
    I = w * Ip + (1 - w) * np.where(okc, Is_, 0.0)
    sig = np.sqrt((w * sp) ** 2 + ((1 - w) * np.where(okc, ss_, 0.0)) ** 2)
    return I, sig, imid, rmeas(Ip, inv, n), rm

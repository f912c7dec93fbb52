"""CPU oracle of the per-block adaptive widths -- TEST INFRASTRUCTURE ONLY.

Restates _adapt_block (xpg/integrate.py:404-461) with its helpers _fit_rocking_width (:1130-1154), _nnls_gram
(:1095-1104) and rc_basis_np (xpg/integrate_kernels.py:144-151), in float64 numpy / scipy, on the arrays the device
version reads: the pass-1 results of one block.

Pinning: _fit_rocking_width recovers the width a noise-free erf rocking curve was generated with (grid resolution);
_nnls_gram equals scipy.optimize.nnls on the full design matrix; the whole chain is exercised against the device
kernels in tests/test_gpu_adapt.py.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

from .kernels import RC_C, RC_W, rc_basis, rc_default_weights

FG_R = 3.0


def fit_rocking_width(fz, iz0, phi, zeta, phi0_deg, osc_deg, sm_guess, lo_fac=0.2, hi_fac=3.0):
    zeta = np.abs(zeta)
    tot = fz.sum(1, keepdims=True)
    ok = tot[:, 0] > 0
    fz, iz0, phi, zeta, tot = fz[ok], iz0[ok], phi[ok], zeta[ok], tot[ok]
    frac = fz / tot
    k = iz0[:, None] + np.arange(fz.shape[1])[None]
    lo = phi0_deg + osc_deg * k.astype(float)
    hi = phi0_deg + osc_deg * (k + 1.0)
    lo, hi = np.minimum(lo, hi), np.maximum(lo, hi)
    best = (np.inf, sm_guess)
    for smv in sm_guess * np.exp(np.linspace(np.log(lo_fac), np.log(hi_fac), 61)):
        s = smv / zeta[:, None] * np.sqrt(2)
        R = 0.5 * (erf((hi - phi[:, None]) / s) - erf((lo - phi[:, None]) / s))
        R = R / np.maximum(R.sum(1, keepdims=True), 1e-9)
        cost = np.sum(tot * (frac - R) ** 2)
        if cost < best[0]:
            best = (cost, smv)
    return float(best[1])


def nnls_gram(A, y):
    from scipy.linalg import LinAlgError, cholesky, solve_triangular
    from scipy.optimize import nnls
    try:
        R = cholesky(A.T @ A)
        return nnls(R, solve_triangular(R.T, A.T @ y, lower=True))[0]
    except (LinAlgError, ValueError):
        return nnls(A, y)[0]


def adapt_block(isum, vsum, n_over, frac_rec, m_d, fz, phi_deg, zeta_k, iz0, fb, Xlen, cosa, sd, sm, phi0_deg,
                osc_deg, qx):
    """(n, sm_new, rcw, sd_new, sd_fit, sm_fit) for one block; n < 100: (n, None, ...)."""
    st = (isum > 0) & (isum / np.sqrt(np.maximum(vsum, 1e-6)) > 15) & (n_over == 0) & (frac_rec > 0.98) & (fb == 1)
    n = int(st.sum())
    if n < 100:
        return n, None, None, None, None, None
    F, PH, ZK, IZ, MD = fz[st].astype(np.float64), phi_deg[st], zeta_k[st], iz0[st], m_d[st]
    sm_fit = np.radians(fit_rocking_width(F, IZ, PH, ZK, phi0_deg, osc_deg, np.degrees(sm), 0.5, 2.0))
    sm_new = float(np.exp((n * np.log(sm_fit) + 300.0 * np.log(sm)) / (n + 300.0)))
    k = IZ[:, None] + np.arange(F.shape[1])[None]
    lo = phi0_deg + osc_deg * k.astype(float)
    hi = phi0_deg + osc_deg * (k + 1.0)
    lo, hi = np.radians(np.minimum(lo, hi)), np.radians(np.maximum(lo, hi))
    Bm = rc_basis(lo, hi, np.radians(PH), ZK, sm_new)
    tot = np.maximum(F.sum(1, keepdims=True), 1e-9)
    A = (Bm * np.sqrt(tot)[:, :, None]).reshape(-1, len(RC_C))
    y = (F / tot * np.sqrt(tot)).reshape(-1)
    w = nnls_gram(A, y)
    lam = 200.0 / n
    w = (w / max(w.sum(), 1e-9) + lam * rc_default_weights()) / (1 + lam)
    t2 = 1.0 - (FG_R ** 2 / 2) * np.exp(-FG_R ** 2 / 2) / (1 - np.exp(-FG_R ** 2 / 2))
    md = MD / t2 - (qx / Xlen[st]) ** 2 * (1 + cosa[st] ** 2) / 24.0
    sd_fit = float(np.clip(np.sqrt(max(np.median(md), 1e-12)), 0.5 * sd, 2.0 * sd))
    sd_new = float(np.exp((n * np.log(sd_fit) + 300.0 * np.log(sd)) / (n + 300.0)))
    return n, sm_new, w.astype(np.float32), sd_new, sd_fit, float(sm_fit)

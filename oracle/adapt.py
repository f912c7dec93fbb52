"""CPU oracle of the per-block adaptive widths -- TEST INFRASTRUCTURE ONLY.

Restates _adapt_block (xpg/integrate.py:404-461) with its helpers _fit_rocking_width (:1130-1154), _nnls_gram
(:1095-1104) and rc_basis_np (xpg/integrate_kernels.py:144-151), in float64 numpy / scipy, on the arrays the device
version reads: the pass-1 results of one block.

Pinning (tests/test_adapt_oracle.py, tests/test_oracle_pins.py): fit_rocking_width recovers the width a noise-free
erf rocking curve was generated with (grid resolution); nnls_gram equals scipy.optimize.nnls on the full design
matrix; the strong-row rule on rows just over and under each threshold; the 300-reflection shrinkage and the
lambda = 200/n weight mixing against hand-computed values; sd_from_moments recovers the width of spots whose second
moments are those of a Gaussian truncated at FG_R sigma, computed by quadrature (not from the closed form used here),
and clips at 0.5x / 2x; the whole chain is exercised against the device kernels in tests/test_gpu_adapt.py.
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


def strong_rows(isum, vsum, n_over, frac_rec, fb):
    """The rows the adaptive fits use: xpg/integrate.py:413-414 (I/sigma > 15, no overload, > 98% recorded, fb 1)."""
    return (isum > 0) & (isum / np.sqrt(np.maximum(vsum, 1e-6)) > 15) & (n_over == 0) & (frac_rec > 0.98) & (fb == 1)


def shrink(fit, run, n, prior=300.0):
    """A block's fitted width pulled towards the run's in log space with a prior worth ``prior`` reflections
    (xpg/integrate.py:436, :457)."""
    return float(np.exp((n * np.log(fit) + prior * np.log(run)) / (n + prior)))


def mix_weights(w, n):
    """NNLS weights normalised and mixed with the Gaussian weights at lambda = 200 / n (xpg/integrate.py:444-447)."""
    lam = 200.0 / n
    return (w / max(w.sum(), 1e-9) + lam * rc_default_weights()) / (1 + lam)


def sd_from_moments(MD, Xlen, cosa, qx, sd):
    """sigma_D of a block (radians): the median truncation-corrected second moment less the pixel footprint, clipped
    to 0.5-2x the run's sigma_D (xpg/integrate.py:448-456)."""
    t2 = 1.0 - (FG_R ** 2 / 2) * np.exp(-FG_R ** 2 / 2) / (1 - np.exp(-FG_R ** 2 / 2))
    md = MD / t2 - (qx / Xlen) ** 2 * (1 + cosa ** 2) / 24.0
    return float(np.clip(np.sqrt(max(np.median(md), 1e-12)), 0.5 * sd, 2.0 * sd))


def adapt_block(isum, vsum, n_over, frac_rec, m_d, fz, phi_deg, zeta_k, iz0, fb, Xlen, cosa, sd, sm, phi0_deg,
                osc_deg, qx):
    """(n, sm_new, rcw, sd_new, sd_fit, sm_fit) for one block; n < 100: (n, None, ...)."""
    st = strong_rows(isum, vsum, n_over, frac_rec, fb)
    n = int(st.sum())
    if n < 100:
        return n, None, None, None, None, None
    F, PH, ZK, IZ, MD = fz[st].astype(np.float64), phi_deg[st], zeta_k[st], iz0[st], m_d[st]
    sm_fit = np.radians(fit_rocking_width(F, IZ, PH, ZK, phi0_deg, osc_deg, np.degrees(sm), 0.5, 2.0))
    sm_new = shrink(sm_fit, sm, n)
    k = IZ[:, None] + np.arange(F.shape[1])[None]
    lo = phi0_deg + osc_deg * k.astype(float)
    hi = phi0_deg + osc_deg * (k + 1.0)
    lo, hi = np.radians(np.minimum(lo, hi)), np.radians(np.maximum(lo, hi))
    Bm = rc_basis(lo, hi, np.radians(PH), ZK, sm_new)
    tot = np.maximum(F.sum(1, keepdims=True), 1e-9)
    A = (Bm * np.sqrt(tot)[:, :, None]).reshape(-1, len(RC_C))
    y = (F / tot * np.sqrt(tot)).reshape(-1)
    w = mix_weights(nnls_gram(A, y), n)
    sd_fit = sd_from_moments(MD, Xlen[st], cosa[st], qx, sd)
    sd_new = shrink(sd_fit, sd, n)
    return n, sm_new, w.astype(np.float32), sd_new, sd_fit, float(sm_fit)

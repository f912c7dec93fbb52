"""Seeded per-block inputs of the adaptive-width fits (pass-1 results of one phi block)."""
import numpy as np
from scipy.special import erf


def make_block(n=800, seed=0, sm_true_deg=0.13, osc_deg=0.1, phi0_deg=0.0, weak_frac=0.3, md_levels=0):
    """md_levels > 0: the second moments take only that many distinct values (many equal sigma_D terms)."""
    rng = np.random.default_rng(seed)
    zeta_k = rng.choice([-1, 1], n) * rng.uniform(0.2, 1.0, n)
    z = rng.uniform(40, 200, n)
    phi_deg = phi0_deg + osc_deg * z
    Dz = 17
    iz0 = np.floor(z).astype(np.int32) - Dz // 2
    k = iz0[:, None] + np.arange(65)[None]
    lo, hi = phi0_deg + osc_deg * k, phi0_deg + osc_deg * (k + 1.0)
    s = sm_true_deg / np.abs(zeta_k)[:, None] * np.sqrt(2)
    R = 0.5 * (erf((hi - phi_deg[:, None]) / s) - erf((lo - phi_deg[:, None]) / s))
    I = np.exp(rng.uniform(np.log(200), np.log(2e5), n))
    fz = I[:, None] * R + rng.normal(0, 1, (n, 65)) * np.sqrt(I[:, None] * R + 20)
    fz[:, Dz:] = 0.0
    fz = fz.astype(np.float32)
    isum = fz.sum(1).astype(np.float64)
    vsum = isum + 400.0
    weak = rng.random(n) < weak_frac
    isum[weak] = rng.uniform(-5, 50, weak.sum())
    n_over = (rng.random(n) < 0.05).astype(np.int32)
    frac_rec = np.where(rng.random(n) < 0.05, 0.9, 0.999)
    fb = np.where(rng.random(n) < 0.03, 2, 1).astype(np.int32)
    m_d = np.abs(rng.normal(2.5e-7, 5e-8, n))
    if md_levels:
        m_d = rng.choice(np.abs(rng.normal(2.5e-7, 5e-8, md_levels)), n)
    Xlen = rng.uniform(150, 220, n)
    cosa = 150.0 / Xlen
    return dict(isum=isum, vsum=vsum, n_over=n_over, frac_rec=frac_rec, m_d=m_d, fz=fz, phi_deg=phi_deg,
                zeta_k=zeta_k, iz0=iz0, fb=fb, Xlen=Xlen, cosa=cosa, phi0_deg=phi0_deg, osc_deg=osc_deg, qx=0.075)

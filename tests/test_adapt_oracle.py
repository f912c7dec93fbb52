"""Pins of the adaptive-width oracle (oracle/adapt.py), CPU only."""
import numpy as np
from scipy.optimize import nnls

from adapt_inputs import make_block
from oracle import adapt as oa


def test_fit_rocking_width_recovers_grid_width():
    """Noise-free erf rocking curves generated at a grid width are fitted to that width exactly."""
    b = make_block(n=300, seed=1)
    guess = 0.1
    grid = guess * np.exp(np.linspace(np.log(0.5), np.log(2.0), 61))
    sm_true = grid[37]
    k = b["iz0"][:, None] + np.arange(65)[None]
    from scipy.special import erf
    lo, hi = b["osc_deg"] * k, b["osc_deg"] * (k + 1.0)
    s = sm_true / np.abs(b["zeta_k"])[:, None] * np.sqrt(2)
    fz = 1e4 * 0.5 * (erf((hi - b["phi_deg"][:, None]) / s) - erf((lo - b["phi_deg"][:, None]) / s))
    got = oa.fit_rocking_width(fz, b["iz0"], b["phi_deg"], b["zeta_k"], 0.0, b["osc_deg"], guess, 0.5, 2.0)
    assert got == sm_true


def test_nnls_gram_equals_full_nnls():
    rng = np.random.default_rng(2)
    A = np.abs(rng.normal(size=(400, 21)))
    y = A @ np.where(rng.random(21) < 0.5, 0, rng.random(21)) + rng.normal(0, 0.01, 400)
    np.testing.assert_allclose(oa.nnls_gram(A, y), nnls(A, y)[0], atol=1e-9)


def test_adapt_block_widths_move_towards_truth():
    b = make_block(n=900, seed=3, sm_true_deg=0.15)
    n, sm_new, w, sd_new, sd_fit, sm_fit = oa.adapt_block(
        b["isum"], b["vsum"], b["n_over"], b["frac_rec"], b["m_d"], b["fz"], b["phi_deg"], b["zeta_k"], b["iz0"],
        b["fb"], b["Xlen"], b["cosa"], np.radians(0.02), np.radians(0.1), b["phi0_deg"], b["osc_deg"], b["qx"])
    assert n >= 100
    assert abs(np.degrees(sm_fit) - 0.15) < 0.01            # the data's own width (grid spacing 2.3%)
    assert np.radians(0.1) < sm_new < sm_fit                 # shrunk towards the run's 0.1 deg
    assert abs(w.sum() - 1) < 1e-5 and (w >= 0).all()

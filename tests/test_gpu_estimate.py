"""estimate_sigmas on the device (xpg/integrate.py:1002-1092): from a wrong starting guess, the beam divergence and
mosaicity estimated from the strong spots' moments move well towards the values the frames were generated with --
the acceptance rule of the reference's own width test (tests/test_pipeline.py:207-208: within half the distance from
the default to the truth)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_estimate_sigmas_recovers_generated_widths():
    from xpg_amd import integrate as I, synth
    dev = torch.device("cuda:0")
    spec = synth.replace(synth.spec_small(), name="estimate", box_xy=0, box_z=0, sigma_d=0.02, sigma_m=0.1, seed=31)
    print("estimate_sigmas", spec.name, flush=True)
    frames, spots = synth.generate(spec, dev)
    exp = spots["exp"]
    guess = I.IntegrationModel(0.05, 0.25)
    model = I.estimate_sigmas(frames, exp, guess, spots["dmin"], (1, spec.nframes), verbose=False)
    print("sigma_D", model.sigma_d, "sigma_M", model.sigma_m, flush=True)
    assert abs(model.sigma_d - spec.sigma_d) < 0.5 * abs(guess.sigma_d - spec.sigma_d)
    assert abs(model.sigma_m - spec.sigma_m) < 0.5 * abs(guess.sigma_m - spec.sigma_m)
    assert len(model.history) >= 1

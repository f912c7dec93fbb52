"""Serving loop on the device: each dataset's results are the same alone, in a mix whose work overlaps on the device,
and under an HBM budget that makes requests wait for admission -- and the same as the offline sweep of the same
frames (bit for bit: the same kernels on the same values)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

KEYS = ("iprf", "vprf", "isum", "vsum", "bg", "corr", "n_over", "n_fg_masked")


def _requests():
    from xpg_amd import serve, synth
    out = []
    for i, (nfr, cell, shape) in enumerate([(73, (52.0, 61.0, 74.0), (1024, 1024)),
                                            (60, (45.0, 45.0, 90.0), (1024, 1024)),
                                            (91, (70.0, 58.0, 49.0), (1100, 1000))]):
        spec = synth.replace(synth.spec_full(), name=f"s{i}", ny=shape[0], nx=shape[1], nframes=nfr,
                             cell=cell + (90.0, 90.0, 90.0), space_group="P 21 21 21", block_frames=20,
                             seed=2000 + i)
        out.append(serve.Request(spec, 0.0, i))
    return out


def _same(ra, rb, tag):
    for k in KEYS:
        assert np.array_equal(ra.refl[k], rb.refl[k], equal_nan=True), (tag, k)
    assert ra.block_params == rb.block_params, tag


def test_results_same_alone_in_the_mix_and_offline():
    from xpg_amd import serve, synth
    from xpg_amd.integrate import Sweep
    dev = torch.device("cuda:0")
    plans = [serve.plan_request(r, batch_frames=10) for r in _requests()]
    print("rings", [p.ring for p in plans], "reflections", [p.hp["refl"]["x"].size for p in plans], flush=True)
    alone = {}
    for p in plans:
        rep = serve.Server(dev, slots=1).run([p], keep_results=True)
        alone[p.req.rid] = rep["results"][p.req.rid]
        assert rep["batches_with_results"] > 0 and rep["p50_ms"] >= 0
    # all at once: three slots, the work of the three datasets overlapping on the device
    rep = serve.Server(dev, slots=3).run(plans, keep_results=True)
    for rid, r in rep["results"].items():
        _same(alone[rid], r, ("mix", rid))
    # an HBM budget that admits one dataset at a time beside a second: requests wait for memory
    need = sorted(serve.dataset_bytes(p) for p in plans)
    rep = serve.Server(dev, slots=3, hbm_budget=need[-1] + need[0] + 1).run(plans, keep_results=True)
    assert rep["datasets"] == 3
    for rid, r in rep["results"].items():
        _same(alone[rid], r, ("budget", rid))
    # the offline sweep of the same frames (the whole dataset resident, batched form)
    p = plans[0]
    spec = p.req.spec
    frames, _ = synth.generate(spec, dev)
    opts = dict(frames=(1, spec.nframes), mask=None, exclude_frames=(), learn_profiles=True, adaptive=True,
                bkg_template=None, moments_only=False, nreg=9, sd=np.radians(p.model.sigma_d),
                sm=np.radians(p.model.sigma_m))
    sw = Sweep(frames, [p.exp] * (len(p.edges) - 1), p.model, p.hp, opts)
    off = sw.finish(sw.run())
    _same(off, alone[0], "offline")
    ok = np.isfinite(off.refl["iprf"])
    assert ok.sum() > 0.8 * ok.size

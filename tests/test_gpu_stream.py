"""Offline sweeps larger than their HBM budget: integrate() on a host ImageSet streams the frames through a ring of
frames on the device (xpg_amd/passes.py), uploading batches through two pinned buffers and running each block as
its last frame arrives.  Under a budget that forces the ring the results are bit-identical to the resident sweep,
and both equal the device frame store's (raw counts coded in the kernels)."""
import numpy as np
import pytest
import torch

from oracle import extract as ox

pytestmark = pytest.mark.gpu

KEYS = ("iprf", "vprf", "isum", "vsum", "bg", "corr", "frac_rec", "n_over", "n_fg_masked")


class HostFrames:
    """An ImageSet-like host source: value-coded float32 frames (xpg/io/images.py:80-96) served by iter_frames."""

    def __init__(self, raw, flag_value, overload):
        self.raw, self.flag_value, self.overload = raw, flag_value, overload
        self.first, self.last = 1, raw.shape[0]
        self.ny, self.nx = raw.shape[1], raw.shape[2]
        self.served = 0

    def iter_frames(self, frames, copy=True):
        for f in frames:
            self.served += 1
            yield f, ox.frame_codes(self.raw[f - 1], flag_value=self.flag_value, overload=self.overload)


def test_streamed_sweep_equals_resident_sweep():
    from xpg_amd import integrate as I, passes, synth
    spec = synth.replace(synth.spec_small(), seed=47)
    print("streamed sweep", spec.name, flush=True)
    dev = torch.device("cuda:0")
    frames, spots = synth.generate(spec, dev)
    exp = spots["exp"]
    edges = synth.block_edges(spec)
    model = I.IntegrationModel(spec.sigma_d, spec.sigma_m, box_xy=spec.box_xy, box_z=spec.box_z)
    blocks = [exp] * (len(edges) - 1)
    host = HostFrames(frames.data.cpu().numpy().view(np.uint32), frames.flag_value, frames.overload)
    hp = I._integrate(host, blocks, edges, model, spots["dmin"], plan_only=True, verbose=False)
    mp = passes.plan_memory(hp, I.sweep_plan(hp, blocks), (1, spec.nframes), spec.ny, spec.nx, 1 << 50)
    assert not mp.streamed and mp.ring < spec.nframes and mp.streamed_bytes < mp.resident_bytes
    budget = (mp.streamed_bytes + mp.resident_bytes) // 2
    print(f"ring {mp.ring} frames: {mp.streamed_bytes / 1e6:.0f} MB streamed, {mp.resident_bytes / 1e6:.0f} MB "
          f"resident, budget {budget / 1e6:.0f} MB", flush=True)
    res = I.integrate(host, blocks, edges, model, spots["dmin"], verbose=False, hbm_budget=1 << 50)
    n_res = host.served
    host.served = 0
    tm = {}
    res_s = I.integrate(host, blocks, edges, model, spots["dmin"], verbose=False, hbm_budget=budget, timing=tm)
    assert host.served == spec.nframes == n_res           # every frame read once either way
    res_d = I.integrate(frames, blocks, edges, model, spots["dmin"], verbose=False)
    ok = np.isfinite(res.refl["iprf"])
    assert ok.sum() > 0.9 * ok.size
    for k in KEYS:
        assert np.array_equal(res_s.refl[k], res.refl[k], equal_nan=True), k
        assert np.array_equal(res_d.refl[k], res.refl[k], equal_nan=True), k
    assert res_s.block_params == res.block_params == res_d.block_params
    with pytest.raises(Exception, match="budget"):
        I.integrate(host, blocks, edges, model, spots["dmin"], verbose=False, hbm_budget=mp.streamed_bytes // 4)

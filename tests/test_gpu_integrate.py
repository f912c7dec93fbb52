"""End-to-end: xpg_amd.integrate on the device against the float64 oracle sweep (oracle/sweep.py) on the same
synthetic frames -- extraction, pass 1, strong-spot profile learning, the device adaptive-width chain, profile
normalisation and pass 2 -- plus determinism of repeated runs."""
import numpy as np
import pytest
import torch

from oracle import extract as ox
from oracle import sweep as osw

pytestmark = pytest.mark.gpu


def _dataset(spec):
    from xpg_amd import integrate as I, synth
    from xpg_amd.geometry import detector_consts
    dev = torch.device("cuda:0")
    frames, spots = synth.generate(spec, dev)
    exp = spots["exp"]
    edges = synth.block_edges(spec)
    model = I.IntegrationModel(spec.sigma_d, spec.sigma_m, box_xy=spec.box_xy, box_z=spec.box_z)
    blocks = [exp] * (len(edges) - 1)
    return frames, exp, edges, model, blocks, spots["dmin"], detector_consts(exp.geom)


def _coded(frames):
    raw = frames.data.cpu().numpy().view(np.uint32)
    return {i: ox.frame_codes(raw[i], flag_value=frames.flag_value, overload=frames.overload) for i in range(len(raw))}


@pytest.mark.parametrize("which", ["small", "medium", "uneven"])
def test_integrate_matches_oracle_sweep(which):
    """small: one fixed shoebox shape; medium: the reference's box rules (several depths and sides, phi-binned wide
    reflections); uneven: drawn odd sides 5-21 and depths 3-15 with sigma_M per detector region, so the batched
    form's size-bin launches mix shapes that the plain form launches one shape at a time."""
    from xpg_amd import integrate as I, synth
    if which == "small":
        spec = synth.spec_small()
    else:
        spec = synth.SynthSpec(name="medium", nx=1536, ny=1536, nframes=40, distance=120.0,
                               cell=(45.0, 50.0, 60.0, 90.0, 90.0, 90.0), space_group="P 21 21 21", bg_centre=3.0,
                               bg_edge=1.0, dead_frac=0.01, block_frames=20, seed=77, mean_i0=5e4)
        if which == "uneven":
            spec = synth.replace(spec, name="uneven", sigma_m_range=(0.05, 0.8), box_draw=((5, 21), (3, 15)), seed=78)
    draw = synth.box_draw_for(spec)
    print(f"integrate {spec.name}", flush=True)
    frames, exp, edges, model, blocks, dmin, det = _dataset(spec)
    res = I.integrate(frames, blocks, edges, model, dmin, verbose=False, _box_draw=draw)
    hp = I._integrate(frames, blocks, edges, model, dmin, plan_only=True, verbose=False, _box_draw=draw)
    coded = _coded(frames)
    out, bp = osw.integrate_sweep(coded, list(range(spec.nframes)), hp, blocks, model, det)
    r = res.refl
    ok = np.isfinite(out["iprf"])
    assert ok.sum() > 0.9 * ok.size and np.array_equal(np.isfinite(r["iprf"]), ok)
    assert np.array_equal(r["n_over"], out["n_over"])
    sig = np.sqrt(out["vprf"][ok])
    dI = np.abs(r["iprf"][ok] - out["iprf"][ok]) / sig
    dS = np.abs(r["isum"][ok] - out["isum"][ok]) / np.sqrt(np.maximum(out["vsum"][ok], 1e-6))
    print(f"{spec.name}: {ok.sum()} reflections, max |dI|/sigma prf {dI.max():.2e} sum {dS.max():.2e}, "
          f"blocks {bp}", flush=True)
    assert dI.max() < 1e-3 and dS.max() < 1e-3
    np.testing.assert_allclose(r["bg"][ok], out["bg"][ok], rtol=1e-6, atol=1e-9)
    # adaptive widths used by the blocks agree with the oracle's
    for b, rec in bp.items():
        assert res.block_params[b]["n"] == rec["n"]
        if rec["n"] >= 100:
            assert abs(res.block_params[b]["sigma_m"] - rec["sigma_m"]) < 1e-9 * rec["sigma_m"]
    # repeated runs are bit-identical (fixed-point profile grids, fixed reduction orders)
    res2 = I.integrate(frames, blocks, edges, model, dmin, verbose=False, _box_draw=draw)
    assert np.array_equal(res2.refl["iprf"], r["iprf"], equal_nan=True)
    # the plain form (one stream, one launch per shoebox shape, voxels read from the frame store) gives the same
    # results as the batched default (size-bin launches over prepared data)
    hp2 = I._integrate(frames, blocks, edges, model, dmin, plan_only=True, verbose=False, _box_draw=draw)
    sw = I.Sweep(frames, blocks, model, hp2, dict(frames=(1, spec.nframes), mask=None, exclude_frames=(),
                                                  learn_profiles=True, adaptive=True, bkg_template=None,
                                                  moments_only=False, nreg=9, sd=np.radians(model.sigma_d),
                                                  sm=np.radians(model.sigma_m), form="plain"))
    if which == "uneven":
        assert len(sw.sp.bins) < len(sw.sp.groups)         # size bins mix shapes in one launch
    res3 = sw.finish(sw.run())
    for k in ("iprf", "vprf", "isum", "bg"):
        assert np.array_equal(res3.refl[k], r[k], equal_nan=True), k
    # the batched form captured in a HIP graph: every replay equals the eager result, with eager runs and
    # synchronisations between the replays
    sw.o["form"] = "batched"
    sw.capture()
    for rep in range(3):
        rr = sw.finish(sw.replay())
        for k in ("iprf", "vprf", "isum", "bg", "corr"):
            assert np.array_equal(rr.refl[k], r[k], equal_nan=True), (rep, k)
        assert rr.block_params == res.block_params, rep
        sw.run()
        torch.cuda.synchronize()


def test_integrate_options_match_oracle_sweep():
    """integrate() with the integrator's own pixel mask, excluded frames, fixed widths (adaptive=False) and the
    Gaussian prior profiles (learn_profiles=False) against the oracle sweep given the same options."""
    from xpg_amd import integrate as I, synth
    spec = synth.replace(synth.spec_small(), seed=41)
    print("integrate options", spec.name, flush=True)
    frames, exp, edges, model, blocks, dmin, det = _dataset(spec)
    rng = np.random.default_rng(5)
    mask = rng.random((spec.ny, spec.nx)) > 0.02          # True = trusted
    excl = (17, 18, 61)
    kw = dict(mask=mask, exclude_frames=excl, adaptive=False, learn_profiles=False)
    res = I.integrate(frames, blocks, edges, model, dmin, verbose=False, **kw)
    hp = I._integrate(frames, blocks, edges, model, dmin, plan_only=True, verbose=False, **kw)
    coded = _coded(frames)
    read = [f for f in range(spec.nframes) if f + 1 not in excl]
    out, _ = osw.integrate_sweep(coded, read, hp, blocks, model, det, adaptive=False, learn_profiles=False, mask=mask)
    r = res.refl
    ok = np.isfinite(out["iprf"])
    assert ok.sum() > 0.8 * ok.size and np.array_equal(np.isfinite(r["iprf"]), ok)
    assert np.array_equal(r["n_over"], out["n_over"])
    dI = np.abs(r["iprf"][ok] - out["iprf"][ok]) / np.sqrt(out["vprf"][ok])
    dS = np.abs(r["isum"][ok] - out["isum"][ok]) / np.sqrt(np.maximum(out["vsum"][ok], 1e-6))
    print(f"options: {ok.sum()} reflections, max |dI|/sigma prf {dI.max():.2e} sum {dS.max():.2e}", flush=True)
    assert dI.max() < 1e-3 and dS.max() < 1e-3
    np.testing.assert_allclose(r["frac_rec"][ok], out["frac_rec"][ok], rtol=1e-9, atol=1e-12)
    assert res.block_params == {}


def test_integrate_lattices_foreign_spots_template_match_oracle():
    """A second lattice integrated simultaneously (its reflections mask or share boxes with the first), foreign spots
    masked in the boxes, and a background template, against the oracle sweep on the same plan."""
    from xpg_amd import integrate as I, synth
    from xpg_amd.geometry import detector_consts
    spec = synth.replace(synth.spec_small(), seed=43)
    print("integrate lattices/foreign/template", spec.name, flush=True)
    frames, exp, edges, model, blocks, dmin, _ = _dataset(spec)
    exp2 = synth.make_experiment(synth.replace(spec, seed=431))           # same cell, another orientation
    rng = np.random.default_rng(9)
    foreign = (rng.uniform(1, spec.nx, 300), rng.uniform(1, spec.ny, 300), rng.uniform(0, spec.nframes, 300))
    yy, xx = np.mgrid[0:spec.ny, 0:spec.nx]
    template = (0.6 + 0.4 * np.hypot(xx - spec.nx / 2, yy - spec.ny / 2) / (0.7 * spec.nx)).astype(np.float32)
    lat2 = [exp2] * len(blocks)
    kw = dict(extra_lattices=[lat2], foreign_spots=foreign, bkg_template=template)
    res = I.integrate(frames, blocks, edges, model, dmin, verbose=False, **kw)
    hp = I._integrate(frames, blocks, edges, model, dmin, plan_only=True, verbose=False, **kw)
    coded = _coded(frames)
    det = detector_consts(exp.geom, template)
    out, _ = osw.integrate_sweep(coded, list(range(spec.nframes)), hp, list(blocks) + lat2, model, det)
    r = res.refl
    assert (r["lattice"] == 1).sum() > 0.3 * r["lattice"].size
    ok = np.isfinite(out["iprf"])
    assert ok.sum() > 0.8 * ok.size and np.array_equal(np.isfinite(r["iprf"]), ok)
    assert np.array_equal(r["n_over"], out["n_over"])
    dI = np.abs(r["iprf"][ok] - out["iprf"][ok]) / np.sqrt(out["vprf"][ok])
    print(f"lattices: {ok.sum()} reflections, max |dI|/sigma prf {dI.max():.2e}", flush=True)
    assert dI.max() < 1e-3
    np.testing.assert_allclose(r["bg"][ok], out["bg"][ok], rtol=1e-6, atol=1e-9)


def test_results_copied_to_pinned_host_memory_per_block():
    """Sweep.enqueue with host buffers: every block's rows reach pinned host memory behind its pass 2 (a copy stream
    overlapping later blocks), eager and captured in the HIP graph, equal to the device results."""
    from xpg_amd import integrate as I, synth
    spec = synth.replace(synth.spec_small(), seed=45)
    print("host copies", spec.name, flush=True)
    frames, exp, edges, model, blocks, dmin, det = _dataset(spec)
    hp = I._integrate(frames, blocks, edges, model, dmin, plan_only=True, verbose=False)
    sw = I.Sweep(frames, blocks, model, hp, dict(frames=(1, spec.nframes), mask=None, exclude_frames=(),
                                                 learn_profiles=True, adaptive=True, bkg_template=None,
                                                 moments_only=False, nreg=9, sd=np.radians(model.sigma_d),
                                                 sm=np.radians(model.sigma_m)))
    assert len(sw.order) >= 2
    host = sw.host_results()
    for t in host.values():
        t.fill_(-7)
    run = sw.enqueue(None, host)
    torch.cuda.synchronize()
    src = {**run["o1"], **run["o2"]}
    for k in sw.RESULT_KEYS:
        assert torch.equal(torch.nan_to_num(host[k], nan=-1.0), torch.nan_to_num(src[k].cpu(), nan=-1.0)), k
    out = sw.capture(host)
    for rep in range(2):
        for t in host.values():
            t.fill_(-7)
        sw.graph.replay()
        torch.cuda.synchronize()
        src = {**out["o1"], **out["o2"]}
        for k in sw.RESULT_KEYS:
            assert torch.equal(torch.nan_to_num(host[k], nan=-1.0), torch.nan_to_num(src[k].cpu(), nan=-1.0)), (rep, k)
        sw.run()
        torch.cuda.synchronize()

"""The refinement's device kernels (xpg_rf_residuals, xpg_rf_normal) and the device-backed refinement, per-block
models and smoothing against the float64 oracle (oracle/refine.py), and the reference's own recovery tests run on the
device (tests/test_core.py:100-114, tests/test_features.py:76-100 with all 200 frames)."""
import ctypes as C
from dataclasses import replace

import numpy as np
import pytest
import torch

from oracle import refine as orf
from xpg_amd import _lib
from xpg_amd import blocks as xb
from xpg_amd import refine as xr
from xpg_amd.geometry import Experiment, predict_reflections
from xpg_amd.models import Crystal

from test_refine_host import close
from test_refine_oracle import A, FLAGS, _geom, _modules, _noisy, drifting_sweep

pytestmark = pytest.mark.gpu


def _case(kind):
    kw = dict(sensor_thickness=0.45, mu=3.7)
    if kind == "segmented":
        kw["segments"] = _modules(False)
        ny, nx = 2162, 2068
        yy, xx = np.mgrid[0:ny, 0:nx]
        kw["geo_corr"] = (0.3 * np.sin(xx / 300.0), -0.2 * np.cos(yy / 250.0))
        kw.update(nx=nx, ny=ny)
    g = _geom(**kw)
    pr = predict_reflections(g, A, 2.3, z_range=(0, 30))
    x, y, z = _noisy(pr, 4)
    exp = Experiment(g, Crystal.from_Astar(A, "m"))
    flags = FLAGS + ["AXIS"] + (["SEGMENT"] if kind == "segmented" else [])
    return exp, pr["hkl"], x, y, z, flags


@pytest.mark.parametrize("kind", ["flat", "segmented"])
def test_residuals_and_normal_equations_match_the_oracle(kind):
    exp, hkl, x, y, z, flags = _case(kind)
    print(f"refine kernels {kind}: {len(x)} spots", flush=True)
    data, valid = orf.spot_data(exp, hkl, x, y, z, 0.5, 0.25)
    opar, par = orf.Param(exp, flags), xr.Parameterisation(exp, flags)
    X = par.x0 + np.random.default_rng(6).normal(scale=0.3, size=(4, par.x0.size))
    spots, setup, valid2 = xr._spot_data(exp, hkl, x, y, z, 0.5, 0.25)
    assert np.array_equal(valid, valid2)
    # residuals of four parameter vectors, two problems sharing one launch (the second: every other spot)
    sub = np.arange(0, len(x), 2)
    p1 = xr._Problem(par, spots, setup)
    p2 = xr._Problem(par, {k: v[sub] for k, v in spots.items()}, setup)
    reqs = [xr._Res(p1, X), xr._Res(p2, X[:2])]
    out = xr._drive_once(reqs)
    want = orf.residuals(opar, X, data)
    ne = par.n_extra
    assert np.abs(out[0][:, :3 * len(x)] - want).max() < 1e-8
    sub3 = np.concatenate([sub, sub + len(x), sub + 2 * len(x)])
    assert np.abs(out[1][:, :3 * sub.size] - want[:2][:, sub3]).max() < 1e-8
    if ne:
        np.testing.assert_allclose(out[0][:, 3 * len(x):], np.array([opar.gauge(v) for v in X]), atol=1e-12)
    # the normal equations at one vector with Cauchy-like weights
    w1 = np.random.default_rng(7).uniform(0.1, 1.0, len(x))
    (Ad, gd, r0), = xr._drive_once([xr._Normal(p1, X[0], w1)])
    assert np.array_equal(r0[0, :3 * len(x)], out[0][0, :3 * len(x)])       # the same kernel, the same vector
    Ao, go = orf.normal(opar, X[0], data, np.concatenate([np.tile(w1, 3), np.ones(ne)]))
    assert np.abs(Ad - Ao).max() < 1e-6 * np.abs(Ao).max()
    assert np.abs(gd - go).max() < 1e-6 * np.abs(go).max()
    assert np.array_equal(Ad, Ad.T)


def test_refinement_recovers_geometry_and_matches_the_oracle():
    g = _geom(sensor_thickness=0.45, mu=3.7)
    pr = predict_reflections(g, A, 2.0, z_range=(0, 80))
    x, y, z = _noisy(pr, 3)
    start = Experiment(replace(g, orgx=513.5, orgy=499.9, distance=151.0),
                       Crystal.from_Astar(A @ np.diag([1.004, 0.998, 1.0]), "m"))
    print(f"refine recovery: {len(x)} spots", flush=True)
    got = xr.refine(start, pr["hkl"], x, y, z, FLAGS)
    assert abs(got.exp.geom.orgx - 512.3) < 0.05 and abs(got.exp.geom.distance - 150.0) < 0.1
    assert np.allclose(got.exp.crystal.cell, [50, 60, 70, 90, 95, 90], atol=0.02) and got.rmsd_xy < 0.2
    close(got, orf.refine(start, pr["hkl"], x, y, z, FLAGS), (pr["hkl"], x, y, z))


def test_segment_refinement_recovers_module_offset_on_all_frames():
    gt = _geom(segments=_modules(False))
    pr = predict_reflections(gt, A, 1.8, z_range=(0, 200))
    x, y, z = _noisy(pr, 1)
    st = Experiment(_geom(segments=_modules()), Crystal.from_Astar(A, "m"))
    print(f"segment refinement: {len(x)} spots", flush=True)
    r = xr.refine(st, pr["hkl"], x, y, z, FLAGS + ["SEGMENT"], max_err_xy=10)
    assert r.rmsd_xy < 0.16
    yy, xx = np.mgrid[20:2150:40, 20:2060:40]
    xx, yy = xx.ravel() + 0.0, yy.ravel() + 0.0
    m = gt.segment_of(xx, yy) >= 0
    xr_, yr_ = r.exp.geom.s_to_pixel(gt.pixel_to_s(xx[m], yy[m]))
    assert np.nanmax(np.hypot(xr_ - xx[m], yr_ - yy[m])) < 0.05


def test_segment_refinement_matches_the_oracle():
    gt = _geom(segments=_modules(False), sensor_thickness=0.32, mu=2.1)
    pr = predict_reflections(gt, A, 2.4, z_range=(0, 40))
    x, y, z = _noisy(pr, 11)
    st = Experiment(_geom(segments=_modules(), sensor_thickness=0.32, mu=2.1), Crystal.from_Astar(A, "m"))
    flags = FLAGS + ["SEGMENT"]
    close(xr.refine(st, pr["hkl"], x, y, z, flags, max_err_xy=10),
          orf.refine(st, pr["hkl"], x, y, z, flags, max_err_xy=10), (pr["hkl"], x, y, z), params=False)


def test_refine_many_equals_one_at_a_time():
    g = _geom()
    probs = []
    for s in range(4):
        pr = predict_reflections(g, A, 2.6, z_range=(10 * s, 10 * s + 20))
        x, y, z = _noisy(pr, 30 + s)
        st = Experiment(replace(g, orgx=512.3 + 0.4 * s), Crystal.from_Astar(A, "m"))
        probs.append(((st, pr["hkl"], x, y, z, ["POSITION", "BEAM", "ORIENTATION"]),
                      dict(prior={"orgx": (512.3, 0.5)})))
    many = xr.refine_many(probs)
    for (a, k), m in zip(probs, many):
        one = xr.refine(*a, **k)
        assert np.array_equal(one.x, m.x) and np.array_equal(one.inliers, m.inliers) and one.rmsd_xy == m.rmsd_xy


def test_block_models_and_smoothing_match_the_oracle():
    exp, spots, rng_, delphi, Us = drifting_sweep(n_blocks=8)
    print(f"block models: {len(spots[0])} spots, 8 blocks", flush=True)
    s1, s2 = {}, {}
    got, e1 = xb.block_models(exp, spots, rng_, delphi, ["POSITION", "BEAM", "ORIENTATION"], state=s1)
    want, e2 = orf.block_models(exp, spots, rng_, delphi, ["POSITION", "BEAM", "ORIENTATION"], state=s2)
    assert np.array_equal(e1, e2) and s1["seed"] == s2["seed"] and s1["n_fallback"] == s2["n_fallback"] == 0
    np.testing.assert_allclose(s1["pos_prior_sd"], s2["pos_prior_sd"], atol=1e-5)
    for a, b, U in zip(got, want, Us):
        np.testing.assert_allclose(a.crystal.U, b.crystal.U, atol=1e-8)
        assert a.geom.orgx == pytest.approx(b.geom.orgx, abs=1e-5)
        assert np.degrees(np.arccos(np.clip((np.trace(a.crystal.U @ U.T) - 1) / 2, -1, 1))) < 0.01
    for a, b in zip(xb.smooth_blocks(got, exp), orf.smooth_blocks(want, exp)):
        np.testing.assert_allclose(a.crystal.U, b.crystal.U, atol=1e-8)
        np.testing.assert_allclose(a.geom.beam, b.geom.beam, atol=1e-8)


def test_entry_points_reject_bad_arguments():
    h = _lib.lib()
    st, su = _lib.RfSpots(), _lib.RfSetup()
    su.qx = su.qy = su.sxy = su.sphi = su.mu = 1.0
    su.n_seg, su.n_models = 1, 3
    assert h.xpg_rf_residuals(None, C.byref(su), None, 0, 0, None, None, None) != 0
    st.n = 5                                           # spots without arrays
    assert h.xpg_rf_residuals(C.byref(st), C.byref(su), None, 0, 0, None, None, None) != 0
    st.n = 0
    su.sxy = 0.0
    assert h.xpg_rf_residuals(C.byref(st), C.byref(su), None, 0, 0, None, None, None) != 0
    su.sxy = 1.0
    assert h.xpg_rf_normal(C.byref(st), C.byref(su), None, 1, None, None, None, None, None, None) != 0
    assert b"job table" in h.xpg_last_error()
    su.n_models = 1
    assert h.xpg_rf_normal(C.byref(st), C.byref(su), None, 0, None, None, None, None, None, None) != 0
    assert b"parameter" in h.xpg_last_error()
    su.n_models = 3
    assert h.xpg_rf_residuals(C.byref(st), C.byref(su), None, 0, 0, None, None, None) == 0   # nothing to do


def test_centroid_refinement_matches_the_oracle_per_block():
    from test_refine_host import _centroid_table
    exp, spots, rng_, delphi, Us = drifting_sweep(n_blocks=6)
    edges = xb.block_edges(exp.geom, rng_, delphi)
    refl = _centroid_table(exp, spots, Us, edges)
    prior = xb.pos_prior(["POSITION"], exp.geom)
    print(f"centroid refinement: {len(refl['xo'])} rows, 6 blocks", flush=True)
    got = xb.refine_blocks_on_centroids([exp] * 6, edges, refl, ["POSITION", "BEAM", "ORIENTATION"], prior=prior)
    ok = (refl["isum"] / np.sqrt(refl["vsum"]) > 10) & (refl["peak"] > 90) & (refl["n_over"] == 0)
    for i in range(6):
        m = ok & (refl["zo"] >= edges[i]) & (refl["zo"] < edges[i + 1])
        want = orf.refine(exp, refl["hkl"][m], refl["xo"][m], refl["yo"][m], refl["zo"][m],
                          ["POSITION", "BEAM", "ORIENTATION"], max_iter=15, n_reject_cycles=2, prior=prior)
        np.testing.assert_allclose(got[i].crystal.U, want.exp.crystal.U, atol=1e-8)
        assert got[i].geom.orgx == pytest.approx(want.exp.geom.orgx, abs=1e-5)

"""The refinement's host side (xpg_amd/refine.py, xpg_amd/blocks.py) on the CPU: the parameterisation's packed device
models, the job tables, the Levenberg-Marquardt generators and their driver, with the two device entry points
replaced by a numpy restatement of their kernels that reads and writes the same buffers.  Checked against the oracle
(oracle/refine.py); the device kernels themselves are checked in tests/test_gpu_refine.py."""
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
from xpg_amd.models import Crystal, b_matrix

from test_refine_oracle import A, FLAGS, _geom, _modules, _noisy, drifting_sweep


def _arr(ptr, n, ct=C.c_double):
    return np.ctypeslib.as_array((ct * int(n)).from_address(ptr.value if hasattr(ptr, "value") else ptr))


class FakeLib:
    """xpg_rf_residuals and xpg_rf_normal of include/xpg_refine.h over host memory, in numpy."""

    def __init__(self):
        self.calls = {"res": 0, "normal": 0}

    @staticmethod
    def _spots(st):
        n = st.n
        return dict(h=_arr(st.h, 3 * n).reshape(n, 3), phi=_arr(st.phi, n), branch=_arr(st.branch, n),
                    xo=_arr(st.xo, n), yo=_arr(st.yo, n), seg=_arr(st.seg, n, C.c_int32))

    @staticmethod
    def _jobs(ptr, n):
        return np.frombuffer(C.string_at(ptr.value, n * xr._JOB.itemsize), xr._JOB)

    def xpg_rf_residuals(self, stp, sup, jobs, n_jobs, total, models, r, stream):
        self.calls["res"] += 1
        st, su = stp._obj, sup._obj
        sp, jb = self._spots(st), self._jobs(jobs, n_jobs)
        stride = _lib.RF_HEAD + _lib.RF_SEG * su.n_seg
        md = _arr(models, (jb["model0"] + jb["n_models"]).max() * stride).reshape(-1, stride)
        out = _arr(r, (jb["r0"] + 3 * jb["n_models"].astype(np.int64) * jb["n_spots"]).max())
        assert int((jb["n_models"].astype(np.int64) * jb["n_spots"]).sum()) == total
        for j in jb:
            i = slice(int(j["spot0"]), int(j["spot0"] + j["n_spots"]))
            for m in range(int(j["n_models"])):
                mm = md[j["model0"] + m]
                s0, ax, Am = mm[0:3], mm[3:6], mm[6:15].reshape(3, 3)
                p0 = sp["h"][i] @ Am.T
                par = (p0 @ ax)[:, None] * ax
                perp, mxp = p0 - par, np.cross(ax, p0)
                a, b = perp @ s0, mxp @ s0
                c = -0.5 * np.sum(p0 * p0, 1) - par @ s0
                ratio = np.clip(c / np.sqrt(a * a + b * b), -1 + 1e-9, 1 - 1e-9)
                phi = np.arctan2(b, a) + sp["branch"][i] * np.arccos(ratio)
                phi = phi + 2 * np.pi * np.rint((sp["phi"][i] - phi) / (2 * np.pi))
                S = par + np.cos(phi)[:, None] * perp + np.sin(phi)[:, None] * mxp + s0
                seg = mm[17:].reshape(-1, 3, 3)[sp["seg"][i]]
                P0, Bx, By = seg[:, 0], seg[:, 1], seg[:, 2]
                nrm = np.cross(Bx, By)
                P = (np.sum(P0 * nrm, 1) / np.sum(S * nrm, 1))[:, None] * S - P0
                xp = mm[15] + np.sum(P * Bx, 1) / np.sum(Bx * Bx, 1)
                yp = mm[16] + np.sum(P * By, 1) / np.sum(By * By, 1)
                e = S / np.linalg.norm(S, axis=1, keepdims=True)
                u1, u2 = Bx / su.qx, By / su.qy
                cosa = np.clip(np.sum(e * np.cross(u1, u2), 1), 1e-3, 1.0)
                L = su.tsens / cosa
                lbar = 1.0 / su.mu - L * np.exp(-su.mu * L) / np.maximum(-np.expm1(-su.mu * L), 1e-12)
                xp = xp + su.plx * lbar * np.sum(e * u1, 1) / su.qx
                yp = yp + su.plx * lbar * np.sum(e * u2, 1) / su.qy
                nj = int(j["n_spots"])
                o = int(j["r0"]) + 3 * m * nj
                out[o:o + nj] = (xp - sp["xo"][i]) / su.sxy
                out[o + nj:o + 2 * nj] = (yp - sp["yo"][i]) / su.sxy
                out[o + 2 * nj:o + 3 * nj] = (phi - sp["phi"][i]) / su.sphi
        return 0

    def xpg_rf_normal(self, stp, sup, jobs, n_jobs, r, h, w, out, r0_out, stream):
        self.calls["normal"] += 1
        st, su = stp._obj, sup._obj
        jb = self._jobs(jobs, n_jobs)
        p = su.n_models - 1
        rr = _arr(r, (jb["r0"] + 3 * jb["n_models"].astype(np.int64) * jb["n_spots"]).max())
        hh, ww = _arr(h, n_jobs * p).reshape(n_jobs, p), _arr(w, (jb["w0"] + jb["n_spots"]).max())
        oo = _arr(out, n_jobs * (p * p + p)).reshape(n_jobs, -1)
        for k, j in enumerate(jb):
            nj = int(j["n_spots"])
            R = rr[int(j["r0"]):int(j["r0"]) + 3 * (p + 1) * nj].reshape(p + 1, 3 * nj)
            J = ((R[1:] - R[0]) / hh[k][:, None]).T
            wj = ww[int(j["w0"]):int(j["w0"]) + nj]
            if su.robust_c > 0:
                rn = np.sqrt((R[0].reshape(3, nj) ** 2).sum(0))
                wj = wj / (1.0 + (rn / su.robust_c) ** 2)
            wv = np.tile(wj, 3)
            if r0_out is not None and r0_out.value:
                _arr(r0_out.value + 8 * 3 * int(j["w0"]), 3 * nj)[:] = R[0]
            JW = J * wv[:, None]
            oo[k, :p * p] = (JW.T @ J).ravel()
            oo[k, p * p:] = JW.T @ R[0]
        return 0


@pytest.fixture
def fake(monkeypatch):
    f = FakeLib()
    monkeypatch.setattr(_lib, "lib", lambda: f)
    monkeypatch.setattr(_lib, "stream_ptr", lambda device=None: None)
    monkeypatch.setattr(xr, "_require_gpu", lambda dev: None)
    return f


def predicted(exp, spots):
    """Residual rows (normalised) of every spot under a refined experiment: what its predictions are."""
    data, _ = orf.spot_data(exp, *spots, 0.5, 0.25)
    par = orf.Param(exp, [])
    return orf.residuals(par, par.x0[None], data)[0]


def close(a, b, spots, params=True):
    """Same inliers and rmsd (1e-6); every spot predicted within 1e-4 px (and 1e-4 of sigma_phi) by the two refined
    models; with params, the parameters within 1e-5 of their units (mrad, px, mm, A, degrees).  The two paths round
    their residuals differently (ulps), which moves where the last Levenberg-Marquardt step stops; in the near-null
    directions of segment refinement (detector pose against segment poses) the parameters move more than the
    predictions."""
    assert a.rmsd_xy == pytest.approx(b.rmsd_xy, rel=1e-6)
    assert a.n_used == b.n_used and np.array_equal(a.inliers, b.inliers)
    assert np.abs(predicted(a.exp, spots) - predicted(b.exp, spots)).max() < 2e-4
    if params:
        np.testing.assert_allclose(a.x, b.x, rtol=0, atol=1e-5)
        np.testing.assert_allclose(a.exp.crystal.U, b.exp.crystal.U, atol=1e-8)


def test_models_pack_the_parameterisation(fake):
    g = _geom(segments=_modules(False), sensor_thickness=0.3, mu=2.0)
    exp = Experiment(g, Crystal.from_Astar(A, "m"))
    par = xr.Parameterisation(exp, FLAGS + ["SEGMENT", "AXIS"])
    opar = orf.Param(exp, FLAGS + ["SEGMENT", "AXIS"])
    assert par.names == opar.names
    X = par.x0 + np.random.default_rng(4).normal(scale=0.2, size=(5, par.x0.size))
    q, oq = par.unpack(X), opar.unpack(X)
    for k in ("Astar", "P0", "Bx", "By", "beam", "axis"):
        np.testing.assert_allclose(q[k], oq[k], atol=1e-12, err_msg=k)
    md = par.models(X)
    assert md.shape == (5, _lib.RF_HEAD + _lib.RF_SEG * 8)
    np.testing.assert_allclose(md[:, 6:15].reshape(5, 3, 3), oq["Astar"], atol=1e-15)
    np.testing.assert_allclose(par.gauge(X), np.array([opar.gauge(x) for x in X]), atol=1e-12)


@pytest.mark.parametrize("case", ["flat", "segmented"])
def test_refine_matches_the_oracle(fake, case):
    if case == "flat":
        g = _geom(sensor_thickness=0.45, mu=3.7)
        start = Experiment(replace(g, orgx=513.5, orgy=499.9, distance=151.0),
                           Crystal.from_Astar(A @ np.diag([1.004, 0.998, 1.0]), "m"))
        pr = predict_reflections(g, A, 2.4, z_range=(0, 40))
        flags, kw = FLAGS, {}
    else:
        g = _geom(segments=_modules(False))
        start = Experiment(_geom(segments=_modules()), Crystal.from_Astar(A, "m"))
        pr = predict_reflections(g, A, 2.4, z_range=(0, 40))
        flags, kw = FLAGS + ["SEGMENT"], dict(max_err_xy=10)
    x, y, z = _noisy(pr, 11)
    want = orf.refine(start, pr["hkl"], x, y, z, flags, **kw)
    got = xr.refine(start, pr["hkl"], x, y, z, flags, device="cpu", **kw)
    close(got, want, (pr["hkl"], x, y, z), params=case == "flat")
    if case == "flat":
        np.testing.assert_allclose(got.esd, want.esd, rtol=1e-5)
    assert fake.calls["normal"] > 3


def test_prefilter_and_classification_paths(fake, monkeypatch):
    monkeypatch.setattr(xr, "PREFILTER_N", 2000)
    monkeypatch.setattr(xr, "REFINE_SUB", 1500)
    monkeypatch.setattr(orf, "PREFILTER_N", 2000)
    monkeypatch.setattr(orf, "REFINE_SUB", 1500)
    g = _geom()
    pr = predict_reflections(g, A, 2.5, z_range=(0, 30))
    x, y, z = _noisy(pr, 2)
    x[::50] += 25.0                                  # outliers the pre-filter drops
    start = Experiment(replace(g, orgx=512.9), Crystal.from_Astar(A, "m"))
    assert pr["x"].size > 3000
    close(xr.refine(start, pr["hkl"], x, y, z, FLAGS, device="cpu"), orf.refine(start, pr["hkl"], x, y, z, FLAGS),
          (pr["hkl"], x, y, z))
    a, b = xr.refine(start, pr["hkl"], x, y, z, [], device="cpu"), orf.refine(start, pr["hkl"], x, y, z, [])
    assert np.array_equal(a.inliers, b.inliers) and a.rmsd_xy == pytest.approx(b.rmsd_xy, rel=1e-9)


def test_refine_many_equals_one_at_a_time(fake):
    g = _geom()
    probs = []
    for s in range(3):
        pr = predict_reflections(g, A, 2.6, z_range=(10 * s, 10 * s + 20))
        x, y, z = _noisy(pr, 20 + s)
        st = Experiment(replace(g, orgx=512.3 + 0.4 * s), Crystal.from_Astar(A, "m"))
        probs.append(((st, pr["hkl"], x, y, z, ["POSITION", "ORIENTATION"]), dict(prior={"orgx": (512.3, 0.5)})))
    fake.calls.update(res=0, normal=0)
    many = xr.refine_many(probs, device="cpu")
    n_batched = dict(fake.calls)
    for (a, k), m in zip(probs, many):
        one = xr.refine(*a, device="cpu", **k)
        assert np.array_equal(one.x, m.x) and np.array_equal(one.inliers, m.inliers) and one.rmsd_xy == m.rmsd_xy
    assert n_batched["normal"] < fake.calls["normal"] - n_batched["normal"]   # launches shared


def test_block_models_and_smoothing_match_the_oracle(fake):
    exp, spots, rng_, delphi, Us = drifting_sweep()
    s1, s2 = {}, {}
    got, e1 = xb.block_models(exp, spots, rng_, delphi, ["POSITION", "BEAM", "ORIENTATION"], state=s1, device="cpu")
    want, e2 = orf.block_models(exp, spots, rng_, delphi, ["POSITION", "BEAM", "ORIENTATION"], state=s2)
    assert np.array_equal(e1, e2) and s1["seed"] == s2["seed"] and s1["n_fallback"] == s2["n_fallback"]
    np.testing.assert_allclose(s1["pos_prior_sd"], s2["pos_prior_sd"], atol=1e-5)     # px, px, mm
    for a, b in zip(got, want):
        np.testing.assert_allclose(a.crystal.U, b.crystal.U, atol=1e-8)
        assert a.geom.orgx == pytest.approx(b.geom.orgx, abs=1e-5)
        assert a.geom.distance == pytest.approx(b.geom.distance, abs=1e-5)
    sm, so = xb.smooth_blocks(got, exp), orf.smooth_blocks(want, exp)
    for a, b in zip(sm, so):
        np.testing.assert_allclose(a.crystal.U, b.crystal.U, atol=1e-8)
        np.testing.assert_allclose(a.geom.beam, b.geom.beam, atol=1e-8)


def test_refinement_needs_the_gpu():
    g = _geom()
    with pytest.raises(_lib.XpgError):
        xr.refine(Experiment(g, Crystal.from_Astar(A, "m")), np.zeros((0, 3)), [], [], [], FLAGS, device="cpu")


def test_one_launch_serves_several_problems(fake):
    g = _geom(segments=_modules(False))
    pr = predict_reflections(g, A, 2.6, z_range=(0, 20))
    x, y, z = _noisy(pr, 8)
    exp = Experiment(g, Crystal.from_Astar(A, "m"))
    flags = FLAGS + ["SEGMENT"]
    par, opar = xr.Parameterisation(exp, flags), orf.Param(exp, flags)
    spots, setup, _ = xr._spot_data(exp, pr["hkl"], x, y, z, 0.5, 0.25)
    data, _ = orf.spot_data(exp, pr["hkl"], x, y, z, 0.5, 0.25)
    X = par.x0 + np.random.default_rng(1).normal(scale=0.2, size=(3, par.x0.size))
    p1, p2 = xr._Problem(par, spots, setup), xr._Problem(par, spots, setup)
    fake.calls.update(res=0, normal=0)
    out = xr._drive_once([xr._Res(p1, X), xr._Normal(p2, X[1], np.ones(len(x))), xr._Res(p1, X[2:])], device="cpu")
    assert fake.calls == {"res": 1, "normal": 1}      # one residual launch for all, one normal launch
    want = orf.residuals(opar, X, data)
    assert np.abs(out[0][:, :want.shape[1]] - want).max() < 1e-9 and np.array_equal(out[2], out[0][2:])
    Ao, go = orf.normal(opar, X[1], data, np.ones(want.shape[1] + 6))
    assert np.abs(out[1][0] - Ao).max() < 1e-6 * np.abs(Ao).max()
    again = xr._drive_once([xr._Res(p2, X[:1]), xr._Res(p1, X[:1])], device="cpu")   # a new pool: problems re-enter it
    assert np.array_equal(again[0], out[0][:1]) and np.array_equal(again[1], out[0][:1])


def test_device_cauchy_weights_equal_the_hosts(fake):
    """A normal request with robust_c weights its spots by the Cauchy form of its own residuals, as the host loop
    would from the same residuals; its residual rows come back with it."""
    g = _geom()
    pr = predict_reflections(g, A, 2.6, z_range=(0, 20))
    x, y, z = _noisy(pr, 9)
    exp = Experiment(replace(g, orgx=513.0), Crystal.from_Astar(A, "m"))
    par, opar = xr.Parameterisation(exp, FLAGS), orf.Param(exp, FLAGS)
    spots, setup, _ = xr._spot_data(exp, pr["hkl"], x, y, z, 0.5, 0.25)
    data, _ = orf.spot_data(exp, pr["hkl"], x, y, z, 0.5, 0.25)
    inl = (np.arange(len(x)) % 7 != 0).astype(float)
    (A_, g_, r), = xr._drive_once([xr._Normal(xr._Problem(par, spots, setup), par.x0, inl, 3.0)], device="cpu")
    ro = orf.full_residuals(opar, par.x0, data)
    assert np.abs(r[0] - ro).max() < 1e-9
    w1 = inl / (1.0 + (np.sqrt((ro.reshape(3, -1) ** 2).sum(0)) / 3.0) ** 2)
    Ao, go = orf.normal(opar, par.x0, data, np.tile(w1, 3))
    assert np.abs(A_ - Ao).max() < 1e-6 * np.abs(Ao).max() and np.abs(g_ - go).max() < 1e-6 * np.abs(go).max()


def _centroid_table(exp, spots, Us, edges):
    """A previous pass's reflection table for the drifting sweep: indices from each block's true orientation,
    observed centroids, and strong / full / clean flags (a few rows failing each criterion)."""
    x, y, z = spots
    hkl = np.zeros((x.size, 3), int)
    for i, U in enumerate(Us):
        m = (z >= edges[i]) & (z < edges[i + 1])
        A_i = U @ b_matrix(exp.crystal.cell)                 # block i's true reciprocal basis
        hkl[m] = np.rint(exp.geom.spot_to_p0(x[m], y[m], z[m]) @ np.linalg.inv(A_i).T)
    rng = np.random.default_rng(3)
    n = x.size
    refl = dict(hkl=hkl, xo=x, yo=y, zo=z, z=z + rng.normal(scale=0.05, size=n), isum=np.full(n, 1e4),
                vsum=np.full(n, 1e4), peak=np.full(n, 99.0), n_over=np.zeros(n, int), rho=np.zeros(n))
    refl["peak"][::9] = 50.0
    refl["isum"][::11] = 5.0
    refl["n_over"][::13] = 1
    return refl


def test_centroid_refinement_matches_the_oracle_per_block(fake):
    exp, spots, rng_, delphi, Us = drifting_sweep(n_blocks=5)
    edges = xb.block_edges(exp.geom, rng_, delphi)
    refl = _centroid_table(exp, spots, Us, edges)
    prior = xb.pos_prior(["POSITION"], exp.geom)
    blocks = [exp] * 5
    got = xb.refine_blocks_on_centroids(blocks, edges, refl, ["POSITION", "BEAM", "ORIENTATION", "CELL"], prior=prior,
                                        device="cpu")
    isg = refl["isum"] / np.sqrt(refl["vsum"])
    ok = (isg > 10) & (refl["peak"] > 90) & (refl["n_over"] == 0)
    for i in range(5):
        m = ok & (refl["zo"] >= edges[i]) & (refl["zo"] < edges[i + 1])
        want = orf.refine(blocks[i], refl["hkl"][m], refl["xo"][m], refl["yo"][m], refl["zo"][m],
                          ["POSITION", "BEAM", "ORIENTATION"], max_iter=15, n_reject_cycles=2, prior=prior)
        np.testing.assert_allclose(got[i].crystal.U, want.exp.crystal.U, atol=1e-8)
        np.testing.assert_allclose(got[i].crystal.U, Us[i], atol=2e-4)

"""The post-integration corrections on the device (xpg_finalise) against the oracle (oracle/finalise.py): random
beams on a flat detector with and without a sensor model, and a two-segment detector with distortion tables and
parallax; then inside the sweep (integrate()'s corrected columns on its own results) and the stage-level pieces on
them (selection of finite reflections, INTEGRATE.HKL, xpg_integrated.npz)."""
import numpy as np
import pytest
import torch

from oracle import finalise as of

pytestmark = pytest.mark.gpu


def _geom(**kw):
    from xpg_amd.geometry import Geometry
    base = dict(wavelength=1.0, beam=np.array([0, 0, 1.0]), axis=np.array([1.0, 0, 0]), orgx=1030, orgy=1080,
                distance=150, ed1=np.array([1.0, 0, 0]), ed2=np.array([0, 1.0, 0]), qx=0.075, qy=0.075, nx=2068,
                ny=2162, phi0=0, osc=0.2, starting_frame=1, polarization_fraction=0.99,
                polarization_normal=np.array([0.0, 1.0, 0.1]))
    base.update(kw)
    return Geometry(**base)


def _two_segments():
    from xpg_amd.geometry import Segment
    th = np.radians(0.4)
    return [Segment(1, 2068, 1, 1000),
            Segment(1, 2068, 1040, 2162, orgx=0.4, orgy=-0.3, distance=0.2,
                    ex=np.array([np.cos(th), np.sin(th), 0.02]), ey=np.array([-np.sin(th), np.cos(th), 0.0]))]


def _inputs(g, n, seed):
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(5, g.nx - 5, n), rng.uniform(5, g.ny - 5, n)
    S = g.pixel_to_s(x, y)
    cols = dict(iprf=rng.normal(500, 300, n), vprf=rng.uniform(-10, 900, n), isum=rng.normal(450, 300, n),
                vsum=rng.uniform(-5, 900, n), p_valid=rng.uniform(0, 1, n), p_total=rng.uniform(0, 1.2, n),
                corr=rng.uniform(-1, 1, n), coverage=rng.uniform(0.3, 1, n).astype(np.float32))
    cols["iprf"][::97] = np.nan
    cols["corr"][::89] = np.nan
    cols["p_total"][::101] = 0.0
    return S, x, y, cols


def _oracle(g, S, x, y, cols):
    from xpg_amd.geometry import mu_linear
    z, P = of.lp_factor(S, g.beam, g.axis, g.polarization_normal, g.polarization_fraction)
    seg = dict(normals=g.panel_normal_at(x, y), lab=g.pixel_to_lab(x, y)) if g.segmented else {}
    ab = of.absorption_factors(S, g.ed3, g.distance, g.sensor_thickness, g.mu, mu_linear(g.wavelength, "AIR"), **seg)
    return of.columns(S, cols["iprf"], cols["vprf"], cols["isum"], cols["vsum"], cols["p_valid"], cols["p_total"],
                      cols["corr"], cols["coverage"].astype(np.float64), z, P, ab)


@pytest.mark.parametrize("kind", ["flat", "flat_sensor", "segmented"])
def test_finalise_kernel_matches_oracle(kind):
    from xpg_amd import finalise as pf
    kw = dict(flat={}, flat_sensor=dict(sensor_thickness=0.45, mu=3.7),
              segmented=dict(sensor_thickness=0.45, mu=3.7, segments=_two_segments()))[kind]
    if kind == "segmented":
        yy, xx = np.mgrid[0:2162, 0:2068]
        kw["geo_corr"] = ((0.5 * np.sin(xx / 150.0)).astype(np.float32), (0.4 * np.cos(yy / 200.0)).astype(np.float32))
    g = _geom(**kw)
    n = 50000
    S, x, y, cols = _inputs(g, n, 21)
    print(f"finalise {kind}: {n} reflections", flush=True)
    got = pf.finalise(dict(S=S, x=x, y=y, **cols), g, "cuda:0")
    want = _oracle(g, S, x, y, cols)
    for k in ("rlp", "I", "sig", "Isum", "sigsum", "peak", "corr"):
        a, b = got[k], want[k]
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        m = np.isfinite(b)
        np.testing.assert_allclose(a[m], b[m], rtol=1e-12, atol=1e-300, err_msg=k)
    assert np.array_equal(got["good"], want["good"])


def test_sweep_columns_and_stage_outputs(tmp_path):
    from xpg_amd import finalise as pf, integrate as I, synth
    spec = synth.spec_small()
    dev = torch.device("cuda:0")
    frames, spots = synth.generate(spec, dev)
    exp = spots["exp"]
    edges = synth.block_edges(spec)
    model = I.IntegrationModel(spec.sigma_d, spec.sigma_m, box_xy=spec.box_xy, box_z=spec.box_z)
    print("sweep finalisation", spec.name, flush=True)
    res = I.integrate(frames, [exp] * (len(edges) - 1), edges, model, spots["dmin"], verbose=False,
                      _box_draw=synth.box_draw_for(spec))
    r = res.refl
    cols = {k: r[k] for k in ("iprf", "vprf", "isum", "vsum", "p_valid", "p_total", "corr")}
    cols["coverage"] = np.asarray(r["coverage"], np.float32)
    want = _oracle(exp.geom, r["S"], r["x"], r["y"], cols)
    for k in pf.FINAL_F64:
        a, b = r[k], want[k]
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        m = np.isfinite(b)
        np.testing.assert_allclose(a[m], b[m], rtol=1e-12, err_msg=k)
    main, extra = pf.finalise_refl(r, exp.geom)
    n_good = int(want["good"].sum())
    assert len(main["I"]) == n_good > 0.9 * len(r["I"]) and extra == {}
    np.testing.assert_allclose(main["corr"], want["corr"][want["good"]], rtol=1e-14)
    p = tmp_path / "INTEGRATE.HKL"
    pf.write_integrate_hkl(p, exp, main, (1, spec.nframes), spec.sigma_m, spec.sigma_d)
    lines = p.read_text().splitlines()
    body = [ln for ln in lines if not ln.startswith("!")]
    assert lines[-1] == "!END_OF_DATA" and len(body) == n_good
    rec = np.array([ln.split() for ln in body], float)
    assert rec.shape[1] == 21
    np.testing.assert_array_equal(rec[:, :3], main["hkl"])
    np.testing.assert_allclose(rec[:, 3], main["I"], rtol=6e-4)
    q = tmp_path / "xpg_integrated.npz"
    pf.write_integrated_npz(q, main)
    with np.load(q) as z:
        assert set(z.files) == {k for k, v in main.items() if isinstance(v, np.ndarray) and k != "fz"}
        np.testing.assert_array_equal(z["I"], main["I"])
